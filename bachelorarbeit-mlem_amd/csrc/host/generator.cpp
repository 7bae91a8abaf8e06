// Synthetic system matrices, phantoms and sinograms.
//
// The reference reads its matrix and counts from files produced by external
// tools that are not part of it; this generator produces inputs of the same
// kind: a cylindrical scanner, Siddon ray-traced intersection lengths (in mm)
// per LOR row, ellipse/ellipsoid phantoms and Poisson-sampled counts.
//
// Determinism: every random number comes from a counter-based stream keyed by
// (seed, row[, sub-ray]), so the output does not depend on the thread count
// or schedule, and the only floating-point library calls are sqrt/log/exp/
// lgamma/sin/cos of glibc.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <vector>

#include "host_common.hpp"

using mlem_host::fail;

namespace {

// SplitMix64 (Steele, Lea, Flood 2014) as a keyed counter-based stream.
inline uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

struct Stream {
    uint64_t s;
    Stream(uint64_t seed, uint64_t a, uint64_t b = 0) {
        s = mix64(seed ^ mix64(a * 0xd1342543de82ef95ull + 0x632be59bd9b4e019ull) ^ mix64(b + 0x1234567ull));
    }
    uint64_t next() {
        s += 0x9e3779b97f4a7c15ull;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
};

// Poisson sampler: sequential-search inversion for small lambda, and the
// transformed rejection with squeeze "PTRS" of W. Hoermann, "The transformed
// rejection method for generating Poisson random variables", Insurance:
// Mathematics and Economics 12 (1993) 39-45, for lambda >= 10.
int64_t poisson(double lam, Stream &st) {
    if (!(lam > 0.0)) return 0;
    if (lam < 10.0) {
        double u = st.uniform();
        double p = std::exp(-lam), F = p;
        int64_t k = 0;
        while (u > F && k < 10000) {
            ++k;
            p *= lam / (double)k;
            F += p;
        }
        return k;
    }
    const double slam = std::sqrt(lam), loglam = std::log(lam);
    const double b = 0.931 + 2.53 * slam;
    const double a = -0.059 + 0.02483 * b;
    const double invalpha = 1.1239 + 1.1328 / (b - 3.4);
    const double vr = 0.9277 - 3.6224 / (b - 2.0);
    for (;;) {
        double U = st.uniform() - 0.5;
        double V = st.uniform();
        double us = 0.5 - std::fabs(U);
        double kf = std::floor((2.0 * a / us + b) * U + lam + 0.43);
        if (us >= 0.07 && V <= vr) return (int64_t)kf;
        if (kf < 0.0 || (us < 0.013 && V > us)) continue;
        if (std::log(V) + std::log(invalpha) - std::log(a / (us * us) + b) <=
            -lam + kf * loglam - std::lgamma(kf + 1.0))
            return (int64_t)kf;
    }
}

struct V3 {
    double x, y, z;
};

struct Grid {
    int n[3];
    double vs[3];
    double lo[3];
};

Grid make_grid(const mlem_geometry &g) {
    Grid G;
    for (int a = 0; a < 3; ++a) {
        G.n[a] = (int)g.grid[a];
        G.vs[a] = g.voxel[a];
        G.lo[a] = -0.5 * g.grid[a] * (double)g.voxel[a];
    }
    return G;
}

// Voxel traversal of Amanatides & Woo ("A fast voxel traversal algorithm for
// ray tracing", Eurographics 1987) with Siddon's exact intersection lengths:
// emits (voxel, length in mm) for every voxel the segment p0-p1 crosses with
// a length above 1e-9 mm.
template <typename Emit>
void siddon(const Grid &G, const V3 &p0, const V3 &p1, Emit emit) {
    const double P0[3] = {p0.x, p0.y, p0.z};
    const double D[3] = {p1.x - p0.x, p1.y - p0.y, p1.z - p0.z};
    const double L = std::sqrt(D[0] * D[0] + D[1] * D[1] + D[2] * D[2]);
    if (!(L > 0.0)) return;
    double t0 = 0.0, t1 = 1.0;
    for (int a = 0; a < 3; ++a) {
        const double lo = G.lo[a], hi = G.lo[a] + G.n[a] * G.vs[a];
        if (D[a] == 0.0) {
            if (P0[a] < lo || P0[a] > hi) return;
        } else {
            double ta = (lo - P0[a]) / D[a], tb = (hi - P0[a]) / D[a];
            if (ta > tb) std::swap(ta, tb);
            t0 = std::max(t0, ta);
            t1 = std::min(t1, tb);
        }
    }
    if (!(t0 < t1)) return;
    int idx[3], step[3];
    double tmax[3], tdelta[3];
    for (int a = 0; a < 3; ++a) {
        double p = P0[a] + t0 * D[a];
        int i = (int)std::floor((p - G.lo[a]) / G.vs[a]);
        // On the entry face the floor may land one past; pick the voxel the
        // ray moves into.
        if (D[a] < 0.0 && p - (G.lo[a] + i * G.vs[a]) == 0.0) i -= 1;
        i = std::min(std::max(i, 0), G.n[a] - 1);
        idx[a] = i;
        if (D[a] > 0.0) {
            step[a] = 1;
            tmax[a] = (G.lo[a] + (i + 1) * G.vs[a] - P0[a]) / D[a];
            tdelta[a] = G.vs[a] / D[a];
        } else if (D[a] < 0.0) {
            step[a] = -1;
            tmax[a] = (G.lo[a] + i * G.vs[a] - P0[a]) / D[a];
            tdelta[a] = -G.vs[a] / D[a];
        } else {
            step[a] = 0;
            tmax[a] = std::numeric_limits<double>::infinity();
            tdelta[a] = std::numeric_limits<double>::infinity();
        }
    }
    double t = t0;
    for (;;) {
        double tn = std::min(std::min(tmax[0], tmax[1]), std::min(tmax[2], t1));
        double len = (tn - t) * L;
        if (len > 1e-9)
            emit((uint32_t)(idx[0] + G.n[0] * (idx[1] + G.n[1] * idx[2])), len);
        if (tn >= t1) break;
        bool out = false;
        for (int a = 0; a < 3; ++a) {
            if (tmax[a] == tn) {
                idx[a] += step[a];
                tmax[a] += tdelta[a];
                if (idx[a] < 0 || idx[a] >= G.n[a]) out = true;
            }
        }
        if (out) break;
        t = tn;
    }
}

// Crystal k of ring r: inner-face centre on the circle of radius R at angle
// 2*pi*k/n, axial centre (r - (rings-1)/2) * ring_pitch.  (ft, fa, fd) in
// [0,1)^3 place a point inside the crystal: tangential, axial, depth.
V3 crystal_point(const mlem_geometry &g, uint32_t crystal, double ft, double fa, double fd) {
    const uint32_t ring = crystal / g.n_per_ring, k = crystal % g.n_per_ring;
    const double ang = 2.0 * M_PI * (double)k / (double)g.n_per_ring;
    const double c = std::cos(ang), s = std::sin(ang);
    const double rad = g.radius + fd * g.crystal[2];
    const double tan_off = (ft - 0.5) * g.crystal[0];
    const double z = ((double)ring - 0.5 * (double)(g.n_rings - 1)) * g.ring_pitch + (fa - 0.5) * g.crystal[1];
    return V3{rad * c - tan_off * s, rad * s + tan_off * c, z};
}

// Row index -> crystal pair (a < b), lexicographic.
inline uint64_t rows_before(uint64_t a, uint64_t n) { return a * n - a * (a + 1) / 2; }

void row_pair(uint64_t row, uint64_t n, uint32_t *a, uint32_t *b) {
    uint64_t lo = 0, hi = n - 1;  // largest a with rows_before(a) <= row
    while (lo < hi) {
        uint64_t mid = (lo + hi + 1) / 2;
        if (rows_before(mid, n) <= row) lo = mid; else hi = mid - 1;
    }
    *a = (uint32_t)lo;
    *b = (uint32_t)(lo + 1 + (row - rows_before(lo, n)));
}

struct Entry {
    uint32_t col;
    float len;
};

// One LOR row: sorted unique columns with float lengths (mean over sub-rays).
void trace_row(const mlem_geometry &g, const Grid &G, uint64_t seed, uint64_t row, std::vector<Entry> &out,
               std::vector<std::pair<uint32_t, double>> &tmp) {
    out.clear();
    tmp.clear();
    const uint64_t n = (uint64_t)g.n_rings * g.n_per_ring;
    uint32_t a, b;
    row_pair(row, n, &a, &b);
    auto emit = [&](uint32_t v, double l) { tmp.emplace_back(v, l); };
    if (g.sub_rays == 0) {
        siddon(G, crystal_point(g, a, 0.5, 0.5, 0.0), crystal_point(g, b, 0.5, 0.5, 0.0), emit);
    } else {
        for (uint32_t s = 0; s < g.sub_rays; ++s) {
            Stream st(seed, row, s);
            double u[6];
            for (double &x : u) x = st.uniform();
            siddon(G, crystal_point(g, a, u[0], u[1], u[2]), crystal_point(g, b, u[3], u[4], u[5]), emit);
        }
    }
    std::sort(tmp.begin(), tmp.end(),
              [](const std::pair<uint32_t, double> &x, const std::pair<uint32_t, double> &y) {
                  return x.first < y.first;
              });
    const double inv = g.sub_rays ? 1.0 / (double)g.sub_rays : 1.0;
    for (size_t i = 0; i < tmp.size();) {
        size_t j = i;
        double sum = 0.0;
        while (j < tmp.size() && tmp[j].first == tmp[i].first) sum += tmp[j++].second;
        out.push_back(Entry{tmp[i].first, (float)(sum * inv)});
        i = j;
    }
}

}  // namespace

extern "C" {

uint32_t mlem_geometry_rows(const mlem_geometry *geo) {
    if (!geo) return 0;
    uint64_t n = (uint64_t)geo->n_rings * geo->n_per_ring;
    return (uint32_t)(n * (n - 1) / 2);
}

uint32_t mlem_geometry_cols(const mlem_geometry *geo) {
    if (!geo) return 0;
    return geo->grid[0] * geo->grid[1] * geo->grid[2];
}

void mlem_csr_host_free(mlem_csr_host *m) {
    if (!m) return;
    std::free(m->rowptr);
    std::free(m->col);
    std::free(m->val);
    m->rowptr = nullptr;
    m->col = nullptr;
    m->val = nullptr;
}

static int check_geometry(const mlem_geometry *geo) {
    if (!geo) return fail(MLEM_ERR_INVALID_ARGUMENT, "geometry: null");
    uint64_t n = (uint64_t)geo->n_rings * geo->n_per_ring;
    uint64_t cols = (uint64_t)geo->grid[0] * geo->grid[1] * geo->grid[2];
    if (n < 2 || n * (n - 1) / 2 > 0xffffffffull || cols == 0 || cols > 0xffffffffull)
        return fail(MLEM_ERR_INVALID_ARGUMENT, "geometry: bad crystal count or grid");
    if (!(geo->radius > 0) || !(geo->voxel[0] > 0) || !(geo->voxel[1] > 0) || !(geo->voxel[2] > 0))
        return fail(MLEM_ERR_INVALID_ARGUMENT, "geometry: radius and voxel sizes must be positive");
    return MLEM_OK;
}

int mlem_gen_system_matrix(const mlem_geometry *geo, uint64_t seed, uint32_t row_begin, uint32_t row_end,
                           int threads, mlem_csr_host *out) {
    if (int st = check_geometry(geo)) return st;
    if (!out) return fail(MLEM_ERR_INVALID_ARGUMENT, "mlem_gen_system_matrix: null output");
    std::memset(out, 0, sizeof(*out));
    const uint32_t total = mlem_geometry_rows(geo);
    if (row_end == 0) row_end = total;
    if (row_begin > row_end || row_end > total)
        return fail(MLEM_ERR_INVALID_ARGUMENT, "mlem_gen_system_matrix: bad row range");
    const uint32_t nrows = row_end - row_begin;
    const Grid G = make_grid(*geo);
    const uint32_t kChunk = 256;
    const uint32_t nchunks = (nrows + kChunk - 1) / kChunk;
    std::vector<std::vector<uint32_t>> ccol(nchunks);
    std::vector<std::vector<float>> cval(nchunks);
    std::vector<uint64_t> rowlen(nrows, 0);
    int nt = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel num_threads(nt)
    {
        std::vector<Entry> ent;
        std::vector<std::pair<uint32_t, double>> tmp;
#pragma omp for schedule(dynamic, 1)
        for (int64_t c = 0; c < (int64_t)nchunks; ++c) {
            uint32_t rb = (uint32_t)c * kChunk, re = std::min(nrows, rb + kChunk);
            std::vector<uint32_t> &cc = ccol[c];
            std::vector<float> &cv = cval[c];
            for (uint32_t r = rb; r < re; ++r) {
                trace_row(*geo, G, seed, (uint64_t)row_begin + r, ent, tmp);
                rowlen[r] = ent.size();
                for (const Entry &e : ent) {
                    cc.push_back(e.col);
                    cv.push_back(e.len);
                }
            }
            cc.shrink_to_fit();
            cv.shrink_to_fit();
        }
    }
    uint64_t nnz = 0;
    std::vector<uint64_t> chunk_off(nchunks + 1, 0);
    for (uint32_t c = 0; c < nchunks; ++c) chunk_off[c + 1] = chunk_off[c] + ccol[c].size();
    nnz = chunk_off[nchunks];
    out->rows = nrows;
    out->cols = mlem_geometry_cols(geo);
    out->nnz = nnz;
    out->rowptr = (uint64_t *)std::malloc(sizeof(uint64_t) * ((size_t)nrows + 1));
    out->col = (uint32_t *)std::malloc(sizeof(uint32_t) * (nnz ? nnz : 1));
    out->val = (float *)std::malloc(sizeof(float) * (nnz ? nnz : 1));
    if (!out->rowptr || !out->col || !out->val) {
        mlem_csr_host_free(out);
        return fail(MLEM_ERR_OUT_OF_MEMORY, "mlem_gen_system_matrix: cannot allocate %llu nnz",
                    (unsigned long long)nnz);
    }
    out->rowptr[0] = 0;
    for (uint32_t r = 0; r < nrows; ++r) out->rowptr[r + 1] = out->rowptr[r] + rowlen[r];
#pragma omp parallel for num_threads(nt) schedule(dynamic, 1)
    for (int64_t c = 0; c < (int64_t)nchunks; ++c) {
        if (!ccol[c].empty()) {
            std::memcpy(out->col + chunk_off[c], ccol[c].data(), ccol[c].size() * sizeof(uint32_t));
            std::memcpy(out->val + chunk_off[c], cval[c].data(), cval[c].size() * sizeof(float));
        }
        std::vector<uint32_t>().swap(ccol[c]);
        std::vector<float>().swap(cval[c]);
    }
    return MLEM_OK;
}

int mlem_gen_phantom(const mlem_geometry *geo, uint64_t seed, uint32_t n_shapes, float a_min, float a_max,
                     float *x_true) {
    if (int st = check_geometry(geo)) return st;
    if (!x_true) return fail(MLEM_ERR_INVALID_ARGUMENT, "mlem_gen_phantom: null output");
    const Grid G = make_grid(*geo);
    const double hx = -G.lo[0], hy = -G.lo[1], hz = -G.lo[2];
    const bool flat = geo->grid[2] == 1;
    struct Shape {
        double cx, cy, cz, ax, ay, az, ca, sa, act;
    };
    std::vector<Shape> shapes(n_shapes);
    // Shapes stay inside the inscribed circle of the ring and of the grid.
    const double rmax = std::min(std::min(hx, hy), (double)geo->radius);
    for (uint32_t i = 0; i < n_shapes; ++i) {
        Stream st(seed, 0xfeedull, i);
        Shape &s = shapes[i];
        double r = 0.55 * rmax * std::sqrt(st.uniform()), phi = 2.0 * M_PI * st.uniform();
        s.cx = r * std::cos(phi);
        s.cy = r * std::sin(phi);
        s.ax = rmax * (0.08 + 0.27 * st.uniform());
        s.ay = rmax * (0.08 + 0.27 * st.uniform());
        double th = M_PI * st.uniform();
        s.ca = std::cos(th);
        s.sa = std::sin(th);
        s.cz = flat ? 0.0 : hz * (st.uniform() - 0.5);
        s.az = flat ? 1.0 : hz * (0.25 + 0.5 * st.uniform());
        s.act = a_min + (a_max - a_min) * st.uniform();
    }
    const uint32_t nx = geo->grid[0], ny = geo->grid[1], nz = geo->grid[2];
#pragma omp parallel for schedule(static)
    for (int64_t z = 0; z < (int64_t)nz; ++z)
        for (uint32_t y = 0; y < ny; ++y)
            for (uint32_t x = 0; x < nx; ++x) {
                double px = G.lo[0] + (x + 0.5) * G.vs[0];
                double py = G.lo[1] + (y + 0.5) * G.vs[1];
                double pz = G.lo[2] + (z + 0.5) * G.vs[2];
                double v = 0.0;
                for (const Shape &s : shapes) {
                    double dx = px - s.cx, dy = py - s.cy;
                    double u = (dx * s.ca + dy * s.sa) / s.ax;
                    double w = (-dx * s.sa + dy * s.ca) / s.ay;
                    double q = u * u + w * w;
                    if (!flat) {
                        double dz = (pz - s.cz) / s.az;
                        q += dz * dz;
                    }
                    if (q <= 1.0) v += s.act;
                }
                x_true[(size_t)x + (size_t)nx * (y + (size_t)ny * z)] = (float)v;
            }
    return MLEM_OK;
}

int mlem_gen_sinogram(uint32_t rows, const uint64_t *rowptr, const uint32_t *col, const float *val, const float *x,
                      double total_counts, uint64_t seed, int threads, int32_t *y, double *expected_out) {
    if (!rowptr || !x || !y || (rowptr[rows] && (!col || !val)))
        return fail(MLEM_ERR_INVALID_ARGUMENT, "mlem_gen_sinogram: null argument");
    if (!(total_counts >= 0.0) || total_counts > 2.0e9)
        return fail(MLEM_ERR_INVALID_ARGUMENT, "mlem_gen_sinogram: total counts must be in [0, 2e9]");
    std::vector<double> ax(rows, 0.0);
    int nt = threads > 0 ? threads : omp_get_max_threads();
    double sum = 0.0;
#pragma omp parallel for num_threads(nt) schedule(dynamic, 1024)
    for (int64_t r = 0; r < (int64_t)rows; ++r) {
        double s = 0.0;
        for (uint64_t k = rowptr[r]; k < rowptr[r + 1]; ++k) s += (double)val[k] * (double)x[col[k]];
        ax[r] = s;
    }
    for (uint32_t r = 0; r < rows; ++r) sum += ax[r];  // fixed order
    const double scale = sum > 0.0 ? total_counts / sum : 0.0;
#pragma omp parallel for num_threads(nt) schedule(static)
    for (int64_t r = 0; r < (int64_t)rows; ++r) {
        double lam = scale * ax[r];
        Stream st(seed, 0x5151ull + (uint64_t)r, 0x9999ull);
        int64_t k = poisson(lam, st);
        y[r] = (int32_t)std::min<int64_t>(k, INT32_MAX);
        if (expected_out) expected_out[r] = lam;
    }
    return MLEM_OK;
}

int64_t mlem_poisson_sample(double lambda, uint64_t seed, uint64_t idx) {
    Stream st(seed, 0x5151ull + idx, 0x9999ull);
    return poisson(lambda, st);
}

}  // extern "C"
