// An order of a system matrix's rows (LORs) by the transaxial line each
// traces through the image, from the matrix and the image grid alone: the
// row order of the batched layouts (mlem_matrix_set_row_order).
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "host_common.hpp"

using mlem_host::fail;

namespace {

constexpr int kAngleBins = 576;   // over [0, pi)
constexpr int kAngleBlock = 32;   // angle bins per block (10 degrees)
constexpr double kOffsetBin = 0.5;  // voxels
constexpr double kPi = 3.14159265358979323846;

struct LorKey {
    int32_t block, offset, angle, z;
    uint32_t row;
    bool operator<(const LorKey &o) const {
        if (block != o.block) return block < o.block;
        if (offset != o.offset) return offset < o.offset;
        if (angle != o.angle) return angle < o.angle;
        if (z != o.z) return z < o.z;
        return row < o.row;
    }
};

}  // namespace

extern "C" {

// Each row's entries are voxels (x fastest, then y, then z) with weights:
// the weighted moments of their x, y give the row's line in the transaxial
// plane (centroid and principal direction: angle phi in [0, pi), signed
// offset s from the grid centre), z its mean slice.  Rows sort by
// (10-degree block of phi, s in half voxels, phi, z in 4-slice bins, row);
// empty rows go last.  LORs through one voxel lie on one sinusoid
// s = x cos(phi) + y sin(phi) (up to sign), so within an angle block they
// take nearby places.  Deterministic for any thread count.
int mlem_lor_order(const uint64_t *rowptr, const uint32_t *col, const float *val, uint32_t rows, uint32_t cols,
                   const uint32_t grid[3], int threads, uint32_t *order) {
    if (!rowptr || !grid || (rows && !order) || (rowptr[rows] && (!col || !val)))
        return fail(MLEM_ERR_INVALID_ARGUMENT, "mlem_lor_order: null argument");
    const uint64_t nx = grid[0], ny = grid[1], nz = grid[2];
    if (nx == 0 || ny == 0 || nz == 0 || nx * ny * nz != cols)
        return fail(MLEM_ERR_INVALID_ARGUMENT, "mlem_lor_order: grid %llu x %llu x %llu does not hold %u columns",
                    (unsigned long long)nx, (unsigned long long)ny, (unsigned long long)nz, cols);
    for (uint32_t r = 0; r < rows; ++r)
        if (rowptr[r + 1] < rowptr[r]) return fail(MLEM_ERR_INVALID_ARGUMENT, "mlem_lor_order: row %u ends before it starts", r);
    std::vector<LorKey> keys(rows);
    int bad = 0;
    const int nt = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 256) num_threads(nt) reduction(| : bad)
    for (int64_t ri = 0; ri < (int64_t)rows; ++ri) {
        const uint32_t r = (uint32_t)ri;
        double w = 0, wx = 0, wy = 0, wxx = 0, wxy = 0, wyy = 0, wz = 0;
        for (uint64_t k = rowptr[r]; k < rowptr[r + 1]; ++k) {
            const uint64_t c = col[k];
            if (c >= cols) {
                bad = 1;
                continue;
            }
            const double a = std::fabs((double)val[k]);
            const double x = (double)(c % nx), y = (double)((c / nx) % ny), z = (double)(c / (nx * ny));
            w += a;
            wx += a * x;
            wy += a * y;
            wxx += a * x * x;
            wxy += a * x * y;
            wyy += a * y * y;
            wz += a * z;
        }
        LorKey &key = keys[r];
        key.row = r;
        if (!(w > 0)) {
            key.block = kAngleBins / kAngleBlock + 1;
            key.offset = key.angle = key.z = 0;
            continue;
        }
        const double cx = wx / w, cy = wy / w;
        const double sxx = wxx / w - cx * cx, sxy = wxy / w - cx * cy, syy = wyy / w - cy * cy;
        double phi = 0.5 * std::atan2(2.0 * sxy, sxx - syy);  // direction of the line, (-pi/2, pi/2]
        if (phi < 0) phi += kPi;
        if (phi >= kPi) phi -= kPi;
        const double s = -(cx - 0.5 * (double)(nx - 1)) * std::sin(phi) + (cy - 0.5 * (double)(ny - 1)) * std::cos(phi);
        int angle = (int)std::floor(phi / kPi * kAngleBins);
        angle = std::min(std::max(angle, 0), kAngleBins - 1);
        key.block = angle / kAngleBlock;
        key.angle = angle;
        key.offset = (int32_t)std::floor(s / kOffsetBin);
        key.z = (int32_t)std::floor(wz / w / 4.0);
    }
    if (bad) return fail(MLEM_ERR_INVALID_ARGUMENT, "mlem_lor_order: a column is >= %u", cols);
    std::sort(keys.begin(), keys.end());
    for (uint32_t p = 0; p < rows; ++p) order[keys[p].row] = p;
    return MLEM_OK;
}

}  // extern "C"
