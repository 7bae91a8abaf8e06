// csr4 system-matrix file and raw vector files.
//
// Format (csr4matrix.cpp:69-148, csr4matrix.hpp:138-149, matrixelement.hpp:37-46):
//   0  char[4]  "PCSR"
//   4  u8       version (== 4)
//   5  u8       flags (x-sym 1, y-sym 2, z-sym 4)
//   8  u32      rows
//   12 u32      columns
//   16 u64      nnz
//   24 u32      scanConfigBytes (>= 60)
//   28 scanner block: u32 nLayers, nLayers x {f32 w,h,d,r,mu}, u32 nBlocks,
//      u32 blocksize, (blocksize-1) x f32, u32 nRings, (nRings-1) x f32,
//      f32 vx,vy,vz, u32 nx,ny,nz    -- must be exactly scanConfigBytes long
//   28+scanConfigBytes: u64[rows] inclusive row-end offsets
//   then nnz x {u32 column, f32 value}
#include <fcntl.h>
#include <omp.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cerrno>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "host_common.hpp"

namespace mlem_host {

static thread_local std::string g_last_error;

int fail(int st, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return st;
}

void clear_error() { g_last_error.clear(); }

}  // namespace mlem_host

using mlem_host::fail;

struct mlem_csr4 {
    int fd = -1;
    std::string path;
    uint64_t file_size = 0;
    uint8_t version = 0, flags = 0;
    uint32_t rows = 0, cols = 0, scan_config_bytes = 0;
    uint64_t nnz = 0;
    mlem_scanner scanner{};
    std::vector<uint64_t> row_ends;
    void *map = nullptr;
    size_t map_len = 0;
    const uint8_t *data = nullptr;  // first element of the data section
};

static const uint32_t kMinHeader = 28;        // csr4matrix.hpp:178
static const uint32_t kMinScanConfig = 60;    // csr4matrix.hpp:179
static const size_t kHeaderRead = 1000;       // csr4matrix.cpp:55

template <typename T>
static T rd(const uint8_t *p) {
    T v;
    std::memcpy(&v, p, sizeof(T));
    return v;
}

// Parses the scanner block that starts at p and may extend to `end`; returns
// the number of bytes consumed or -1 if it runs past `end` or the caps.
static long parse_scanner(const uint8_t *p, const uint8_t *end, mlem_scanner *s) {
    const uint8_t *q = p;
    auto need = [&](size_t n) { return (size_t)(end - q) >= n; };
    std::memset(s, 0, sizeof(*s));
    if (!need(4)) return -1;
    s->n_layers = rd<uint32_t>(q); q += 4;
    if (s->n_layers > MLEM_MAX_LAYERS || !need(20ull * s->n_layers)) return -1;
    for (uint32_t i = 0; i < s->n_layers; ++i)
        for (int k = 0; k < 5; ++k) { s->layer[i][k] = rd<float>(q); q += 4; }
    if (!need(8)) return -1;
    s->n_blocks = rd<uint32_t>(q); q += 4;
    s->blocksize = rd<uint32_t>(q); q += 4;
    // blocksize-1 gaps (csr4matrix.cpp:123); blocksize 0 would be 2^32-1 gaps.
    if (s->blocksize == 0 || s->blocksize - 1 > MLEM_MAX_GAPS || !need(4ull * (s->blocksize - 1))) return -1;
    for (uint32_t i = 0; i + 1 < s->blocksize; ++i) { s->block_gap[i] = rd<float>(q); q += 4; }
    if (!need(4)) return -1;
    s->n_rings = rd<uint32_t>(q); q += 4;
    if (s->n_rings == 0 || s->n_rings - 1 > MLEM_MAX_GAPS || !need(4ull * (s->n_rings - 1))) return -1;
    for (uint32_t i = 0; i + 1 < s->n_rings; ++i) { s->ring_gap[i] = rd<float>(q); q += 4; }
    if (!need(24)) return -1;
    for (int k = 0; k < 3; ++k) { s->voxel_size[k] = rd<float>(q); q += 4; }
    for (int k = 0; k < 3; ++k) { s->grid[k] = rd<uint32_t>(q); q += 4; }
    return (long)(q - p);
}

extern "C" {

const char *mlem_host_last_error(void) { return mlem_host::g_last_error.c_str(); }

const char *mlem_host_status_string(int status) {
    switch (status) {
        case MLEM_OK: return "ok";
        case MLEM_ERR_INVALID_ARGUMENT: return "invalid argument";
        case MLEM_ERR_HIP: return "HIP runtime error";
        case MLEM_ERR_OUT_OF_MEMORY: return "out of memory";
        case MLEM_ERR_UNSUPPORTED: return "unsupported";
        case MLEM_ERR_NO_DEVICE: return "no device";
        case MLEM_ERR_IO: return "I/O error";
        case MLEM_ERR_FORMAT: return "format error";
        default: return "unknown status";
    }
}

void mlem_host_free(void *p) { std::free(p); }

int mlem_csr4_open(const char *path, mlem_csr4 **out) {
    if (!path || !out) return fail(MLEM_ERR_INVALID_ARGUMENT, "mlem_csr4_open: null argument");
    *out = nullptr;
    auto *m = new mlem_csr4();
    m->path = path;
    auto bail = [&](int st) {
        if (m->fd >= 0) close(m->fd);
        delete m;
        return st;
    };
    m->fd = open(path, O_RDONLY);
    if (m->fd < 0) return bail(fail(MLEM_ERR_IO, "Cannot open file %s", path));
    off_t sz = lseek(m->fd, 0, SEEK_END);
    if (sz < 0) return bail(fail(MLEM_ERR_IO, "Cannot seek file %s", path));
    m->file_size = (uint64_t)sz;
    if (m->file_size < kMinHeader) return bail(fail(MLEM_ERR_FORMAT, "File too short %s", path));
    size_t hlen = m->file_size < kHeaderRead ? (size_t)m->file_size : kHeaderRead;
    std::vector<uint8_t> hdr(hlen);
    if (pread(m->fd, hdr.data(), hlen, 0) < (ssize_t)hlen)
        return bail(fail(MLEM_ERR_IO, "Cannot read header %s", path));
    m->version = hdr[4];
    m->flags = hdr[5];
    m->rows = rd<uint32_t>(&hdr[8]);
    m->cols = rd<uint32_t>(&hdr[12]);
    m->nnz = rd<uint64_t>(&hdr[16]);
    m->scan_config_bytes = rd<uint32_t>(&hdr[24]);
    if (std::memcmp(hdr.data(), "PCSR", 4) != 0)
        return bail(fail(MLEM_ERR_FORMAT, "Input file '%s' is not a CSR matrix", path));
    if (m->version != 4)
        return bail(fail(MLEM_ERR_FORMAT, "Input file '%s' has wrong CSR version (need v4)", path));
    // csr4matrix.cpp:95-102 computes this in uint32_t (vacuous above 4 GiB);
    // here it is exact.
    unsigned __int128 need = (unsigned __int128)kMinHeader + m->scan_config_bytes +
                             (unsigned __int128)m->rows * 8u + (unsigned __int128)m->nnz * 8u;
    if ((unsigned __int128)m->file_size < need)
        return bail(fail(MLEM_ERR_FORMAT, "Input file '%s' is too short: %llu B instead of at least %llu B",
                         path, (unsigned long long)m->file_size, (unsigned long long)need));
    if (m->scan_config_bytes < kMinScanConfig)
        return bail(fail(MLEM_ERR_FORMAT, "Cannot find scanner geometry configuration"));
    long used = parse_scanner(hdr.data() + kMinHeader, hdr.data() + hlen, &m->scanner);
    if (used < 0 && kMinHeader + (uint64_t)m->scan_config_bytes > hlen)
        return bail(fail(MLEM_ERR_FORMAT, "Internal Error: headerLength too small"));
    if (used < 0 || (uint32_t)used != m->scan_config_bytes)
        return bail(fail(MLEM_ERR_FORMAT, "Error reading scanner geometry configuration"));
    // Row index (csr4matrix.cpp:156-168).
    m->row_ends.resize(m->rows);
    size_t rlen = (size_t)m->rows * 8u;
    if (rlen && pread(m->fd, m->row_ends.data(), rlen, kMinHeader + m->scan_config_bytes) < (ssize_t)rlen)
        return bail(fail(MLEM_ERR_IO, "Cannot read row indexes %s", path));
    // Row ends must be monotone and end at nnz, else the data section would
    // be indexed out of range (the reference only asserts in debug builds).
    uint64_t prev = 0;
    for (uint32_t r = 0; r < m->rows; ++r) {
        if (m->row_ends[r] < prev || m->row_ends[r] > m->nnz)
            return bail(fail(MLEM_ERR_FORMAT, "Input file '%s': row index %u not monotone / beyond nnz", path, r));
        prev = m->row_ends[r];
    }
    if (m->rows && prev != m->nnz)
        return bail(fail(MLEM_ERR_FORMAT, "Input file '%s': last row end %llu != nnz %llu", path,
                         (unsigned long long)prev, (unsigned long long)m->nnz));
    if (m->rows == 0 && m->nnz != 0)
        return bail(fail(MLEM_ERR_FORMAT, "Input file '%s': nnz without rows", path));
    // Map the data section once (mapRows(0, rows), page-aligned start).
    uint64_t data_off = kMinHeader + (uint64_t)m->scan_config_bytes + (uint64_t)m->rows * 8u;
    uint64_t data_len = m->nnz * 8u;
    if (data_len) {
        long page = sysconf(_SC_PAGE_SIZE);
        uint64_t map_start = data_off & ~(uint64_t)(page - 1);
        uint64_t off = data_off - map_start;
        m->map_len = (size_t)(data_len + off);
        m->map = mmap(nullptr, m->map_len, PROT_READ, MAP_PRIVATE, m->fd, (off_t)map_start);
        if (m->map == MAP_FAILED) {
            m->map = nullptr;
            return bail(fail(MLEM_ERR_IO, "Cannot map file %s", path));
        }
        madvise(m->map, m->map_len, MADV_SEQUENTIAL);
        m->data = (const uint8_t *)m->map + off;
    }
    *out = m;
    return MLEM_OK;
}

int mlem_csr4_close(mlem_csr4 *m) {
    if (!m) return MLEM_OK;
    if (m->map) munmap(m->map, m->map_len);
    if (m->fd >= 0) close(m->fd);
    delete m;
    return MLEM_OK;
}

int mlem_csr4_info(const mlem_csr4 *m, uint32_t *rows, uint32_t *cols, uint64_t *nnz,
                   uint8_t *flags, uint32_t *scan_config_bytes) {
    if (!m) return fail(MLEM_ERR_INVALID_ARGUMENT, "mlem_csr4_info: null matrix");
    if (rows) *rows = m->rows;
    if (cols) *cols = m->cols;
    if (nnz) *nnz = m->nnz;
    if (flags) *flags = m->flags;
    if (scan_config_bytes) *scan_config_bytes = m->scan_config_bytes;
    return MLEM_OK;
}

int mlem_csr4_scanner(const mlem_csr4 *m, mlem_scanner *out) {
    if (!m || !out) return fail(MLEM_ERR_INVALID_ARGUMENT, "mlem_csr4_scanner: null argument");
    *out = m->scanner;
    return MLEM_OK;
}

const uint64_t *mlem_csr4_row_ends(const mlem_csr4 *m) { return m ? m->row_ends.data() : nullptr; }

const void *mlem_csr4_data(const mlem_csr4 *m) { return m ? m->data : nullptr; }

uint64_t mlem_csr4_elements_in_row(const mlem_csr4 *m, uint32_t row) {
    if (!m || row >= m->rows) return 0;
    return row == 0 ? m->row_ends[0] : m->row_ends[row] - m->row_ends[row - 1];
}

int mlem_csr4_to_csr(const mlem_csr4 *m, uint64_t *rowptr, uint32_t *col, float *val, int threads) {
    if (!m || !rowptr || (m->nnz && (!col || !val)))
        return fail(MLEM_ERR_INVALID_ARGUMENT, "mlem_csr4_to_csr: null argument");
    // Row pointers: a serial prefix in 64 bits (main.cu:519-530 uses 32-bit
    // counters there, which wrap above 2^32 nnz).
    rowptr[0] = 0;
    for (uint32_t r = 0; r < m->rows; ++r) rowptr[r + 1] = m->row_ends[r];
    // Split the AoS copy over nnz (not rows) so heavy rows do not serialise.
    const uint8_t *src = m->data;
    const uint64_t nnz = m->nnz;
    int nt = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel num_threads(nt)
    {
        int t = omp_get_thread_num(), T = omp_get_num_threads();
        uint64_t b = nnz * (uint64_t)t / (uint64_t)T, e = nnz * (uint64_t)(t + 1) / (uint64_t)T;
        for (uint64_t k = b; k < e; ++k) {
            std::memcpy(&col[k], src + 8 * k, 4);
            std::memcpy(&val[k], src + 8 * k + 4, 4);
        }
    }
    return MLEM_OK;
}

int mlem_csr4_write(const char *path, uint32_t rows, uint32_t cols, uint64_t nnz, const uint64_t *rowptr,
                    const uint32_t *col, const float *val, const mlem_scanner *scanner, uint8_t flags) {
    if (!path || !rowptr || (nnz && (!col || !val)))
        return fail(MLEM_ERR_INVALID_ARGUMENT, "mlem_csr4_write: null argument");
    if (rowptr[0] != 0 || rowptr[rows] != nnz)
        return fail(MLEM_ERR_INVALID_ARGUMENT, "mlem_csr4_write: rowptr must run from 0 to nnz");
    mlem_scanner s{};
    if (scanner) {
        s = *scanner;
    } else {
        s.n_layers = 1;
        s.n_blocks = 1;
        s.blocksize = 1;
        s.n_rings = 1;
        s.voxel_size[0] = s.voxel_size[1] = s.voxel_size[2] = 1.0f;
        s.grid[0] = cols; s.grid[1] = 1; s.grid[2] = 1;
    }
    if (s.n_layers > MLEM_MAX_LAYERS || s.blocksize == 0 || s.blocksize - 1 > MLEM_MAX_GAPS ||
        s.n_rings == 0 || s.n_rings - 1 > MLEM_MAX_GAPS)
        return fail(MLEM_ERR_INVALID_ARGUMENT, "mlem_csr4_write: scanner block out of range");
    std::vector<uint8_t> blk;
    auto put = [&](const void *p, size_t n) { blk.insert(blk.end(), (const uint8_t *)p, (const uint8_t *)p + n); };
    put(&s.n_layers, 4);
    for (uint32_t i = 0; i < s.n_layers; ++i) put(s.layer[i], 20);
    put(&s.n_blocks, 4);
    put(&s.blocksize, 4);
    put(s.block_gap, 4ull * (s.blocksize - 1));
    put(&s.n_rings, 4);
    put(s.ring_gap, 4ull * (s.n_rings - 1));
    put(s.voxel_size, 12);
    put(s.grid, 12);
    uint8_t hdr[28] = {0};
    std::memcpy(hdr, "PCSR", 4);
    hdr[4] = 4;
    hdr[5] = flags;
    uint32_t scb = (uint32_t)blk.size();
    std::memcpy(hdr + 8, &rows, 4);
    std::memcpy(hdr + 12, &cols, 4);
    std::memcpy(hdr + 16, &nnz, 8);
    std::memcpy(hdr + 24, &scb, 4);
    FILE *fp = std::fopen(path, "wb");
    if (!fp) return fail(MLEM_ERR_IO, "Cannot open file %s; reason: %s", path, std::strerror(errno));
    bool ok = std::fwrite(hdr, 1, 28, fp) == 28 && std::fwrite(blk.data(), 1, blk.size(), fp) == blk.size();
    if (ok && rows) ok = std::fwrite(rowptr + 1, 8, rows, fp) == rows;
    // Interleave {col, val} in chunks.
    std::vector<uint32_t> buf;
    const uint64_t chunk = 1u << 20;
    for (uint64_t k = 0; ok && k < nnz; k += chunk) {
        uint64_t n = nnz - k < chunk ? nnz - k : chunk;
        buf.resize(2 * n);
        for (uint64_t e = 0; e < n; ++e) {
            buf[2 * e] = col[k + e];
            std::memcpy(&buf[2 * e + 1], &val[k + e], 4);
        }
        ok = std::fwrite(buf.data(), 8, n, fp) == n;
    }
    if (std::fclose(fp) != 0) ok = false;
    if (!ok) return fail(MLEM_ERR_IO, "Cannot write file %s", path);
    return MLEM_OK;
}

int mlem_read_raw(const char *path, size_t elem_bytes, void **data, uint64_t *n) {
    if (!path || !data || !n || elem_bytes == 0)
        return fail(MLEM_ERR_INVALID_ARGUMENT, "mlem_read_raw: bad argument");
    *data = nullptr;
    *n = 0;
    int fd = open(path, O_RDONLY);
    if (fd < 0) return fail(MLEM_ERR_IO, "Cannot open file %s", path);
    struct stat st;
    if (fstat(fd, &st) != 0) {
        close(fd);
        return fail(MLEM_ERR_IO, "Cannot open file %s", path);
    }
    uint64_t bytes = (uint64_t)st.st_size;
    // vector.hpp:96-97: a size that is not a multiple of the element size fails.
    if (bytes % elem_bytes != 0) {
        close(fd);
        return fail(MLEM_ERR_FORMAT, "Cannot open file %s (size %llu not a multiple of %zu)", path,
                    (unsigned long long)bytes, elem_bytes);
    }
    void *buf = std::malloc(bytes ? bytes : 1);
    if (!buf) {
        close(fd);
        return fail(MLEM_ERR_OUT_OF_MEMORY, "Cannot allocate %llu bytes for %s", (unsigned long long)bytes, path);
    }
    uint64_t got = 0;
    while (got < bytes) {
        ssize_t r = pread(fd, (char *)buf + got, bytes - got, (off_t)got);
        if (r <= 0) {
            std::free(buf);
            close(fd);
            return fail(MLEM_ERR_IO, "Cannot read file %s", path);
        }
        got += (uint64_t)r;
    }
    close(fd);
    *data = buf;
    *n = bytes / elem_bytes;
    return MLEM_OK;
}

int mlem_write_raw(const char *path, const void *data, size_t elem_bytes, uint64_t n) {
    if (!path || (n && !data) || elem_bytes == 0)
        return fail(MLEM_ERR_INVALID_ARGUMENT, "mlem_write_raw: bad argument");
    FILE *fp = std::fopen(path, "wb");
    if (!fp) return fail(MLEM_ERR_IO, "Cannot open file %s; reason: %s", path, std::strerror(errno));
    bool ok = n == 0 || std::fwrite(data, elem_bytes, n, fp) == n;
    if (std::fclose(fp) != 0) ok = false;
    if (!ok) return fail(MLEM_ERR_IO, "Cannot write file %s", path);
    return MLEM_OK;
}

}  // extern "C"
