// mlem -- the reference's command line (main.cu:22-143, help() :614-625) on
// one MI355X:
//
//   mlem <matrix.csr4> <sinogram> <iters> <nGPU> <secsize_fp> <secsize_bp>
//        <trans: 0 = BP on A^T, 1 = scatter> <spmv: 0 = merge, 1 = CSR-vector>
//        [--out=image.f32] [--sens=gpu64|host32] [--transpose=host|gpu]
//        [--x0=ref|one] [--graph=0|1] [--threads=N] [--loglik] [--verbose]
//        [--values=f32|bf16|fp16|uf16] [--resident=csr|tiles]
//
// --resident=tiles keeps only the tiles of A and A^T on the GPU (no CSR
// copies; for matrices whose two CSR copies leave no room for the tiles,
// such as the 100 GB class); it needs trans = 0 and spmv = 0.
//
// Same checks as main.cu:43-62 (exactly 8 positional arguments, iters, nGPU
// and both section sizes parsed with strtol and != 0, trans and spmv in
// {0,1}, nGPU <= device count); violations print help() to stdout and exit
// with EXIT_FAILURE.  stdout ends with "\nSum f: %f\n" (main.cu:460-463).
// Keyword options may only follow the positional arguments.
//
// Exit codes: 0 ok, 1 usage (help), 2 matrix file, 3 sinogram file,
// 4 GPU / engine error, 5 output file.
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "mlem_hip.h"
#include "mlem_host.h"

static void help() {
    printf("\nWrong parameters! Please read the following instructions:\n\n");
    printf("total amount of parameters: 9\n");
    printf("argv[1]: path of matrix file\n");
    printf("argv[2]: path of image  file\n");
    printf("argv[3]: iteration times\n");
    printf("argv[4]: number of GPUs to be used\n");
    printf("argv[5]: section size for forward  projection (merge-based SpMV)\n");
    printf("argv[6]: section size for backward projection (merge-based SpMV)\n");
    printf("argv[7]: BP uses transposed matrix?      0: yes                1: no\n");
    printf("argv[8]: which SpMV algorithm to use?    0: merge-based SpMV   1: csr-vector SpMV\n");
    printf("options after the 8 arguments: --out=PATH --sens=gpu64|host32 --transpose=host|gpu\n");
    printf("                               --x0=ref|one --graph=0|1 --threads=N --loglik --verbose\n");
    printf("                               --values=f32|bf16|fp16|uf16 (stored matrix values, merge-based only)\n");
    printf("                               --resident=csr|tiles (tiles: no CSR copies on the GPU; trans 0, spmv 0)\n\n");
}

struct Timer {
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    double lap() {
        auto t = std::chrono::steady_clock::now();
        double s = std::chrono::duration<double>(t - t0).count();
        t0 = t;
        return s;
    }
};

int main(int argc, char **argv) {
    if (argc < 9) {
        help();
        return EXIT_FAILURE;
    }
    for (int a = 9; a < argc; ++a)
        if (std::strncmp(argv[a], "--", 2) != 0) {
            help();
            return EXIT_FAILURE;
        }
    const std::string matrixPath(argv[1]);
    const std::string imagePath(argv[2]);
    unsigned int iterations = (unsigned int)strtol(argv[3], nullptr, 10);
    if (iterations == 0) { help(); return EXIT_FAILURE; }
    unsigned int device_numbers = (unsigned int)strtol(argv[4], nullptr, 10);
    if (device_numbers == 0) { help(); return EXIT_FAILURE; }
    unsigned int secsize_fp = (unsigned int)strtol(argv[5], nullptr, 10);
    if (secsize_fp == 0) { help(); return EXIT_FAILURE; }
    unsigned int secsize_bp = (unsigned int)strtol(argv[6], nullptr, 10);
    if (secsize_bp == 0) { help(); return EXIT_FAILURE; }
    unsigned int using_trans = (unsigned int)strtol(argv[7], nullptr, 10);
    if (using_trans != 0 && using_trans != 1) { help(); return EXIT_FAILURE; }
    unsigned int matrix_vector_mul = (unsigned int)strtol(argv[8], nullptr, 10);
    if (matrix_vector_mul != 0 && matrix_vector_mul != 1) { help(); return EXIT_FAILURE; }

    std::string out_path, sens_mode = "gpu64", trans_mode = "host", x0_mode = "ref", values_mode = "f32";
    std::string resident = "csr";
    int graph = 0, threads = 0;
    bool verbose = false, want_ll = false;
    for (int a = 9; a < argc; ++a) {
        std::string s(argv[a]);
        auto val = [&](const char *key) -> const char * {
            size_t n = std::strlen(key);
            return s.compare(0, n, key) == 0 ? argv[a] + n : nullptr;
        };
        if (const char *v = val("--out=")) out_path = v;
        else if (const char *v = val("--sens=")) sens_mode = v;
        else if (const char *v = val("--transpose=")) trans_mode = v;
        else if (const char *v = val("--x0=")) x0_mode = v;
        else if (const char *v = val("--graph=")) graph = std::atoi(v);
        else if (const char *v = val("--threads=")) threads = std::atoi(v);
        else if (const char *v = val("--values=")) values_mode = v;
        else if (const char *v = val("--resident=")) resident = v;
        else if (s == "--verbose") verbose = true;
        else if (s == "--loglik") want_ll = true;
        else { help(); return EXIT_FAILURE; }
    }
    const char *value_names[] = {"f32", "bf16", "fp16", "uf16"};
    int values = -1;
    for (int k = 0; k < 4; ++k)
        if (values_mode == value_names[k]) values = k;
    if ((sens_mode != "gpu64" && sens_mode != "host32") || (trans_mode != "host" && trans_mode != "gpu") ||
        (x0_mode != "ref" && x0_mode != "one") || values < 0 || (resident != "csr" && resident != "tiles")) {
        help();
        return EXIT_FAILURE;
    }
    if (threads > 0) omp_set_num_threads(threads);

    int available = 0;
    if (mlem_device_count(&available) != MLEM_OK) available = 0;
    if ((unsigned int)available < device_numbers) {
        help();
        return EXIT_FAILURE;
    }

    Timer T;
    // With the transpose and the sensitivity on the GPU no host CSR is needed:
    // the csr4 data section streams straight to the device.
    const bool tiles_only = resident == "tiles";
    if (tiles_only && (using_trans != 0 || matrix_vector_mul != 0)) {
        fprintf(stderr, "error: --resident=tiles needs the transposed, merge-based path (argv[7] = 0, argv[8] = 0)\n");
        return EXIT_FAILURE;
    }
    const bool stream_load = trans_mode == "gpu" && sens_mode == "gpu64" && !tiles_only;
    uint32_t rows = 0, cols = 0;
    uint64_t nnz = 0;
    std::vector<uint64_t> rowptr;
    std::vector<uint32_t> col;
    std::vector<float> val;
    mlem_matrix *M = nullptr;
    int st = MLEM_OK;
    double t_load = 0, t_upload = 0, t_trans = 0;
    if (stream_load) {
        mlem_load_stats ls{};
        if (mlem_matrix_load_csr4(0, matrixPath.c_str(), threads, &M, &ls) != MLEM_OK) {
            fprintf(stderr, "error: %s\n", mlem_last_error());
            return 2;
        }
        int has_t = 0;
        uint64_t dev_bytes = 0;
        mlem_matrix_info(M, &rows, &cols, &nnz, &has_t, &dev_bytes);
        t_load = T.lap();
        if (verbose)
            fprintf(stderr, "streamed %.2f GB of csr4 data to the GPU in %.3f s (%.1f GB/s)\n", ls.bytes / 1e9,
                    ls.stream_s, ls.bytes / 1e9 / std::max(ls.stream_s, 1e-9));
    } else {
        // Matrix (Csr4Matrix + mapRows + csr_format_for_cuda, main.cu:71-79).
        mlem_csr4 *file = nullptr;
        if (mlem_csr4_open(matrixPath.c_str(), &file) != MLEM_OK) {
            fprintf(stderr, "error: %s\n", mlem_host_last_error());
            return 2;
        }
        mlem_csr4_info(file, &rows, &cols, &nnz, nullptr, nullptr);
        rowptr.resize((size_t)rows + 1);
        col.resize(nnz);
        val.resize(nnz);
        if (mlem_csr4_to_csr(file, rowptr.data(), col.data(), val.data(), threads) != MLEM_OK) {
            fprintf(stderr, "error: %s\n", mlem_host_last_error());
            return 2;
        }
        mlem_csr4_close(file);
        t_load = T.lap();
    }

    // Sinogram (Vector<int>, main.cu:87-90).
    void *graw = nullptr;
    uint64_t gn = 0;
    if (mlem_read_raw(imagePath.c_str(), sizeof(int32_t), &graw, &gn) != MLEM_OK) {
        fprintf(stderr, "error: %s\n", mlem_host_last_error());
        return 3;
    }
    if (gn < rows) {
        // The reference reads past the end here (main.cu:89-90, 289).
        fprintf(stderr, "error: sinogram %s has %llu entries, the matrix has %u rows\n", imagePath.c_str(),
                (unsigned long long)gn, rows);
        mlem_host_free(graw);
        return 3;
    }
    const int32_t *g = static_cast<const int32_t *>(graw);

    if (!stream_load) {
        // Resident matrix on GPU 0; the transpose on the host (sptrans) or GPU.
        std::vector<uint64_t> rowptrT;
        std::vector<uint32_t> colT;
        std::vector<float> valT;
        if (using_trans == 0 && trans_mode == "host" && !tiles_only) {
            rowptrT.resize((size_t)cols + 1);
            colT.resize(nnz);
            valT.resize(nnz);
            if (mlem_transpose(rows, cols, nnz, rowptr.data(), col.data(), val.data(), rowptrT.data(), colT.data(),
                               valT.data(), threads) != MLEM_OK) {
                fprintf(stderr, "error: %s\n", mlem_host_last_error());
                return 4;
            }
            t_trans = T.lap();
        }
        if (tiles_only)  // A^T's tiles from column ranges of A on the GPU, A's from row chunks of these arrays
            st = mlem_matrix_create_tiled(0, rows, cols, nnz, rowptr.data(), col.data(), val.data(), 0, 0,
                                          (uint32_t)values, 1, 0, &M);
        else
            st = mlem_matrix_create(0, rows, cols, nnz, rowptr.data(), col.data(), val.data(),
                                    rowptrT.empty() ? nullptr : rowptrT.data(), colT.empty() ? nullptr : colT.data(),
                                    valT.empty() ? nullptr : valT.data(), &M);
        if (st != MLEM_OK) {
            fprintf(stderr, "error: %s\n", mlem_last_error());
            return 4;
        }
        t_upload = T.lap();
    }
    if (using_trans == 0 && trans_mode == "gpu" && !tiles_only) {
        if (mlem_matrix_build_transpose(M) != MLEM_OK) {
            fprintf(stderr, "error: %s\n", mlem_last_error());
            return 4;
        }
        t_trans = T.lap();
    }

    // Sensitivity (calcColumnSums, main.cu:80-84): FP64 on the GPU, or the
    // reference's float sums on the host.
    std::vector<float> norm(cols);
    if (sens_mode == "gpu64")
        st = mlem_matrix_sensitivity(M, norm.data(), nullptr);
    else
        st = mlem_colsum_f32(rowptr.data(), col.data(), val.data(), rows, cols, norm.data());
    if (st != MLEM_OK) {
        fprintf(stderr, "error: %s\n", sens_mode == "gpu64" ? mlem_last_error() : mlem_host_last_error());
        return 4;
    }
    double t_sens = T.lap();

    // Initial image (main.cu:92-97), or x0 = 1.
    std::vector<float> f0(cols, 1.0f), f(cols);
    float init = 1.0f;
    if (x0_mode == "ref") mlem_initial_image(g, rows, norm.data(), cols, f0.data(), &init);

    if (verbose && !rowptr.empty()) {
        std::vector<uint32_t> seg(device_numbers + 1), srows(device_numbers), snnz(device_numbers);
        std::vector<uint64_t> off(device_numbers);
        mlem_partition(rowptr.data(), nnz, rows, device_numbers, seg.data(), srows.data(), snnz.data(), nullptr,
                       off.data());
        fprintf(stderr, "matrix %u x %u, %llu nnz; init %g\n", rows, cols, (unsigned long long)nnz, init);
        for (unsigned int d = 0; d < device_numbers; ++d)
            fprintf(stderr, "  partition %u: rows [%u, %u) nnz %u offset %llu\n", d, seg[d], seg[d + 1], snnz[d],
                    (unsigned long long)off[d]);
    }
    if (device_numbers > 1)
        fprintf(stderr,
                "note: the whole matrix stays resident on GPU 0 (288 GB HBM); the %u-way row partition is not "
                "used\n",
                device_numbers);

    mlem_opts o;
    mlem_opts_default(&o);
    o.spmv = matrix_vector_mul;
    o.using_trans = using_trans;
    o.secsize_fp = secsize_fp;
    o.secsize_bp = secsize_bp;
    o.use_graph = graph ? 1 : 0;
    o.values = (uint32_t)values;  // 16-bit values: the quantized matrix's own sensitivity is used
    mlem_recon *R = nullptr;
    st = mlem_recon_create(M, g, norm.data(), f0.data(), &o, &R);
    if (st == MLEM_OK) {
        T.lap();
        st = mlem_recon_iterate(R, iterations);
        if (st == MLEM_OK) st = mlem_recon_sync(R);
    }
    double t_loop = T.lap();
    if (st == MLEM_OK) st = mlem_recon_get_image(R, f.data());
    double ll = 0;
    if (st == MLEM_OK && want_ll) st = mlem_recon_loglik(R, &ll);
    if (st != MLEM_OK) {
        fprintf(stderr, "error: %s\n", mlem_last_error());
        return 4;
    }
    mlem_recon_destroy(R);
    mlem_matrix_destroy(M);
    mlem_host_free(graw);

    if (verbose)
        fprintf(stderr,
                "load %.3f s, transpose %.3f s, upload %.3f s, sensitivity %.3f s, %u iterations %.3f s "
                "(%.1f it/s)\n",
                t_load, t_trans, t_upload, t_sens, iterations, t_loop, iterations / t_loop);
    if (want_ll) printf("log-likelihood: %.10e\n", ll);
    if (!out_path.empty() && mlem_write_raw(out_path.c_str(), f.data(), sizeof(float), cols) != MLEM_OK) {
        fprintf(stderr, "error: %s\n", mlem_host_last_error());
        return 5;
    }
    float sum = 0;
    for (unsigned int i = 0; i < cols; i++) sum += f[i];
    printf("\nSum f: %f\n", sum);
    return EXIT_SUCCESS;
}
