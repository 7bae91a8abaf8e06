// Wraps the reference's own ScanTrans (sptrans.h, found through -I on the
// reference checkout, never copied) in a C entry point, so the tests can
// compare the oracle's restatement and the engine's transpose against the
// original code.  Built into oracle/_ref/ by `make -C oracle ref`; used only
// to validate and to write tests/golden fixtures.
#include <omp.h>

#include "sptrans.h"

extern "C" int ref_sptrans(unsigned int m, unsigned int n, unsigned long nnz, unsigned long *rowptr,
                           unsigned int *col, float *val, unsigned int *rowT, unsigned long *colptrT, float *valT,
                           int threads) {
    if (threads > 0) omp_set_num_threads(threads);
    // sptrans_scanTrans_specialized expects a zeroed colptrT (main.cu:101).
    for (unsigned long i = 0; i <= n; ++i) colptrT[i] = 0;
    sptrans_scanTrans_specialized(m, n, nnz, rowptr, col, val, rowT, colptrT, valT);
    return 0;
}
