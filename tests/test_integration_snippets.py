"""The code the integration guide at the repository root (DOC) gives a
maintainer compiles against the headers it names: every C++ block is
compiled (g++ -fsyntax-only) inside a function whose parameters are the
variables the block uses, with the types the reference
gives them (main.cu:65-68, 167-172: unsigned long row pointers, unsigned int
columns, int counts, float values), and the ctypes stub is valid Python with
a field list equal to mlem_opts in include/mlem_hip.h.  A changed signature
or a renamed entry point fails here, not in a maintainer's build.  No GPU."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DOC = os.path.join(ROOT, "INTEGRATION.md")

# Parameters of the function each C++ block is compiled in, in the order the blocks appear.
HARNESS = [
    # the host driver, in main.cu's mlem() (main.cu:167-172)
    "unsigned long *csr_Rows, unsigned int *csr_Cols, float *csr_Vals, unsigned long *csr_Rows_Trans, "
    "unsigned int *csr_Cols_Trans, float *csr_Vals_Trans, int *g, float *norm, float *f, float *result_f, "
    "unsigned int rows, unsigned int cols, unsigned long nnzs, unsigned int iterations, unsigned int device_numbers, "
    "unsigned int matrix_vector_mul, unsigned int secsize_fp, unsigned int secsize_bp, unsigned int using_trans",
    # loading a csr4 file
    "",
    # a batch of 32 sinograms
    "mlem_matrix *M, const int32_t *const *sino, const float *f0, float *const *image",
    # the row order
    "mlem_matrix *M, uint32_t rows, uint32_t cols, const uint64_t *rowptr, const uint32_t *col, const float *val, "
    "uint32_t nx, uint32_t ny, uint32_t nz",
    # a tiles-only matrix, one reconstruction
    "uint32_t rows, uint32_t cols, uint64_t nnz, const uint64_t *rowptr, const uint32_t *col, const float *val, "
    "const int32_t *g, const float *f0, float *image",
    # a tiles-only matrix for batches
    "mlem_matrix *M, mlem_batch *B, uint32_t rows, uint32_t cols, uint64_t nnz, const uint64_t *rowptr, "
    "const uint32_t *col, const float *val",
    # the request loop
    "mlem_batch_opts bo, mlem_matrix *M, const int32_t *sino, const float *f0, float *image",
]


def blocks(lang):
    text = open(DOC).read()
    return re.findall(r"```" + lang + r"\n(.*?)```", text, flags=re.S)


@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
def test_cpp_blocks_compile_against_the_headers(tmp_path):
    cpp = blocks("cpp")
    assert len(cpp) == len(HARNESS), "a C++ block was added to or removed from the integration guide: update HARNESS"
    src = ['#include <cstdint>', '#include <cstdio>', '#include <vector>', '#include "mlem_hip.h"',
           '#include "mlem_host.h"']
    for k, (body, params) in enumerate(zip(cpp, HARNESS)):
        lines = [ln for ln in body.splitlines() if not ln.startswith("#include")]
        src.append(f"int snippet_{k}({params}) {{")
        src.append("{")
        src.extend(lines)
        src.append("}")
        src.append("return 0;")
        src.append("}")
    path = tmp_path / "integration_snippets.cpp"
    path.write_text("\n".join(src) + "\n")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wno-unused-variable",
                        "-Wno-unused-but-set-variable", "-I", os.path.join(ROOT, "include"), str(path)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]


def test_python_stub_is_valid_and_mirrors_mlem_opts():
    py = blocks("python")
    assert len(py) == 1
    code = compile(py[0], "INTEGRATION.md", "exec")  # syntax only: the stub loads the GPU library
    assert code is not None
    fields = re.search(r'for n in \((.*?)\)\]', py[0], flags=re.S).group(1)
    stub = re.findall(r'"(\w+)"', fields)
    header = open(os.path.join(ROOT, "include", "mlem_hip.h")).read()
    body = re.search(r"typedef struct mlem_opts \{(.*?)\} mlem_opts;", header, flags=re.S).group(1)
    declared = re.findall(r"uint32_t\s+(\w+);", body)
    assert stub == declared
