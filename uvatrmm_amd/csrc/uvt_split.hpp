// uvt_split.hpp -- batched split-precision stages, shared by the single-problem
// entry point (uvt_split.hip), the batched runner (uvt_batch.hip) and the
// server (uvt_serve.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "uvt_internal.h"

struct uvt_split_problem {
    const double* A;
    int64_t lda;
    const double* B;
    int64_t ldb;
    double* C;
    int64_t ldc;
    int64_t n, m;
};

// Kernel stages of the split path (per-stage workgroup counts and lookup
// tables are arrays of this length).
constexpr int UVT_SPLIT_STAGES = 11;

// Whether an n x n triangle can run on the int8 path with nmod moduli: the
// int32 accumulation bound (n <= 131072) and at least 6 bits left per scaled
// operand at that depth.
UVT_INTERNAL bool uvt_split_supported(int64_t n, int nmod);
// Bytes of the device descriptor table for `count` problems.
UVT_INTERNAL size_t uvt_split_desc_bytes(int count);
// Workspace bytes (residues, scales, error-bound terms) and lookup-table
// entries of a batch.
UVT_INTERNAL uvt_status uvt_split_batch_sizes(const uvt_split_problem* probs, int count, int nmod,
                                              size_t* work_bytes, int64_t* lut_entries);
// Host images of the descriptor table and lookup tables for a batch whose
// workspace starts at dwork; err[p] receives the device address of problem
// p's error statement (two doubles: bound, estimate; see uvt_trmm_split_error).
UVT_INTERNAL uvt_status uvt_split_batch_build(const uvt_split_problem* probs, int count, int nmod,
                                              void* dwork, std::vector<char>& desc_img,
                                              std::vector<int>& lut_img,
                                              int64_t totals[UVT_SPLIT_STAGES],
                                              int64_t lut_off[UVT_SPLIT_STAGES],
                                              std::vector<double*>* err);
// Stage groups.  The diagonal blocks store C; the reconstruction adds to C,
// so it must run after them.
enum {
    UVT_SPLIT_DIAG = 1,  // FP64 diagonal blocks
    UVT_SPLIT_PREP = 2,  // scales, residues, error-bound terms of A and B
    UVT_SPLIT_GEMM = 4,  // int8 residue products
    UVT_SPLIT_REC = 8,   // reconstruction and the error statement
    UVT_SPLIT_ALL = 15
};
// Enqueues the selected stages on stream s.  ddesc == nullptr: `single` is
// one problem launched directly (workspace at dwork, totals and lut_off
// unused); otherwise every stage reads its problem from the device
// descriptor table ddesc through the lookup tables at dlut.
UVT_INTERNAL uvt_status uvt_split_batch_launch(const uvt_split_problem* single, int count, int nmod,
                                               void* dwork, const void* ddesc, const int* dlut,
                                               const int64_t totals[UVT_SPLIT_STAGES],
                                               const int64_t lut_off[UVT_SPLIT_STAGES],
                                               hipStream_t s, int stages);
// Byte offset of the error statement inside the workspace of one problem
// run by uvt_trmm_split (n, m, nmod as passed to it).
UVT_INTERNAL size_t uvt_split_err_offset(int64_t n, int64_t m, int nmod);
