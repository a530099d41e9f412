// uvt_split.hpp -- batched split-precision stages, shared by the single-problem
// entry point (uvt_split.hip) and the batched runner (uvt_batch.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "uvt.h"

struct uvt_split_problem {
    const double* A;
    int64_t lda;
    const double* B;
    int64_t ldb;
    double* C;
    int64_t ldc;
    int64_t n, m;
};

// Bytes of the device descriptor table for `count` problems.
size_t uvt_split_desc_bytes(int count);
// Workspace bytes (residues, scales) and lookup-table entries of a batch.
uvt_status uvt_split_batch_sizes(const uvt_split_problem* probs, int count, int nmod,
                                 size_t* work_bytes, int64_t* lut_entries);
// Host images of the descriptor table and lookup tables for a batch whose
// workspace starts at dwork.
uvt_status uvt_split_batch_build(const uvt_split_problem* probs, int count, int nmod, void* dwork,
                                 std::vector<char>& desc_img, std::vector<int>& lut_img,
                                 int64_t totals[8], int64_t lut_off[8]);
// Stage groups.  The diagonal blocks store C; the reconstruction adds to C,
// so it must run after them.
enum {
    UVT_SPLIT_DIAG = 1,  // FP64 diagonal blocks
    UVT_SPLIT_PREP = 2,  // scales and residues
    UVT_SPLIT_GEMM = 4,  // int8 residue products
    UVT_SPLIT_REC = 8,   // reconstruction
    UVT_SPLIT_BOX = UVT_SPLIT_PREP | UVT_SPLIT_GEMM,
    UVT_SPLIT_ALL = 15
};
// Enqueues the selected stages of the batch on stream s.
uvt_status uvt_split_batch_launch(const uvt_split_problem* single, int count, int nmod,
                                  void* dwork, const void* ddesc, const int* dlut,
                                  const int64_t totals[8], const int64_t lut_off[8], hipStream_t s,
                                  int stages);
