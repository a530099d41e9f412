// uvt_split.hip -- split-precision TRMM (placeholder until the int8 path lands).
#include "uvt_runtime.hpp"

extern "C" uvt_status uvt_trmm_split_workspace(int64_t n, int64_t m, int32_t slices, size_t* bytes)
{
    return UVT_ERR_UNSUPPORTED;
}

extern "C" uvt_status uvt_trmm_split(uvt_handle h, int64_t n, int64_t m, int32_t slices,
                                     const double* dA, int64_t lda, const double* dB, int64_t ldb,
                                     double* dC, int64_t ldc, void* dwork, size_t work_bytes,
                                     void* stream)
{
    return UVT_ERR_UNSUPPORTED;
}
