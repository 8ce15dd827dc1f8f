// Device-wide stable radix sorts (rocPRIM) with temporary storage kept in the context.  In
// debug-sync mode each sort is waited for and named, like the library's own kernels.
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "pma_internal.h"

namespace pma {

void radix_sort_pairs_u64(pma_ctx* c, const uint64_t* keys_in, uint64_t* keys_out, const uint32_t* vals_in,
                          uint32_t* vals_out, size_t n, int begin_bit, int end_bit) {
    if (n == 0) return;
    size_t bytes = 0;
    PMA_HIP(rocprim::radix_sort_pairs(nullptr, bytes, keys_in, keys_out, vals_in, vals_out, n, begin_bit, end_bit,
                                      c->stream));
    void* tmp = c->buf("rsort_tmp", bytes);
    PMA_HIP(rocprim::radix_sort_pairs(tmp, bytes, keys_in, keys_out, vals_in, vals_out, n, begin_bit, end_bit,
                                      c->stream));
    PMA_CHECKED(c, "rocprim::radix_sort", nullptr);
}

void radix_sort_pairs_u64_u64(pma_ctx* c, const uint64_t* keys_in, uint64_t* keys_out, const uint64_t* vals_in,
                              uint64_t* vals_out, size_t n, int begin_bit, int end_bit) {
    if (n == 0) return;
    size_t bytes = 0;
    PMA_HIP(rocprim::radix_sort_pairs(nullptr, bytes, keys_in, keys_out, vals_in, vals_out, n, begin_bit, end_bit,
                                      c->stream));
    void* tmp = c->buf("rsort_tmp", bytes);
    PMA_HIP(rocprim::radix_sort_pairs(tmp, bytes, keys_in, keys_out, vals_in, vals_out, n, begin_bit, end_bit,
                                      c->stream));
    PMA_CHECKED(c, "rocprim::radix_sort", nullptr);
}

void radix_sort_pairs_u32(pma_ctx* c, const uint32_t* keys_in, uint32_t* keys_out, const uint32_t* vals_in,
                          uint32_t* vals_out, size_t n, int begin_bit, int end_bit) {
    if (n == 0) return;
    size_t bytes = 0;
    PMA_HIP(rocprim::radix_sort_pairs(nullptr, bytes, keys_in, keys_out, vals_in, vals_out, n, begin_bit, end_bit,
                                      c->stream));
    void* tmp = c->buf("rsort_tmp", bytes);
    PMA_HIP(rocprim::radix_sort_pairs(tmp, bytes, keys_in, keys_out, vals_in, vals_out, n, begin_bit, end_bit,
                                      c->stream));
    PMA_CHECKED(c, "rocprim::radix_sort", nullptr);
}

void radix_sort_pairs_u32_u64(pma_ctx* c, const uint32_t* keys_in, uint32_t* keys_out, const uint64_t* vals_in,
                              uint64_t* vals_out, size_t n, int begin_bit, int end_bit) {
    if (n == 0) return;
    size_t bytes = 0;
    PMA_HIP(rocprim::radix_sort_pairs(nullptr, bytes, keys_in, keys_out, vals_in, vals_out, n, begin_bit, end_bit,
                                      c->stream));
    void* tmp = c->buf("rsort_tmp", bytes);
    PMA_HIP(rocprim::radix_sort_pairs(tmp, bytes, keys_in, keys_out, vals_in, vals_out, n, begin_bit, end_bit,
                                      c->stream));
    PMA_CHECKED(c, "rocprim::radix_sort", nullptr);
}

void radix_sort_keys_u64(pma_ctx* c, const uint64_t* keys_in, uint64_t* keys_out, size_t n, int begin_bit,
                         int end_bit) {
    if (n == 0) return;
    size_t bytes = 0;
    PMA_HIP(rocprim::radix_sort_keys(nullptr, bytes, keys_in, keys_out, n, begin_bit, end_bit, c->stream));
    void* tmp = c->buf("rsort_tmp", bytes);
    PMA_HIP(rocprim::radix_sort_keys(tmp, bytes, keys_in, keys_out, n, begin_bit, end_bit, c->stream));
    PMA_CHECKED(c, "rocprim::radix_sort", nullptr);
}

}  // namespace pma
