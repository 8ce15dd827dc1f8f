// Exact deduplication of ordered pore lists (throat keys).
//
// The reference keys shared pixel types by the exact ordered list of pore IDs
// (std::map<std::vector<PoreThroatIdType>, ...>, BasePoreAndThroatComputer.cpp:112-133), and
// numbers a key by the moment its list is first formed.  Here every formation is one record
// (list, time, weight); records are sorted by list content with three stable radix
// passes (no hashing, so distinct lists are never merged), equal runs give one distinct
// list with its earliest formation time and the summed weight (voxels whose final list it is).
#include "pma_internal.h"

namespace pma {
namespace {

__global__ void iota_u32(uint32_t* p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = (uint32_t)i;
}

__global__ void gather_u64(const uint64_t* __restrict__ src, const uint32_t* __restrict__ perm,
                           uint64_t* __restrict__ dst, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[perm[i]];
}

__global__ void mark_heads(const uint64_t* __restrict__ w0, const uint64_t* __restrict__ w1,
                           const uint64_t* __restrict__ w2, const uint32_t* __restrict__ perm,
                           uint32_t* __restrict__ head, size_t n) {
    size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    uint32_t a = perm[s];
    uint32_t h = 1;
    if (s > 0) {
        uint32_t b = perm[s - 1];
        h = (w0[a] != w0[b]) || (w1[a] != w1[b]) || (w2[a] != w2[b]);
    }
    head[s] = h;
}

__global__ void reduce_runs(const uint64_t* __restrict__ w0, const uint64_t* __restrict__ w1,
                            const uint64_t* __restrict__ w2, const uint64_t* __restrict__ time,
                            const uint32_t* __restrict__ weight, const uint32_t* __restrict__ perm,
                            const uint32_t* __restrict__ head, const uint32_t* __restrict__ head_scan,
                            uint64_t* __restrict__ dwords, unsigned long long* __restrict__ dtime,
                            unsigned long long* __restrict__ dfinal, uint32_t* __restrict__ rec_distinct, size_t n) {
    size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    uint32_t r = perm[s];
    uint32_t d = head_scan[s] + head[s] - 1;
    rec_distinct[r] = d;
    if (head[s]) {
        dwords[3 * (size_t)d] = w0[r];
        dwords[3 * (size_t)d + 1] = w1[r];
        dwords[3 * (size_t)d + 2] = w2[r];
    }
    atomicMin(&dtime[d], (unsigned long long)time[r]);
    if (weight[r]) atomicAdd(&dfinal[d], (unsigned long long)weight[r]);
}

}  // namespace

std::vector<uint32_t> list_members(const uint64_t* w) {
    std::vector<uint32_t> out;
    for (int k = 0; k < kListMax; ++k) {
        uint64_t v = (w[k / 3] >> (21 * (k % 3))) & 0x1FFFFFull;
        if (v == 0) break;
        out.push_back((uint32_t)(v - 1));
    }
    return out;
}

ListTable dedupe_lists(pma_ctx* c, const uint64_t* w0, const uint64_t* w1, const uint64_t* w2, const uint64_t* time,
                       const uint32_t* weight, size_t R, uint32_t* rec_distinct) {
    ListTable t;
    if (R == 0) return t;
    uint32_t* perm = c->buf<uint32_t>("list_perm", R);
    uint32_t* perm2 = c->buf<uint32_t>("list_perm2", R);
    uint64_t* keys = c->buf<uint64_t>("list_keys", R);
    uint64_t* keys2 = c->buf<uint64_t>("list_keys2", R);
    iota_u32<<<grid_for(R, 256), 256, 0, c->stream>>>(perm, R);
    PMA_LAUNCHED(c);
    const uint64_t* words[3] = {w0, w1, w2};
    for (int k = 2; k >= 0; --k) {  // least significant word first: stable passes give lexicographic order
        gather_u64<<<grid_for(R, 256), 256, 0, c->stream>>>(words[k], perm, keys, R);
        PMA_LAUNCHED(c);
        radix_sort_pairs_u64(c, keys, keys2, perm, perm2, R, 0, 63);
        std::swap(perm, perm2);
    }
    uint32_t* head = c->buf<uint32_t>("list_head", R);
    uint32_t* hscan = c->buf<uint32_t>("list_hscan", R);
    mark_heads<<<grid_for(R, 256), 256, 0, c->stream>>>(w0, w1, w2, perm, head, R);
    PMA_LAUNCHED(c);
    uint64_t D = exclusive_scan_u32(c, head, hscan, R);
    uint64_t* dwords = c->buf<uint64_t>("list_dwords", 3 * D);
    unsigned long long* dtime = c->buf<unsigned long long>("list_dtime", D);
    unsigned long long* dfinal = c->buf<unsigned long long>("list_dfinal", D);
    PMA_HIP(hipMemsetAsync(dtime, 0xFF, D * sizeof(uint64_t), c->stream));
    PMA_HIP(hipMemsetAsync(dfinal, 0, D * sizeof(uint64_t), c->stream));
    reduce_runs<<<grid_for(R, 256), 256, 0, c->stream>>>(w0, w1, w2, time, weight, perm, head, hscan, dwords, dtime,
                                                         dfinal, rec_distinct, R);
    PMA_LAUNCHED(c);
    t.distinct = (uint32_t)D;
    t.words.resize(3 * D);
    t.min_time.resize(D);
    t.finals.resize(D);
    PMA_HIP(hipMemcpyAsync(t.words.data(), dwords, 3 * D * sizeof(uint64_t), hipMemcpyDeviceToHost, c->stream));
    PMA_HIP(hipMemcpyAsync(t.min_time.data(), dtime, D * sizeof(uint64_t), hipMemcpyDeviceToHost, c->stream));
    PMA_HIP(hipMemcpyAsync(t.finals.data(), dfinal, D * sizeof(uint64_t), hipMemcpyDeviceToHost, c->stream));
    PMA_HIP(hipStreamSynchronize(c->stream));
    return t;
}

}  // namespace pma
