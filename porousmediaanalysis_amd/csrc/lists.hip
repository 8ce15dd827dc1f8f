// Exact deduplication of ordered pore lists (throat keys) of any length.
//
// The reference keys shared pixel types by the exact ordered list of pore IDs
// (std::map<std::vector<PoreThroatIdType>, ...>, BasePoreAndThroatComputer.cpp:112-133), and
// every list it ever forms is a prefix of the final list of some voxel (a list only grows by
// push_back, :95-99).  So the distinct lists are the nodes of a prefix trie over the final
// lists: the list of length L + 1 is (its length-L prefix, its last element).  The trie is
// built one length at a time: at length L + 1 every sequence longer than L contributes the
// key (node of its length-L prefix << 32 | its element L); a radix sort and a run split give
// the distinct nodes of that length.  No hashing is involved, so distinct lists are never
// merged.
#include "pma_internal.h"

#include <algorithm>

namespace pma {
namespace {

constexpr uint32_t kNodeBit = kTrieNode;

__global__ void length_keys(const uint32_t* __restrict__ len, size_t S, uint32_t* __restrict__ key,
                            uint32_t* __restrict__ val, unsigned int* __restrict__ max_len) {
    size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    key[s] = 0xFFFFFFFFu - len[s];  // ascending key = descending length
    val[s] = (uint32_t)s;
    atomicMax(max_len, len[s]);
}

// Counts per length: each workgroup counts its grid-stride share in LDS (lengths below
// `lds_bins`), then adds its non-zero bins once to the global histogram; longer lengths (rare) go
// to the global histogram directly.
__global__ void __launch_bounds__(256) length_histogram(const uint32_t* __restrict__ len, size_t S, uint32_t max_len,
                                                        uint32_t lds_bins, unsigned int* __restrict__ hist) {
    extern __shared__ uint32_t lhist[];
    for (uint32_t i = threadIdx.x; i < lds_bins; i += blockDim.x) lhist[i] = 0;
    __syncthreads();
    for (size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x; s < S; s += (size_t)gridDim.x * blockDim.x) {
        const uint32_t l = len[s];
        if (l < lds_bins) atomicAdd(&lhist[l], 1u);
        else if (l <= max_len) atomicAdd(&hist[l], 1u);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < lds_bins; i += blockDim.x)
        if (lhist[i]) atomicAdd(&hist[i], lhist[i]);
}

// Kernel ids of the in-kernel checks (code = id << 8 | check).
enum : uint32_t { kTrieStart = 0x41, kTrieKeys, kTrieAssign };

__global__ void trie_start(const uint32_t* __restrict__ act, size_t A, size_t S, const uint32_t* __restrict__ elems,
                           const uint64_t* __restrict__ off, uint64_t elem_cap, uint32_t* __restrict__ cur,
                           uint32_t* __restrict__ err) {
    size_t a = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= A) return;
    const uint32_t s = act[a];
    if (s >= S || off[s] >= elem_cap) {
        fail_check(err, kTrieStart << 8 | 1);
        cur[a] = 0;
        return;
    }
    cur[a] = elems[off[s]];
}

__global__ void trie_keys(const uint32_t* __restrict__ act, size_t A, size_t S, const uint32_t* __restrict__ elems,
                          const uint64_t* __restrict__ off, uint64_t elem_cap, const uint32_t* __restrict__ cur,
                          uint32_t L, uint64_t* __restrict__ key, uint32_t* __restrict__ val,
                          uint32_t* __restrict__ err) {
    size_t a = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= A) return;
    const uint32_t s = act[a];
    val[a] = (uint32_t)a;
    if (s >= S || off[s] + L >= elem_cap) {
        fail_check(err, kTrieKeys << 8 | 1);
        key[a] = 0;
        return;
    }
    key[a] = ((uint64_t)cur[a] << 32) | elems[off[s] + L];
}

__global__ void trie_heads(const uint64_t* __restrict__ key, size_t A, uint32_t* __restrict__ head) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < A) head[i] = (i == 0 || key[i] != key[i - 1]) ? 1u : 0u;
}

__global__ void trie_assign(const uint64_t* __restrict__ key, const uint32_t* __restrict__ val,
                            const uint32_t* __restrict__ head, const uint32_t* __restrict__ hscan, size_t A, size_t S,
                            uint32_t dbase, uint32_t L, const uint32_t* __restrict__ act,
                            const uint32_t* __restrict__ len, const uint64_t* __restrict__ rbase, uint64_t R,
                            const uint64_t* __restrict__ rec_time, const uint32_t* __restrict__ weight,
                            uint32_t* __restrict__ cur, uint32_t* __restrict__ rec_node,
                            uint64_t* __restrict__ node_key, unsigned long long* __restrict__ node_time,
                            unsigned long long* __restrict__ node_final, uint32_t* __restrict__ err) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A) return;
    const uint32_t a = val[i];
    const uint64_t d = (uint64_t)dbase + hscan[i] + head[i] - 1;
    const uint32_t s = a < A ? act[a] : kNodeBit;
    const uint64_t r = s < S ? rbase[s] + (L - 1) : R;  // record of the prefix of length L + 1
    if (d >= R || r >= R) {
        fail_check(err, kTrieAssign << 8 | 1);
        return;
    }
    cur[a] = kNodeBit | (uint32_t)d;
    rec_node[r] = (uint32_t)d;
    if (head[i]) node_key[d] = key[i];
    if (rec_time) atomicMin(&node_time[d], (unsigned long long)rec_time[r]);
    if (L + 1 == len[s] && weight) atomicAdd(&node_final[d], (unsigned long long)weight[s]);
}

}  // namespace

std::vector<uint32_t> PrefixTrie::members(uint32_t d) const {
    std::vector<uint32_t> rev;
    uint64_t k = key[d];
    while (true) {
        rev.push_back((uint32_t)k);
        const uint32_t parent = (uint32_t)(k >> 32);
        if (!(parent & kNodeBit)) {
            rev.push_back(parent);
            break;
        }
        k = key[parent & ~kNodeBit];
    }
    std::reverse(rev.begin(), rev.end());
    return rev;
}

PrefixTrie build_prefix_trie(pma_ctx* c, const uint32_t* elems, uint64_t elem_cap, const uint64_t* seq_off,
                             const uint32_t* seq_len, const uint64_t* rbase, const uint64_t* rec_time,
                             const uint32_t* seq_weight, size_t S, uint64_t R, uint32_t* rec_node) {
    PrefixTrie t;
    if (S == 0 || R == 0) return t;
    PMA_REQUIRE(R < kNodeBit, PMA_ERR_CAPACITY, "more than 2^31 throat-key formations");
    // sequences by descending length: the ones longer than L are a prefix of this order
    uint32_t* lkey = c->buf<uint32_t>("trie_lkey", S);
    uint32_t* lkey2 = c->buf<uint32_t>("trie_lkey2", S);
    uint32_t* lval = c->buf<uint32_t>("trie_lval", S);
    uint32_t* act = c->buf<uint32_t>("trie_act", S);
    unsigned int* dmax = c->buf<unsigned int>("trie_max_len", 1);
    uint32_t* err = c->buf<uint32_t>("trie_err", 1);
    PMA_HIP(hipMemsetAsync(dmax, 0, 4, c->stream));
    PMA_HIP(hipMemsetAsync(err, 0, 4, c->stream));
    length_keys<<<grid_for(S, 256), 256, 0, c->stream>>>(seq_len, S, lkey, lval, dmax);
    PMA_CHECKED(c, "length_keys", err);
    radix_sort_pairs_u32(c, lkey, lkey2, lval, act, S);
    const uint32_t max_len = read_scalar(c, dmax);
    unsigned int* dhist = c->buf<unsigned int>("trie_hist", (size_t)max_len + 1);
    PMA_HIP(hipMemsetAsync(dhist, 0, ((size_t)max_len + 1) * 4, c->stream));
    const uint32_t lds_bins = std::min<uint32_t>(max_len + 1, 8192);
    length_histogram<<<std::min(grid_for(S, 256), 1024u), 256, lds_bins * 4, c->stream>>>(seq_len, S, max_len,
                                                                                          lds_bins, dhist);
    PMA_CHECKED(c, "length_histogram", err);
    std::vector<uint32_t> hist((size_t)max_len + 1);
    PMA_HIP(hipMemcpyAsync(hist.data(), dhist, hist.size() * 4, hipMemcpyDeviceToHost, c->stream));
    PMA_HIP(hipStreamSynchronize(c->stream));
    // longer[L] = number of sequences with length > L
    std::vector<uint64_t> longer((size_t)max_len + 1, 0);
    for (size_t L = max_len; L-- > 0;) longer[L] = longer[L + 1] + hist[L + 1];
    if (max_len < 2) return t;

    const size_t A1 = longer[1];
    uint32_t* cur = c->buf<uint32_t>("trie_cur", A1);
    uint64_t* key = c->buf<uint64_t>("trie_key", A1);
    uint64_t* key2 = c->buf<uint64_t>("trie_key2", A1);
    uint32_t* val = c->buf<uint32_t>("trie_val", A1);
    uint32_t* val2 = c->buf<uint32_t>("trie_val2", A1);
    uint32_t* head = c->buf<uint32_t>("trie_head", A1);
    uint32_t* hscan = c->buf<uint32_t>("trie_hscan", A1);
    uint64_t* node_key = c->buf<uint64_t>("trie_node_key", R);
    unsigned long long* node_time = c->buf<unsigned long long>("trie_node_time", R);
    unsigned long long* node_final = c->buf<unsigned long long>("trie_node_final", R);
    PMA_HIP(hipMemsetAsync(node_time, 0xFF, R * 8, c->stream));
    PMA_HIP(hipMemsetAsync(node_final, 0, R * 8, c->stream));
    PMA_REQUIRE(A1 <= S, PMA_ERR_INTERNAL, "prefix trie: more active sequences than sequences");
    trie_start<<<grid_for(A1, 256), 256, 0, c->stream>>>(act, A1, S, elems, seq_off, elem_cap, cur, err);
    PMA_CHECKED(c, "trie_start", err);
    uint64_t D = 0;
    for (uint32_t L = 1; L < max_len; ++L) {
        const size_t A = longer[L];
        if (A == 0) break;
        trie_keys<<<grid_for(A, 256), 256, 0, c->stream>>>(act, A, S, elems, seq_off, elem_cap, cur, L, key, val, err);
        PMA_CHECKED(c, "trie_keys", err);
        radix_sort_pairs_u64(c, key, key2, val, val2, A);
        trie_heads<<<grid_for(A, 256), 256, 0, c->stream>>>(key2, A, head);
        PMA_CHECKED(c, "trie_heads", err);
        const uint64_t DL = exclusive_scan_u32(c, head, hscan, A);
        PMA_REQUIRE(D + DL <= R, PMA_ERR_INTERNAL, "prefix trie: more nodes than formations");
        trie_assign<<<grid_for(A, 256), 256, 0, c->stream>>>(key2, val2, head, hscan, A, S, (uint32_t)D, L, act,
                                                             seq_len, rbase, R, rec_time, seq_weight, cur, rec_node,
                                                             node_key, node_time, node_final, err);
        PMA_CHECKED(c, "trie_assign", err);
        D += DL;
    }
    PMA_REQUIRE(read_scalar(c, err) == 0, PMA_ERR_INTERNAL,
                "prefix trie: in-kernel check " + std::to_string(read_scalar(c, err)) + " failed");
    t.distinct = (uint32_t)D;
    t.key.resize(D);
    t.min_time.resize(D);
    t.finals.resize(D);
    PMA_HIP(hipMemcpyAsync(t.key.data(), node_key, D * 8, hipMemcpyDeviceToHost, c->stream));
    PMA_HIP(hipMemcpyAsync(t.min_time.data(), node_time, D * 8, hipMemcpyDeviceToHost, c->stream));
    PMA_HIP(hipMemcpyAsync(t.finals.data(), node_final, D * 8, hipMemcpyDeviceToHost, c->stream));
    PMA_HIP(hipStreamSynchronize(c->stream));
    return t;
}

}  // namespace pma
