// Exact GPU emulation of libstdc++ std::sort (introsort) with comp(a, b) = a.key > b.key.
//
// The reference orders its maximum balls with std::sort (PoreAndThroatComputer.cpp:121-122
// via StlUtilities.h:95-98); the order of equal radii decides pore seeds and IDs, so the
// permutation itself must be reproduced, not just sortedness.  libstdc++'s algorithm
// (bits/stl_algo.h: __introsort_loop, __unguarded_partition_pivot, __move_median_to_first,
// __final_insertion_sort, threshold 16, depth limit 2 * floor(log2 n), heap sort fallback)
// is restated level by level:
//   * every segment longer than 16 is partitioned at once.  Hoare's unguarded partition of
//     [first+1, last) around the median moved to *first swaps the k-th "left stop"
//     (ascending positions with key <= P) with the k-th "right stop" (descending positions
//     in [first, last) with key >= P) for every k with Lo_k < Ro_k, and returns
//     min(Lo_{K+1}, Ro_K).  Stops are ranked with one prefix sum per level.
//   * segments whose depth limit reaches 0 are heap-sorted by one thread with libstdc++'s
//     make_heap / sort_heap steps.
//   * the final insertion sort never moves an element out of its leaf (leaves are mutually
//     ordered and the insertion is strict), so it is an insertion sort of each leaf.
// Records are u64 with the key in the high 32 bits and the payload in the low 32 bits.
#include "pma_internal.h"

namespace pma {
namespace {

constexpr uint32_t kInactive = 0xFFFFFFFFu;
constexpr int kThreshold = 16;

// Kernel ids of the in-kernel checks (code = id << 8 | check); every index read from memory
// (segment bounds, stop positions, ranks, leaves) is tested before it is used.
enum : uint32_t { kHeapSort = 0x21, kChoosePivots, kStopFlags, kRankStops, kFindCuts, kApplySwaps, kCountChildren,
                  kWriteChildren, kRelabel, kInsertion };

__device__ __forceinline__ bool comp(uint64_t a, uint64_t b) { return (a >> 32) > (b >> 32); }

struct Seg {
    uint32_t first, last;
    int depth;
    uint32_t pad;
};

__device__ __forceinline__ bool seg_ok(const Seg& g, size_t M) { return g.first <= g.last && g.last <= M; }

__device__ void push_heap_(uint64_t* a, long hole, long top, uint64_t v) {
    long p = (hole - 1) / 2;
    while (hole > top && comp(a[p], v)) {
        a[hole] = a[p];
        hole = p;
        p = (hole - 1) / 2;
    }
    a[hole] = v;
}

__device__ void adjust_heap_(uint64_t* a, long hole, long len, uint64_t v) {
    const long top = hole;
    long sc = hole;
    while (sc < (len - 1) / 2) {
        sc = 2 * (sc + 1);
        if (comp(a[sc], a[sc - 1])) sc--;
        a[hole] = a[sc];
        hole = sc;
    }
    if ((len & 1) == 0 && sc == (len - 2) / 2) {
        sc = 2 * (sc + 1);
        a[hole] = a[sc - 1];
        hole = sc - 1;
    }
    push_heap_(a, hole, top, v);
}

// std::__partial_sort(first, last, last): __make_heap then __sort_heap.
__global__ void heap_sort_segments(uint64_t* __restrict__ A, size_t M, const Seg* __restrict__ segs, uint32_t nseg,
                                   const uint32_t* __restrict__ heap_flag, uint32_t* __restrict__ err) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nseg || !heap_flag[s]) return;
    if (!seg_ok(segs[s], M)) {
        fail_check(err, kHeapSort << 8 | 1);
        return;
    }
    uint64_t* a = A + segs[s].first;
    long n = (long)segs[s].last - segs[s].first;
    if (n >= 2) {
        long p = (n - 2) / 2;
        while (true) {
            adjust_heap_(a, p, n, a[p]);
            if (p == 0) break;
            --p;
        }
    }
    long last = n;
    while (last > 1) {
        --last;
        uint64_t v = a[last];
        a[last] = a[0];
        adjust_heap_(a, 0, last, v);
    }
}

// Median of three to *first (or a heap-sort flag when the depth limit is exhausted).
__global__ void choose_pivots(uint64_t* __restrict__ A, size_t M, const Seg* __restrict__ segs, uint32_t nseg,
                              uint32_t* __restrict__ pivot, uint32_t* __restrict__ heap_flag,
                              uint32_t* __restrict__ err) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nseg) return;
    Seg g = segs[s];
    if (!seg_ok(g, M) || g.last - g.first <= (uint32_t)kThreshold) {
        fail_check(err, kChoosePivots << 8 | 1);
        heap_flag[s] = 1;  // a heap flag keeps the later kernels away from this segment
        return;
    }
    if (g.depth == 0) {
        heap_flag[s] = 1;
        return;
    }
    heap_flag[s] = 0;
    uint32_t first = g.first, last = g.last, mid = first + (last - first) / 2;
    uint32_t a = first + 1, b = mid, c = last - 1, m;
    uint64_t va = A[a], vb = A[b], vc = A[c];
    if (comp(va, vb)) m = comp(vb, vc) ? b : (comp(va, vc) ? c : a);
    else if (comp(va, vc)) m = a;
    else m = comp(vb, vc) ? c : b;
    uint64_t t = A[first];
    A[first] = A[m];
    A[m] = t;
    pivot[s] = (uint32_t)(A[first] >> 32);
}

// Packed stop flags: low 32 bits = left stop, high 32 bits = right stop.
__global__ void stop_flags(const uint64_t* __restrict__ A, const uint32_t* __restrict__ seg_of,
                           const Seg* __restrict__ segs, uint32_t nseg, const uint32_t* __restrict__ pivot,
                           const uint32_t* __restrict__ heap_flag, unsigned long long* __restrict__ flags, size_t M,
                           uint32_t* __restrict__ err) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i > M) return;
    unsigned long long f = 0;
    if (i < M) {
        uint32_t s = seg_of[i];
        if (s != kInactive && s >= nseg) {
            fail_check(err, kStopFlags << 8 | 1);
        } else if (s != kInactive && !heap_flag[s]) {
            uint32_t key = (uint32_t)(A[i] >> 32), P = pivot[s];
            bool L = i > segs[s].first && key <= P;
            bool R = key >= P;
            f = (unsigned long long)L | ((unsigned long long)R << 32);
        }
    }
    flags[i] = f;
}

__global__ void rank_stops(const unsigned long long* __restrict__ flags, const unsigned long long* __restrict__ pre,
                           const uint32_t* __restrict__ seg_of, const Seg* __restrict__ segs, uint32_t nseg,
                           uint32_t* __restrict__ posL, uint32_t* __restrict__ posR, uint32_t* __restrict__ rankL,
                           size_t M, uint32_t* __restrict__ err) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    unsigned long long f = flags[i];
    rankL[i] = 0;
    if (!f) return;
    uint32_t s = seg_of[i];
    if (s >= nseg || !seg_ok(segs[s], M) || i < segs[s].first || i >= segs[s].last) {
        fail_check(err, kRankStops << 8 | 1);
        return;
    }
    Seg g = segs[s];
    unsigned long long base = pre[g.first], here = pre[i], end = pre[g.last];
    uint32_t inclL = (uint32_t)((here - base) & 0xFFFFFFFFull) + (uint32_t)(f & 1ull);
    uint32_t inclR = (uint32_t)((here - base) >> 32) + (uint32_t)(f >> 32);
    uint32_t totR = (uint32_t)((end - base) >> 32);
    const uint32_t width = g.last - g.first;
    if (inclL > width || inclR > totR || totR > width || ((f & 1ull) && inclL == 0) || ((f >> 32) && inclR == 0)) {
        fail_check(err, kRankStops << 8 | 2);
        return;
    }
    if (f & 1ull) {
        posL[g.first + inclL - 1] = (uint32_t)i;
        rankL[i] = inclL;
    }
    if (f >> 32) posR[g.first + (totR - inclR)] = (uint32_t)i;  // rank from the right = totR - inclR + 1
}

__global__ void find_cuts(const unsigned long long* __restrict__ pre, size_t M, const Seg* __restrict__ segs,
                          uint32_t nseg, const uint32_t* __restrict__ heap_flag, const uint32_t* __restrict__ posL,
                          const uint32_t* __restrict__ posR, uint32_t* __restrict__ swaps, uint32_t* __restrict__ cut,
                          uint32_t* __restrict__ err) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nseg || heap_flag[s]) return;
    Seg g = segs[s];
    unsigned long long tot = pre[g.last] - pre[g.first];
    uint32_t totL = (uint32_t)(tot & 0xFFFFFFFFull), totR = (uint32_t)(tot >> 32);
    if (totL > g.last - g.first || totR > g.last - g.first) {
        fail_check(err, kFindCuts << 8 | 1);
        swaps[s] = 0;
        cut[s] = g.first + 1;
        return;
    }
    uint32_t lo = 0, hi = min(totL, totR);  // largest K in [lo, hi] with posL[K-1] < posR[K-1]
    while (lo < hi) {
        uint32_t k = (lo + hi + 1) / 2;
        if (posL[g.first + k - 1] < posR[g.first + k - 1]) lo = k;
        else hi = k - 1;
    }
    uint32_t K = lo;
    uint32_t c = K < totL ? posL[g.first + K] : 0xFFFFFFFFu;
    if (K >= 1) c = min(c, posR[g.first + K - 1]);
    if (c <= g.first || c >= g.last) {  // Hoare's cut lies strictly inside the segment
        fail_check(err, kFindCuts << 8 | 2);
        K = 0;
        c = g.first + 1;
    }
    swaps[s] = K;
    cut[s] = c;
}

__global__ void apply_swaps(uint64_t* __restrict__ A, const uint32_t* __restrict__ rankL,
                            const uint32_t* __restrict__ seg_of, const Seg* __restrict__ segs, uint32_t nseg,
                            const uint32_t* __restrict__ swaps, const uint32_t* __restrict__ posR, size_t M,
                            uint32_t* __restrict__ err) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    uint32_t r = rankL[i];
    if (r == 0) return;
    uint32_t s = seg_of[i];
    if (s >= nseg) {
        fail_check(err, kApplySwaps << 8 | 1);
        return;
    }
    if (r > swaps[s]) return;
    const size_t pr = (size_t)segs[s].first + r - 1;
    const uint32_t j = pr < M ? posR[pr] : kInactive;
    if (j >= M) {
        fail_check(err, kApplySwaps << 8 | 2);
        return;
    }
    uint64_t t = A[i];
    A[i] = A[j];
    A[j] = t;
}

// Children of each partitioned segment: count active ones (> 16) and record leaves.
__global__ void count_children(const Seg* __restrict__ segs, uint32_t nseg, const uint32_t* __restrict__ heap_flag,
                               const uint32_t* __restrict__ cut, uint32_t* __restrict__ nchild,
                               uint2* __restrict__ leaves, uint32_t* __restrict__ nleaves, size_t leaf_cap,
                               uint32_t* __restrict__ err) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nseg) return;
    uint32_t k = 0;
    if (!heap_flag[s]) {
        Seg g = segs[s];
        uint32_t c = cut[s];
        if (c - g.first > (uint32_t)kThreshold) {
            ++k;
        } else {
            const uint32_t l = atomicAdd(nleaves, 1u);
            if (l < leaf_cap) leaves[l] = make_uint2(g.first, c);
            else fail_check(err, kCountChildren << 8 | 1);
        }
        if (g.last - c > (uint32_t)kThreshold) {
            ++k;
        } else {
            const uint32_t l = atomicAdd(nleaves, 1u);
            if (l < leaf_cap) leaves[l] = make_uint2(c, g.last);
            else fail_check(err, kCountChildren << 8 | 1);
        }
    }
    nchild[s] = k;
}

__global__ void write_children(const Seg* __restrict__ segs, uint32_t nseg, const uint32_t* __restrict__ heap_flag,
                               const uint32_t* __restrict__ cut, const uint32_t* __restrict__ child_off,
                               Seg* __restrict__ next, size_t next_cap, uint32_t* __restrict__ err) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nseg || heap_flag[s]) return;
    Seg g = segs[s];
    uint32_t c = cut[s], o = child_off[s];
    const uint32_t need = (c - g.first > (uint32_t)kThreshold) + (g.last - c > (uint32_t)kThreshold);
    if ((size_t)o + need > next_cap) {
        fail_check(err, kWriteChildren << 8 | 1);
        return;
    }
    if (c - g.first > (uint32_t)kThreshold) next[o++] = Seg{g.first, c, g.depth - 1, 0};
    if (g.last - c > (uint32_t)kThreshold) next[o++] = Seg{c, g.last, g.depth - 1, 0};
}

__global__ void relabel_elements(uint32_t* __restrict__ seg_of, const Seg* __restrict__ segs, uint32_t nseg,
                                 const uint32_t* __restrict__ heap_flag, const uint32_t* __restrict__ cut,
                                 const uint32_t* __restrict__ child_off, size_t M, uint32_t* __restrict__ err) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    uint32_t s = seg_of[i];
    if (s == kInactive) return;
    if (s >= nseg) {
        fail_check(err, kRelabel << 8 | 1);
        seg_of[i] = kInactive;
        return;
    }
    if (heap_flag[s]) {
        seg_of[i] = kInactive;
        return;
    }
    Seg g = segs[s];
    uint32_t c = cut[s], o = child_off[s];
    bool left_active = c - g.first > (uint32_t)kThreshold;
    if (i < c) seg_of[i] = left_active ? o : kInactive;
    else seg_of[i] = (g.last - c > (uint32_t)kThreshold) ? o + (left_active ? 1 : 0) : kInactive;
}

__global__ void insertion_sort_leaves(uint64_t* __restrict__ A, size_t M, const uint2* __restrict__ leaves, uint32_t n,
                                      uint32_t* __restrict__ err) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    uint2 L = leaves[t];
    if (L.x > L.y || L.y > M) {
        fail_check(err, kInsertion << 8 | 1);
        return;
    }
    for (uint32_t i = L.x + 1; i < L.y; ++i) {
        uint64_t v = A[i];
        uint32_t j = i;
        while (j > L.x && comp(v, A[j - 1])) {
            A[j] = A[j - 1];
            --j;
        }
        A[j] = v;
    }
}

__global__ void fill_u32(uint32_t* p, uint32_t v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace

void sort_like_libstdcxx_desc(pma_ctx* c, uint64_t* A, size_t M) {
    if (M <= 1) return;
    PMA_REQUIRE(M < 0xFFFFFFF0ull, PMA_ERR_INVALID_ARGUMENT, "sort: too many elements");
    const size_t leaf_cap = M + 2;
    uint2* leaves = c->buf<uint2>("sort_leaves", leaf_cap);
    uint32_t* nleaves = c->buf<uint32_t>("sort_nleaves", 1);
    uint32_t* err = c->buf<uint32_t>("sort_err", 1);
    PMA_HIP(hipMemsetAsync(nleaves, 0, sizeof(uint32_t), c->stream));
    PMA_HIP(hipMemsetAsync(err, 0, sizeof(uint32_t), c->stream));
    if (M <= (size_t)kThreshold) {
        uint2 one = make_uint2(0, (uint32_t)M);
        upload(c, leaves, &one, sizeof one);
        insertion_sort_leaves<<<1, 64, 0, c->stream>>>(A, M, leaves, 1, err);
        PMA_CHECKED(c, "insertion_sort_leaves", err);
        PMA_REQUIRE(read_scalar(c, err) == 0, PMA_ERR_INTERNAL, "sort: in-kernel check failed");
        return;
    }
    int lg = 0;
    for (size_t v = M; v > 1; v >>= 1) ++lg;
    // every active segment holds more than 16 elements and active segments are disjoint
    const size_t cap = M / (kThreshold + 1) + 2;
    Seg* segs = c->buf<Seg>("sort_segs_a", cap);
    Seg* next = c->buf<Seg>("sort_segs_b", cap);
    uint32_t* seg_of = c->buf<uint32_t>("sort_seg_of", M);
    uint32_t* pivot = c->buf<uint32_t>("sort_pivot", cap);
    uint32_t* heap_flag = c->buf<uint32_t>("sort_heap", cap);
    uint32_t* swaps = c->buf<uint32_t>("sort_swaps", cap);
    uint32_t* cut = c->buf<uint32_t>("sort_cut", cap);
    uint32_t* nchild = c->buf<uint32_t>("sort_nchild", cap);
    uint32_t* child_off = c->buf<uint32_t>("sort_child_off", cap);
    unsigned long long* flags = c->buf<unsigned long long>("sort_flags", M + 1);
    unsigned long long* pre = c->buf<unsigned long long>("sort_pre", M + 1);
    uint32_t* posL = c->buf<uint32_t>("sort_posL", M);
    uint32_t* posR = c->buf<uint32_t>("sort_posR", M);
    uint32_t* rankL = c->buf<uint32_t>("sort_rankL", M);

    Seg root{0, (uint32_t)M, 2 * lg, 0};
    upload(c, segs, &root, sizeof root);
    fill_u32<<<grid_for(M, 256), 256, 0, c->stream>>>(seg_of, 0u, M);
    PMA_CHECKED(c, "fill_u32", err);
    uint32_t nseg = 1;
    const unsigned eb = grid_for(M + 1, 256);
    while (nseg > 0) {
        PMA_REQUIRE(nseg <= cap, PMA_ERR_INTERNAL, "sort: more active segments than capacity");
        const unsigned sb = grid_for(nseg, 256);
        choose_pivots<<<sb, 256, 0, c->stream>>>(A, M, segs, nseg, pivot, heap_flag, err);
        PMA_CHECKED(c, "choose_pivots", err);
        heap_sort_segments<<<sb, 256, 0, c->stream>>>(A, M, segs, nseg, heap_flag, err);
        PMA_CHECKED(c, "heap_sort_segments", err);
        stop_flags<<<eb, 256, 0, c->stream>>>(A, seg_of, segs, nseg, pivot, heap_flag, flags, M, err);
        PMA_CHECKED(c, "stop_flags", err);
        exclusive_scan_u64(c, reinterpret_cast<const uint64_t*>(flags), reinterpret_cast<uint64_t*>(pre), M + 1);
        rank_stops<<<eb, 256, 0, c->stream>>>(flags, pre, seg_of, segs, nseg, posL, posR, rankL, M, err);
        PMA_CHECKED(c, "rank_stops", err);
        find_cuts<<<sb, 256, 0, c->stream>>>(pre, M, segs, nseg, heap_flag, posL, posR, swaps, cut, err);
        PMA_CHECKED(c, "find_cuts", err);
        apply_swaps<<<eb, 256, 0, c->stream>>>(A, rankL, seg_of, segs, nseg, swaps, posR, M, err);
        PMA_CHECKED(c, "apply_swaps", err);
        count_children<<<sb, 256, 0, c->stream>>>(segs, nseg, heap_flag, cut, nchild, leaves, nleaves, leaf_cap, err);
        PMA_CHECKED(c, "count_children", err);
        uint64_t total = exclusive_scan_u32(c, nchild, child_off, nseg);
        PMA_REQUIRE(total <= cap, PMA_ERR_INTERNAL, "sort: more child segments than capacity");
        write_children<<<sb, 256, 0, c->stream>>>(segs, nseg, heap_flag, cut, child_off, next, cap, err);
        PMA_CHECKED(c, "write_children", err);
        relabel_elements<<<grid_for(M, 256), 256, 0, c->stream>>>(seg_of, segs, nseg, heap_flag, cut, child_off, M,
                                                                  err);
        PMA_CHECKED(c, "relabel_elements", err);
        std::swap(segs, next);
        nseg = (uint32_t)total;
        PMA_REQUIRE(read_scalar(c, err) == 0, PMA_ERR_INTERNAL,
                    "sort: in-kernel check " + std::to_string(read_scalar(c, err)) + " failed");
    }
    uint32_t nl = read_scalar(c, nleaves);
    PMA_REQUIRE(nl <= leaf_cap, PMA_ERR_INTERNAL, "sort: more leaves than capacity");
    if (nl) {
        insertion_sort_leaves<<<grid_for(nl, 256), 256, 0, c->stream>>>(A, M, leaves, nl, err);
        PMA_CHECKED(c, "insertion_sort_leaves", err);
    }
    PMA_REQUIRE(read_scalar(c, err) == 0, PMA_ERR_INTERNAL,
                "sort: in-kernel check " + std::to_string(read_scalar(c, err)) + " failed");
}

}  // namespace pma

// Test hook (not part of the public header): sorts host records in place on the device.
extern "C" int pma_testing_sort_desc(pma_ctx* c, uint64_t* host, size_t n) {
    try {
        PMA_HIP(hipSetDevice(c->device));
        uint64_t* d = c->buf<uint64_t>("testing_sort", std::max<size_t>(n, 1));
        PMA_HIP(hipMemcpyAsync(d, host, n * sizeof(uint64_t), hipMemcpyHostToDevice, c->stream));
        pma::sort_like_libstdcxx_desc(c, d, n);
        PMA_HIP(hipMemcpyAsync(host, d, n * sizeof(uint64_t), hipMemcpyDeviceToHost, c->stream));
        PMA_HIP(hipStreamSynchronize(c->stream));
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "pma_testing_sort_desc: %s\n", e.what());
        return 1;
    }
}
