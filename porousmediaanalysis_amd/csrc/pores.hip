// Pore seeds and pore propagation (PoreAndThroatComputer.cpp:81-269 with
// BasePoreAndThroatComputer.cpp:37-133), reproduced exactly.
//
// Reference (one active area): maximum balls (mask 255) are listed in z, x, y order and
// sorted by r2 with std::sort (:93-123); walking that order, a still-unlabelled centre opens
// pore 127, 128, ... (+127 after a multiple of 255) and a DFS adds the pore to every voxel
// reachable through void neighbours of non-increasing CR (:192-269).  A voxel reached by
// several pores carries the key of the ordered list [p2, p1, p3, ...]; keys are created in
// order of first DFS contact and numbered down from UINT_MAX - 1; zero-volume keys are erased.
//
// What determines each output, and how it is computed here:
//   * seeds: the pores are exactly the regional-maximum plateaus of CR (26- or 6-connected
//     components of equal CR with no higher neighbour); a plateau's seed is its element that
//     comes first in the sorted list, and pores are numbered in that order.  (A non-maximal
//     plateau is reachable, through non-increasing CR, from a regional maximum of higher CR
//     whose seed sorts earlier; a maximal plateau cannot be entered from outside.)  The sort
//     is the exact libstdc++ permutation (sort.hip).
//   * pore sets: set(v) = {p : v reachable from seed(p)} is constant on a plateau and is the
//     union of the sets of the higher plateaus adjacent to it, plus its own pore; it is
//     computed over the plateau DAG level by level in descending CR.
//   * key numbering: the key for set S is created while the largest pore of S runs its DFS,
//     at the first visit of a voxel whose earlier-pore class is S minus that pore; every such
//     prefix class is one creation.  The order of first visits inside one pore's DFS is
//     replayed by one wave per pore (a recursive DFS with neighbours in reverse loop order is
//     the same visit order as the reference's stack DFS), stopping once all its classes have
//     been seen.
//   * volumes = voxels per final set; boundary = some voxel of the pore's reach lies on the
//     image border.
//
// Every kernel is total: an index read from memory is tested against the length of the buffer
// it selects from (lengths are kernel arguments), and a failed test reports a code
// (kernel << 8 | check) to the stage's error flag instead of making the access.  Loops over
// parent chains and merge cursors are bounded by the counts that bound them in a consistent
// state.  So whatever values a kernel reads, it touches only its own buffers.
#include "pma_internal.h"

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <numeric>

namespace pma {
namespace {

constexpr uint32_t kInv = 0xFFFFFFFFu;
#ifndef PMA_PORES_STACK_CAP
#define PMA_PORES_STACK_CAP (1ull << 19)  // DFS frames per replaying wave in the first pass
#endif
#ifndef PMA_PORES_TRAV_CAP
#define PMA_PORES_TRAV_CAP (1ull << 20)  // plateaus one skip traversal may list (then it gives up)
#endif
#ifndef PMA_PORES_VIS_PAGES
#define PMA_PORES_VIS_PAGES 2048u  // visited-voxel pages per replaying wave in the first pass
#endif
#ifndef PMA_PORES_VIS_MIN_BITS
#define PMA_PORES_VIS_MIN_BITS 16u  // voxels per visited-voxel page: 2^bits, at least
#endif
#ifndef PMA_PORES_RING
#define PMA_PORES_RING 512u  // DFS frames of a replaying wave kept in LDS
#endif
#ifndef PMA_PORES_REPLAY_MEM
#define PMA_PORES_REPLAY_MEM (128ull << 30)  // device bytes the replaying waves may hold
#endif
#ifndef PMA_PORES_POOL_MIN
#define PMA_PORES_POOL_MIN 8ull     // pore-set pool entries per plateau, at least
#define PMA_PORES_POOL_MAX 64ull    // ... and at most, while free memory allows
#define PMA_PORES_POOL_FLOOR 1024ull
#endif
#ifndef PMA_UNION_LDS_MAX
#define PMA_UNION_LDS_MAX 65536u  // LDS bytes a workgroup of the bitmap set union may use
#endif

// Kernel ids of the in-kernel checks (the code's high bits).
enum : uint32_t {
    kWriteCentres = 1, kCcUnion, kCcFlatten, kCcNotMax, kFirstInSorted, kSeedFlags, kWriteSeeds, kPlateauIndex,
    kPlateauOfVoxels, kSeedPlateaus, kPlateauEdges, kWriteEdges, kEdgeRanges, kLevelStarts, kSetUnion,
    kReachSizes, kDfsReplay, kWriteLabels, kSingleVolumes, kBelowMin, kPoreLists, kRecordsInOrder
};
constexpr uint32_t kNoClassU = 0xFFFF;  // class of a pore on a plateau whose first member it is
const char* const kKernelNames[] = {"?", "write_centres", "cc_union", "cc_flatten", "cc_not_max",
                                    "first_in_sorted", "seed_flags", "write_seeds", "plateau_index",
                                    "plateau_of_voxels", "seed_plateaus", "plateau_edges", "write_edges",
                                    "edge_ranges", "level_starts", "set_union_level", "reach_sizes",
                                    "dfs_replay", "write_labels", "single_volumes", "below_min", "pore_lists",
                                    "records_in_order"};

__device__ __forceinline__ uint32_t ld_relaxed(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_relaxed(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void decode(uint32_t i, int X, int Y, int& x, int& y, int& z) {
    y = (int)(i % (uint32_t)Y);
    uint32_t r = i / (uint32_t)Y;
    x = (int)(r % (uint32_t)X);
    z = (int)(r / (uint32_t)X);
}

__device__ __forceinline__ bool in_stencil(int k, int d3q7) {
    const int dx = k / 9 - 1, dy = (k / 3) % 3 - 1, dz = k % 3 - 1;
    const int m = abs(dx) + abs(dy) + abs(dz);
    return m != 0 && (!d3q7 || m == 1);
}

// Neighbour index k in the reference's loop order (x outer, y, z inner).
__device__ __forceinline__ bool neighbour(int k, int x, int y, int z, int X, int Y, int Z, uint32_t& j) {
    const int nx = x + k / 9 - 1, ny = y + (k / 3) % 3 - 1, nz = z + k % 3 - 1;
    if (nx < 0 || nx >= X || ny < 0 || ny >= Y || nz < 0 || nz >= Z) return false;
    j = ((uint32_t)nz * (uint32_t)X + (uint32_t)nx) * (uint32_t)Y + (uint32_t)ny;
    return true;
}

__device__ __forceinline__ uint32_t pore_id_of(uint32_t k) { return 127u + 255u * (k / 129u) + (k % 129u); }

// ---- seeds ------------------------------------------------------------------------------
constexpr int kChunk = 16;

__global__ void __launch_bounds__(256) count_centres(const uint8_t* __restrict__ mask, size_t n,
                                                     uint32_t* __restrict__ block_count) {
    __shared__ uint32_t sum;
    if (threadIdx.x == 0) sum = 0;
    __syncthreads();
    size_t base = (size_t)blockIdx.x * 256 * kChunk + (size_t)threadIdx.x * kChunk;
    uint32_t c = 0;
    for (int k = 0; k < kChunk; ++k)
        if (base + k < n) c += mask[base + k] == 255;
    atomicAdd(&sum, c);
    __syncthreads();
    if (threadIdx.x == 0) block_count[blockIdx.x] = sum;
}

__global__ void __launch_bounds__(256) write_centres(const uint8_t* __restrict__ mask, const uint32_t* __restrict__ e,
                                                     size_t n, const uint32_t* __restrict__ block_off,
                                                     unsigned long long* __restrict__ rec, size_t M,
                                                     uint32_t* __restrict__ err) {
    __shared__ uint32_t wsum[4];
    size_t base = (size_t)blockIdx.x * 256 * kChunk + (size_t)threadIdx.x * kChunk;
    uint32_t c = 0;
    for (int k = 0; k < kChunk; ++k)
        if (base + k < n) c += mask[base + k] == 255;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t incl = c;
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t o = __shfl_up(incl, off);
        if (lane >= off) incl += o;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    size_t o = (size_t)block_off[blockIdx.x] + incl - c;
    for (int w = 0; w < wave; ++w) o += wsum[w];
    for (int k = 0; k < kChunk; ++k) {
        size_t i = base + k;
        if (i < n && mask[i] == 255) {
            if (o >= M) {
                fail_check(err, kWriteCentres << 8 | 1);
                return;
            }
            rec[o++] = ((unsigned long long)e[i] << 32) | (unsigned long long)i;
        }
    }
}

// Records (e << 32 | voxel) for maximum balls given in processing order (the test hook's order).
// The order must be non-increasing in r2, as every std::sort by r2 leaves it: only the order of
// equal radii may differ.
__global__ void __launch_bounds__(256) records_in_order(const uint32_t* __restrict__ order, size_t M,
                                                        const uint8_t* __restrict__ mask,
                                                        const uint32_t* __restrict__ e, size_t n,
                                                        unsigned long long* __restrict__ rec,
                                                        uint32_t* __restrict__ err) {
    const size_t k = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= M) return;
    const uint32_t i = order[k];
    if (i >= n || mask[i] != 255) {
        fail_check(err, kRecordsInOrder << 8 | 1);
        return;
    }
    if (k > 0) {
        const uint32_t prev = order[k - 1];
        if (prev < n && e[prev] < e[i]) {
            fail_check(err, kRecordsInOrder << 8 | 2);
            return;
        }
    }
    rec[k] = ((unsigned long long)e[i] << 32) | (unsigned long long)i;
}

__global__ void cc_init(const uint32_t* __restrict__ cr, size_t n, uint32_t* __restrict__ parent) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) parent[i] = cr[i] ? (uint32_t)i : kInv;
}

// Root of x in the plateau forest.  Parents always point to smaller indices (a root is hooked
// under a smaller root, and path halving replaces a parent by one of its ancestors), so a
// chain ends within n links; a parent that does not decrease is reported and ends the walk
// with kInv.  Parents are read and written with agent-scope atomics, so no value is taken
// from a cache line another compute unit has since changed.
__device__ uint32_t find_root(uint32_t* parent, uint32_t x, uint32_t n, uint32_t* err, uint32_t code) {
    for (uint32_t step = 0; step <= n; ++step) {
        const uint32_t p = ld_relaxed(&parent[x]);
        if (p == x) return x;
        if (p > x) {
            fail_check(err, code);
            return kInv;
        }
        const uint32_t g = ld_relaxed(&parent[p]);
        if (g > p) {
            fail_check(err, code);
            return kInv;
        }
        if (g != p) st_relaxed(&parent[x], g);  // path halving: g is an ancestor of x
        x = p;
    }
    fail_check(err, code | 1);
    return kInv;
}

__device__ void unite(uint32_t* parent, uint32_t a, uint32_t b, uint32_t n, uint32_t* err) {
    // Every failed hook means another thread hooked the root, which happens at most n times.
    for (uint32_t step = 0; step <= n; ++step) {
        a = find_root(parent, a, n, err, kCcUnion << 8 | 2);
        b = find_root(parent, b, n, err, kCcUnion << 8 | 2);
        if (a == kInv || b == kInv || a == b) return;
        if (a < b) {
            const uint32_t t = a;
            a = b;
            b = t;
        }
        const uint32_t old = atomicCAS(&parent[a], a, b);  // hook the larger root under the smaller
        if (old == a) return;
        a = old;
    }
    fail_check(err, kCcUnion << 8 | 4);
}

__global__ void __launch_bounds__(256) cc_union(const uint32_t* __restrict__ cr, uint32_t* parent, int X, int Y, int Z,
                                                int d3q7, uint32_t* __restrict__ err) {
    const size_t n = (size_t)X * Y * Z;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t c = cr[i];
    if (!c) return;
    int x, y, z;
    decode((uint32_t)i, X, Y, x, y, z);
    for (int k = 0; k < 13; ++k) {  // the half of the stencil with smaller linear index
        if (!in_stencil(k, d3q7)) continue;
        // k < 13 in loop order means dx < 0, or dx == 0 and (dy < 0 or (dy == 0 and dz < 0))
        uint32_t j;
        if (!neighbour(k, x, y, z, X, Y, Z, j)) continue;
        if (cr[j] == c) unite(parent, (uint32_t)i, j, (uint32_t)n, err);
    }
}

// Root of x without writing any parent.  cc_flatten walks with this: there, thread i stores
// the root into parent[i] while other threads still walk chains through i, and a path-halving
// store by such a thread (parent[i] := the grandparent it read before thread i's store) could
// land after it and put a non-root back into parent[i].  Concurrent root stores only shorten
// the chains this walk follows, since each replaces a parent by the same root.
__device__ uint32_t root_of(const uint32_t* parent, uint32_t x, uint32_t n, uint32_t* err, uint32_t code) {
    for (uint32_t step = 0; step <= n; ++step) {
        const uint32_t p = ld_relaxed(&parent[x]);
        if (p == x) return x;
        if (p > x) {
            fail_check(err, code);
            return kInv;
        }
        x = p;
    }
    fail_check(err, code | 1);
    return kInv;
}

// parent[i] := root of i; only thread i writes parent[i] in this launch.
__global__ void cc_flatten(uint32_t* parent, size_t n, uint32_t* __restrict__ err) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || ld_relaxed(&parent[i]) == kInv) return;
    const uint32_t r = root_of(parent, (uint32_t)i, (uint32_t)n, err, kCcFlatten << 8 | 2);
    if (r != kInv) st_relaxed(&parent[i], r);
}

__global__ void __launch_bounds__(256) cc_not_max(const uint32_t* __restrict__ cr, const uint32_t* __restrict__ root,
                                                  uint8_t* __restrict__ not_max, int X, int Y, int Z, int d3q7,
                                                  uint32_t* __restrict__ err) {
    const size_t n = (size_t)X * Y * Z;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t c = cr[i];
    if (!c) return;
    int x, y, z;
    decode((uint32_t)i, X, Y, x, y, z);
    for (int k = 0; k < 27; ++k) {
        if (!in_stencil(k, d3q7)) continue;
        uint32_t j;
        if (!neighbour(k, x, y, z, X, Y, Z, j)) continue;
        if (cr[j] > c) {
            const uint32_t r = root[i];
            if (r >= n) fail_check(err, kCcNotMax << 8 | 1);
            else not_max[r] = 1;
            return;
        }
    }
}

__global__ void first_in_sorted(const unsigned long long* __restrict__ rec, size_t M, const uint32_t* __restrict__ root,
                                size_t n, const uint8_t* __restrict__ not_max, uint32_t* __restrict__ seedpos,
                                uint32_t* __restrict__ err) {
    size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= M) return;
    const uint32_t v = (uint32_t)rec[s];
    const uint32_t r = v < n ? root[v] : kInv;
    if (r >= n) {
        fail_check(err, kFirstInSorted << 8 | 1);
        return;
    }
    if (!not_max[r]) atomicMin(&seedpos[r], (uint32_t)s);
}

__global__ void seed_flags(const unsigned long long* __restrict__ rec, size_t M, const uint32_t* __restrict__ root,
                           size_t n, const uint32_t* __restrict__ seedpos, uint32_t* __restrict__ flag,
                           uint32_t* __restrict__ err) {
    size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= M) return;
    const uint32_t v = (uint32_t)rec[s];
    const uint32_t r = v < n ? root[v] : kInv;
    if (r >= n) {
        fail_check(err, kSeedFlags << 8 | 1);
        flag[s] = 0;
        return;
    }
    flag[s] = seedpos[r] == (uint32_t)s;
}

__global__ void write_seeds(const unsigned long long* __restrict__ rec, size_t M, const uint32_t* __restrict__ flag,
                            const uint32_t* __restrict__ off, unsigned long long* __restrict__ seeds,
                            uint32_t* __restrict__ err) {
    size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= M || !flag[s]) return;
    if (off[s] >= M) fail_check(err, kWriteSeeds << 8 | 1);
    else seeds[off[s]] = rec[s];
}

// ---- plateaus ----------------------------------------------------------------------------
__global__ void root_flags(const uint32_t* __restrict__ root, size_t n, uint32_t* __restrict__ flag) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flag[i] = root[i] == (uint32_t)i;
}

__global__ void plateau_index(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ off, size_t n,
                              uint32_t* __restrict__ pidx, const uint32_t* __restrict__ cr,
                              uint32_t* __restrict__ plat_cr, size_t P, uint32_t* __restrict__ err) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !flag[i]) return;
    if (off[i] >= P) {
        fail_check(err, kPlateauIndex << 8 | 1);
        return;
    }
    pidx[i] = off[i];
    plat_cr[off[i]] = cr[i];
}

// root -> plateau index for every void voxel (in place), plus plateau size, border flag and
// the voxel's rank inside its plateau.
__global__ void plateau_of_voxels(uint32_t* root, const uint32_t* __restrict__ pidx, int X, int Y, int Z, size_t P,
                                  uint32_t* __restrict__ plat_size, uint32_t* __restrict__ plat_border,
                                  uint32_t* __restrict__ vrank, uint32_t* __restrict__ err) {
    const size_t n = (size_t)X * Y * Z;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t r = root[i];
    if (r == kInv) return;
    const uint32_t q = r < n ? pidx[r] : kInv;  // only this thread reads or writes root[i]
    if (q >= P) {
        fail_check(err, kPlateauOfVoxels << 8 | 1);
        root[i] = kInv;
        return;
    }
    root[i] = q;
    vrank[i] = atomicAdd(&plat_size[q], 1u);  // rank of the voxel inside its plateau (any order)
    int x, y, z;
    decode((uint32_t)i, X, Y, x, y, z);
    if (x == 0 || y == 0 || z == 0 || x == X - 1 || y == Y - 1 || z == Z - 1) plat_border[q] = 1;
}

__device__ __forceinline__ unsigned long long mix64(unsigned long long k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return k;
}

// Distinct plateau edges (higher CR -> lower CR) in an open-addressing set.  A voxel skips a
// source plateau it has already inserted for its previous neighbours (at most 26).  The set is
// kept about half full: a wave whose insertions take the fill count past half the capacity (or a
// probe sequence longer than kEdgeProbes) flags an overflow, threads that start later skip their
// work, and the host retries with a table four times larger -- so probe sequences stay short
// instead of sweeping a nearly full table.
constexpr unsigned long long kEdgeProbes = 1024;
__global__ void __launch_bounds__(256) plateau_edges(const uint32_t* __restrict__ cr, const uint32_t* __restrict__ plat,
                                                     int X, int Y, int Z, int d3q7, size_t P,
                                                     unsigned long long* __restrict__ table, unsigned long long mask,
                                                     unsigned long long* __restrict__ fill, uint32_t* __restrict__ overflow,
                                                     uint32_t* __restrict__ err) {
    const size_t n = (size_t)X * Y * Z;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    // every lane reaches the count at the end: no early return
    uint32_t inserted = 0;
    const uint32_t c = i < n ? cr[i] : 0u;
    const uint32_t dst = c ? plat[i] : 0u;
    bool go = c != 0 && !ld_relaxed(overflow);  // after an overflow the host retries with a larger table
    if (go && dst >= P) {
        fail_check(err, kPlateauEdges << 8 | 1);
        go = false;
    }
    if (go) {
        int x, y, z;
        decode((uint32_t)i, X, Y, x, y, z);
        uint32_t done[26];
        int nd = 0;
        for (int k = 0; k < 27 && go; ++k) {
            if (!in_stencil(k, d3q7)) continue;
            uint32_t j;
            if (!neighbour(k, x, y, z, X, Y, Z, j)) continue;
            if (cr[j] <= c) continue;
            const uint32_t src = plat[j];
            if (src >= P) {
                fail_check(err, kPlateauEdges << 8 | 2);
                go = false;
                break;
            }
            bool seen = false;
            for (int q = 0; q < nd; ++q) seen |= done[q] == src;
            if (seen) continue;
            if (nd < 26) done[nd++] = src;
            const unsigned long long key = ((unsigned long long)src << 32) | dst;
            unsigned long long h = mix64(key) & mask;
            for (unsigned long long probe = 0;; ++probe) {
                if (probe > mask || probe >= kEdgeProbes) {
                    atomicOr(overflow, 1u);
                    go = false;
                    break;
                }
                unsigned long long cur = table[h];
                if (cur == key) break;
                if (cur == ~0ull) {
                    unsigned long long old = atomicCAS(&table[h], ~0ull, key);
                    if (old == ~0ull) ++inserted;
                    if (old == ~0ull || old == key) break;
                }
                h = (h + 1) & mask;
            }
        }
    }
    // one atomic per wave on the fill count; passing half the capacity is an overflow
    uint32_t sum = inserted;
    for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off);
    if ((threadIdx.x & 63) == 0 && sum) {
        if (atomicAdd(fill, (unsigned long long)sum) + sum > (mask + 1) / 2) atomicOr(overflow, 1u);
    }
}

__global__ void slot_flags(const unsigned long long* __restrict__ table, size_t cap, uint32_t* __restrict__ flag) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cap) flag[i] = table[i] != ~0ull;
}

// Edge keys re-packed dst-major so a sort groups them by target.
__global__ void write_edges(const unsigned long long* __restrict__ table, size_t cap, const uint32_t* __restrict__ flag,
                            const uint32_t* __restrict__ off, size_t E, unsigned long long* __restrict__ edges,
                            uint32_t* __restrict__ err) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cap || !flag[i]) return;
    if (off[i] >= E) {
        fail_check(err, kWriteEdges << 8 | 1);
        return;
    }
    unsigned long long k = table[i];
    edges[off[i]] = (k << 32) | (k >> 32);
}

__global__ void edge_ranges(const unsigned long long* __restrict__ edges, size_t E, size_t P,
                            uint32_t* __restrict__ in_start, uint32_t* __restrict__ in_count,
                            uint32_t* __restrict__ err) {
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    uint32_t dst = (uint32_t)(edges[e] >> 32);
    if (dst >= P) {
        fail_check(err, kEdgeRanges << 8 | 1);
        return;
    }
    if (e == 0 || (uint32_t)(edges[e - 1] >> 32) != dst) in_start[dst] = (uint32_t)e;
    atomicAdd(&in_count[dst], 1u);
}

__global__ void level_keys(const uint32_t* __restrict__ plat_cr, size_t P, unsigned long long* __restrict__ keys,
                           uint32_t* __restrict__ vals) {
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= P) return;
    keys[q] = ((unsigned long long)(0xFFFFFFFFu - plat_cr[q]) << 32) | q;
    vals[q] = (uint32_t)q;
}

__global__ void level_heads(const unsigned long long* __restrict__ keys, size_t P, uint32_t* __restrict__ head) {
    size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s < P) head[s] = (s == 0 || (keys[s] >> 32) != (keys[s - 1] >> 32));
}

__global__ void level_starts(const uint32_t* __restrict__ head, const uint32_t* __restrict__ off, size_t P,
                             uint32_t* __restrict__ start, size_t L, uint32_t* __restrict__ err) {
    size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= P || !head[s]) return;
    if (off[s] >= L) fail_check(err, kLevelStarts << 8 | 1);
    else start[off[s]] = (uint32_t)s;
}

// Sizes the pore-set stage reads its inputs with (each index read from memory is tested
// against one of them).
struct SetBounds {
    unsigned long long E, P, npores, pool_cap, nsub;
};

// Room for n members in the sub-pool of plateau q: the pool is cut into b.nsub equal sub-pools,
// each with its own counter, so a level's allocations do not all meet on one address.  Returns
// the offset, or ~0 when the sub-pool is full (the host repeats the run with a larger pool).
__device__ __forceinline__ unsigned long long pool_take(unsigned long long* tops, uint32_t q, unsigned long long n,
                                                        const SetBounds& b) {
    const unsigned long long sub = b.pool_cap / b.nsub, sp = q % b.nsub;
    const unsigned long long o = atomicAdd(&tops[sp], n);
    return o + n > sub ? ~0ull : sp * sub + o;
}

// k-way merge of the sorted pore sets of a plateau's higher neighbours and its own pore.
// cur/end hold one cursor per in-edge; out (if not null, with room for out_cap) receives the
// union in ascending order.  Returns the union's length, or kInv after a failed check.  Each
// step emits one member and advances at least one cursor or retires the own pore, so the
// loop ends within the total length of the inputs plus two steps.
__device__ uint32_t merge_sets(uint32_t own, uint32_t e0, uint32_t ne, const unsigned long long* __restrict__ edges,
                               const unsigned long long* __restrict__ set_off, const uint32_t* __restrict__ set_len,
                               const uint32_t* __restrict__ pool, unsigned long long* __restrict__ cur,
                               unsigned long long* __restrict__ end, uint32_t* __restrict__ out, uint32_t out_cap,
                               const SetBounds& b, uint32_t* __restrict__ err) {
    unsigned long long total = 1;
    for (uint32_t e = e0; e < e0 + ne; ++e) {
        const uint32_t src = (uint32_t)edges[e];
        if (src >= b.P || set_off[src] + set_len[src] > b.pool_cap) {
            fail_check(err, kSetUnion << 8 | 3);
            return kInv;
        }
        cur[e] = set_off[src];
        end[e] = set_off[src] + set_len[src];
        total += set_len[src];
    }
    uint32_t m = 0;
    uint32_t own_left = own;
    for (unsigned long long step = 0; step <= total; ++step) {
        uint32_t best = own_left;
        for (uint32_t e = e0; e < e0 + ne; ++e)
            if (cur[e] < end[e]) best = min(best, pool[cur[e]]);
        if (best == kInv) return m;
        if (out) {
            if (m >= out_cap) {
                fail_check(err, kSetUnion << 8 | 4);
                return kInv;
            }
            out[m] = best;
        }
        ++m;
        if (own_left == best) own_left = kInv;
        for (uint32_t e = e0; e < e0 + ne; ++e)
            if (cur[e] < end[e] && pool[cur[e]] == best) ++cur[e];
    }
    fail_check(err, kSetUnion << 8 | 5);
    return kInv;
}

// Pore sets of the plateaus of one CR level (all their sources are on higher levels).
__global__ void set_union_level(const uint32_t* __restrict__ order, uint32_t lo, uint32_t hi,
                                const uint32_t* __restrict__ in_start, const uint32_t* __restrict__ in_count,
                                const unsigned long long* __restrict__ edges, const uint32_t* __restrict__ plat_pore,
                                unsigned long long* __restrict__ set_off, uint32_t* __restrict__ set_len,
                                uint32_t* __restrict__ pool, unsigned long long* __restrict__ pool_top,
                                unsigned long long* __restrict__ cur, unsigned long long* __restrict__ end,
                                SetBounds b, uint32_t* __restrict__ overflow, uint32_t* __restrict__ err) {
    const uint32_t s = lo + blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= hi) return;
    const uint32_t q = order[s];
    if (q >= b.P) {
        fail_check(err, kSetUnion << 8 | 1);
        return;
    }
    const uint32_t own = plat_pore[q];
    const uint32_t e0 = in_start[q], ne = in_count[q];
    if ((own != kInv && own >= b.npores) || (unsigned long long)e0 + ne > b.E) {
        fail_check(err, kSetUnion << 8 | 2);
        return;
    }
    const uint32_t m = merge_sets(own, e0, ne, edges, set_off, set_len, pool, cur, end, nullptr, 0, b, err);
    if (m == kInv) return;
    const unsigned long long o = m ? pool_take(pool_top, q, m, b) : 0ull;
    if (o == ~0ull) {
        atomicOr(overflow, 2u);
        set_off[q] = 0;
        set_len[q] = 0;
        return;
    }
    const uint32_t m2 = merge_sets(own, e0, ne, edges, set_off, set_len, pool, cur, end, pool + o, m, b, err);
    if (m2 != m) {
        fail_check(err, kSetUnion << 8 | 6);
        set_off[q] = 0;
        set_len[q] = 0;
        return;
    }
    set_off[q] = o;
    set_len[q] = m;
}

// The same sets with one wave per plateau: the members of the sets of the plateau's in-edges
// (lane per in-edge) and its own pore are or-ed into a bitmap over pores in LDS, then written
// in ascending order from the bitmap (lane per run of words, offsets by a wave prefix sum).
// Work per plateau is the total length of its inputs plus the bitmap, not their product with the
// in-degree as in the merge.  Waves use only their own LDS slice (no workgroup barrier).
constexpr int kUnionWaves = 4;
__global__ void __launch_bounds__(64 * kUnionWaves)
    set_union_bitmap(const uint32_t* __restrict__ order, uint32_t lo, uint32_t hi,
                     const uint32_t* __restrict__ in_start, const uint32_t* __restrict__ in_count,
                     const unsigned long long* __restrict__ edges, const uint32_t* __restrict__ plat_pore,
                     unsigned long long* __restrict__ set_off, uint32_t* __restrict__ set_len,
                     uint32_t* __restrict__ pool, unsigned long long* __restrict__ pool_top, SetBounds b,
                     uint32_t words, uint32_t* __restrict__ overflow, uint32_t* __restrict__ err) {
    extern __shared__ uint32_t union_bits[];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t* bm = union_bits + (size_t)wave * words;
    const uint32_t s = lo + blockIdx.x * kUnionWaves + wave;
    if (s >= hi) return;  // the whole wave leaves
    const uint32_t q = order[s];
    if (q >= b.P) {
        if (lane == 0) fail_check(err, kSetUnion << 8 | 1);
        return;
    }
    const uint32_t own = plat_pore[q];
    const uint32_t e0 = in_start[q], ne = in_count[q];
    if ((own != kInv && own >= b.npores) || (unsigned long long)e0 + ne > b.E) {
        if (lane == 0) fail_check(err, kSetUnion << 8 | 2);
        return;
    }
    if (ne <= 1) {
        // the source's set with the own pore inserted: no bitmap needed
        uint32_t src = kInv, len = 0;
        unsigned long long off = 0;
        if (ne == 1) {
            src = (uint32_t)edges[e0];
            if (src >= b.P || set_off[src] + set_len[src] > b.pool_cap) {
                if (lane == 0) fail_check(err, kSetUnion << 8 | 3);
                return;
            }
            off = set_off[src];
            len = set_len[src];
        }
        // members below the own pore keep their place, the own pore follows them, the rest shift by one
        uint32_t below_own = 0;
        bool dup = false, bad = false;
        for (uint32_t t = lane; t < len; t += 64) {
            const uint32_t v = pool[off + t];
            bad |= v >= b.npores;
            below_own += v < own;
            dup |= v == own;
        }
        if (__ballot(bad)) {
            if (lane == 0) fail_check(err, kSetUnion << 8 | 3);
            return;
        }
        for (int d = 32; d >= 1; d >>= 1) below_own += __shfl_xor(below_own, d);
        const bool ins = own != kInv && !__ballot(dup);
        const uint32_t total = len + (ins ? 1u : 0u);
        unsigned long long o = 0;
        if (lane == 0 && total) o = pool_take(pool_top, q, total, b);
        o = __shfl(o, 0);
        if (o == ~0ull) {
            if (lane == 0) {
                atomicOr(overflow, 2u);
                set_off[q] = 0;
                set_len[q] = 0;
            }
            return;
        }
        for (uint32_t t = lane; t < len; t += 64) {
            const uint32_t v = pool[off + t];
            pool[o + t + (ins && t >= below_own ? 1u : 0u)] = v;
        }
        if (lane == 0) {
            if (ins) pool[o + below_own] = own;
            set_off[q] = total ? o : 0;
            set_len[q] = total;
        }
        return;
    }
    for (uint32_t w = lane; w < words; w += 64) bm[w] = 0;
    __builtin_amdgcn_wave_barrier();
    bool bad = false;
    if (lane == 0 && own != kInv) atomicOr(&bm[own >> 5], 1u << (own & 31));
    for (uint32_t e = e0 + lane; e < e0 + ne; e += 64) {
        const uint32_t src = (uint32_t)edges[e];
        if (src >= b.P || set_off[src] + set_len[src] > b.pool_cap) {
            bad = true;
            break;
        }
        const unsigned long long off = set_off[src];
        const uint32_t len = set_len[src];
        for (uint32_t t = 0; t < len; ++t) {
            const uint32_t v = pool[off + t];
            if (v >= b.npores) {
                bad = true;
                break;
            }
            atomicOr(&bm[v >> 5], 1u << (v & 31));
        }
    }
    __builtin_amdgcn_wave_barrier();
    if (__ballot(bad)) {
        if (lane == 0) fail_check(err, kSetUnion << 8 | 3);
        return;
    }
    // lane l owns words [l * per, (l + 1) * per)
    const uint32_t per = (words + 63) / 64;
    const uint32_t w0 = min(words, lane * per), w1 = min(words, w0 + per);
    uint32_t cnt = 0;
    for (uint32_t w = w0; w < w1; ++w) cnt += __popc(bm[w]);
    uint32_t incl = cnt;
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t o = __shfl_up(incl, d);
        if (lane >= d) incl += o;
    }
    const uint32_t total = __shfl(incl, 63);
    unsigned long long o = 0;
    if (lane == 0 && total) o = pool_take(pool_top, q, total, b);
    o = __shfl(o, 0);
    if (o == ~0ull) {
        if (lane == 0) {
            atomicOr(overflow, 2u);
            set_off[q] = 0;
            set_len[q] = 0;
        }
        return;
    }
    unsigned long long at = o + incl - cnt;
    for (uint32_t w = w0; w < w1; ++w) {
        uint32_t bits = bm[w];
        while (bits) {
            const int k = __ffs(bits) - 1;
            bits &= bits - 1;
            pool[at++] = w * 32 + k;
        }
    }
    if (lane == 0) {
        set_off[q] = total ? o : 0;
        set_len[q] = total;
    }
}

// ---- replay skip data ----------------------------------------------------------------------
// Edge words with their halves swapped: the dst-major words become src-major, so that a sort
// groups a plateau's out-edges (to lower plateaus).
__global__ void swap_halves(const unsigned long long* __restrict__ in, size_t E, unsigned long long* __restrict__ out) {
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < E) out[e] = (in[e] << 32) | (in[e] >> 32);
}

// below[q] = the smallest pore that reaches some plateau below q (q included) without reaching q,
// or kInv.  Sets only grow along an edge, so for a pore p of set(q) with below[q] > p, every plateau
// below q has the same pores smaller than p as q: all of them hold q's class of p.  Computed in
// ascending CR (a level's out-edges lead to levels already done): below[q] is the minimum over
// its out-edges d of below[d] and the first member of set(d) that set(q) lacks.  One wave per
// plateau, lane per out-edge.
__global__ void __launch_bounds__(256) below_min(const uint32_t* __restrict__ order, uint32_t lo, uint32_t hi,
                                                 const uint32_t* __restrict__ out_start,
                                                 const uint32_t* __restrict__ out_count,
                                                 const unsigned long long* __restrict__ oedges,
                                                 const unsigned long long* __restrict__ set_off,
                                                 const uint32_t* __restrict__ set_len, const uint32_t* __restrict__ pool,
                                                 SetBounds b, uint32_t* __restrict__ below, uint32_t* __restrict__ err) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t s = lo + blockIdx.x * 4 + wave;
    if (s >= hi) return;
    const uint32_t q = order[s];
    const bool okq = q < b.P && set_off[q] + set_len[q] <= b.pool_cap && out_start[q] + (unsigned long long)out_count[q] <= b.E;
    if (!okq) {
        if (lane == 0) fail_check(err, kBelowMin << 8 | 1);
        return;
    }
    const unsigned long long offq = set_off[q];
    const uint32_t lenq = set_len[q];
    const uint32_t e0 = out_start[q], ne = out_count[q];
    uint32_t m = kInv;
    bool bad = false;
    for (uint32_t e = e0 + lane; e < e0 + ne; e += 64) {
        const uint32_t d = (uint32_t)oedges[e];
        if (d >= b.P || set_off[d] + set_len[d] > b.pool_cap) {
            bad = true;
            break;
        }
        m = min(m, below[d]);
        const unsigned long long offd = set_off[d];
        const uint32_t lend = set_len[d];
        uint32_t i = 0;
        for (uint32_t t = 0; t < lend; ++t) {  // set(d) contains set(q): the first mismatch is the first extra
            const uint32_t v = pool[offd + t];
            if (i < lenq && pool[offq + i] == v) {
                ++i;
            } else {
                m = min(m, v);
                break;
            }
        }
    }
    if (__ballot(bad)) {
        if (lane == 0) fail_check(err, kBelowMin << 8 | 2);
        return;
    }
    for (int d = 32; d >= 1; d >>= 1) m = min(m, (uint32_t)__shfl_xor(m, d));
    if (lane == 0) below[q] = m;
}

// Per replayed pore, its (plateau << 32 | class) pairs: every plateau whose set holds the pore, with
// the pore's class there (its position among the key events it forms, kNoClass where it is the
// plateau's first member).  Counted, then written at per-pore offsets (any order).
__global__ void list_counts(const unsigned long long* __restrict__ set_off, const uint32_t* __restrict__ set_len,
                            const uint32_t* __restrict__ pool, SetBounds b, const uint32_t* __restrict__ replayed,
                            unsigned long long* __restrict__ cnt, uint32_t* __restrict__ err) {
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= b.P) return;
    const unsigned long long off = set_off[q];
    const uint32_t len = set_len[q];
    if (off + len > b.pool_cap) {
        fail_check(err, kPoreLists << 8 | 1);
        return;
    }
    for (uint32_t j = 0; j < len; ++j) {
        const uint32_t v = pool[off + j];
        if (v >= b.npores) {
            fail_check(err, kPoreLists << 8 | 2);
            return;
        }
        if (replayed[v]) atomicAdd(&cnt[v], 1ull);
    }
}

__global__ void list_fill(const unsigned long long* __restrict__ set_off, const uint32_t* __restrict__ set_len,
                          const uint32_t* __restrict__ pool, const unsigned long long* __restrict__ rbase,
                          const uint32_t* __restrict__ rec_node, const uint32_t* __restrict__ ev_local, SetBounds b,
                          unsigned long long R, uint32_t D, const uint32_t* __restrict__ replayed,
                          const unsigned long long* __restrict__ list_off, unsigned long long* __restrict__ cursor,
                          unsigned long long* __restrict__ lists, unsigned long long lists_len,
                          uint32_t* __restrict__ err) {
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= b.P) return;
    const unsigned long long off = set_off[q];
    const uint32_t len = set_len[q];
    if (off + len > b.pool_cap) {
        fail_check(err, kPoreLists << 8 | 3);
        return;
    }
    for (uint32_t j = 0; j < len; ++j) {
        const uint32_t v = pool[off + j];
        if (v >= b.npores) {
            fail_check(err, kPoreLists << 8 | 4);
            return;
        }
        if (!replayed[v]) continue;
        uint32_t l = kNoClassU;
        if (j >= 1) {
            const unsigned long long r = rbase[q] + j - 1;
            const uint32_t d = r < R ? rec_node[r] : kInv;
            l = d < D ? ev_local[d] : kInv;
            if (l >= kNoClassU) {
                fail_check(err, kPoreLists << 8 | 5);
                return;
            }
        }
        const unsigned long long pos = list_off[v] + atomicAdd(&cursor[v], 1ull);
        if (pos >= lists_len) {
            fail_check(err, kPoreLists << 8 | 6);
            return;
        }
        lists[pos] = ((unsigned long long)q << 32) | l;
    }
}

// ---- key records -------------------------------------------------------------------------
// A plateau whose set is s0 < s1 < ... < s(m-1) is first visited by pore s_j (j >= 1) while
// its earlier-pore class is {s0 .. s(j-1)}: that visit forms the list of {s0 .. s_j}.
__global__ void prefix_counts(const uint32_t* __restrict__ set_len, size_t P, unsigned long long* __restrict__ cnt) {
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < P) cnt[q] = set_len[q] >= 2 ? set_len[q] - 1 : 0;
}

__global__ void reach_sizes(const unsigned long long* __restrict__ set_off, const uint32_t* __restrict__ set_len,
                            const uint32_t* __restrict__ pool, const uint32_t* __restrict__ plat_size,
                            const uint32_t* __restrict__ plat_border, SetBounds b,
                            unsigned long long* __restrict__ reach, uint32_t* __restrict__ boundary,
                            uint32_t* __restrict__ err) {
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= b.P) return;
    if (set_off[q] + set_len[q] > b.pool_cap) {
        fail_check(err, kReachSizes << 8 | 1);
        return;
    }
    const uint32_t* s = pool + set_off[q];
    for (uint32_t t = 0; t < set_len[q]; ++t) {
        if (s[t] >= b.npores) {
            fail_check(err, kReachSizes << 8 | 2);
            return;
        }
        atomicAdd(&reach[s[t]], (unsigned long long)plat_size[q]);
        if (plat_border[q]) boundary[s[t]] = 1;
    }
}

// ---- per-pore sums with counters in LDS ---------------------------------------------------
// The same three walks over the plateaus' pore sets as list_counts, list_fill and reach_sizes, for
// up to kLdsPores pores: each workgroup adds into its own LDS counters (one slot per pore) and
// adds its non-zero totals to the global ones once at the end, instead of one global atomic per
// (plateau, pore) pair on a few thousand counters.  Workgroups take plateaus grid-stride.  A bad
// index fails the stage (the walk of that plateau stops; the workgroup still reaches its barriers).
constexpr uint32_t kLdsPores = 8064;  // u64 counters plus a flag bit per pore within 64 KB
constexpr uint32_t kLdsSumGrid = 512;  // workgroups of the LDS walks (two fit a CU)

__device__ __forceinline__ bool set_in_pool(const unsigned long long* set_off, const uint32_t* set_len, size_t q,
                                            const SetBounds& b) {
    return set_off[q] + set_len[q] <= b.pool_cap;
}

__global__ void __launch_bounds__(256) reach_sizes_lds(const unsigned long long* __restrict__ set_off,
                                                       const uint32_t* __restrict__ set_len,
                                                       const uint32_t* __restrict__ pool,
                                                       const uint32_t* __restrict__ plat_size,
                                                       const uint32_t* __restrict__ plat_border, SetBounds b,
                                                       unsigned long long* __restrict__ reach,
                                                       uint32_t* __restrict__ boundary, uint32_t* __restrict__ err) {
    extern __shared__ unsigned long long lsum[];  // npores sums, then (npores + 31) / 32 flag words
    const uint32_t np = (uint32_t)b.npores;
    uint32_t* lflag = reinterpret_cast<uint32_t*>(lsum + np);
    for (uint32_t i = threadIdx.x; i < np; i += blockDim.x) lsum[i] = 0;
    for (uint32_t i = threadIdx.x; i < (np + 31) / 32; i += blockDim.x) lflag[i] = 0;
    __syncthreads();
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < b.P; q += (size_t)gridDim.x * blockDim.x) {
        if (!set_in_pool(set_off, set_len, q, b)) {
            fail_check(err, kReachSizes << 8 | 1);
            continue;
        }
        const uint32_t* s = pool + set_off[q];
        const unsigned long long size = plat_size[q];
        const bool border = plat_border[q] != 0;
        for (uint32_t t = 0; t < set_len[q]; ++t) {
            const uint32_t v = s[t];
            if (v >= np) {
                fail_check(err, kReachSizes << 8 | 2);
                break;
            }
            atomicAdd(&lsum[v], size);
            if (border) atomicOr(&lflag[v >> 5], 1u << (v & 31));
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < np; i += blockDim.x) {
        if (lsum[i]) atomicAdd(&reach[i], lsum[i]);
        if ((lflag[i >> 5] >> (i & 31)) & 1u) boundary[i] = 1;
    }
}

__global__ void __launch_bounds__(256) list_counts_lds(const unsigned long long* __restrict__ set_off,
                                                       const uint32_t* __restrict__ set_len,
                                                       const uint32_t* __restrict__ pool, SetBounds b,
                                                       const uint32_t* __restrict__ replayed,
                                                       unsigned long long* __restrict__ cnt, uint32_t* __restrict__ err) {
    extern __shared__ unsigned long long lcount[];  // npores counts
    const uint32_t np = (uint32_t)b.npores;
    for (uint32_t i = threadIdx.x; i < np; i += blockDim.x) lcount[i] = 0;
    __syncthreads();
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < b.P; q += (size_t)gridDim.x * blockDim.x) {
        if (!set_in_pool(set_off, set_len, q, b)) {
            fail_check(err, kPoreLists << 8 | 1);
            continue;
        }
        const uint32_t* s = pool + set_off[q];
        for (uint32_t j = 0; j < set_len[q]; ++j) {
            const uint32_t v = s[j];
            if (v >= np) {
                fail_check(err, kPoreLists << 8 | 2);
                break;
            }
            if (replayed[v]) atomicAdd(&lcount[v], 1ull);
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < np; i += blockDim.x)
        if (lcount[i]) atomicAdd(&cnt[i], lcount[i]);
}

// The workgroup counts its entries per pore, reserves that many places in each pore's list with
// one global atomic, then walks its plateaus again and fills the places (the order of a pore's
// entries is free: the replay scatters them into a table by plateau).
__global__ void __launch_bounds__(256) list_fill_lds(const unsigned long long* __restrict__ set_off,
                                                     const uint32_t* __restrict__ set_len,
                                                     const uint32_t* __restrict__ pool,
                                                     const unsigned long long* __restrict__ rbase,
                                                     const uint32_t* __restrict__ rec_node,
                                                     const uint32_t* __restrict__ ev_local, SetBounds b,
                                                     unsigned long long R, uint32_t D,
                                                     const uint32_t* __restrict__ replayed,
                                                     const unsigned long long* __restrict__ list_off,
                                                     unsigned long long* __restrict__ cursor,
                                                     unsigned long long* __restrict__ lists,
                                                     unsigned long long lists_len, uint32_t* __restrict__ err) {
    extern __shared__ unsigned long long lpos[];  // per pore: the count, then the next place
    const uint32_t np = (uint32_t)b.npores;
    for (uint32_t i = threadIdx.x; i < np; i += blockDim.x) lpos[i] = 0;
    __syncthreads();
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < b.P; q += stride) {
        if (!set_in_pool(set_off, set_len, q, b)) {
            fail_check(err, kPoreLists << 8 | 3);
            continue;
        }
        const uint32_t* s = pool + set_off[q];
        for (uint32_t j = 0; j < set_len[q]; ++j) {
            const uint32_t v = s[j];
            if (v >= np) {
                fail_check(err, kPoreLists << 8 | 4);
                break;
            }
            if (replayed[v]) atomicAdd(&lpos[v], 1ull);
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < np; i += blockDim.x)
        if (lpos[i]) lpos[i] = list_off[i] + atomicAdd(&cursor[i], lpos[i]);
    __syncthreads();
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < b.P; q += stride) {
        if (!set_in_pool(set_off, set_len, q, b)) continue;  // flagged above
        const uint32_t* s = pool + set_off[q];
        for (uint32_t j = 0; j < set_len[q]; ++j) {
            const uint32_t v = s[j];
            if (v >= np) break;  // flagged above
            if (!replayed[v]) continue;
            uint32_t l = kNoClassU;
            if (j >= 1) {
                const unsigned long long r = rbase[q] + j - 1;
                const uint32_t d = r < R ? rec_node[r] : kInv;
                l = d < D ? ev_local[d] : kInv;
                if (l >= kNoClassU) {
                    fail_check(err, kPoreLists << 8 | 5);
                    break;
                }
            }
            const unsigned long long pos = atomicAdd(&lpos[v], 1ull);
            if (pos >= lists_len) {
                fail_check(err, kPoreLists << 8 | 6);
                break;
            }
            lists[pos] = ((unsigned long long)q << 32) | l;
        }
    }
}

// ---- DFS replay: a persistent pool of waves, one pore at a time per wave --------------------
// What the replay reproduces.  Plateaus (connected sets of equal CR) are strongly connected and
// edges between plateaus lead to lower CR, so the pore's DFS enters each plateau it reaches once,
// at one voxel, and the whole plateau is visited inside that voxel's subtree.  A step from a
// voxel into a plateau c not yet entered starts an excursion that visits exactly the plateaus
// below c that are not yet visited (a plateau visited earlier was closed together with
// everything below it), before the DFS goes on where it was.  Only the first entry into each
// class matters, and the classes met in such an excursion that are not ranked yet are exactly
// the unranked classes of the plateaus below c (a plateau of an unranked class has never been
// entered).  So when at most one unranked class lies below c, the excursion ranks that class
// and nothing else: it is skipped, its plateaus are marked entered and closed (their voxels are
// visited), and the walk goes on.  Two tests decide this:
//   * static: below[c] > p means every plateau below c holds c's class (O(1));
//   * dynamic: a wave-parallel traversal of the plateaus below c that are not entered, not
//     descending below static plateaus, stops at the second distinct unranked class.  A
//     traversal that runs out of room answers "two" (the step is taken: always exact).
// Plateaus below a static plateau are left unmarked; when the walk meets one it is static too.
//
// Classes are numbered per pore (0 .. ncls - 1, kNoClass where the pore is the plateau's first
// member).  A task's plateau classes come from its list of (plateau, class) pairs, scattered at
// the start of the task into the slot's dense table, so a lookup is one load; the bitmap of
// ranked classes lives in LDS.
constexpr uint16_t kNoClass = (uint16_t)kNoClassU;
constexpr uint32_t kRing = PMA_PORES_RING;  // DFS frames kept in LDS per wave (a power of two)
constexpr uint16_t kNoPage = 0xFFFF;
constexpr uint32_t kAllTests = (1u << 27) - 1u;  // neighbour indices 0..26 (13, the voxel itself, fails the stencil)

struct ReplayTask {
    uint32_t pore, seed, ncls, rank_base;  // ranks of the pore's classes at rank_slot[rank_base + class]
    unsigned long long list_off;           // the pore's (plateau << 32 | class) pairs
    uint32_t list_len, pad;
};

// Lengths of the buffers the replay selects from with values read from memory.
struct ReplayBounds {  // 32-bit where the host has checked the value fits: fewer live scalar registers
    uint32_t n, P, E, trav_cap;
    unsigned long long lists_len;
    uint32_t ranks_len;
};

// Lane `src`'s value as a wave-uniform scalar: the loop conditions built from broadcasts are then
// uniform to the compiler as well.
__device__ __forceinline__ uint32_t bcast(uint32_t v, int src) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, src);
}

// The replaying wave's LDS copy of recently met plateaus' state: entries (plateau << 32) | class << 16
// | closed << 1 | entered, one per slot, ~0 when empty.  The wave is the only writer of its
// entered and closed bitmaps and of its class table, so the copy is updated wherever they are.
#ifndef PMA_PORES_PCACHE
#define PMA_PORES_PCACHE 256u
#endif
constexpr uint32_t kPCache = PMA_PORES_PCACHE;  // a power of two
__device__ __forceinline__ uint32_t pcache_slot(uint32_t q) { return (q * 0x9E3779B1u) >> 24 & (kPCache - 1u); }
__device__ __forceinline__ unsigned long long pcache_entry(uint32_t q, uint32_t cls, bool closed, bool entered) {
    return ((unsigned long long)q << 32) | ((cls & 0xFFFFu) << 16) | (closed ? 2u : 0u) | (entered ? 1u : 0u);
}

// The replaying wave's LDS copy of recently read visited words, by halves: entries
// (voxel index / 32 << 32) | the 32 visited bits of those voxels, ~0 when empty.  Lane 0 sets a voxel's bit in the
// copy when it marks the voxel, so the copy is never behind the wave's own bitmap.
#ifndef PMA_PORES_VCACHE
#define PMA_PORES_VCACHE 512u
#endif
constexpr uint32_t kVCache = PMA_PORES_VCACHE;  // a power of two
__device__ __forceinline__ uint32_t vcache_slot(uint32_t h) { return (h * 0x9E3779B1u) >> 16 & (kVCache - 1u); }

__device__ __forceinline__ bool bit_of(const unsigned long long* words, uint32_t i) {
    return (__hip_atomic_load(&words[i >> 6], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) >> (i & 63)) & 1ull;
}

// A pool of waves (workgroup b uses slot b of the per-slot buffers: a visited bitmap over all
// voxels, bitmaps over plateaus of entered and closed plateaus, the dense class table, the DFS stack and the traversal stack and list) takes pores from a queue, largest reach
// first.  The code at the boundary of the loop over pores is executed by every lane (the fetch
// adds 1 in lane 0 and 0 elsewhere, the last lane's value is read back): with lane-0-only blocks
// there the compiler merged the end of one pore with the fetch of the next and let lanes 1..63
// start a pore while lane 0, which owns the stack, was still fetching.  Lanes 0..26 test the 26
// neighbours of the current voxel with independent loads (radius, plateau, visited word); lane 0
// owns the stack and the counters.  Frames are (voxel, next neighbour index); a recursive DFS
// that tries neighbours in reverse loop order visits voxels in the reference's stack-DFS order.
// The replay of a pore stops once all but one of its classes are ranked: the last one is
// necessarily found after them.  Words written by one lane and read by another go through
// workgroup-scope atomics: every per-slot buffer belongs to this one wave, so they need not be
// coherent beyond its CU (an agent-scope fence writes back and invalidates the XCD's L2, and an
// agent-scope load skips it -- both slow every wave on the XCD).  Dynamic LDS: the ranked-class bitmap (seen_words words), then the
// block-to-page table of the visited voxels.
__global__ void __launch_bounds__(64) dfs_replay(
    const ReplayTask* __restrict__ tasks, uint32_t T, uint32_t* __restrict__ queue, const uint32_t* __restrict__ cr,
    const uint32_t* __restrict__ plat, const unsigned long long* __restrict__ lists,
    const uint32_t* __restrict__ below, const uint32_t* __restrict__ out_start, const uint32_t* __restrict__ out_count,
    const unsigned long long* __restrict__ oedges, unsigned long long* vis_all, uint32_t vis_pages, uint32_t vis_bits,
    unsigned long long* pbits_all, unsigned long long pent_words, uint16_t* cls_all, uint32_t seen_words,
    uint4* __restrict__ stack_all, unsigned long long stack_cap,
    uint32_t* trav_all, uint32_t* __restrict__ rank_slot, uint32_t* __restrict__ task_status,
    unsigned long long* __restrict__ task_steps, int X, int Y, int Z, int d3q7, ReplayBounds b,
    uint32_t* __restrict__ err) {
    extern __shared__ uint32_t seen[];  // ranked classes of the current pore
    __shared__ uint32_t tcount[2];      // traversal: stack length, list length
    __shared__ uint4 ring[kRing];       // the newest DFS frames (voxel, plateau, radius, next index)
    __shared__ unsigned long long pcache[kPCache];  // recently met plateaus' state
    __shared__ unsigned long long vcache[kVCache];  // recently read visited half-words
    const int lane = threadIdx.x;
    // visited voxels: blocks of 2^vis_bits voxels, each on an 8 KB-per-65536-voxel page of this
    // slot's pool taken on the block's first visit (vtab in LDS: block -> page, kNoPage if none)
    const uint32_t page_words = 1u << (vis_bits - 6), nblocks = (uint32_t)((b.n + (1ull << vis_bits) - 1) >> vis_bits);
    unsigned long long* vpool = vis_all + (size_t)blockIdx.x * vis_pages * page_words;
    uint16_t* vtab = reinterpret_cast<uint16_t*>(seen + seen_words);
    unsigned long long* pent = pbits_all + (size_t)blockIdx.x * 2 * pent_words;
    unsigned long long* pclosed = pent + pent_words;
    uint16_t* cls = cls_all + (size_t)blockIdx.x * b.P;
    uint4* gstack = stack_all + (size_t)blockIdx.x * stack_cap;
    uint32_t* tstack = trav_all + (size_t)blockIdx.x * 2 * b.trav_cap;
    uint32_t* tlist = tstack + b.trav_cap;

    while (true) {
        // every lane adds (lane 0 adds 1): the returned values are the old count and old count + 1
        const uint32_t mine = atomicAdd(queue, lane == 0 ? 1u : 0u);
        const uint32_t t = bcast(mine, 0) == bcast(mine, 63) ? bcast(mine, 0) : bcast(mine, 63) - 1;
        if (t >= T) return;
        const ReplayTask task = tasks[t];
        const uint32_t p = task.pore;
        if (task.seed >= b.n || task.ncls > seen_words * 32 || task.list_off + task.list_len > b.lists_len ||
            (unsigned long long)task.rank_base + task.ncls > b.ranks_len) {
            if (lane == 0) fail_check(err, kDfsReplay << 8 | 1);
            return;
        }
        for (uint32_t i = lane; i < nblocks; i += 64) vtab[i] = kNoPage;
        uint32_t npages = 0;
        for (unsigned long long i = lane; i < 2 * pent_words; i += 64) pent[i] = 0;  // entered, closed
        for (uint32_t i = lane; i < kPCache; i += 64) pcache[i] = ~0ull;
        for (uint32_t i = lane; i < kVCache; i += 64) vcache[i] = ~0ull;
        for (uint32_t i = lane; i < seen_words; i += 64) seen[i] = 0;
        bool bad_list = false;
        for (uint32_t i = lane; i < task.list_len; i += 64) {
            const unsigned long long e = lists[task.list_off + i];
            const uint32_t q = (uint32_t)(e >> 32), l = (uint32_t)e;
            if (q >= b.P || (l != kNoClass && l >= task.ncls)) bad_list = true;
            else cls[q] = (uint16_t)l;
        }
        if (__ballot(bad_list)) {
            if (lane == 0) fail_check(err, kDfsReplay << 8 | 2);
            return;
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");  // the cleared words precede the atomics
        __builtin_amdgcn_wave_barrier();  // every lane has cleared its share before lane 0 marks the seed

        // Lane 0: ranks class l if it is new; returns the new count of ranked classes.
        auto rank_class = [&](uint32_t l, uint32_t found) -> uint32_t {
            if (l == kNoClass) return found;
            const uint32_t old = atomicOr(&seen[l >> 5], 1u << (l & 31));
            if (old & (1u << (l & 31))) return found;
            rank_slot[task.rank_base + l] = found;
            return found + 1;
        };
        // Wave-wide: the page of voxel v's block, taken and zeroed on first use; false when the
        // pool is full (the host repeats the pore with a larger pool).
        auto ensure_page = [&](uint32_t v) -> bool {
            const uint32_t blk = v >> vis_bits;
            if (vtab[blk] != kNoPage) return true;
            if (npages >= vis_pages) return false;
            unsigned long long* pg = vpool + (size_t)npages * page_words;
            for (uint32_t i = lane; i < page_words; i += 64) pg[i] = 0;
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");  // zeros land before the marks
            __builtin_amdgcn_wave_barrier();
            if (lane == 0) vtab[blk] = (uint16_t)npages;
            ++npages;
            __builtin_amdgcn_wave_barrier();
            return true;
        };
        uint32_t u = task.seed;
        uint32_t uq = plat[u];
        uint32_t cu = cr[u];
        if (uq >= b.P) {
            if (lane == 0) fail_check(err, kDfsReplay << 8 | 3);
            return;
        }
        if (!ensure_page(u)) {  // a pool of one page at least: never taken
            if (lane == 0) fail_check(err, kDfsReplay << 8 | 11);
            return;
        }
        const uint32_t seed_cls = cls[uq];
        uint32_t found = 0;
        if (lane == 0) {
            atomicOr(&vpool[(size_t)vtab[u >> vis_bits] * page_words + ((u & ((1u << vis_bits) - 1)) >> 6)],
                     1ull << (u & 63));
            atomicOr(&pent[uq >> 6], 1ull << (uq & 63));
            found = rank_class(seed_cls, 0);
        }
        found = bcast(found, 0);
        // The stack: frames [0, gtop) in global memory, [gtop, top) in the LDS ring at index % kRing.
        // `tests` holds the neighbour indices still to test at u: all 27 at a new voxel; at a resumed
        // frame (or after a skip) only those below the taken index that were candidates then -- a
        // candidate can only stop being one (visited and closed marks are never cleared).
        uint32_t top = 0, gtop = 0, steps = 0;  // at most the pore's reach, below 2^32
        uint32_t tests = kAllTests;
        bool overflow = false;
        while (found + 1 < task.ncls) {
            bool cand = false, entered = true;
            uint32_t j = 0, qn = 0, cn = 0, ql = kNoClass;
            uint32_t vbits = 0;  // the neighbour's visited half-word as it is now
            int x = 0, y = 0, z = 0;
            if (tests != 0) decode(u, X, Y, x, y, z);  // a frame with nothing left to test pops at once
            if (lane < 27 && ((tests >> lane) & 1u) && in_stencil(lane, d3q7) && neighbour(lane, x, y, z, X, Y, Z, j)) {
                // visited first: a neighbour the LDS copy shows as visited needs no load at all (the
                // copy is keyed by the voxel's half-word index, so a hit needs no page lookup)
                const uint32_t vh = j >> 5;
                const uint32_t slot = vcache_slot(vh);
                const unsigned long long e = vcache[slot];
                const bool hit = (uint32_t)(e >> 32) == vh;
                bool visited = hit && (((uint32_t)e >> (j & 31)) & 1u);
                vbits = (uint32_t)e;
                if (!visited) {
                    cn = cr[j];
                    qn = plat[j];
                    if (!hit) {
                        const uint32_t pg = vtab[j >> vis_bits];
                        uint32_t bits = 0;  // a block without a page has nothing visited
                        if (pg != kNoPage) {
                            const uint32_t off = j & ((1u << vis_bits) - 1);
                            const unsigned long long w = __hip_atomic_load(
                                &vpool[(size_t)pg * page_words + (off >> 6)], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                            bits = (uint32_t)(w >> ((off >> 5) & 1u) * 32u);
                        }
                        vcache[slot] = ((unsigned long long)vh << 32) | bits;
                        visited = (bits >> (j & 31)) & 1u;
                        vbits = bits;
                    }
                }
                cand = cn != 0 && cn <= cu && !visited;
                // neighbours in other plateaus: closed plateaus are visited; the entered test and the
                // class are for the lane whose neighbour is taken
                if (cand && qn != uq && qn < b.P) {
                    const uint32_t slot = pcache_slot(qn);
                    const unsigned long long e = pcache[slot];
                    bool closed;
                    if ((uint32_t)(e >> 32) == qn) {
                        entered = e & 1ull;
                        closed = (e >> 1) & 1ull;
                        ql = (uint32_t)(e >> 16) & 0xFFFFu;
                    } else {
                        entered = bit_of(pent, qn);
                        closed = bit_of(pclosed, qn);
                        ql = cls[qn];
                        pcache[slot] = pcache_entry(qn, ql, closed, entered);
                    }
                    if (closed) cand = false;
                }
            }
            const unsigned long long bal = __ballot(cand);
            if (!bal) {
                if (top == 0) break;  // stack empty: the pore's reach is exhausted
                if (top - 1 < gtop) {
                    // the ring is empty: bring back the newest half ring of frames from global memory
                    const uint32_t h = min(gtop, kRing / 2);
                    for (unsigned long long i = lane; i < h; i += 64) {
                        const unsigned long long at = gtop - h + i;
                        ring[at % kRing] = gstack[at];
                    }
                    gtop -= h;
                    __builtin_amdgcn_wave_barrier();
                }
                --top;  // pop: every lane reads the frame from LDS
                const uint4 f = ring[top % kRing];
                u = f.x;
                uq = f.y;
                cu = f.z;
                tests = f.w;
                if (u >= b.n || uq >= b.P || (tests & ~kAllTests)) {
                    if (lane == 0) fail_check(err, kDfsReplay << 8 | 5);
                    return;
                }
                continue;
            }
            const int kk = 63 - __clzll((long long)bal);  // the highest candidate index
            const uint32_t nvox = bcast(j, kk);
            const uint32_t nq = bcast(qn, kk);
            const uint32_t ncr = bcast(cn, kk);
            const bool is_new = !bcast(entered ? 1u : 0u, kk);  // first entry into plateau nq
            if (nq >= b.P) {
                if (lane == 0) fail_check(err, kDfsReplay << 8 | 4);
                return;
            }
            if (is_new) {
                // an excursion into nq: skipped when it holds at most one unranked class
                bool skip = false;
                uint32_t skip_cls = kNoClass;
                if (below[nq] > p) {
                    skip = true;
                    skip_cls = bcast(ql, kk);
                    if (lane == 0) {
                        atomicOr(&pent[nq >> 6], 1ull << (nq & 63));
                        atomicOr(&pclosed[nq >> 6], 1ull << (nq & 63));
                        pcache[pcache_slot(nq)] = pcache_entry(nq, skip_cls, true, true);
                    }
                    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");  // marks land before the next loads
                } else {
                    // traversal: stack tstack[0 .. tcount[0]), every plateau it marks listed in tlist.
                    // It marks plateaus in the entered bitmap itself (the walk waits meanwhile), so one
                    // atomic both tests "entered or already met" and marks; a skip keeps the marks, a
                    // step takes them back.
                    if (lane == 0) {
                        atomicOr(&pent[nq >> 6], 1ull << (nq & 63));
                        __hip_atomic_store(&tstack[0], nq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                        __hip_atomic_store(&tlist[0], nq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                        tcount[0] = 1;
                        tcount[1] = 1;
                    }
                    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
                    __builtin_amdgcn_wave_barrier();
                    uint32_t nun = 0, k1 = kNoClass;
                    bool room = true;
                    for (uint32_t ts = 1; ts > 0;) {
                        // pop up to 64 plateaus, one per lane
                        const uint32_t nb = min(ts, 64u);
                        uint32_t bq = kInv, l = kNoClass, cnt = 0, st = 0;
                        bool bad = false, expand = false;
                        if ((uint32_t)lane < nb) {
                            bq = __hip_atomic_load(&tstack[ts - 1 - lane], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                            if (bq >= b.P) {
                                bad = true;
                            } else {
                                l = cls[bq];
                                expand = below[bq] <= p;  // below a static plateau all share its class
                                bad = l != kNoClass && l >= task.ncls;
                                st = out_start[bq];
                                cnt = out_count[bq];
                                if ((unsigned long long)st + cnt > b.E) bad = true;
                            }
                        }
                        if (__ballot(bad)) {
                            if (lane == 0) fail_check(err, kDfsReplay << 8 | 8);
                            return;
                        }
                        const bool unr = l != kNoClass && !((seen[l >> 5] >> (l & 31)) & 1u);
                        const unsigned long long ub = __ballot(unr);
                        if (ub) {
                            if (k1 == kNoClass) k1 = bcast(l, __ffsll((long long)ub) - 1);
                            nun = __ballot(unr && l != k1) ? 2u : 1u;
                        }
                        if (nun >= 2) break;
                        if (!expand) cnt = 0;
                        // out-edges of the popped plateaus, four per lane per round: inclusive prefix of counts
                        uint32_t incl = cnt;
                        for (int d = 1; d < 64; d <<= 1) {
                            const uint32_t o = __shfl_up(incl, d);
                            if (lane >= d) incl += o;
                        }
                        const uint32_t total = bcast(incl, 63);
                        const uint32_t first = incl - cnt;
                        if (lane == 0) tcount[0] = ts - nb;  // the popped entries are in registers
                        __builtin_amdgcn_wave_barrier();
                        bool lane_room = true;
                        for (uint32_t base = 0; base < total; base += 256) {
                            uint32_t dd[4];
                            bool have[4];
                            unsigned long long old[4];
#pragma unroll
                            for (int u = 0; u < 4; ++u) {
                                const uint32_t i = base + 64 * u + lane;
                                // owner lane of edge i: the number of lanes whose inclusive count is <= i
                                int o = 0;
                                for (int s = 32; s >= 1; s >>= 1)
                                    if (__shfl(incl, o + s - 1) <= i) o += s;
                                const uint32_t o_start = __shfl(st, o), o_first = __shfl(first, o);
                                have[u] = i < total;
                                dd[u] = have[u] ? (uint32_t)oedges[o_start + (i - o_first)] : kInv;
                            }
#pragma unroll
                            for (int u = 0; u < 4; ++u) {
                                old[u] = ~0ull;
                                if (have[u] && dd[u] < b.P) old[u] = atomicOr(&pent[dd[u] >> 6], 1ull << (dd[u] & 63));
                                else if (have[u]) bad = true;
                            }
#pragma unroll
                            for (int u = 0; u < 4; ++u) {
                                const uint32_t d = dd[u];
                                if (!have[u] || d >= b.P || ((old[u] >> (d & 63)) & 1ull)) continue;  // entered or met
                                const uint32_t w = atomicAdd(&tcount[1], 1u);
                                if (w >= b.trav_cap) {  // not listed: take the mark back
                                    atomicAnd(&pent[d >> 6], ~(1ull << (d & 63)));
                                    lane_room = false;
                                    continue;
                                }
                                __hip_atomic_store(&tlist[w], d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                                const uint32_t at = atomicAdd(&tcount[0], 1u);  // at <= w: every entry is listed
                                __hip_atomic_store(&tstack[at], d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                            }
                        }
                        if (__ballot(bad)) {
                            if (lane == 0) fail_check(err, kDfsReplay << 8 | 10);
                            return;
                        }
                        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
                        __builtin_amdgcn_wave_barrier();
                        if (__ballot(!lane_room)) {
                            room = false;
                            break;
                        }
                        ts = bcast(tcount[0], 0);
                    }
                    __builtin_amdgcn_wave_barrier();
                    const uint32_t wn = min(bcast(tcount[1], 0), (uint32_t)b.trav_cap);
                    skip = room && nun <= 1;
                    for (uint32_t i = lane; i < wn; i += 64) {
                        const uint32_t d = __hip_atomic_load(&tlist[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                        if (d >= b.P) continue;
                        const unsigned long long m1 = 1ull << (d & 63);
                        if (skip) {
                            atomicOr(&pclosed[d >> 6], m1);
                            // a listed plateau was not entered before the walk: a copy of its state
                            // says so and now becomes entered and closed (one lane per plateau)
                            const uint32_t slot = pcache_slot(d);
                            const unsigned long long e = pcache[slot];
                            if ((uint32_t)(e >> 32) == d) pcache[slot] = e | 3ull;
                        } else {
                            atomicAnd(&pent[d >> 6], ~m1);
                        }
                    }
                    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
                    __builtin_amdgcn_wave_barrier();
                    skip_cls = nun == 1 ? k1 : kNoClass;
                }
                if (skip) {
                    if (lane == 0) found = rank_class(skip_cls, found);
                    found = bcast(found, 0);
                    tests = (uint32_t)bal & ((1u << kk) - 1u);  // go on with the candidates below kk
                    continue;
                }
            }
            if (top - gtop == kRing) {
                // the ring is full: move its oldest half to global memory
                if (gtop + kRing / 2 > stack_cap) {  // no room: the host repeats the pore with a larger stack
                    overflow = true;
                    break;
                }
                for (unsigned long long i = lane; i < kRing / 2; i += 64) gstack[gtop + i] = ring[(gtop + i) % kRing];
                gtop += kRing / 2;
                __builtin_amdgcn_wave_barrier();
            }
            // The lane that tested nvox holds its visited half-word as it is now (no visited bit
            // changes between the test and here); a block without a page yet takes one, all zero.
            uint32_t page = vtab[nvox >> vis_bits], nbits = bcast(vbits, kk);
            if (page == kNoPage) {
                const uint32_t taken = npages;
                if (!ensure_page(nvox)) {  // no page left: the host repeats the pore with a larger pool
                    overflow = true;
                    break;
                }
                if (npages != taken) {  // a new page, all zero
                    page = taken;
                    nbits = 0;
                } else {  // the block had a page: read the word
                    page = vtab[nvox >> vis_bits];
                    const uint32_t off = nvox & ((1u << vis_bits) - 1);
                    const unsigned long long w = __hip_atomic_load(&vpool[(size_t)page * page_words + (off >> 6)],
                                                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    nbits = (uint32_t)(w >> ((off >> 5) & 1u) * 32u);
                }
            }
            const uint32_t ncls_new = bcast(ql, kk);
            if (lane == 0) {
                ring[top % kRing] = make_uint4(u, uq, cu, (uint32_t)bal & ((1u << kk) - 1u));  // resumes below kk
                const uint32_t off = nvox & ((1u << vis_bits) - 1);
                atomicOr(&vpool[(size_t)page * page_words + (off >> 6)], 1ull << (nvox & 63));
                vcache[vcache_slot(nvox >> 5)] = ((unsigned long long)(nvox >> 5) << 32) | nbits | (1u << (nvox & 31));
                if (is_new) {
                    atomicOr(&pent[nq >> 6], 1ull << (nq & 63));
                    pcache[pcache_slot(nq)] = pcache_entry(nq, ncls_new, false, true);
                    found = rank_class(ncls_new, found);
                }
            }
            if (is_new) found = bcast(found, 0);
            ++top;
            ++steps;
            u = nvox;
            uq = nq;
            cu = ncr;
            tests = kAllTests;
        }
        // every lane stores the same values
        task_steps[t] = steps;
        task_status[t] = overflow ? 1u : 0u;
        if (!overflow && found + 1 < task.ncls) fail_check(err, kDfsReplay << 8 | 6);  // reach exhausted too early
    }
}

// ---- labels ------------------------------------------------------------------------------------
struct LabelBounds {
    unsigned long long P, pool_cap, R, D, npores;
};

__global__ void write_labels(const uint32_t* __restrict__ plat, const unsigned long long* __restrict__ set_off,
                             const uint32_t* __restrict__ set_len, const uint32_t* __restrict__ pool,
                             const unsigned long long* __restrict__ rbase, const uint32_t* __restrict__ rec_node,
                             const uint32_t* __restrict__ key_id, size_t n, LabelBounds b, uint32_t* __restrict__ pt,
                             uint32_t* __restrict__ err) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t q = plat[i];
    pt[i] = 0;
    if (q == kInv) return;
    if (q >= b.P) {
        fail_check(err, kWriteLabels << 8 | 1);
        return;
    }
    const uint32_t m = set_len[q];
    if (m == 1) {
        if (set_off[q] >= b.pool_cap || pool[set_off[q]] >= b.npores) fail_check(err, kWriteLabels << 8 | 2);
        else pt[i] = pore_id_of(pool[set_off[q]]);
    } else if (m >= 2) {
        const unsigned long long r = rbase[q] + m - 2;
        const uint32_t d = r < b.R ? rec_node[r] : kInv;
        if (d >= b.D) fail_check(err, kWriteLabels << 8 | 3);
        else pt[i] = key_id[d];
    }
}

__global__ void single_volumes(const unsigned long long* __restrict__ set_off, const uint32_t* __restrict__ set_len,
                               const uint32_t* __restrict__ pool, const uint32_t* __restrict__ plat_size, SetBounds b,
                               unsigned long long* __restrict__ vol, uint32_t* __restrict__ err) {
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= b.P || set_len[q] != 1) return;
    const uint32_t k = set_off[q] < b.pool_cap ? pool[set_off[q]] : kInv;
    if (k >= b.npores) fail_check(err, kSingleVolumes << 8 | 1);
    else atomicAdd(&vol[k], (unsigned long long)plat_size[q]);
}

__global__ void seed_plateaus(const uint32_t* __restrict__ plat, const uint32_t* __restrict__ seed_voxel, uint32_t n,
                              size_t nvox, uint32_t nplat, uint32_t* __restrict__ out, uint32_t* __restrict__ plat_pore,
                              uint32_t* __restrict__ err) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t v = seed_voxel[k];
    const uint32_t q = v < nvox ? plat[v] : kInv;
    if (q >= nplat) {
        fail_check(err, kSeedPlateaus << 8 | 1);
        return;
    }
    out[k] = q;
    plat_pore[q] = k;
}

__global__ void fill_u32(uint32_t* p, uint32_t v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Raises PMA_ERR_INTERNAL naming the kernel whose check failed, if any did.
void check_stage(pma_ctx* c, const uint32_t* err) {
    const uint32_t v = read_scalar(c, err);
    if (!v) return;
    const uint32_t kid = v >> 8;
    const char* name = kid < sizeof kKernelNames / sizeof kKernelNames[0] ? kKernelNames[kid] : "?";
    throw Error(PMA_ERR_INTERNAL, std::string("pma_pores: in-kernel check ") + std::to_string(v & 0xFF) + " of " +
                                      name + " failed (inconsistent intermediate state)");
}

}  // namespace

// With PMA_DEBUG_SYNC, prints the time of each phase of the stage (after waiting for the stream).
struct PhaseClock {
    pma_ctx* c;
    std::chrono::steady_clock::time_point last = std::chrono::steady_clock::now();
    void mark(const char* what) {
        if (!c->debug_sync) return;
        c->sync();
        const auto now = std::chrono::steady_clock::now();
        std::fprintf(stderr, "[pma] pores phase %-34s %9.1f ms\n", what,
                     std::chrono::duration<double, std::milli>(now - last).count());
        last = now;
    }
};

void run_pores(pma_ctx* c) {
    PhaseClock phase{c};
    PMA_REQUIRE(c->valid[PMA_VOL_MASK] && c->valid[PMA_VOL_EDT2] && c->valid[PMA_VOL_CR2], PMA_ERR_STATE,
                "pma_pores: the maximum-ball mask, EDT and containing radii are needed");
    const int X = c->dims.x, Y = c->dims.y, Z = c->dims.z;
    const size_t n = c->n;
    PMA_REQUIRE(n < 0xFFFFFFFFull, PMA_ERR_INVALID_ARGUMENT, "pma_pores: volume too large");
    const int d3q7 = c->stencil == PMA_STENCIL_D3Q7;
    const uint8_t* mask = c->v<uint8_t>(PMA_VOL_MASK);
    const uint32_t* e = c->v<uint32_t>(PMA_VOL_EDT2);
    const uint32_t* cr = c->v<uint32_t>(PMA_VOL_CR2);
    c->ensure_volume(PMA_VOL_PT);
    uint32_t* pt = c->v<uint32_t>(PMA_VOL_PT);
    const unsigned nb = grid_for(n, 256);
    uint32_t* err = c->buf<uint32_t>("pt_err", 1);
    uint32_t* of = c->buf<uint32_t>("pt_overflow", 1);
    PMA_HIP(hipMemsetAsync(err, 0, 4, c->stream));

    // 1. maximum balls in z, x, y order (memory order), then the libstdc++ permutation.
    const size_t cblocks = (n + 256 * kChunk - 1) / (256 * kChunk);
    uint32_t* bcount = c->buf<uint32_t>("pt_bcount", cblocks);
    uint32_t* boff = c->buf<uint32_t>("pt_boff", cblocks);
    count_centres<<<(unsigned)cblocks, 256, 0, c->stream>>>(mask, n, bcount);
    PMA_CHECKED(c, "count_centres", err);
    const uint64_t M = exclusive_scan_u32(c, bcount, boff, cblocks);
    unsigned long long* rec = c->buf<unsigned long long>("pt_rec", std::max<uint64_t>(M, 1));
    write_centres<<<(unsigned)cblocks, 256, 0, c->stream>>>(mask, e, n, boff, rec, M, err);
    PMA_CHECKED(c, "write_centres", err);
    if (c->testing_ball_order.empty()) {
        sort_like_libstdcxx_desc(c, reinterpret_cast<uint64_t*>(rec), M);
    } else {
        std::vector<uint32_t> order;
        order.swap(c->testing_ball_order);
        PMA_REQUIRE(order.size() == M, PMA_ERR_INVALID_ARGUMENT,
                    "pma_pores: the ball order does not list every maximum ball once");
        uint32_t* d_order = c->buf<uint32_t>("pt_testing_order", std::max<uint64_t>(M, 1));
        upload(c, d_order, order.data(), M * 4ull);
        if (M) {
            records_in_order<<<grid_for(M, 256), 256, 0, c->stream>>>(d_order, M, mask, e, n, rec, err);
            PMA_CHECKED(c, "records_in_order", err);
        }
    }

    phase.mark("1 maximum balls and their order");
    // 2. plateaus of CR and regional maxima.
    uint32_t* root = c->buf<uint32_t>("pt_root", n);
    uint8_t* not_max = c->buf<uint8_t>("pt_not_max", n);
    uint32_t* seedpos = c->buf<uint32_t>("pt_seedpos", n);
    cc_init<<<nb, 256, 0, c->stream>>>(cr, n, root);
    PMA_CHECKED(c, "cc_init", err);
    {
        PMA_TIMED(c, "cc_union");
        cc_union<<<nb, 256, 0, c->stream>>>(cr, root, X, Y, Z, d3q7, err);
        PMA_CHECKED(c, "cc_union", err);
    }
    cc_flatten<<<nb, 256, 0, c->stream>>>(root, n, err);
    PMA_CHECKED(c, "cc_flatten", err);
    PMA_HIP(hipMemsetAsync(not_max, 0, n, c->stream));
    PMA_HIP(hipMemsetAsync(seedpos, 0xFF, n * sizeof(uint32_t), c->stream));
    {
        PMA_TIMED(c, "cc_not_max");
        cc_not_max<<<nb, 256, 0, c->stream>>>(cr, root, not_max, X, Y, Z, d3q7, err);
        PMA_CHECKED(c, "cc_not_max", err);
    }

    phase.mark("2 plateaus");
    // 3. seeds: the first sorted maximum ball of every regional-maximum plateau.
    uint32_t npores = 0;
    unsigned long long* seeds = c->buf<unsigned long long>("pt_seeds", std::max<uint64_t>(M, 1));
    if (M) {
        first_in_sorted<<<grid_for(M, 256), 256, 0, c->stream>>>(rec, M, root, n, not_max, seedpos, err);
        PMA_CHECKED(c, "first_in_sorted", err);
        uint32_t* sflag = c->buf<uint32_t>("pt_sflag", M);
        uint32_t* soff = c->buf<uint32_t>("pt_soff", M);
        seed_flags<<<grid_for(M, 256), 256, 0, c->stream>>>(rec, M, root, n, seedpos, sflag, err);
        PMA_CHECKED(c, "seed_flags", err);
        npores = (uint32_t)exclusive_scan_u32(c, sflag, soff, M);
        write_seeds<<<grid_for(M, 256), 256, 0, c->stream>>>(rec, M, sflag, soff, seeds, err);
        PMA_CHECKED(c, "write_seeds", err);
    }
    check_stage(c, err);
    PMA_REQUIRE(npores <= M, PMA_ERR_INTERNAL, "pma_pores: more seeds than maximum balls");
    std::vector<unsigned long long> seeds_h(npores);
    if (npores) {
        PMA_HIP(hipMemcpyAsync(seeds_h.data(), seeds, npores * 8ull, hipMemcpyDeviceToHost, c->stream));
        PMA_HIP(hipStreamSynchronize(c->stream));
    }
    for (uint32_t k = 0; k < npores; ++k)
        PMA_REQUIRE((uint32_t)seeds_h[k] < n, PMA_ERR_INTERNAL, "pma_pores: a seed lies outside the volume");

    phase.mark("3 seeds");
    // 4. plateau indices, sizes, border flags.
    uint32_t* rflag = c->buf<uint32_t>("pt_rflag", n);
    uint32_t* roff = c->buf<uint32_t>("pt_roff", n);
    root_flags<<<nb, 256, 0, c->stream>>>(root, n, rflag);
    PMA_CHECKED(c, "root_flags", err);
    const uint64_t P = exclusive_scan_u32(c, rflag, roff, n);
    PMA_REQUIRE(P <= n, PMA_ERR_INTERNAL, "pma_pores: more plateaus than voxels");
    uint32_t* pidx = seedpos;  // seed positions are no longer needed
    const size_t P1 = std::max<uint64_t>(P, 1);
    uint32_t* plat_cr = c->buf<uint32_t>("pt_plat_cr", P1);
    uint32_t* plat_size = c->buf<uint32_t>("pt_plat_size", P1);
    uint32_t* plat_border = c->buf<uint32_t>("pt_plat_border", P1);
    uint32_t* plat_pore = c->buf<uint32_t>("pt_plat_pore", P1);
    plateau_index<<<nb, 256, 0, c->stream>>>(rflag, roff, n, pidx, cr, plat_cr, P, err);
    PMA_CHECKED(c, "plateau_index", err);
    PMA_HIP(hipMemsetAsync(plat_size, 0, P1 * 4, c->stream));
    PMA_HIP(hipMemsetAsync(plat_border, 0, P1 * 4, c->stream));
    fill_u32<<<grid_for(P1, 256), 256, 0, c->stream>>>(plat_pore, kInv, P1);
    PMA_CHECKED(c, "fill_u32", err);
    uint32_t* vrank = c->buf<uint32_t>("pt_vrank", n);
    plateau_of_voxels<<<nb, 256, 0, c->stream>>>(root, pidx, X, Y, Z, P, plat_size, plat_border, vrank, err);
    PMA_CHECKED(c, "plateau_of_voxels", err);
    uint32_t* plat = root;  // now the plateau index of every void voxel
    if (npores) {
        // plat_pore[plat[seed_k]] = k
        std::vector<uint32_t> sv(npores);
        for (uint32_t k = 0; k < npores; ++k) sv[k] = (uint32_t)seeds_h[k];
        uint32_t* d_sv = c->buf<uint32_t>("pt_seed_voxels", npores);
        upload(c, d_sv, sv.data(), npores * 4ull);
        uint32_t* d_sp = c->buf<uint32_t>("pt_seed_plat", npores);
        seed_plateaus<<<grid_for(npores, 256), 256, 0, c->stream>>>(plat, d_sv, npores, n, (uint32_t)P, d_sp,
                                                                    plat_pore, err);
        PMA_CHECKED(c, "seed_plateaus", err);
    }
    check_stage(c, err);

    phase.mark("4 plateau indices");
    // 5. plateau DAG edges (distinct), grouped by target.
    size_t cap = 1024;
    while (cap < P * 16) cap <<= 1;
    unsigned long long* table = nullptr;
    for (int attempt = 0;; ++attempt) {
        table = c->buf<unsigned long long>("pt_edge_table", cap);
        unsigned long long* fill = c->buf<unsigned long long>("pt_edge_fill", 1);
        PMA_HIP(hipMemsetAsync(table, 0xFF, cap * 8, c->stream));
        PMA_HIP(hipMemsetAsync(of, 0, 4, c->stream));
        PMA_HIP(hipMemsetAsync(fill, 0, 8, c->stream));
        {
            PMA_TIMED(c, "plateau_edges");
            plateau_edges<<<nb, 256, 0, c->stream>>>(cr, plat, X, Y, Z, d3q7, P, table, cap - 1, fill, of, err);
            PMA_CHECKED(c, "plateau_edges", err);
        }
        check_stage(c, err);
        if (read_scalar(c, of) == 0) break;
        PMA_REQUIRE(attempt < 8, PMA_ERR_INTERNAL, "pma_pores: plateau edge table overflow");
        cap <<= 2;
    }
    uint32_t* tflag = c->buf<uint32_t>("pt_tflag", cap);
    uint32_t* toff = c->buf<uint32_t>("pt_toff", cap);
    slot_flags<<<grid_for(cap, 256), 256, 0, c->stream>>>(table, cap, tflag);
    PMA_CHECKED(c, "slot_flags", err);
    const uint64_t E = exclusive_scan_u32(c, tflag, toff, cap);
    PMA_REQUIRE(E <= cap && E < 0xFFFFFFFFull, PMA_ERR_INTERNAL, "pma_pores: edge count out of range");
    const size_t E1 = std::max<uint64_t>(E, 1);
    unsigned long long* edges = c->buf<unsigned long long>("pt_edges", E1);
    unsigned long long* edges2 = c->buf<unsigned long long>("pt_edges2", E1);
    write_edges<<<grid_for(cap, 256), 256, 0, c->stream>>>(table, cap, tflag, toff, E, edges, err);
    PMA_CHECKED(c, "write_edges", err);
    uint32_t* dummy = c->buf<uint32_t>("pt_dummy", E1);
    uint32_t* dummy2 = c->buf<uint32_t>("pt_dummy2", E1);
    radix_sort_pairs_u64(c, reinterpret_cast<uint64_t*>(edges), reinterpret_cast<uint64_t*>(edges2), dummy, dummy2, E);
    if (E) std::swap(edges, edges2);
    uint32_t* in_start = c->buf<uint32_t>("pt_in_start", P1);
    uint32_t* in_count = c->buf<uint32_t>("pt_in_count", P1);
    PMA_HIP(hipMemsetAsync(in_start, 0, P1 * 4, c->stream));
    PMA_HIP(hipMemsetAsync(in_count, 0, P1 * 4, c->stream));
    if (E) {
        edge_ranges<<<grid_for(E, 256), 256, 0, c->stream>>>(edges, E, P, in_start, in_count, err);
        PMA_CHECKED(c, "edge_ranges", err);
    }

    phase.mark("5 plateau edges");
    // 6. pore sets, level by level in descending CR.
    unsigned long long* lkeys = c->buf<unsigned long long>("pt_lkeys", P1);
    unsigned long long* lkeys2 = c->buf<unsigned long long>("pt_lkeys2", P1);
    uint32_t* lvals = c->buf<uint32_t>("pt_lvals", P1);
    uint32_t* order = c->buf<uint32_t>("pt_order", P1);
    unsigned long long* set_off = c->buf<unsigned long long>("pt_set_off", P1);
    uint32_t* set_len = c->buf<uint32_t>("pt_set_len", P1);
    unsigned long long* cur = c->buf<unsigned long long>("pt_merge_cur", E1);
    unsigned long long* cend = c->buf<unsigned long long>("pt_merge_end", E1);
    PMA_HIP(hipMemsetAsync(set_off, 0, P1 * 8, c->stream));
    PMA_HIP(hipMemsetAsync(set_len, 0, P1 * 4, c->stream));
    std::vector<uint32_t> lstart_h;
    if (P) {
        level_keys<<<grid_for(P, 256), 256, 0, c->stream>>>(plat_cr, P, lkeys, lvals);
        PMA_CHECKED(c, "level_keys", err);
        radix_sort_pairs_u64(c, reinterpret_cast<uint64_t*>(lkeys), reinterpret_cast<uint64_t*>(lkeys2), lvals, order, P);
        uint32_t* lhead = c->buf<uint32_t>("pt_lhead", P);
        uint32_t* loff = c->buf<uint32_t>("pt_loff", P);
        level_heads<<<grid_for(P, 256), 256, 0, c->stream>>>(lkeys2, P, lhead);
        PMA_CHECKED(c, "level_heads", err);
        const uint64_t L = exclusive_scan_u32(c, lhead, loff, P);
        PMA_REQUIRE(L >= 1 && L <= P, PMA_ERR_INTERNAL, "pma_pores: level count out of range");
        uint32_t* lstart = c->buf<uint32_t>("pt_lstart", L);
        level_starts<<<grid_for(P, 256), 256, 0, c->stream>>>(lhead, loff, P, lstart, L, err);
        PMA_CHECKED(c, "level_starts", err);
        check_stage(c, err);
        lstart_h.resize(L + 1);
        PMA_HIP(hipMemcpyAsync(lstart_h.data(), lstart, L * 4ull, hipMemcpyDeviceToHost, c->stream));
        PMA_HIP(hipStreamSynchronize(c->stream));
        lstart_h[L] = (uint32_t)P;
        for (size_t l = 0; l < L; ++l)
            PMA_REQUIRE(lstart_h[l] < lstart_h[l + 1], PMA_ERR_INTERNAL, "pma_pores: level starts out of order");
    }
    // Room for the sets: a run that overflows is repeated in full, so start with what the free
    // device memory allows (up to PMA_PORES_POOL_MAX members per plateau), at least
    // PMA_PORES_POOL_MIN per plateau.
    unsigned long long pool_cap = std::max<unsigned long long>(PMA_PORES_POOL_MIN * P, PMA_PORES_POOL_FLOOR);
    {
        size_t free_b = 0, total_b = 0;
        PMA_HIP(hipMemGetInfo(&free_b, &total_b));
        const unsigned long long room = free_b / 8 / sizeof(uint32_t);
        pool_cap = std::max<unsigned long long>(pool_cap, std::min<unsigned long long>(PMA_PORES_POOL_MAX * P, room));
    }
    uint32_t* pool = nullptr;
    constexpr unsigned long long kSubPools = 64;
    unsigned long long* pool_top = c->buf<unsigned long long>("pt_pool_top", kSubPools);
    for (int attempt = 0;; ++attempt) {
        pool = c->buf<uint32_t>("pt_pool", pool_cap);
        PMA_HIP(hipMemsetAsync(pool_top, 0, kSubPools * 8, c->stream));
        PMA_HIP(hipMemsetAsync(of, 0, 4, c->stream));
        // sub-pools of at least 4096 members
        const unsigned long long nsub = std::max<unsigned long long>(1, std::min(kSubPools, pool_cap / 4096));
        const SetBounds sb{E, P, npores, pool_cap, nsub};
        // one wave per plateau over an LDS bitmap of the pores while the bitmaps of a workgroup's
        // waves fit in PMA_UNION_LDS_MAX bytes; the per-thread merge otherwise
        const uint32_t words = (npores + 31) / 32;
        const size_t union_lds = (size_t)kUnionWaves * std::max(words, 1u) * 4;
        const bool bitmap = union_lds <= PMA_UNION_LDS_MAX;
        for (size_t l = 0; l + 1 < lstart_h.size(); ++l) {
            uint32_t lo = lstart_h[l], hi = lstart_h[l + 1];
            PMA_TIMED(c, "set_union_level");
            if (bitmap) {
                set_union_bitmap<<<grid_for(hi - lo, kUnionWaves), 64 * kUnionWaves, union_lds, c->stream>>>(
                    order, lo, hi, in_start, in_count, edges, plat_pore, set_off, set_len, pool, pool_top, sb,
                    std::max(words, 1u), of, err);
                PMA_CHECKED(c, "set_union_bitmap", err);
            } else {
                set_union_level<<<grid_for(hi - lo, 64), 64, 0, c->stream>>>(
                    order, lo, hi, in_start, in_count, edges, plat_pore, set_off, set_len, pool, pool_top, cur, cend,
                    sb, of, err);
                PMA_CHECKED(c, "set_union_level", err);
            }
        }
        check_stage(c, err);
        const uint32_t o = read_scalar(c, of);
        if (!(o & 2u)) break;
        // After an overflow the lower levels merged truncated sets, so the space taken is only a
        // lower bound of what is needed: grow geometrically and start again.
        PMA_REQUIRE(attempt < 8, PMA_ERR_INTERNAL, "pma_pores: pore set pool overflow");
        unsigned long long tops[kSubPools] = {}, taken = 0;
        PMA_HIP(hipMemcpyAsync(tops, pool_top, sizeof tops, hipMemcpyDeviceToHost, c->stream));
        PMA_HIP(hipStreamSynchronize(c->stream));
        for (unsigned long long t : tops) taken += t;
        pool_cap = std::max<unsigned long long>(4 * pool_cap, 2 * taken);
    }
    const SetBounds sb{E, P, npores, pool_cap, std::max<unsigned long long>(1, std::min(kSubPools, pool_cap / 4096))};

    phase.mark("6 pore sets");
    // 7. key formations (prefix records) and their distinct lists.
    unsigned long long* pcnt = c->buf<unsigned long long>("pt_pcnt", P1);
    unsigned long long* rbase = c->buf<unsigned long long>("pt_rbase", P1);
    uint64_t R = 0;
    if (P) {
        prefix_counts<<<grid_for(P, 256), 256, 0, c->stream>>>(set_len, P, pcnt);
        PMA_CHECKED(c, "prefix_counts", err);
        R = exclusive_scan_u64(c, reinterpret_cast<const uint64_t*>(pcnt), reinterpret_cast<uint64_t*>(rbase), P);
    }
    PMA_REQUIRE(R <= pool_cap, PMA_ERR_INTERNAL, "pma_pores: more key formations than set members");
    uint32_t* rd = c->buf<uint32_t>("pt_rd", std::max<uint64_t>(R, 1));
    PrefixTrie trie;
    if (R)
        trie = build_prefix_trie(c, pool, pool_cap, reinterpret_cast<const uint64_t*>(set_off), set_len,
                                 reinterpret_cast<const uint64_t*>(rbase), nullptr, plat_size, P, R, rd);
    // the pore that forms a list is its largest member, the last one in ascending order
    const uint32_t D = trie.distinct;
    std::vector<uint32_t> ev_owner(D), ncls(npores, 0);
    for (uint32_t d = 0; d < D; ++d) {
        ev_owner[d] = (uint32_t)trie.key[d];
        PMA_REQUIRE(ev_owner[d] < npores, PMA_ERR_INTERNAL, "pma_pores: list member out of range");
        ncls[ev_owner[d]]++;
    }

    // reach sizes and boundary flags
    unsigned long long* reach = c->buf<unsigned long long>("pt_reach", std::max(npores, 1u));
    uint32_t* bflag = c->buf<uint32_t>("pt_bflag", std::max(npores, 1u));
    PMA_HIP(hipMemsetAsync(reach, 0, std::max(npores, 1u) * 8ull, c->stream));
    PMA_HIP(hipMemsetAsync(bflag, 0, std::max(npores, 1u) * 4ull, c->stream));
    if (P) {
        if (npores <= kLdsPores) {
            reach_sizes_lds<<<std::min(grid_for(P, 256), kLdsSumGrid), 256, npores * 8 + (npores + 31) / 32 * 4,
                              c->stream>>>(set_off, set_len, pool, plat_size, plat_border, sb, reach, bflag, err);
        } else {
            reach_sizes<<<grid_for(P, 256), 256, 0, c->stream>>>(set_off, set_len, pool, plat_size, plat_border, sb,
                                                                 reach, bflag, err);
        }
        PMA_CHECKED(c, "reach_sizes", err);
    }
    check_stage(c, err);
    std::vector<unsigned long long> reach_h(npores);
    std::vector<uint32_t> bflag_h(npores);
    if (npores) {
        PMA_HIP(hipMemcpyAsync(reach_h.data(), reach, npores * 8ull, hipMemcpyDeviceToHost, c->stream));
        PMA_HIP(hipMemcpyAsync(bflag_h.data(), bflag, npores * 4ull, hipMemcpyDeviceToHost, c->stream));
        PMA_HIP(hipStreamSynchronize(c->stream));
    }
    for (uint32_t k = 0; k < npores; ++k)
        PMA_REQUIRE(reach_h[k] >= 1 && reach_h[k] <= n, PMA_ERR_INTERNAL, "pma_pores: pore reach out of range");

    phase.mark("7 key formations");
    // 8. DFS replays for pores that form two or more lists: rank of each list inside its pore.
    std::vector<uint32_t> rank_h(D, kInv);
    uint64_t replay_tasks = 0, replay_steps = 0;
    {
        std::vector<ReplayTask> all;
        for (uint32_t k = 0; k < npores; ++k) {
            if (ncls[k] < 2) continue;
            ReplayTask t{};
            t.pore = k;
            t.seed = (uint32_t)seeds_h[k];
            t.ncls = ncls[k];
            all.push_back(t);
        }
        // largest reach first, so that the longest replays start at once
        std::stable_sort(all.begin(), all.end(),
                         [&](const ReplayTask& a, const ReplayTask& b2) { return reach_h[a.pore] > reach_h[b2.pore]; });
        replay_tasks = all.size();
        if (!all.empty()) {
            const size_t T = all.size();
            // skip data: out-edges by source, below[] (ascending CR), the class event of each set member
            unsigned long long* oe_in = c->buf<unsigned long long>("pt_oedges_in", E1);
            unsigned long long* oedges = c->buf<unsigned long long>("pt_oedges", E1);
            uint32_t* out_start = c->buf<uint32_t>("pt_out_start", P1);
            uint32_t* out_count = c->buf<uint32_t>("pt_out_count", P1);
            PMA_HIP(hipMemsetAsync(out_start, 0, P1 * 4, c->stream));
            PMA_HIP(hipMemsetAsync(out_count, 0, P1 * 4, c->stream));
            if (E) {
                swap_halves<<<grid_for(E, 256), 256, 0, c->stream>>>(edges, E, oe_in);
                PMA_CHECKED(c, "swap_halves", err);
                radix_sort_keys_u64(c, reinterpret_cast<uint64_t*>(oe_in), reinterpret_cast<uint64_t*>(oedges), E);
                edge_ranges<<<grid_for(E, 256), 256, 0, c->stream>>>(oedges, E, P, out_start, out_count, err);
                PMA_CHECKED(c, "edge_ranges", err);
            }
            uint32_t* below = c->buf<uint32_t>("pt_below", P1);
            for (size_t l = lstart_h.size() - 1; l-- > 0;) {
                const uint32_t lo = lstart_h[l], hi = lstart_h[l + 1];
                PMA_TIMED(c, "below_min");
                below_min<<<grid_for(hi - lo, 4), 256, 0, c->stream>>>(order, lo, hi, out_start, out_count, oedges,
                                                                      set_off, set_len, pool, sb, below, err);
                PMA_CHECKED(c, "below_min", err);
            }
            // per replayed pore: its (plateau, class) pairs, and where its class ranks go
            std::vector<uint32_t> ev_local(D), seen_cls(npores, 0), replayed_h(npores, 0);
            for (uint32_t d = 0; d < D; ++d) ev_local[d] = seen_cls[ev_owner[d]]++;
            std::vector<uint64_t> rank_base(npores, 0);
            uint64_t ranks_len = 0;
            uint32_t max_ncls = 1;
            for (const ReplayTask& t : all) {
                replayed_h[t.pore] = 1;
                rank_base[t.pore] = ranks_len;
                ranks_len += t.ncls;
                max_ncls = std::max(max_ncls, t.ncls);
            }
            PMA_REQUIRE(max_ncls < kNoClassU && ranks_len < 0xFFFFFFFFull && (max_ncls + 31) / 32 * 4 <= 8192,
                        PMA_ERR_CAPACITY,
                        "pma_pores: a pore forms 65535 or more throat keys (the DFS replay numbers them in 16 bits)");
            uint32_t* d_replayed = c->buf<uint32_t>("pt_replayed", npores);
            uint32_t* d_evl = c->buf<uint32_t>("pt_ev_local", std::max(D, 1u));
            unsigned long long* lcnt = c->buf<unsigned long long>("pt_list_count", npores);
            unsigned long long* loff = c->buf<unsigned long long>("pt_list_off", npores);
            unsigned long long* lcur = c->buf<unsigned long long>("pt_list_cursor", npores);
            upload(c, d_replayed, replayed_h.data(), npores * 4ull);
            upload(c, d_evl, ev_local.data(), D * 4ull);
            PMA_HIP(hipMemsetAsync(lcnt, 0, npores * 8ull, c->stream));
            PMA_HIP(hipMemsetAsync(lcur, 0, npores * 8ull, c->stream));
            if (npores <= kLdsPores)
                list_counts_lds<<<std::min(grid_for(P, 256), kLdsSumGrid), 256, npores * 8, c->stream>>>(
                    set_off, set_len, pool, sb, d_replayed, lcnt, err);
            else
                list_counts<<<grid_for(P, 256), 256, 0, c->stream>>>(set_off, set_len, pool, sb, d_replayed, lcnt, err);
            PMA_CHECKED(c, "list_counts", err);
            const uint64_t lists_len = exclusive_scan_u64(c, reinterpret_cast<const uint64_t*>(lcnt),
                                                          reinterpret_cast<uint64_t*>(loff), npores);
            unsigned long long* lists = c->buf<unsigned long long>("pt_lists", std::max<uint64_t>(lists_len, 1));
            if (npores <= kLdsPores)
                list_fill_lds<<<std::min(grid_for(P, 256), kLdsSumGrid), 256, npores * 8, c->stream>>>(
                    set_off, set_len, pool, rbase, rd, d_evl, sb, R, D, d_replayed, loff, lcur, lists, lists_len, err);
            else
                list_fill<<<grid_for(P, 256), 256, 0, c->stream>>>(set_off, set_len, pool, rbase, rd, d_evl, sb, R, D,
                                                                 d_replayed, loff, lcur, lists, lists_len, err);
            PMA_CHECKED(c, "list_fill", err);
            check_stage(c, err);
            std::vector<unsigned long long> lcnt_h(npores), loff_h(npores);
            PMA_HIP(hipMemcpyAsync(lcnt_h.data(), lcnt, npores * 8ull, hipMemcpyDeviceToHost, c->stream));
            PMA_HIP(hipMemcpyAsync(loff_h.data(), loff, npores * 8ull, hipMemcpyDeviceToHost, c->stream));
            PMA_HIP(hipStreamSynchronize(c->stream));
            for (ReplayTask& t : all) {
                PMA_REQUIRE(lcnt_h[t.pore] < 0xFFFFFFFFull, PMA_ERR_INTERNAL, "pma_pores: pore list too long");
                t.list_off = loff_h[t.pore];
                t.list_len = (uint32_t)lcnt_h[t.pore];
                t.rank_base = (uint32_t)rank_base[t.pore];
            }

            uint32_t* d_rank = c->buf<uint32_t>("pt_rank_slot", std::max<uint64_t>(ranks_len, 1));
            uint32_t* status = c->buf<uint32_t>("pt_task_status", T);
            unsigned long long* steps = c->buf<unsigned long long>("pt_task_steps", T);
            uint32_t* queue = c->buf<uint32_t>("pt_queue", 1);
            ReplayTask* d_tasks = c->buf<ReplayTask>("pt_tasks", T);
            PMA_HIP(hipMemsetAsync(d_rank, 0xFF, ranks_len * 4ull, c->stream));
            // Per wave: pages of visited voxels, two bitmaps over plateaus (entered, closed), the
            // dense class table, a stack and the traversal stack
            // and list.  The number of waves follows from the device memory free for them.
            // visited voxels by blocks of 2^vis_bits (at most 16384 blocks: the table is in LDS)
            uint32_t vis_bits = PMA_PORES_VIS_MIN_BITS;
            while (vis_bits < 32 && ((n + (1ull << vis_bits) - 1) >> vis_bits) > 16384) ++vis_bits;
            const uint32_t nblocks = (uint32_t)((n + (1ull << vis_bits) - 1) >> vis_bits);
            const unsigned long long page_words = 1ull << (vis_bits - 6);
            const unsigned long long pent_words = P / 64 + 1;
            const uint32_t seen_words = (max_ncls + 31) / 32;
            const unsigned long long trav_cap = std::min<unsigned long long>(P + 1, PMA_PORES_TRAV_CAP);
            unsigned long long max_reach = 0;
            for (const ReplayTask& t : all) max_reach = std::max<unsigned long long>(max_reach, reach_h[t.pore]);
            // the replay buffers of an earlier call are kept (allocating ~100 GB again costs more
            // than a second) and count as free here; any allocation that finds the device full
            // frees them first
            static const char* const kReplayBuffers[] = {"pt_replay_vis", "pt_replay_plateaus", "pt_replay_cls",
                                                         "pt_stack", "pt_replay_trav"};
            for (const char* name : kReplayBuffers) {
                c->persistent.insert(name);
                c->evictable.insert(name);
            }
            size_t free_b = 0, total_b = 0;
            PMA_HIP(hipMemGetInfo(&free_b, &total_b));
            free_b += c->evictable_bytes();
            for (const char* name : kReplayBuffers) c->evictable.erase(name);  // in use until the replays end
            const unsigned long long budget =
                std::min<unsigned long long>(c->replay_mem, free_b > (40ull << 30)
                                                                ? std::min<unsigned long long>(free_b - (24ull << 30),
                                                                                               PMA_PORES_REPLAY_MEM)
                                                                : free_b / 2);
            std::vector<ReplayTask> pending = all;
            for (int pass = 0; !pending.empty(); ++pass) {
                PMA_REQUIRE(pass < 2, PMA_ERR_INTERNAL, "pma_pores: DFS replay stack overflow with full capacity");
                // pass 0: stacks of up to PMA_PORES_STACK_CAP frames; pass 1: room for the whole reach
                const unsigned long long cap =
                    pass == 0 ? std::min<unsigned long long>(max_reach + 2, PMA_PORES_STACK_CAP) : max_reach + 2;
                // pass 1: a page for every block
                const uint32_t vis_pages = pass == 0 ? std::min<uint32_t>(nblocks, PMA_PORES_VIS_PAGES) : nblocks;
                const unsigned long long per_wave =
                    (vis_pages * page_words + 2 * pent_words) * 8 + P * 2 + cap * 16 + trav_cap * 8;
                unsigned long long W =
                    std::max<unsigned long long>(1, std::min<unsigned long long>({pending.size(), budget / per_wave, 1024}));
                {
                    // waves the kept buffers already hold: take them rather than grow by a little
                    const unsigned long long need[5] = {vis_pages * page_words * 8, 2 * pent_words * 8, P * 2,
                                                        cap * 16, trav_cap * 8};
                    unsigned long long fit = ~0ull;
                    for (int k = 0; k < 5; ++k) {
                        fit = std::min(fit, (unsigned long long)c->buffer_bytes(kReplayBuffers[k]) / std::max(need[k], 1ull));
                    }
                    // held buffers cost nothing to use: up to the knob's cap, even where the free
                    // memory alone (other stages hold theirs by now) would allow fewer waves
                    const unsigned long long cap_w = std::max<unsigned long long>(
                        1, std::min<unsigned long long>({pending.size(), c->replay_mem / per_wave, 1024}));
                    if (fit >= W - W / 4) W = std::max<unsigned long long>(1, std::min(fit, cap_w));
                }
                unsigned long long* vis = c->buf<unsigned long long>("pt_replay_vis", W * vis_pages * page_words);
                unsigned long long* pbits = c->buf<unsigned long long>("pt_replay_plateaus", W * 2 * pent_words);
                uint16_t* cls = c->buf<uint16_t>("pt_replay_cls", W * P);
                uint4* stack = c->buf<uint4>("pt_stack", W * cap);
                uint32_t* trav = c->buf<uint32_t>("pt_replay_trav", W * 2 * trav_cap);
                upload(c, d_tasks, pending.data(), pending.size() * sizeof(ReplayTask));
                PMA_HIP(hipMemsetAsync(status, 0, pending.size() * 4, c->stream));
                PMA_REQUIRE(E < 0xFFFFFFFFull && trav_cap < 0xFFFFFFFFull && ranks_len < 0xFFFFFFFFull, PMA_ERR_CAPACITY,
                            "pma_pores: plateau edges, traversal room or class ranks beyond 32 bits");
                const ReplayBounds rb{(uint32_t)n, (uint32_t)P, (uint32_t)E, (uint32_t)trav_cap, lists_len,
                                      (uint32_t)ranks_len};
                PMA_HIP(hipMemsetAsync(queue, 0, 4, c->stream));
                phase.mark("8 replay set-up and buffers");
                const auto t0 = std::chrono::steady_clock::now();
                {
                    PMA_TIMED(c, "dfs_replay");
                    dfs_replay<<<(unsigned)W, 64, seen_words * 4 + (nblocks * 2 + 3) / 4 * 4, c->stream>>>(
                        d_tasks, (uint32_t)pending.size(), queue, cr, plat, lists, below, out_start, out_count, oedges,
                        vis, vis_pages, vis_bits, pbits, pent_words, cls, seen_words, stack, cap, trav, d_rank,
                        status, steps, X, Y, Z, d3q7, rb, err);
                    PMA_CHECKED(c, "dfs_replay", err);
                }
                if (c->debug_sync)
                    std::fprintf(stderr, "[pma] dfs_replay pass %d on %llu waves: %.1f ms\n", pass, W,
                                 std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
                check_stage(c, err);
                std::vector<uint32_t> st(pending.size());
                std::vector<unsigned long long> sp(pending.size());
                PMA_HIP(hipMemcpyAsync(st.data(), status, pending.size() * 4ull, hipMemcpyDeviceToHost, c->stream));
                PMA_HIP(hipMemcpyAsync(sp.data(), steps, pending.size() * 8ull, hipMemcpyDeviceToHost, c->stream));
                PMA_HIP(hipStreamSynchronize(c->stream));
                unsigned long long smax = 0;
                for (size_t i = 0; i < pending.size(); ++i) {
                    replay_steps += sp[i];
                    smax = std::max(smax, sp[i]);
                }
                if (c->debug_sync)
                    std::fprintf(stderr, "[pma] dfs_replay pass %d: %zu pores, %llu steps, longest %llu\n", pass,
                                 pending.size(), (unsigned long long)replay_steps, smax);
                std::vector<ReplayTask> failed;
                for (size_t i = 0; i < pending.size(); ++i)
                    if (st[i]) failed.push_back(pending[i]);
                pending.swap(failed);
            }
            std::vector<uint32_t> rank_slot(ranks_len);
            PMA_HIP(hipMemcpyAsync(rank_slot.data(), d_rank, ranks_len * 4ull, hipMemcpyDeviceToHost, c->stream));
            PMA_HIP(hipStreamSynchronize(c->stream));
            for (uint32_t d = 0; d < D; ++d)
                if (replayed_h[ev_owner[d]]) rank_h[d] = rank_slot[rank_base[ev_owner[d]] + ev_local[d]];
            for (const char* name : kReplayBuffers) c->evictable.insert(name);
            c->drop("pt_lists");
            phase.mark("8 replay passes and release");
        }
        c->stage[PMA_STAGE_PORES].work_items = replay_steps;
        (void)replay_tasks;
    }
    // A replay stops with one class of its pore unranked: that class comes last.  Pores with
    // one class rank it 0.
    {
        std::vector<uint32_t> unranked(npores, 0);
        for (uint32_t d = 0; d < D; ++d)
            if (rank_h[d] == kInv) {
                PMA_REQUIRE(++unranked[ev_owner[d]] == 1, PMA_ERR_INTERNAL, "pma_pores: a replay left two classes unranked");
                rank_h[d] = ncls[ev_owner[d]] - 1;
            }
    }
    // creation order: by pore, then by first visit inside the pore's DFS
    std::vector<uint32_t> ord(D);
    std::iota(ord.begin(), ord.end(), 0u);
    std::sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t b) {
        if (ev_owner[a] != ev_owner[b]) return ev_owner[a] < ev_owner[b];
        return rank_h[a] < rank_h[b];
    });
    std::vector<uint32_t> key_id(std::max(D, 1u), 0);
    for (uint32_t r = 0; r < D; ++r) key_id[ord[r]] = 0xFFFFFFFFu - 1u - r;

    phase.mark("8 DFS replays");
    // 9. labels.
    uint32_t* d_kid = c->buf<uint32_t>("pt_key_id", std::max(D, 1u));
    upload(c, d_kid, key_id.data(), D * 4ull);
    {
        PMA_TIMED(c, "write_labels");
        const LabelBounds lb{P, pool_cap, R, D, npores};
        write_labels<<<nb, 256, 0, c->stream>>>(plat, set_off, set_len, pool, rbase, rd, d_kid, n, lb, pt, err);
        PMA_CHECKED(c, "write_labels", err);
    }
    unsigned long long* pvol = c->buf<unsigned long long>("pt_pvol", std::max(npores, 1u));
    PMA_HIP(hipMemsetAsync(pvol, 0, std::max(npores, 1u) * 8ull, c->stream));
    if (P) {
        single_volumes<<<grid_for(P, 256), 256, 0, c->stream>>>(set_off, set_len, pool, plat_size, sb, pvol, err);
        PMA_CHECKED(c, "single_volumes", err);
    }
    check_stage(c, err);
    std::vector<unsigned long long> pvol_h(npores);
    if (npores) PMA_HIP(hipMemcpyAsync(pvol_h.data(), pvol, npores * 8ull, hipMemcpyDeviceToHost, c->stream));
    PMA_HIP(hipStreamSynchronize(c->stream));

    phase.mark("9 labels");
    // 10. tables: pores in ascending ID, keys with volume in ascending ID, lists in the
    // reference's order [p2, p1, p3, ...].
    c->pores.assign(npores, PoreRow{});
    for (uint32_t k = 0; k < npores; ++k) {
        PoreRow& r = c->pores[k];
        r.id = 127u + 255u * (k / 129u) + (k % 129u);
        uint32_t v = (uint32_t)seeds_h[k];
        r.y = (int)(v % (uint32_t)Y);
        r.x = (int)((v / (uint32_t)Y) % (uint32_t)X);
        r.z = (int)(v / (uint32_t)Y / (uint32_t)X);
        r.r2 = (uint32_t)(seeds_h[k] >> 32);
        r.volume = pvol_h[k];
        r.boundary = bflag_h[k] != 0;
    }
    c->throats.clear();
    for (uint32_t r = D; r-- > 0;) {
        uint32_t d = ord[r];
        if (trie.finals[d] == 0) continue;
        std::vector<uint32_t> mem = trie.members(d);
        PMA_REQUIRE(mem.size() >= 2, PMA_ERR_INTERNAL, "pma_pores: a throat key with fewer than two pores");
        ThroatRow t;
        t.id = key_id[d];
        t.volume = trie.finals[d];
        t.pores.push_back(127u + 255u * (mem[1] / 129u) + (mem[1] % 129u));
        t.pores.push_back(127u + 255u * (mem[0] / 129u) + (mem[0] % 129u));
        for (size_t q = 2; q < mem.size(); ++q) t.pores.push_back(127u + 255u * (mem[q] / 129u) + (mem[q] % 129u));
        c->throats.push_back(t);
    }
    phase.mark("10 tables");
    c->have_pores = true;
    c->valid[PMA_VOL_PT] = true;
}

}  // namespace pma
