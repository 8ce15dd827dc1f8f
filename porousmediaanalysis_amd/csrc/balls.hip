// Containing balls and the maximum-ball mask (ContainingBallsAssigner.cpp:101-250).
//
// Reference semantics with one thread: for every void centre c with r2 = e(c), every void
// voxel v with |v - c|^2 <= r2 gets CR(v) = max r2; among equal maxima the first centre in
// x, y, z scan order wins (strict '<' at :170).  So the ball that a voxel takes is the first
// ball covering it in the order (r2 descending, scan(c) ascending), scan(c) = (x * Y + y) * Z + z.
//
// Work plan:
//   1. far tables: far_k(a) = max over lattice p with |p|^2 <= a of |p - d_k|^2 for the face,
//      edge and corner offsets d_k and the offsets (2, 1, 0) and (2, 1, 1) (up to signs and
//      order).  A centre c is skipped when a neighbour q at one of these offsets has
//      e(q) > e(c) and far_k(e(c)) <= e(q): then B(c) is a subset of B(q), every voxel of B(c)
//      has a ball of larger r2, and c can never be taken (exact pruning).  (Offsets (2, 0, 0),
//      (2, 2, 0), (2, 2, 1) and (2, 2, 2) are left out: they remove few or no further centres.)
//   2. the surviving centres, emitted column by column in scan order, are sorted stably by r2
//      descending: their rank is the order in which the reference would let them win.
//   3. every centre lists the pair cells (4 x 2 tiles of 8 x 8 x 64 voxels along z, x) its ball
//      reaches, as one 64-bit word holding the ball packed relative to the cell with the cell
//      number in its middle bits (or, when these do not fit, as (cell, packed ball) pairs, or
//      (cell, rank) pairs when a ball does not fit 64 bits); they are sorted stably by cell, so
//      each cell sees its balls in rank order.
//   4. one wave per tile, walking its cell's balls and skipping those that cannot reach the tile's
//      still-uncovered voxels: lane l owns the row (z, x) = (l / 8, l % 8) of the tile and a 64-bit mask
//      of the row's voxels that are solid or already covered.  For each ball in rank order a lane
//      takes the row's chord [y - h, y + h] of the ball; its uncovered voxels take this ball.
//      Every voxel is written once, the first time it is covered, so no atomics are needed.
//   5. every ball that wins a voxel marks its centre 255 in the mask
//      (SetMaximumBallsMaskInActiveArea).
#include "pma_internal.h"

#include <type_traits>

namespace pma {
namespace {

constexpr int TZ = 8, TX = 8, TY = 64;  // tile: 64 rows (z, x) of 64 voxels along y
#ifndef PMA_CB_SLAB_PAIRS
#define PMA_CB_SLAB_PAIRS (1ull << 30)  // most (cell, ball) pairs sorted at once
#endif
#ifndef PMA_CB_GZ
#define PMA_CB_GZ 4  // tiles per pair cell along z (4 x 2 measured best of 1..4 x 1..4 at 1024^3)
#endif
#ifndef PMA_CB_GX
#define PMA_CB_GX 2  // tiles per pair cell along x
#endif
#ifndef PMA_CB_PACKED
#define PMA_CB_PACKED 1  // 0 keeps ranks in the pair sort and gathers the balls afterwards
#endif
#ifndef PMA_CB_PRUNE_FAR_LDS
#define PMA_CB_PRUNE_FAR_LDS 2048u  // largest far-table length L that prune_centres copies to LDS
#endif
#ifndef PMA_CB_FUSED
#define PMA_CB_FUSED 1  // 0 sorts packed balls as values of separate cell keys even where they fit one word
#endif
#ifndef PMA_CB_SMALL_R2
#define PMA_CB_SMALL_R2 (1u << 22)  // r2 below this use isqrt_small (its exactness bound)
#endif
#ifndef PMA_CB_WIN16_MAX
#define PMA_CB_WIN16_MAX 0xFFFFu  // longest tile run whose winners fit 16-bit slots
#endif
constexpr uint8_t kCandidate = 126;    // temporary mask value of a surviving centre
constexpr uint32_t kXcds = 8;          // accelerator complex dies; workgroups go round-robin over them

struct alignas(16) Ball {
    int x, y, z;
    uint32_t r2;
};

__device__ __forceinline__ uint32_t isqrt_u32(uint32_t v) {
    uint32_t s = (uint32_t)sqrtf((float)v);
    while ((uint64_t)s * s > v) --s;
    while ((uint64_t)(s + 1) * (s + 1) <= v) ++s;
    return s;
}

// Lattice points of the non-negative octant with |p|^2 <= max_e; each records, per offset
// class, the largest |p - d|^2 reachable by a signed permutation of p.
__global__ void far_shells(uint32_t* __restrict__ shell, uint32_t max_e, int R) {
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    int b = blockIdx.y;
    if (a > R || b > R) return;
    uint32_t ab = (uint32_t)(a * a + b * b);
    size_t L = (size_t)max_e + 1;
    for (int c = 0; ab + (uint32_t)(c * c) <= max_e; ++c) {
        uint32_t p2 = ab + (uint32_t)(c * c);
        int hi = max(a, max(b, c)), lo = min(a, min(b, c));
        int s = a + b + c, mid = s - hi - lo;
        // |p' - d|^2 = p2 - 2 p'.d + |d|^2, largest for the image p' of p whose largest
        // coordinates point away from d
        atomicMax(&shell[p2], p2 + 2u * hi + 1u);                                    // (1, 0, 0)
        atomicMax(&shell[L + p2], p2 + 2u * (uint32_t)(s - lo) + 2u);                // (1, 1, 0)
        atomicMax(&shell[2 * L + p2], p2 + 2u * (uint32_t)s + 3u);                   // (1, 1, 1)
        atomicMax(&shell[3 * L + p2], p2 + 2u * (uint32_t)(2 * hi + mid) + 5u);      // (2, 1, 0)
        atomicMax(&shell[4 * L + p2], p2 + 2u * (uint32_t)(hi + s) + 6u);            // (2, 1, 1)
    }
}

// Running maximum over each class's table (one block per class).
__global__ void __launch_bounds__(1024) far_prefix_max(uint32_t* __restrict__ shell, size_t L) {
    __shared__ uint32_t part[1024];
    uint32_t* t = shell + blockIdx.x * L;
    size_t chunk = (L + 1023) / 1024;
    size_t b = threadIdx.x * chunk, e = min(L, b + chunk);
    uint32_t m = 0;
    for (size_t i = b; i < e; ++i) m = max(m, t[i]);
    part[threadIdx.x] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int i = 0; i < 1024; ++i) {
            uint32_t v = part[i];
            part[i] = run;
            run = max(run, v);
        }
    }
    __syncthreads();
    uint32_t run = part[threadIdx.x];
    for (size_t i = b; i < e; ++i) {
        run = max(run, t[i]);
        t[i] = run;
    }
}

// Block = 4 rows of 64 voxels along y; each thread walks its (x, y) column through every plane.
// Each plane passes once through a 2-slot LDS ring (with a two-voxel halo in x and y; outside the
// image reads 0): at step z the ring holds plane z + 2 while plane z + 3, loaded one step
// earlier, goes into the other slot and plane z + 4 is loaded into registers.  From plane z + 2
// the thread takes, for its own column, the maxima over the in-plane offset sets
//   M0 {(0,0)}, Ma {(0,±1),(±1,0)}, Mb {(±1,±1)}, Mc {(±1,±2),(±2,±1)}, Md {(0,±2),(±2,0)}
// and keeps them for planes z - 2 .. z + 2 in registers; the class maxima of voxel z are then
//   (1,0,0): Ma[z], M0[z±1]            (1,1,0): Mb[z], Ma[z±1]       (1,1,1): Mb[z±1]
//   (2,1,0): Mc[z], Md[z±1], Ma[z±2]   (2,1,1): Mc[z±1], Mb[z±2]
// (every offset of a class exactly once: 6, 12, 8, 24 and 24 offsets).  With FarLds the far
// tables are copied to LDS first (kFarClasses L words of dynamic shared memory).  Marks solid 0,
// void 127 and surviving centres kCandidate, and writes the column's count of surviving centres
// at col_count[x * Y + y] (the columns in the reference's scan order).
// c is dropped when a neighbour q of offset class k has q > c and q >= far_k(c), i.e. when the
// class maximum of q reaches max(c + 1, far_k(c)).
constexpr int kPruneHalo = 2, kPruneW = 64 + 2 * kPruneHalo, kPruneH = 4 + 2 * kPruneHalo;
constexpr int kPrunePlane = kPruneW * kPruneH;
constexpr int kPruneLoads = (kPrunePlane + 255) / 256;
constexpr int kFarClasses = 5;

constexpr uint32_t kPruneFarLds = PMA_CB_PRUNE_FAR_LDS;

template <bool FarLds>
__global__ void __launch_bounds__(256) prune_centres(const uint32_t* __restrict__ e, uint8_t* __restrict__ mask,
                                                     const uint32_t* __restrict__ far_global, uint32_t L,
                                                     uint32_t* __restrict__ col_count, int X, int Y, int Z) {
    __shared__ uint32_t ring[2][kPrunePlane];
    extern __shared__ uint32_t far_lds[];
    if (FarLds)
        for (uint32_t k = threadIdx.x; k < kFarClasses * L; k += blockDim.x) far_lds[k] = far_global[k];
    const uint32_t* far = FarLds ? far_lds : far_global;
    const int lane = threadIdx.x & 63, row = threadIdx.x >> 6;
    const int y = blockIdx.x * 64 + lane, x = blockIdx.y * 4 + row;
    auto fetch = [&](int z, uint32_t (&r)[kPruneLoads]) {
#pragma unroll
        for (int j = 0; j < kPruneLoads; ++j) {
            const int k = threadIdx.x + 256 * j;
            const int xx = blockIdx.y * 4 - kPruneHalo + k / kPruneW, yy = blockIdx.x * 64 - kPruneHalo + k % kPruneW;
            r[j] = (k < kPrunePlane && z >= 0 && z < Z && xx >= 0 && xx < X && yy >= 0 && yy < Y)
                       ? e[((size_t)z * X + xx) * Y + yy]
                       : 0u;
        }
    };
    auto store = [&](int z, const uint32_t (&r)[kPruneLoads]) {
#pragma unroll
        for (int j = 0; j < kPruneLoads; ++j) {
            const int k = threadIdx.x + 256 * j;
            if (k < kPrunePlane) ring[z & 1][k] = r[j];
        }
    };
    // window[s][t]: set s (M0, Ma, Mb, Mc, Md) of plane z - 2 + t
    uint32_t w0[5] = {}, wa[5] = {}, wb[5] = {}, wc[5] = {}, wd[5] = {};
    uint32_t r[kPruneLoads], ahead[kPruneLoads];
    fetch(-2, r);
    store(-2, r);
    fetch(-1, r);
    __syncthreads();
    uint32_t kept = 0;
    const int o = (row + kPruneHalo) * kPruneW + lane + kPruneHalo;
    for (int z = -4; z < Z; ++z) {
        // plane z + 4 is loaded a whole step before it is stored, so its latency is hidden
        fetch(z + 4, ahead);
        {
            const uint32_t* pl = ring[(z + 2) & 1];
            auto at = [&](int dx, int dy) { return pl[o + dx * kPruneW + dy]; };
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                w0[t] = w0[t + 1];
                wa[t] = wa[t + 1];
                wb[t] = wb[t + 1];
                wc[t] = wc[t + 1];
                wd[t] = wd[t + 1];
            }
            w0[4] = at(0, 0);
            wa[4] = max(max(at(0, -1), at(0, 1)), max(at(-1, 0), at(1, 0)));
            wb[4] = max(max(at(-1, -1), at(-1, 1)), max(at(1, -1), at(1, 1)));
            wc[4] = max(max(max(at(-1, -2), at(-1, 2)), max(at(1, -2), at(1, 2))),
                        max(max(at(-2, -1), at(-2, 1)), max(at(2, -1), at(2, 1))));
            wd[4] = max(max(at(0, -2), at(0, 2)), max(at(-2, 0), at(2, 0)));
        }
        if (z >= 0 && x < X && y < Y) {
            const size_t i = ((size_t)z * X + x) * Y + y;
            const uint32_t c = w0[2];
            if (c == 0) {
                mask[i] = 0;
            } else {
                const uint32_t m[kFarClasses] = {
                    max(wa[2], max(w0[1], w0[3])),
                    max(wb[2], max(wa[1], wa[3])),
                    max(wb[1], wb[3]),
                    max(max(wc[2], max(wd[1], wd[3])), max(wa[0], wa[4])),
                    max(max(wc[1], wc[3]), max(wb[0], wb[4])),
                };
                bool keep = true;
#pragma unroll
                for (int k = 0; k < kFarClasses; ++k) keep = keep && m[k] < max(c + 1, far[(size_t)k * L + c]);
                mask[i] = keep ? kCandidate : 127;
                kept += keep;
            }
        }
        store(z + 3, r);  // the slot of plane z + 1, last read in the previous step
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kPruneLoads; ++j) r[j] = ahead[j];
    }
    if (x < X && y < Y) col_count[(size_t)x * Y + y] = kept;
}

// Sort key of every surviving centre: (maxE - r2) << 32 | scan.  Each thread writes its column's
// centres in z order from the column's offset (exclusive scan of col_count), so the keys come out
// in scan order and a stable sort on the r2 bits alone gives the win order.
__global__ void __launch_bounds__(256) emit_centres(const uint32_t* __restrict__ e, uint8_t* __restrict__ mask,
                                                    const uint32_t* __restrict__ col_off,
                                                    unsigned long long* __restrict__ keys, uint32_t max_e, int X, int Y,
                                                    int Z) {
    const int y = blockIdx.x * 64 + (threadIdx.x & 63), x = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= X || y >= Y) return;
    uint32_t pos = col_off[(size_t)x * Y + y];
    const uint32_t scan0 = (uint32_t)(((uint64_t)x * Y + y) * (uint64_t)Z);
    const size_t plane = (size_t)X * Y;
    size_t i = (size_t)x * Y + y;
    int z = 0;
    // 8 planes at a time, their loads issued together (the rows they touch are read anyway)
    for (; z + 8 <= Z; z += 8, i += 8 * plane) {
        uint8_t m[8];
        uint32_t v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            m[j] = mask[i + j * plane];
            v[j] = e[i + j * plane];
        }
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if (m[j] == kCandidate) {
                keys[pos++] = ((unsigned long long)(max_e - v[j]) << 32) | (scan0 + (uint32_t)(z + j));
                mask[i + j * plane] = 127;
            }
    }
    for (; z < Z; ++z, i += plane) {
        if (mask[i] != kCandidate) continue;
        keys[pos++] = ((unsigned long long)(max_e - e[i]) << 32) | (scan0 + (uint32_t)z);
        mask[i] = 127;
    }
}

__global__ void decode_centres(const unsigned long long* __restrict__ keys, size_t n, uint32_t max_e, int Y, int Z,
                               Ball* __restrict__ balls) {
    size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const unsigned long long k = keys[r];
    const uint32_t scan = (uint32_t)k;
    Ball b;
    b.z = (int)(scan % (uint32_t)Z);
    const uint32_t xy = scan / (uint32_t)Z;
    b.y = (int)(xy % (uint32_t)Y);
    b.x = (int)(xy / (uint32_t)Y);
    b.r2 = max_e - (uint32_t)(k >> 32);
    balls[r] = b;
}

// Pair cells: gz x gx tiles in (z, x) share one pair list, so a ball is listed once per cell
// instead of once per tile; each tile's wave still skips the balls that miss its own voxels.
struct Cells {
    int sz, sx;    // cell edge in voxels along z and x (the y edge is TY)
    int gz, gx;    // tiles per cell along z and x
    int NCX, NCY;  // cells along x and y
};

// A ball packed relative to the origin (z0, x0, y0) of a cell it reaches, so that the pair
// sort carries the ball itself: r2 << r2_shift | (z - z0 + off) << 22 | (x - x0 + off) << 11 |
// (y - y0 + off).  A ball that reaches a cell has its centre within sqrt(r2) <= off - 1 of it,
// so each field is in [1, cell edge + 2 off - 2]; the host uses this form only when that is
// below 2^11 and r2 < 2^(64 - r2_shift).  Bits [33, r2_shift) hold the cell number when the word
// is its own sort key (r2_shift = 33 + the cell bits), and are absent when the cell is a separate
// key (r2_shift = 33).
__device__ __forceinline__ uint64_t pack_ball(const Ball& b, int z0, int x0, int y0, int off, int r2_shift) {
    return (uint64_t)b.r2 << r2_shift | (uint64_t)(b.z - z0 + off) << 22 | (uint64_t)(b.x - x0 + off) << 11 |
           (uint64_t)(b.y - y0 + off);
}

__device__ __forceinline__ Ball unpack_ball(uint64_t v, int z0, int x0, int y0, int off, int r2_shift) {
    Ball b;
    b.y = (int)(v & 2047u) - off + y0;
    b.x = (int)((v >> 11) & 2047u) - off + x0;
    b.z = (int)((v >> 22) & 2047u) - off + z0;
    b.r2 = (uint32_t)(v >> r2_shift);
    return b;
}

// Cells of the slab [cz0, cz1) that ball b reaches (box distance^2 <= r2): counted (pv == null)
// or listed from `at` as (cell - cell_base, value) pairs, the value being the ball's rank
// (V = uint32_t) or the ball packed relative to the cell (V = uint64_t); with V = uint64_t and
// pk null, as single words that carry the cell in bits [33, r2_shift).
template <class V>
__device__ uint32_t ball_tiles(const Ball& b, int X, int Y, int Z, Cells g, int cz0, int cz1, uint32_t cell_base,
                               uint32_t rank, int off, int r2_shift, uint32_t* __restrict__ pk,
                               V* __restrict__ pv, unsigned long long at) {
    const int R = (int)isqrt_u32(b.r2);
    const int za = max(b.z - R, 0) / g.sz, zb = min(b.z + R, Z - 1) / g.sz;
    const int xa = max(b.x - R, 0) / g.sx, xb = min(b.x + R, X - 1) / g.sx;
    const int ya = max(b.y - R, 0) / TY, yb = min(b.y + R, Y - 1) / TY;
    uint32_t n = 0;
    for (int tz = max(za, cz0); tz <= min(zb, cz1 - 1); ++tz) {
        const int z0 = tz * g.sz, z1 = min(z0 + g.sz, Z) - 1;
        const int dz = max(max(z0 - b.z, b.z - z1), 0);
        for (int tx = xa; tx <= xb; ++tx) {
            const int x0 = tx * g.sx, x1 = min(x0 + g.sx, X) - 1;
            const int dx = max(max(x0 - b.x, b.x - x1), 0);
            for (int ty = ya; ty <= yb; ++ty) {
                const int y0 = ty * TY, y1 = min(y0 + TY, Y) - 1;
                const int dy = max(max(y0 - b.y, b.y - y1), 0);
                if ((uint32_t)(dz * dz + dx * dx + dy * dy) > b.r2) continue;
                if (pv) {
                    const uint32_t cell = (uint32_t)((tz * g.NCX + tx) * g.NCY + ty) - cell_base;
                    if constexpr (sizeof(V) == 8) {
                        const uint64_t w = pack_ball(b, z0, x0, y0, off, r2_shift);
                        if (pk) {
                            pk[at + n] = cell;
                            pv[at + n] = w;
                        } else {
                            pv[at + n] = w | (uint64_t)cell << 33;
                        }
                    } else {
                        pk[at + n] = cell;
                        pv[at + n] = rank;
                    }
                }
                ++n;
            }
        }
    }
    return n;
}

__global__ void count_pairs(const Ball* __restrict__ balls, size_t n, int X, int Y, int Z, Cells g, int cz0, int cz1,
                            unsigned long long* __restrict__ cnt) {
    size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    cnt[r] = ball_tiles<uint32_t>(balls[r], X, Y, Z, g, cz0, cz1, 0, 0, 0, 33, nullptr, nullptr, 0);
}

template <class V>
__global__ void emit_pairs(const Ball* __restrict__ balls, size_t n, const unsigned long long* __restrict__ at,
                           int X, int Y, int Z, Cells g, int cz0, int cz1, uint32_t cell_base, int off,
                           int r2_shift, uint32_t* __restrict__ pk, V* __restrict__ pv) {
    size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    ball_tiles<V>(balls[r], X, Y, Z, g, cz0, cz1, cell_base, (uint32_t)r, off, r2_shift, pk, pv, at[r]);
}

// [start, end) of every cell's run in the cell-sorted (cell, rank) pair list, and the balls
// themselves in that order (so that a tile reads its cell's balls contiguously).
__global__ void tile_ranges(const uint32_t* __restrict__ pk, const uint32_t* __restrict__ rank, size_t P,
                            const Ball* __restrict__ balls, Ball* __restrict__ cell_balls,
                            uint32_t* __restrict__ start, uint32_t* __restrict__ end) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    cell_balls[i] = balls[rank[i]];
    const uint32_t t = pk[i];
    if (i == 0 || pk[i - 1] != t) start[t] = (uint32_t)i;
    if (i + 1 == P || pk[i + 1] != t) end[t] = (uint32_t)(i + 1);
}

// The same when no gather is needed, by binary search (one thread per cell): the list is sorted
// by cell, so a cell's run starts at the first entry whose cell is not below it.  The cells come
// from the words' bits [33, 33 + cell_bits) (Fused) or from separate keys.
template <bool Fused>
__global__ void cell_bounds(const uint64_t* __restrict__ words, const uint32_t* __restrict__ keys, size_t P,
                            uint32_t cell_mask, uint32_t n_cells, uint32_t* __restrict__ start,
                            uint32_t* __restrict__ end) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_cells) return;
    auto cell = [&](size_t j) -> uint32_t {
        if constexpr (Fused) return (uint32_t)(words[j] >> 33) & cell_mask;
        else return keys[j];
    };
    auto first_not_below = [&](uint32_t v) {
        size_t lo = 0, hi = P;
        while (lo < hi) {
            const size_t mid = lo + (hi - lo) / 2;
            if (cell(mid) < v) lo = mid + 1;
            else hi = mid;
        }
        return (uint32_t)lo;
    };
    start[t] = first_not_below(t);
    end[t] = first_not_below(t + 1);
}

// floor(sqrt(v)) from one correctly rounded float square root and a one-step correction.
__device__ __forceinline__ uint32_t isqrt_fast(uint32_t v) {
    uint32_t s = (uint32_t)sqrtf((float)v);
    if ((uint64_t)s * s > v) --s;
    else if ((uint64_t)(s + 1) * (s + 1) <= v) ++s;
    return s;
}

// floor(sqrt(v)) for 0 <= v < 2^22: the hardware square root is within 1 ulp, so the truncated
// root is off by at most one, and one correction each way (in 32-bit integers) makes it exact.
__device__ __forceinline__ int isqrt_small(int v) {
    int s = (int)__builtin_amdgcn_sqrtf((float)v);
    s -= __mul24(s, s) > v;
    s += __mul24(s + 1, s + 1) <= v;
    return s;
}

// How cover_tiles takes the chord half-width floor(sqrt(rem)).
enum ChordRoot : int {
    kRootGeneral = 0,  // isqrt_fast (any r2)
    kRootSmall = 1,    // isqrt_small (r2 < 2^22)
    kRootTable = 2,    // a table of floor(sqrt(v)) in LDS (r2 < kRootTableSize - 1)
};
constexpr int kRootTableSize = 2048;

// Largest pair run of any cell of the slab (decides the width of the per-voxel winner slot).
__global__ void __launch_bounds__(256) max_run(const uint32_t* __restrict__ start, const uint32_t* __restrict__ end,
                                               size_t n, unsigned int* __restrict__ out) {
    uint32_t m = 0;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
        m = max(m, end[t] - start[t]);
    for (int off = 32; off > 0; off >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, off));
    if ((threadIdx.x & 63) == 0 && m) atomicMax(out, m);
}

// One wave per tile (work plan, step 4), walking the pair run of the tile's cell.  The winner of
// a voxel is kept in LDS as the position of its ball in that run (16 bits when every run is
// shorter than 65536).  Workgroups are dispatched round-robin over the 8 XCDs, so the block
// index is mapped such that the tiles of one cell run on the same XCD and share its L2.
template <class W, int Root, bool Packed>
__global__ void __launch_bounds__(64) cover_tiles(const Ball* __restrict__ cell_balls,
                                                  const uint64_t* __restrict__ cell_packed, int off, int r2_shift,
                                                  const uint32_t* __restrict__ start, const uint32_t* __restrict__ end,
                                                  const uint32_t* __restrict__ e, uint32_t* __restrict__ cr,
                                                  uint32_t* __restrict__ cx, uint32_t* __restrict__ cy,
                                                  uint32_t* __restrict__ cz, uint8_t* __restrict__ mask, int X, int Y,
                                                  int Z, Cells g, uint32_t cell_base, uint32_t slab_cells,
                                                  uint32_t* __restrict__ error_flag) {
    __shared__ W win[TZ * TX * TY];
    // root_tab[v + 1] = floor(sqrt(v)) and root_tab[0] = -1 (a row the ball misses)
    __shared__ int8_t root_tab[Root == kRootTable ? kRootTableSize : 1];
    const int lane = threadIdx.x;
    const uint32_t per_cell = (uint32_t)(g.gz * g.gx);
    const uint32_t k = blockIdx.x / kXcds, sub = k % per_cell;
    const uint32_t local = (k / per_cell) * kXcds + blockIdx.x % kXcds;  // cell within the slab
    if (local >= slab_cells) return;
    const uint32_t cell = cell_base + local;
    const int ty = (int)(cell % (uint32_t)g.NCY), ccx = (int)((cell / (uint32_t)g.NCY) % (uint32_t)g.NCX);
    const int ccz = (int)(cell / ((uint32_t)g.NCY * (uint32_t)g.NCX));
    const int tz = ccz * g.gz + (int)(sub / (uint32_t)g.gx), tx = ccx * g.gx + (int)(sub % (uint32_t)g.gx);
    const int z = tz * TZ + lane / TX, x = tx * TX + lane % TX, y0 = ty * TY;
    const int oz = ccz * g.sz, ox = ccx * g.sx;  // origin of the cell (packed balls are relative to it)
    using Raw = std::conditional_t<Packed, uint64_t, Ball>;
    auto raw_at = [&](uint32_t i) -> Raw {
        if constexpr (Packed) return cell_packed[i];
        else return cell_balls[i];
    };
    auto ball_of = [&](const Raw& r) -> Ball {
        if constexpr (Packed) return unpack_ball(r, oz, ox, y0, off, r2_shift);
        else return r;
    };
    if (Root == kRootTable)
        for (int v = lane; v < kRootTableSize; v += 64) root_tab[v] = v == 0 ? (int8_t)-1 : (int8_t)isqrt_small(v - 1);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();

    // coverage of this lane's row: solid voxels and voxels outside the image count as covered
    uint64_t cov = ~0ull;
    if (z < Z && x < X) {
        const uint32_t* row = e + ((size_t)z * X + x) * Y + y0;
        if ((Y & 3) == 0 && y0 + TY <= Y) {  // 16-byte aligned full row
            const uint4* row4 = reinterpret_cast<const uint4*>(row);
            uint4 v[TY / 4];
#pragma unroll
            for (int q = 0; q < TY / 4; ++q) v[q] = row4[q];
            uint64_t open = 0;
#pragma unroll
            for (int q = 0; q < TY / 4; ++q)
                open |= (uint64_t)((v[q].x != 0) | (v[q].y != 0) << 1 | (v[q].z != 0) << 2 | (v[q].w != 0) << 3)
                        << (4 * q);
            cov = ~open;
        } else {
            for (int j = 0; j < TY && y0 + j < Y; ++j)
                if (row[j] != 0) cov &= ~(1ull << j);
        }
    }

    const uint32_t s = start[local], en = end[local];
    Raw next{};
    if (s + lane < en) next = raw_at(s + lane);
    // bounding box of the voxels still uncovered (a ball that cannot reach it covers nothing new),
    // taken again only after a batch that covered something
    int zlo = 0, zhi = 0, xlo = 0, xhi = 0, ylo = 0, yhi = 0;
    bool dirty = true;  // uniform
    for (uint32_t base = s; base < en; base += 64) {
        if (dirty) {
            const bool open_row = cov != ~0ull;
            const uint64_t rows = __ballot(open_row);
            if (rows == 0) break;  // every void voxel of the tile has its ball
            const uint64_t unc = open_row ? ~cov : 0ull;
            const uint64_t ys = (uint64_t)__ockl_wfred_or_u32((uint32_t)unc) |
                                (uint64_t)__ockl_wfred_or_u32((uint32_t)(unc >> 32)) << 32;
            uint32_t xs = 0;
            for (int r = 0; r < TZ; ++r) xs |= (uint32_t)(rows >> (r * TX)) & 0xFFu;
            zlo = tz * TZ + __builtin_ctzll(rows) / TX;
            zhi = tz * TZ + (63 - __builtin_clzll(rows)) / TX;
            xlo = tx * TX + __builtin_ctz(xs);
            xhi = tx * TX + 31 - __builtin_clz(xs);
            ylo = y0 + __builtin_ctzll(ys);
            yhi = y0 + 63 - __builtin_clzll(ys);
            dirty = false;
        }

        const Raw cur = next;
        const Ball mb = ball_of(cur);  // this batch; the next one is loaded while it is processed
        if (base + 64 + lane < en) next = raw_at(base + 64 + lane);
        bool reach = false;
        if (base + lane < en) {
            const int ez = max(max(zlo - mb.z, mb.z - zhi), 0), ex = max(max(xlo - mb.x, mb.x - xhi), 0);
            const int ey = max(max(ylo - mb.y, mb.y - yhi), 0);
            reach = (uint32_t)(ez * ez + ex * ex + ey * ey) <= mb.r2;
        }
        uint64_t todo = __ballot(reach), won = 0;
        while (todo) {
            const int j = __builtin_ctzll(todo);
            todo &= todo - 1;
            // the ball's fields are uniform: read them into scalar registers (a packed ball takes
            // two reads and is unpacked on the scalar unit)
            int bx, by, bz, r2;
            if constexpr (Packed) {
                const uint64_t pv = (uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(cur >> 32), j) << 32 |
                                    (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)cur, j);
                const Ball u = unpack_ball(pv, oz, ox, y0, off, r2_shift);
                bx = u.x;
                by = u.y;
                bz = u.z;
                r2 = (int)u.r2;
            } else {
                bx = __builtin_amdgcn_readlane(mb.x, j);
                by = __builtin_amdgcn_readlane(mb.y, j);
                bz = __builtin_amdgcn_readlane(mb.z, j);
                r2 = __builtin_amdgcn_readlane((int)mb.r2, j);
            }
            const int dz = z - bz, dx = x - bx;
            const int rem = r2 - (__mul24(dz, dz) + __mul24(dx, dx));  // |dz|, |dx| < 2^16
            // chord [by - h, by + h] of the ball on this lane's row; h = -1 when the row misses it
            int h;
            if (Root == kRootTable) h = root_tab[max(rem, -1) + 1];  // rem <= r2 < kRootTableSize - 1
            else h = rem < 0 ? -1 : (Root == kRootSmall ? isqrt_small(rem) : (int)isqrt_fast((uint32_t)rem));
            const int by_t = __builtin_amdgcn_readfirstlane(by - y0);  // the centre's y within the tile (scalar)
            const int lo = max(by_t - h, 0), hi = min(by_t + h, 63);
            uint64_t nw = lo <= hi ? (((~0ull) >> (63 - (hi - lo))) << lo) & ~cov : 0ull;
            if (__ballot(nw != 0) == 0) continue;
            won |= 1ull << j;
            dirty = true;
            if (nw) {
                cov |= nw;
                const W slot = (W)(base - s + (uint32_t)j);
                do {
                    const int bit = __ffsll((long long)nw) - 1;
                    win[lane * TY + bit] = slot;
                    nw &= nw - 1;
                } while (nw);
            }
        }
        // a ball that won a voxel of this tile is a maximal ball
        if ((won >> lane) & 1ull) mask[((size_t)mb.z * X + mb.x) * Y + mb.y] = 255;
    }

    // outputs, one row of 64 voxels per step (lanes along y)
    for (int r = 0; r < 64; ++r) {
        const uint64_t row_cov = __shfl(cov, r);
        const int zr = tz * TZ + r / TX, xr = tx * TX + r % TX, yr = y0 + lane;
        if (zr >= Z || xr >= X || yr >= Y) continue;
        const size_t i = ((size_t)zr * X + xr) * Y + yr;
        if (e[i] == 0) {
            cr[i] = 0u;
            cx[i] = cy[i] = cz[i] = 0xFFFFFFFFu;
            continue;
        }
        if (!((row_cov >> lane) & 1ull)) {
            atomicOr(error_flag, 1u);
            continue;
        }
        const Ball b = ball_of(raw_at(s + (uint32_t)win[r * TY + lane]));
        cr[i] = b.r2;
        cx[i] = (uint32_t)b.x;
        cy[i] = (uint32_t)b.y;
        cz[i] = (uint32_t)b.z;
    }
}

struct CoverArgs {
    uint32_t blocks;
    const Ball* cell_balls;
    const uint64_t* cell_packed;
    int off, r2_shift;
    const uint32_t *start, *end, *e;
    uint32_t *cr, *cx, *cy, *cz;
    uint8_t* mask;
    int X, Y, Z;
    Cells g;
    uint32_t cell_base, slab_cells;
    uint32_t* err;
};

template <class W, int Root, bool Packed>
void launch_cover(hipStream_t stream, const CoverArgs& a) {
    cover_tiles<W, Root, Packed><<<a.blocks, 64, 0, stream>>>(a.cell_balls, a.cell_packed, a.off, a.r2_shift, a.start,
                                                              a.end, a.e,
                                                              a.cr, a.cx, a.cy, a.cz, a.mask, a.X, a.Y, a.Z, a.g,
                                                              a.cell_base, a.slab_cells, a.err);
}

using CoverLaunch = void (*)(hipStream_t, const CoverArgs&);
// [16-bit winner slots][chord root mode][packed balls]
constexpr CoverLaunch kCoverLaunch[2][3][2] = {
    {{launch_cover<uint32_t, kRootGeneral, false>, launch_cover<uint32_t, kRootGeneral, true>},
     {launch_cover<uint32_t, kRootSmall, false>, launch_cover<uint32_t, kRootSmall, true>},
     {launch_cover<uint32_t, kRootTable, false>, launch_cover<uint32_t, kRootTable, true>}},
    {{launch_cover<uint16_t, kRootGeneral, false>, launch_cover<uint16_t, kRootGeneral, true>},
     {launch_cover<uint16_t, kRootSmall, false>, launch_cover<uint16_t, kRootSmall, true>},
     {launch_cover<uint16_t, kRootTable, false>, launch_cover<uint16_t, kRootTable, true>}},
};

// Bytes of w equal to the byte b (b replicated in all four bytes): exact count of the zero
// bytes of w ^ b (a zero byte, and only a zero byte, leaves 0x80 in its place).
__device__ __forceinline__ uint32_t equal_bytes(uint32_t w, uint32_t b) {
    const uint32_t t = w ^ b;
    return (uint32_t)__popc(~(((t & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | t | 0x7F7F7F7Fu));
}

// Count of bytes equal to `value`, 16 bytes per thread and step (the buffer is 16-byte aligned).
__global__ void __launch_bounds__(256) count_equal_u8(const uint8_t* __restrict__ v, size_t n, uint8_t value,
                                                      unsigned long long* __restrict__ out) {
    unsigned long long local = 0;
    const uint32_t b = 0x01010101u * value;
    const size_t n16 = n / 16;
    const uint4* v4 = reinterpret_cast<const uint4*>(v);
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (size_t)gridDim.x * blockDim.x) {
        const uint4 q = v4[i];
        local += equal_bytes(q.x, b) + equal_bytes(q.y, b) + equal_bytes(q.z, b) + equal_bytes(q.w, b);
    }
    for (size_t i = n16 * 16 + (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        local += v[i] == value;
    for (int off = 32; off > 0; off >>= 1) local += __shfl_xor(local, off);
    if ((threadIdx.x & 63) == 0 && local) atomicAdd(out, local);
}

}  // namespace

void run_containing_balls(pma_ctx* c) {
    PMA_REQUIRE(c->valid[PMA_VOL_EDT2] && c->have_edt_stats, PMA_ERR_STATE,
                "pma_containing_balls: the EDT and its statistics have not been computed or loaded");
    const int X = c->dims.x, Y = c->dims.y, Z = c->dims.z;
    PMA_REQUIRE((double)X * Y * Z < 4294967295.0, PMA_ERR_INVALID_ARGUMENT,
                "pma_containing_balls: the volume must have fewer than 2^32 voxels");
    for (pma_volume w : {PMA_VOL_CR2, PMA_VOL_CX, PMA_VOL_CY, PMA_VOL_CZ, PMA_VOL_MASK}) c->ensure_volume(w);
    const uint32_t max_e = c->max_e;
    const int R = (int)std::floor(std::sqrt((double)max_e));
    const size_t L = (size_t)max_e + 1;
    uint32_t* far = c->buf<uint32_t>("cb_far", kFarClasses * L);
    PMA_HIP(hipMemsetAsync(far, 0, kFarClasses * L * sizeof(uint32_t), c->stream));
    far_shells<<<dim3(grid_for(R + 1, 64), R + 1), 64, 0, c->stream>>>(far, max_e, R);
    PMA_LAUNCHED(c);
    far_prefix_max<<<kFarClasses, 1024, 0, c->stream>>>(far, L);
    PMA_LAUNCHED(c);

    const uint32_t* e = c->v<uint32_t>(PMA_VOL_EDT2);
    uint8_t* mask = c->v<uint8_t>(PMA_VOL_MASK);
    const dim3 cgrid(grid_for(Y, 64), grid_for(X, 4));  // one thread per (x, y) column

    // 1-2. surviving centres in scan order, then sorted by r2 (stable) into win order.
    const size_t ncols = (size_t)X * Y;
    uint32_t* col_count = c->buf<uint32_t>("cb_col_count", ncols);
    uint32_t* col_off = c->buf<uint32_t>("cb_col_off", ncols);
    {
        PMA_TIMED(c, "prune_centres");
        if (L <= kPruneFarLds)
            prune_centres<true><<<cgrid, 256, kFarClasses * L * sizeof(uint32_t), c->stream>>>(e, mask, far, (uint32_t)L,
                                                                                     col_count, X, Y, Z);
        else
            prune_centres<false><<<cgrid, 256, 0, c->stream>>>(e, mask, far, (uint32_t)L, col_count, X, Y, Z);
        PMA_LAUNCHED(c);
    }
    // fewer than 2^32 voxels, so the offsets fit 32 bits
    const unsigned long long n_cand = exclusive_scan_u32(c, col_count, col_off, ncols);
    const size_t N1 = std::max<unsigned long long>(n_cand, 1);
    unsigned long long* keys = c->buf<unsigned long long>("cb_keys", N1);
    unsigned long long* keys2 = c->buf<unsigned long long>("cb_keys2", N1);
    {
        PMA_TIMED(c, "emit_centres");
        emit_centres<<<cgrid, 256, 0, c->stream>>>(e, mask, col_off, keys, max_e, X, Y, Z);
        PMA_LAUNCHED(c);
    }
    int hi_bits = 0;
    while (hi_bits < 32 && (max_e >> hi_bits) != 0) ++hi_bits;
    const int r2_bits = std::max(hi_bits, 1);  // bits of any r2 <= max_e
    radix_sort_keys_u64(c, reinterpret_cast<const uint64_t*>(keys), reinterpret_cast<uint64_t*>(keys2), n_cand, 32,
                        32 + std::max(hi_bits, 1));
    Ball* balls = c->buf<Ball>("cb_balls", N1);
    if (n_cand) {
        decode_centres<<<grid_for(n_cand, 256), 256, 0, c->stream>>>(keys2, n_cand, max_e, Y, Z, balls);
        PMA_LAUNCHED(c);
    }

    // 3-4. (cell, rank) pairs sorted by cell, then one wave per tile; z-slabs of cells keep the
    // pair list below kSlabPairs.
    Cells g;
    g.gz = PMA_CB_GZ;
    g.gx = PMA_CB_GX;
    g.sz = TZ * g.gz;
    g.sx = TX * g.gx;
    const int NCZ = (Z + g.sz - 1) / g.sz;
    g.NCX = (X + g.sx - 1) / g.sx;
    g.NCY = (Y + TY - 1) / TY;
    const uint32_t per_cell = (uint32_t)(g.gz * g.gx);
    const uint32_t layer_cells = (uint32_t)g.NCX * (uint32_t)g.NCY;
    constexpr unsigned long long kSlabPairs = PMA_CB_SLAB_PAIRS;
    unsigned long long* pcount = c->buf<unsigned long long>("cb_pair_count", N1);
    unsigned long long* poff = c->buf<unsigned long long>("cb_pair_off", N1);
    uint32_t* err = c->buf<uint32_t>("cb_error", 1);
    PMA_HIP(hipMemsetAsync(err, 0, sizeof(uint32_t), c->stream));
    // balls travel packed in the pair sort when they fit 64 bits relative to their cell
    const int off = (int)std::floor(std::sqrt((double)max_e)) + 1;
    const bool packed = PMA_CB_PACKED && max_e < (1u << 31) && std::max(std::max(g.sz, g.sx), TY) + 2 * off < 2048;
    uint64_t total_pairs = 0;
    int slabs = 1;
    for (int tz0 = 0; tz0 < NCZ;) {
        int tz1 = std::min(NCZ, tz0 + (NCZ + slabs - 1) / slabs);
        uint64_t P = 0;
        while (true) {
            if (n_cand) {
                count_pairs<<<grid_for(n_cand, 256), 256, 0, c->stream>>>(balls, n_cand, X, Y, Z, g, tz0, tz1, pcount);
                PMA_LAUNCHED(c);
                P = exclusive_scan_u64(c, reinterpret_cast<const uint64_t*>(pcount), reinterpret_cast<uint64_t*>(poff),
                                       n_cand);
            }
            if (P <= kSlabPairs || tz1 - tz0 == 1) break;
            tz1 = tz0 + (tz1 - tz0 + 1) / 2;  // halve the slab until its pairs fit
        }
        PMA_REQUIRE(P <= 0xFFFFFFFFull, PMA_ERR_CAPACITY, "pma_containing_balls: one tile layer reaches too many balls");
        const uint32_t cell_base = (uint32_t)tz0 * layer_cells;
        const uint32_t slab_cells = (uint32_t)(tz1 - tz0) * layer_cells;
        const size_t P1 = std::max<uint64_t>(P, 1);
        int cell_bits = 1;
        while (cell_bits < 32 && (slab_cells >> cell_bits) != 0) ++cell_bits;
        // one sort key per pair: the packed ball with its cell in bits [33, 33 + cell_bits)
        const bool fused = packed && PMA_CB_FUSED && 33 + cell_bits + r2_bits <= 64;
        const int r2_shift = fused ? 33 + cell_bits : 33;
        uint32_t* pk = nullptr;
        uint32_t* pk2 = nullptr;
        if (!fused) {
            pk = c->buf<uint32_t>("cb_pair_tile", P1);
            pk2 = c->buf<uint32_t>("cb_pair_tile2", P1);
        }
        uint32_t* pv = nullptr;
        uint32_t* pv2 = nullptr;
        Ball* cell_balls = nullptr;
        uint64_t* pb = nullptr;
        uint64_t* pb2 = nullptr;
        if (packed) {
            pb = c->buf<uint64_t>("cb_pair_ball", P1);
            pb2 = c->buf<uint64_t>("cb_pair_ball2", P1);
        } else {
            pv = c->buf<uint32_t>("cb_pair_rank", P1);
            pv2 = c->buf<uint32_t>("cb_pair_rank2", P1);
            cell_balls = c->buf<Ball>("cb_cell_balls", P1);
        }
        uint32_t* tstart = c->buf<uint32_t>("cb_cell_start", slab_cells);
        uint32_t* tend = c->buf<uint32_t>("cb_cell_end", slab_cells);
        PMA_HIP(hipMemsetAsync(tstart, 0, slab_cells * sizeof(uint32_t), c->stream));
        PMA_HIP(hipMemsetAsync(tend, 0, slab_cells * sizeof(uint32_t), c->stream));
        if (P) {
            if (fused) {
                emit_pairs<uint64_t><<<grid_for(n_cand, 256), 256, 0, c->stream>>>(balls, n_cand, poff, X, Y, Z, g, tz0,
                                                                                   tz1, cell_base, off, r2_shift,
                                                                                   nullptr, pb);
                PMA_LAUNCHED(c);
                radix_sort_keys_u64(c, pb, pb2, P, 33, 33 + cell_bits);
                cell_bounds<true><<<grid_for(slab_cells, 256), 256, 0, c->stream>>>(
                    pb2, nullptr, P, (1u << cell_bits) - 1u, slab_cells, tstart, tend);
                PMA_LAUNCHED(c);
            } else if (packed) {
                emit_pairs<uint64_t><<<grid_for(n_cand, 256), 256, 0, c->stream>>>(balls, n_cand, poff, X, Y, Z, g, tz0,
                                                                                   tz1, cell_base, off, r2_shift, pk,
                                                                                   pb);
                PMA_LAUNCHED(c);
                radix_sort_pairs_u32_u64(c, pk, pk2, pb, pb2, P, 0, cell_bits);
                cell_bounds<false><<<grid_for(slab_cells, 256), 256, 0, c->stream>>>(nullptr, pk2, P, 0u, slab_cells,
                                                                                     tstart, tend);
                PMA_LAUNCHED(c);
            } else {
                emit_pairs<uint32_t><<<grid_for(n_cand, 256), 256, 0, c->stream>>>(balls, n_cand, poff, X, Y, Z, g, tz0,
                                                                                   tz1, cell_base, off, r2_shift, pk,
                                                                                   pv);
                PMA_LAUNCHED(c);
                radix_sort_pairs_u32(c, pk, pk2, pv, pv2, P, 0, cell_bits);
                tile_ranges<<<grid_for(P, 256), 256, 0, c->stream>>>(pk2, pv2, P, balls, cell_balls, tstart, tend);
                PMA_LAUNCHED(c);
            }
        }
        unsigned int longest = 0;
        if (P) {
            unsigned int* dmax = c->buf<unsigned int>("cb_max_run", 1);
            PMA_HIP(hipMemsetAsync(dmax, 0, 4, c->stream));
            max_run<<<std::min(grid_for(slab_cells, 256), 1024u), 256, 0, c->stream>>>(tstart, tend, slab_cells, dmax);
            PMA_LAUNCHED(c);
            longest = read_scalar(c, dmax);
        }
        {
            PMA_TIMED(c, "cover_tiles");
            uint32_t* cr = c->v<uint32_t>(PMA_VOL_CR2);
            uint32_t* cx = c->v<uint32_t>(PMA_VOL_CX);
            uint32_t* cy = c->v<uint32_t>(PMA_VOL_CY);
            uint32_t* cz = c->v<uint32_t>(PMA_VOL_CZ);
            // one workgroup per tile, cell-major within each XCD's share (see cover_tiles)
            const uint32_t blocks = (slab_cells + kXcds - 1) / kXcds * kXcds * per_cell;
            // chord widths: a table below kRootTableSize, the short integer root below its
            // exactness bound, else the general one
            const int root = max_e < (uint32_t)std::min<uint64_t>(kRootTableSize - 1, PMA_CB_SMALL_R2) ? kRootTable
                             : max_e < PMA_CB_SMALL_R2                                            ? kRootSmall
                                                                                                   : kRootGeneral;
            const bool w16 = longest <= PMA_CB_WIN16_MAX;
            CoverArgs a{blocks, cell_balls, pb2, off, r2_shift, tstart, tend, e, cr, cx, cy, cz, mask, X, Y, Z, g,
                        cell_base, slab_cells, err};
            kCoverLaunch[w16 ? 1 : 0][root][packed ? 1 : 0](c->stream, a);
            PMA_LAUNCHED(c);
        }
        total_pairs += P;
        tz0 = tz1;
        slabs = std::max(slabs, (int)((total_pairs + kSlabPairs - 1) / kSlabPairs));
    }

    unsigned long long* cnt = c->buf<unsigned long long>("cb_count", 1);
    PMA_HIP(hipMemsetAsync(cnt, 0, sizeof(unsigned long long), c->stream));
    {
        PMA_TIMED(c, "count_equal_u8");
        count_equal_u8<<<1024, 256, 0, c->stream>>>(mask, c->n, 255, cnt);
        PMA_LAUNCHED(c);
    }
    PMA_REQUIRE(read_scalar(c, err) == 0, PMA_ERR_INTERNAL, "pma_containing_balls: a void voxel was left uncovered");
    c->n_max_balls = read_scalar(c, cnt);
    c->stage[PMA_STAGE_CONTAINING_BALLS].work_items = total_pairs;
    for (pma_volume w : {PMA_VOL_CR2, PMA_VOL_CX, PMA_VOL_CY, PMA_VOL_CZ, PMA_VOL_MASK}) c->valid[w] = true;
}

}  // namespace pma
