// Geodesic distances and cavities (ShortestPathComputer.cpp:39-511).
//
// The reference runs a float32 Dijkstra: d(v) = min over settled neighbours u of
// fl(d(u) + w) with w = sqrtf(1|2|3) by |offset|_1 (1.0f with D3Q7).  fl(a + w) is isotone
// in a and never below a, so the result is the least fixed point
//     d(v) = min(d(v), min_u fl(d(u) + w)),  i.e. the minimum over paths of the left-fold
// float32 sum, whatever the settling order.  It is computed here by block-synchronous label
// correcting: the volume is cut into 8 x 8 x 8 blocks; an active block is taken by one wave,
// which stages it and a one-voxel halo in LDS (+inf for solid voxels and voxels outside the
// image), relaxes its own voxels by pulling from their neighbours until nothing changes, writes
// back what improved, and activates the neighbour blocks across any face, edge or corner where
// a border voxel improved.  An active block's key is the smallest improved border distance
// that woke it; a round processes the active blocks whose key is within kDelta of the smallest
// key and carries the others, so blocks are visited roughly in distance order (fewer
// re-visits).  Rounds run back to back on the device until no block is active.  Every add is
// one correctly rounded fp32 add (no contraction).
//
// Start voxels follow SelectStartingPosition / FindVoidPixelOnBoxBoundaries (:396-511):
// the centre min((int)(0.5 * size + 0.5), size - 1) if void and unreached, else the first
// void unreached voxel on box shells of growing half size h, scanned face by face
// (constant x low/high, then y, then z; the other two axes outer-to-inner).  The scan
// position of a voxel is computed in closed form and the first one is found by atomicMin.
#include "pma_internal.h"

#include <cfloat>
#include <cmath>
#include <mutex>

namespace pma {
namespace {

constexpr uint32_t kFltMaxBits = 0x7F7FFFFFu;

// Four voxels per thread: solid -> distance -1 and cavity 0, void -> FLT_MAX and ~0.
__global__ void geo_init(const uint8_t* __restrict__ solid, float* __restrict__ dist, uint32_t* __restrict__ cav,
                         size_t n) {
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t i = q * 4;
    if (i + 4 <= n) {
        const uchar4 s = reinterpret_cast<const uchar4*>(solid)[q];
        reinterpret_cast<float4*>(dist)[q] = make_float4(s.x ? -1.0f : FLT_MAX, s.y ? -1.0f : FLT_MAX,
                                                         s.z ? -1.0f : FLT_MAX, s.w ? -1.0f : FLT_MAX);
        reinterpret_cast<uint4*>(cav)[q] = make_uint4(s.x ? 0u : 0xFFFFFFFFu, s.y ? 0u : 0xFFFFFFFFu,
                                                      s.z ? 0u : 0xFFFFFFFFu, s.w ? 0u : 0xFFFFFFFFu);
    } else {
        for (size_t k = i; k < n; ++k) {
            const bool sk = solid[k] != 0;
            dist[k] = sk ? -1.0f : FLT_MAX;
            cav[k] = sk ? 0u : 0xFFFFFFFFu;
        }
    }
}

// Smallest scan key of voxel v among the shells h in [lo_h, hi_h), or ~0 if none.
__device__ unsigned long long shell_key(const int v[3], const int c[3], const int S[3], int lo_h, int hi_h) {
    unsigned long long best = ~0ull;
    for (int k = 0; k < 3; ++k) {
        const int a0 = k == 0 ? 1 : 0, a1 = k == 2 ? 1 : 2;
        int lower = max(lo_h, max(abs(v[a0] - c[a0]), abs(v[a1] - c[a1])));
        int hexact = abs(v[k] - c[k]);
        int h = -1;
        if (hexact >= lower) h = hexact;
        else {
            // a face clamped to the image border stays on v[k] for every larger h
            if (v[k] == 0 && lower >= c[k]) h = lower;
            if (v[k] == S[k] - 1 && lower >= S[k] - 1 - c[k]) h = lower;
        }
        if (h < 0 || h >= hi_h) continue;
        int f0 = max(c[k] - h, 0);
        int fi = v[k] == f0 ? 0 : 1;
        unsigned long long key = ((unsigned long long)h << 44) | ((unsigned long long)k << 42) |
                                 ((unsigned long long)fi << 41) | ((unsigned long long)v[a0] << 20) |
                                 (unsigned long long)v[a1];
        best = min(best, key);
    }
    return best;
}

// Grid-stride; one atomicMin per block.
__global__ void __launch_bounds__(256) find_start(const uint8_t* __restrict__ solid, const float* __restrict__ dist,
                                                  int X, int Y, int Z, int cx, int cy, int cz, int lo_h, int hi_h,
                                                  unsigned long long* __restrict__ out) {
    __shared__ unsigned long long part[4];
    unsigned long long best = ~0ull;
    const size_t n = (size_t)X * Y * Z;
    const int c[3] = {cx, cy, cz}, S[3] = {X, Y, Z};
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        if (solid[i] != 0 || __float_as_uint(dist[i]) != kFltMaxBits) continue;
        int v[3];
        v[1] = (int)(i % Y);
        size_t r = i / Y;
        v[0] = (int)(r % X);
        v[2] = (int)(r / X);
        best = min(best, shell_key(v, c, S, lo_h, hi_h));
    }
    for (int off = 32; off > 0; off >>= 1) best = min(best, (unsigned long long)__shfl_xor(best, off));
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
        best = min(min(part[0], part[1]), min(part[2], part[3]));
        if (best != ~0ull) atomicMin(out, best);
    }
}


constexpr int kB = 8;           // block edge
constexpr float kDelta = 16.0f;  // default bucket width of a round (distance units; 4..64 measured within 10%)
#ifndef PMA_GEO_RELAX_GRID
// one-wave workgroups of a relax launch: 64 per CU, twice what fits at once, so that waves that
// finish early are replaced (16384 measured best of 2048..16384 at 1024^3)
#define PMA_GEO_RELAX_GRID 16384u
#endif
#ifndef PMA_GEO_PUSH_BUF
#define PMA_GEO_PUSH_BUF 256u  // per-wave buffer of activated blocks (more than 27)
#endif
constexpr uint32_t kRelaxGrid = PMA_GEO_RELAX_GRID;
constexpr uint32_t kPushBuf = PMA_GEO_PUSH_BUF;
static_assert(kPushBuf > 27, "a block activates up to 26 neighbours at once");
constexpr int kH = kB + 2;      // with halo
constexpr int kHalo = kH * kH * kH;

// Rounds run back to back on the device: every kernel of a round reads the list lengths it
// needs from the round states, so the host does not wait between rounds.
struct RoundState {
    uint32_t count;  // length of the active list
    uint32_t kmin;   // smallest key over it (float bits; +inf when empty)
    uint32_t proc;   // length of the round's processing list
    uint32_t pad;
};

__global__ void fill_key(uint32_t* __restrict__ p, uint32_t v, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Splits the active list: keys within `delta` of its smallest key go to `proc` (their key is
// cleared, so that activations during the round start it afresh), the others are carried into
// the next list (tagged so that a new activation does not list them twice).  Grid-stride over
// the list, whose length is cur->count, 256 entries per workgroup step; each step reserves its
// places in the two lists with one atomic per list.
__global__ void __launch_bounds__(256) partition_active(const uint32_t* __restrict__ active,
                                                        RoundState* __restrict__ cur, uint32_t* __restrict__ block_key,
                                                        uint32_t* __restrict__ proc, uint32_t* __restrict__ next,
                                                        uint32_t* __restrict__ block_tag, uint32_t tag,
                                                        RoundState* __restrict__ nxt, float delta) {
    __shared__ uint32_t wave_now[4], wave_later[4], base_now, base_later;
    const uint32_t n = cur->count;
    const float thr = __uint_as_float(cur->kmin) + delta;
    const uint32_t threshold = isfinite(thr) ? __float_as_uint(thr) : 0xFFFFFFFFu;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned long long below = (1ull << lane) - 1;
    for (uint32_t step = blockIdx.x * blockDim.x; step < n; step += gridDim.x * blockDim.x) {  // uniform
        const uint32_t t = step + threadIdx.x;
        uint32_t b = 0, key = 0;
        bool now = false, later = false;
        if (t < n) {
            b = active[t];
            key = block_key[b];
            now = key <= threshold;
            later = !now;
        }
        if (now) block_key[b] = kFltMaxBits;
        if (later) (void)atomicExch(&block_tag[b], tag);
        const unsigned long long bn = __ballot(now), bl = __ballot(later);
        if (lane == 0) {
            wave_now[wave] = (uint32_t)__popcll(bn);
            wave_later[wave] = (uint32_t)__popcll(bl);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            const uint32_t tn = wave_now[0] + wave_now[1] + wave_now[2] + wave_now[3];
            const uint32_t tl = wave_later[0] + wave_later[1] + wave_later[2] + wave_later[3];
            base_now = tn ? atomicAdd(&cur->proc, tn) : 0u;
            base_later = tl ? atomicAdd(&nxt->count, tl) : 0u;
        }
        __syncthreads();
        uint32_t off_n = base_now, off_l = base_later;
        for (int w = 0; w < wave; ++w) {
            off_n += wave_now[w];
            off_l += wave_later[w];
        }
        if (now) proc[off_n + (uint32_t)__popcll(bn & below)] = b;
        if (later) next[off_l + (uint32_t)__popcll(bl & below)] = b;
        __syncthreads();  // the counts are read before the next step overwrites them
    }
}

// Smallest key over the next list (length nxt->count, read on the device), one atomic per
// workgroup.  Workgroup 0 also adds the round's block visits to *visits and prepares the two
// states for the next round, in which they swap roles: cur (read by nothing after this
// kernel) becomes the empty next list, and nxt's processing count starts at 0.
__global__ void __launch_bounds__(256) next_min_key(const uint32_t* __restrict__ next,
                                                    const uint32_t* __restrict__ block_key,
                                                    RoundState* __restrict__ nxt, RoundState* __restrict__ cur,
                                                    unsigned long long* __restrict__ visits) {
    __shared__ uint32_t part[4];
    const uint32_t n = nxt->count;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *visits += cur->proc;
        cur->count = 0;
        cur->kmin = 0x7F800000u;  // +inf
        nxt->proc = 0;
    }
    uint32_t m = 0x7F800000u;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        m = min(m, block_key[next[i]]);
    m = __ockl_wfred_min_u32(m);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        m = min(min(part[0], part[1]), min(part[2], part[3]));
        if (m != 0x7F800000u) atomicMin(&nxt->kmin, m);
    }
}

// One wave relaxes one block of this round's processing list (length cur->proc) at a time,
// striding over the list by the grid size.  Lane l owns the z-column (x, y) = (l / 8, l % 8) of the block
// and keeps its 8 values in registers; a sweep walks every column along z (up and down in turn),
// so a value crosses the whole column in one sweep and one column per sweep sideways.  Sweeps
// repeat until one changes nothing.  The block and a one-voxel halo sit in LDS with +inf for
// solid voxels and voxels outside the image: inf + w stays inf, and an unreached FLT_MAX + w
// rounds back to FLT_MAX, which never lowers a void voxel, so every neighbour is one add and one
// min.  The voxel `src` (or none, ~0) is set to 0 while staging and counts as improved, so the
// blocks around it wake up too.
// Rounding is monotone, so fl(min(a, b) + w) = min(fl(a + w), fl(b + w)): the neighbours of one
// weight are reduced first and the weight is added once per class (faces 1, edges sqrt 2,
// corners sqrt 3), which gives the same value as adding it to each neighbour.
template <bool D3Q7, bool UP>
__device__ __forceinline__ bool sweep_columns(float* sd, int col, float (&v)[kB], uint32_t open) {
    const float w2 = sqrtf(2.0f), w3 = sqrtf(3.0f);
    bool changed = false;
#pragma unroll
    for (int s = 0; s < kB; ++s) {
        const int z = UP ? s : kB - 1 - s;
        if (!((open >> z) & 1u)) continue;
        const int c = (z + 1) * kH * kH + col;
        // own column: registers inside the block, LDS at its ends
        float face = fminf(z > 0 ? v[z > 0 ? z - 1 : 0] : sd[c - kH * kH],
                           z < kB - 1 ? v[z < kB - 1 ? z + 1 : 0] : sd[c + kH * kH]);
        float edge = INFINITY, corner = INFINITY;
#pragma unroll
        for (int k = 0; k < 27; ++k) {
            const int dz = k / 9 - 1, dx = (k / 3) % 3 - 1, dy = k % 3 - 1;
            if (dx == 0 && dy == 0) continue;  // own column, above
            const int m = abs(dz) + abs(dx) + abs(dy);
            if (D3Q7 && m > 1) continue;
            const float u = sd[c + (dz * kH + dx) * kH + dy];
            if (m == 1) face = fminf(face, u);
            else if (m == 2) edge = fminf(edge, u);
            else corner = fminf(corner, u);
        }
        float best = fminf(v[z], __fadd_rn(face, 1.0f));
        if (!D3Q7) best = fminf(best, fminf(__fadd_rn(edge, w2), __fadd_rn(corner, w3)));
        if (best < v[z]) {
            v[z] = best;
            sd[c] = best;
            changed = true;
        }
    }
    return changed;
}

// A voxel first lowered from FLT_MAX (unreached) belongs to the current cavity: it takes the
// cavity id, and the workgroup adds it and the boundary planes it lies on (bits axis * 2 +
// side) to its own slot of part_count / part_walls, which the host clears per cavity.
template <bool D3Q7>
__global__ void __launch_bounds__(64) relax_blocks(float* __restrict__ dist, uint32_t* __restrict__ cav, uint32_t id,
                                                   const uint32_t* __restrict__ proc, uint32_t* __restrict__ next,
                                                   uint32_t* __restrict__ block_tag, uint32_t tag,
                                                   uint32_t* __restrict__ block_key,
                                                   const RoundState* __restrict__ cur, RoundState* __restrict__ nxt,
                                                   unsigned long long* __restrict__ part_count,
                                                   uint32_t* __restrict__ part_walls, int X, int Y, int Z, int NBX,
                                                   int NBY, int NBZ, unsigned long long src) {
    __shared__ float sd[kHalo];
    __shared__ uint32_t pushbuf[kPushBuf];  // activated blocks, reserved in `next` a batch at a time
    const int lane = threadIdx.x;
    const int cx = lane >> 3, cy = lane & 7;
    const int col = (cx + 1) * kH + cy + 1;  // the lane's column within a z-layer of sd
    const uint32_t count = cur->proc;
    // the short staging path needs lz * X * Y in 24-bit operands and a 32-bit offset
    const bool fast_plane = (uint64_t)X * (uint64_t)Y < (1ull << 24);
    uint32_t npush = 0;                      // entries in pushbuf (uniform)
    uint32_t reached = 0, walls = 0;         // this lane's voxels new to the cavity, and their planes
    auto flush = [&]() {
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(&nxt->count, npush);
        base = (uint32_t)__shfl((int)base, 0);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        for (uint32_t k = lane; k < npush; k += 64) next[base + k] = pushbuf[k];
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        npush = 0;
    };
    for (uint32_t item = blockIdx.x; item < count; item += gridDim.x) {
        const uint32_t b = proc[item];
        const int by = (int)(b % (uint32_t)NBY), bx = (int)((b / (uint32_t)NBY) % (uint32_t)NBX);
        const int bz = (int)(b / ((uint32_t)NBY * (uint32_t)NBX));
        const int z0 = bz * kB - 1, x0 = bx * kB - 1, y0 = by * kB - 1;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");  // the previous block's LDS use is over
        __builtin_amdgcn_wave_barrier();
        // all loads first, then all stores, so that the loads are in flight together.  The lane
        // number is passed through an empty asm so that the compiler recomputes the staging
        // offsets per block instead of keeping them live across the whole loop.
        int slane = lane;
        asm volatile("" : "+v"(slane));
        constexpr int kStage = (kHalo + 63) / 64;
        float staged[kStage];
        // a block whose halo lies inside the image, in a round without a start voxel, takes the
        // short path: 32-bit offsets from the halo's first voxel and no bounds tests (k / 100 and
        // r / 10 by 24-bit multiplies, exact for k < 1000 and r < 100)
        const bool interior = src == ~0ull && fast_plane && z0 >= 0 && x0 >= 0 && y0 >= 0 && z0 + kH <= Z &&
                              x0 + kH <= X && y0 + kH <= Y;  // uniform
        if (interior) {
            const float* hb = dist + ((size_t)z0 * X + x0) * Y + y0;
#pragma unroll
            for (int j = 0; j < kStage; ++j) {
                const uint32_t k = (uint32_t)slane + 64u * (uint32_t)j;
                const uint32_t lz = __umul24(k, 656u) >> 16, r = k - __umul24(lz, 100u);
                const uint32_t lx = __umul24(r, 6554u) >> 16, ly = r - __umul24(lx, 10u);
                const uint32_t rel = __umul24(lz, (uint32_t)X * (uint32_t)Y) + __umul24(lx, (uint32_t)Y) + ly;
                const float v = k < (uint32_t)kHalo ? hb[rel] : -1.0f;
                staged[j] = v < 0.0f ? INFINITY : v;
            }
        } else {
#pragma unroll
            for (int j = 0; j < kStage; ++j) {
                const int k = slane + 64 * j;
                const int lz = k / (kH * kH), lx = (k / kH) % kH, ly = k % kH;
                const int z = z0 + lz, x = x0 + lx, y = y0 + ly;
                const bool in = k < kHalo && z >= 0 && z < Z && x >= 0 && x < X && y >= 0 && y < Y;
                const size_t gi = ((size_t)z * X + x) * Y + y;
                const float v = in ? dist[gi] : -1.0f;
                staged[j] = (in && gi == src) ? 0.0f : (v < 0.0f ? INFINITY : v);
            }
        }
#pragma unroll
        for (int j = 0; j < kStage; ++j)
            if (slane + 64 * j < kHalo) sd[slane + 64 * j] = staged[j];
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();

        const int gx = x0 + 1 + cx, gy = y0 + 1 + cy;
        float v[kB], orig[kB];
        uint32_t open = 0;  // bit z: the voxel is void and inside the image
#pragma unroll
        for (int z = 0; z < kB; ++z) {
            const float cur = sd[(z + 1) * kH * kH + col];
            v[z] = cur;
            orig[z] = cur;
            open |= (cur != INFINITY ? 1u : 0u) << z;
        }
        if (src != ~0ull) {  // uniform: the start voxel was staged as 0 and counts as improved
#pragma unroll
            for (int z = 0; z < kB; ++z) {
                const int gz = z0 + 1 + z;
                if (gz < Z && gx < X && gy < Y && ((size_t)gz * X + gx) * Y + gy == src) orig[z] = FLT_MAX;
            }
        }
        for (bool up = true;; up = !up) {
            const bool changed = up ? sweep_columns<D3Q7, true>(sd, col, v, open)
                                    : sweep_columns<D3Q7, false>(sd, col, v, open);
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");  // this sweep's writes precede the next reads
            __builtin_amdgcn_wave_barrier();
            if (__ballot(changed) == 0) break;
        }

        // write back what improved; border voxels that improved wake the blocks across
        uint32_t touch = 0, bmin = kFltMaxBits;
        const int xlo = cx == 0 ? -1 : 0, xhi = cx == kB - 1 ? 1 : 0;
        const int ylo = cy == 0 ? -1 : 0, yhi = cy == kB - 1 ? 1 : 0;
        const uint32_t side_xy = (gx == 0 ? 1u : 0u) | (gx == X - 1 ? 2u : 0u) | (gy == 0 ? 4u : 0u) |
                                 (gy == Y - 1 ? 8u : 0u);
#pragma unroll
        for (int z = 0; z < kB; ++z) {
            if (!(v[z] < orig[z])) continue;
            const int gz = z0 + 1 + z;
            const size_t gi = ((size_t)gz * X + gx) * Y + gy;
            dist[gi] = v[z];
            if (__float_as_uint(orig[z]) == kFltMaxBits) {
                cav[gi] = id;
                ++reached;
                walls |= side_xy | (gz == 0 ? 16u : 0u) | (gz == Z - 1 ? 32u : 0u);
            }
            const int zlo = z == 0 ? -1 : 0, zhi = z == kB - 1 ? 1 : 0;
            if (!(zlo | zhi | xlo | xhi | ylo | yhi)) continue;
            for (int dz = zlo; dz <= zhi; ++dz)
                for (int dx = xlo; dx <= xhi; ++dx)
                    for (int dy = ylo; dy <= yhi; ++dy)
                        if (dz | dx | dy) touch |= 1u << ((dz + 1) * 9 + (dx + 1) * 3 + (dy + 1));
            bmin = min(bmin, __float_as_uint(v[z]));
        }
        const uint32_t wtouch = __ockl_wfred_or_u32(touch);
        const uint32_t wbmin = __ockl_wfred_min_u32(bmin);
        bool pushed = false;
        uint32_t nb = 0;
        if (lane < 27 && ((wtouch >> lane) & 1u)) {
            const int dz = lane / 9 - 1, dx = (lane / 3) % 3 - 1, dy = lane % 3 - 1;
            const int nz = bz + dz, nx = bx + dx, ny = by + dy;
            if (nz >= 0 && nz < NBZ && nx >= 0 && nx < NBX && ny >= 0 && ny < NBY) {
                nb = ((uint32_t)nz * (uint32_t)NBX + (uint32_t)nx) * (uint32_t)NBY + (uint32_t)ny;
                atomicMin(&block_key[nb], wbmin);
                pushed = atomicExch(&block_tag[nb], tag) != tag;
            }
        }
        const uint64_t pm = __ballot(pushed);
        if (pushed) pushbuf[npush + (uint32_t)__popcll(pm & ((1ull << lane) - 1))] = nb;
        npush += (uint32_t)__popcll(pm);
        if (npush > kPushBuf - 27) flush();
    }
    if (npush) flush();
    reached = __ockl_wfred_add_u32(reached);
    walls = __ockl_wfred_or_u32(walls);
    if (lane == 0) {
        part_count[blockIdx.x] += reached;
        part_walls[blockIdx.x] |= walls;
    }
}

// Sums the workgroups' slots of the cavity just finished.
__global__ void __launch_bounds__(256) cavity_totals(const unsigned long long* __restrict__ part_count,
                                                     const uint32_t* __restrict__ part_walls, uint32_t n,
                                                     unsigned long long* __restrict__ count,
                                                     uint32_t* __restrict__ walls) {
    __shared__ unsigned long long pc[4];
    __shared__ uint32_t pw[4];
    unsigned long long c = 0;
    uint32_t w = 0;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        c += part_count[i];
        w |= part_walls[i];
    }
    for (int off = 32; off > 0; off >>= 1) {
        c += __shfl_xor(c, off);
        w |= (uint32_t)__shfl_xor((int)w, off);
    }
    if ((threadIdx.x & 63) == 0) {
        pc[threadIdx.x >> 6] = c;
        pw[threadIdx.x >> 6] = w;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        *count = pc[0] + pc[1] + pc[2] + pc[3];
        *walls = pw[0] | pw[1] | pw[2] | pw[3];
    }
}

// Distinct boundary planes as bits: a size-1 axis has one plane (FillAllBoundaryPlanes).
uint32_t plane_bits(uint32_t wb, const int S[3]) {
    uint32_t out = 0;
    for (int a = 0; a < 3; ++a) {
        uint32_t lo = (wb >> (2 * a)) & 1u, hi = (wb >> (2 * a + 1)) & 1u;
        if (S[a] == 1) out |= (lo | hi) << (2 * a);
        else out |= (lo << (2 * a)) | (hi << (2 * a + 1));
    }
    return out;
}

__global__ void fill_u64(unsigned long long* __restrict__ p, unsigned long long v, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Scratch of one run: per 8^3 block the two active lists, the processing list, the round tag
// and the wake key (NB entries each); the two round states; counters; one cavity slot per
// relaxing workgroup.
struct GeoScratch {
    uint32_t *fa, *fb, *fp, *block_tag, *block_key;
    RoundState* rs;
    unsigned long long *visits, *key, *vcount;
    uint32_t* walls;
    unsigned long long* part_count;
    uint32_t* part_walls;
};

// How the run reaches the host: a context (its staging buffer, kernel timing and launch
// counts), or, without one, plain copies on the caller's stream.
struct GeoHost {
    hipStream_t stream;
    pma_ctx* ctx;  // null for the allocation-free entry point
    template <class T> T read(const T* p) {
        if (ctx) return read_scalar(ctx, p);
        T v;
        PMA_HIP(hipMemcpyAsync(&v, p, sizeof(T), hipMemcpyDeviceToHost, stream));
        PMA_HIP(hipStreamSynchronize(stream));
        return v;
    }
    void put(void* dst, const void* src, size_t bytes) {
        PMA_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream));
        PMA_HIP(hipStreamSynchronize(stream));
    }
    void launched(const char* name) {
        if (ctx) {
            PMA_CHECKED(ctx, name, nullptr);
            return;
        }
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) throw Error(PMA_ERR_HIP, std::string("launch of ") + name + ": " + hipGetErrorString(e));
    }
};

// ShortestPathComputer::ComputeShortestPathDistances on device buffers: distances and cavity ids
// into dist and cav, one row per cavity handed to emit(row) (which returns false to stop with
// PMA_ERR_CAPACITY).  Returns the number of block relaxations.  Clears its scratch with kernels
// (no memset nodes), so the launches it makes are all kernels and small copies.
template <class Emit>
unsigned long long geodesic_core(GeoHost& h, const uint8_t* solid, float* dist, uint32_t* cav, int X, int Y, int Z,
                                 int d3q7, const GeoScratch& g, Emit emit) {
    const size_t n = (size_t)X * Y * Z;
    hipStream_t st = h.stream;
    geo_init<<<grid_for((n + 3) / 4, 256), 256, 0, st>>>(solid, dist, cav, n);
    h.launched("geo_init");
    const int NBZ = (Z + kB - 1) / kB, NBX = (X + kB - 1) / kB, NBY = (Y + kB - 1) / kB;
    const size_t NB = (size_t)NBZ * NBX * NBY;
    fill_u64<<<1, 64, 0, st>>>(g.visits, 0ull, 1);
    h.launched("fill_u64");
    fill_key<<<grid_for(NB, 256), 256, 0, st>>>(g.block_tag, 0xFFFFFFFFu, NB);
    h.launched("fill_key");
    fill_key<<<grid_for(NB, 256), 256, 0, st>>>(g.block_key, kFltMaxBits, NB);
    h.launched("fill_key");

    const int S[3] = {X, Y, Z};
    int ctr[3];
    for (int a = 0; a < 3; ++a) ctr[a] = std::min((int)(0.5 * S[a] + 0.5), S[a] - 1);
    int max_half = 0;
    for (int a = 0; a < 3; ++a) max_half = std::max(max_half, std::max(S[a] - 1 - ctr[a], ctr[a]));
    uint32_t all_walls = plane_bits(0x3Fu, S);

    int half = 0;
    uint32_t id = 0, round = 0;
    while (true) {
        ++half;
        int start[3];
        bool found = false;
        const size_t ci = ((size_t)ctr[2] * X + ctr[0]) * Y + ctr[1];
        uint8_t cs = h.read(solid + ci);
        float cd = h.read(dist + ci);
        if (cs == 0 && cd == FLT_MAX) {
            for (int a = 0; a < 3; ++a) start[a] = ctr[a];
            found = true;
        } else if (half < max_half) {
            fill_u64<<<1, 64, 0, st>>>(g.key, ~0ull, 1);
            h.launched("fill_u64");
            find_start<<<std::min(grid_for(n, 256), 8192u), 256, 0, st>>>(solid, dist, X, Y, Z, ctr[0], ctr[1],
                                                                        ctr[2], half, max_half, g.key);
            h.launched("find_start");
            unsigned long long k = h.read(g.key);
            if (k != ~0ull) {
                int hh = (int)(k >> 44), ax = (int)((k >> 42) & 3), fi = (int)((k >> 41) & 1);
                int a0 = ax == 0 ? 1 : 0, a1 = ax == 2 ? 1 : 2;
                start[a0] = (int)((k >> 20) & 0xFFFFF);
                start[a1] = (int)(k & 0xFFFFF);
                start[ax] = fi == 0 ? std::max(ctr[ax] - hh, 0) : std::min(ctr[ax] + hh, S[ax] - 1);
                half = hh;
                found = true;
            } else {
                half = max_half;
            }
        }
        if (!found) break;
        id += 127;
        const size_t si = ((size_t)start[2] * X + start[0]) * Y + start[1];
        const uint32_t b0 = ((uint32_t)(start[2] / kB) * (uint32_t)NBX + (uint32_t)(start[0] / kB)) * (uint32_t)NBY +
                            (uint32_t)(start[1] / kB);
        h.put(g.fa, &b0, sizeof b0);
        const uint32_t zero = 0;
        h.put(g.block_key + b0, &zero, sizeof zero);
        fill_u64<<<grid_for(kRelaxGrid, 256), 256, 0, st>>>(g.part_count, 0ull, kRelaxGrid);
        h.launched("fill_u64");
        fill_key<<<grid_for(kRelaxGrid, 256), 256, 0, st>>>(g.part_walls, 0u, kRelaxGrid);
        h.launched("fill_key");
        // round 0 processes the start block (key 0); the host checks for an empty list after
        // rounds 1, 2, 4, 8 and then every 8 rounds, and never waits between the others
        const RoundState start_states[2] = {{1u, 0u, 0u, 0u}, {0u, 0x7F800000u, 0u, 0u}};
        h.put(g.rs, start_states, sizeof start_states);
        uint32_t r = 0, check_at = 1, step = 1;
        while (true) {
            RoundState* cur_s = g.rs + (r & 1);
            RoundState* nxt_s = g.rs + ((r + 1) & 1);
            uint32_t* cur = (r & 1) ? g.fb : g.fa;
            uint32_t* nxt = (r & 1) ? g.fa : g.fb;
            partition_active<<<std::min<size_t>(grid_for(NB, 256), 256u), 256, 0, st>>>(
                cur, cur_s, g.block_key, g.fp, nxt, g.block_tag, round, nxt_s, h.ctx ? h.ctx->geo_delta : kDelta);
            h.launched("partition_active");
            {
                PMA_TIMED(h.ctx, "relax_blocks");
                const unsigned long long src = r == 0 ? (unsigned long long)si : ~0ull;
                if (d3q7)
                    relax_blocks<true><<<kRelaxGrid, 64, 0, st>>>(dist, cav, id, g.fp, nxt, g.block_tag, round,
                                                                  g.block_key, cur_s, nxt_s, g.part_count,
                                                                  g.part_walls, X, Y, Z, NBX, NBY, NBZ, src);
                else
                    relax_blocks<false><<<kRelaxGrid, 64, 0, st>>>(dist, cav, id, g.fp, nxt, g.block_tag, round,
                                                                   g.block_key, cur_s, nxt_s, g.part_count,
                                                                   g.part_walls, X, Y, Z, NBX, NBY, NBZ, src);
                h.launched("relax_blocks");
            }
            next_min_key<<<256, 256, 0, st>>>(nxt, g.block_key, nxt_s, cur_s, g.visits);
            h.launched("next_min_key");
            ++round;
            ++r;
            if (r == check_at) {
                if (h.read(&nxt_s->count) == 0) break;
                step = std::min(step * 2, 8u);
                check_at += step;
            }
        }
        cavity_totals<<<1, 256, 0, st>>>(g.part_count, g.part_walls, kRelaxGrid, g.vcount, g.walls);
        h.launched("cavity_totals");
        CavityRow row;
        row.id = id;
        row.x = start[0];
        row.y = start[1];
        row.z = start[2];
        int d2 = 0;
        for (int a = 0; a < 3; ++a) d2 += (start[a] - ctr[a]) * (start[a] - ctr[a]);
        row.distance_to_center = std::sqrt((double)d2);
        row.void_count = h.read(g.vcount);
        uint32_t wb = plane_bits(h.read(g.walls), S);
        row.reached_walls = __builtin_popcount(wb);
        row.reaches_all = wb == all_walls;
        PMA_REQUIRE(emit(row), PMA_ERR_CAPACITY, "pma_geodesic: more cavities than the caller's table holds");
        if (row.reaches_all) break;
    }
    return h.read(g.visits);
}

void check_geodesic_dims(int X, int Y, int Z) {
    PMA_REQUIRE(X > 0 && Y > 0 && Z > 0, PMA_ERR_INVALID_ARGUMENT, "pma_geodesic: dimensions must be positive");
    PMA_REQUIRE((double)X * Y * Z < 4294967295.0 && X < (1 << 20) && Y < (1 << 20) && Z < (1 << 20),
                PMA_ERR_INVALID_ARGUMENT, "pma_geodesic: volume too large");
}

size_t geodesic_blocks(int X, int Y, int Z) {
    return (size_t)((Z + kB - 1) / kB) * ((X + kB - 1) / kB) * ((Y + kB - 1) / kB);
}

// ---- the allocation-free entry point's scratch ---------------------------------------------
// One array per device for pma_geodesic_noalloc, held until the process ends.  Host code uses
// the per-device address that hipGetSymbolAddress returns, never the array's name.
constexpr size_t kNoAllocScratch = (size_t)PMA_GEODESIC_NOALLOC_SCRATCH_BYTES;
__device__ __align__(256) unsigned char scratch[kNoAllocScratch];

constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Bytes of scratch a run over NB blocks takes, in the order carve() lays it out.
size_t geodesic_scratch_bytes(size_t NB) {
    return 5 * align256(NB * 4) + align256(2 * sizeof(RoundState)) + 3 * align256(8) + align256(4) +
           align256(kRelaxGrid * 8ull) + align256(kRelaxGrid * 4ull);
}

GeoScratch carve(unsigned char* base, size_t NB) {
    GeoScratch g;
    size_t o = 0;
    auto take = [&](size_t bytes) {
        unsigned char* p = base + o;
        o += align256(bytes);
        return p;
    };
    g.fa = (uint32_t*)take(NB * 4);
    g.fb = (uint32_t*)take(NB * 4);
    g.fp = (uint32_t*)take(NB * 4);
    g.block_tag = (uint32_t*)take(NB * 4);
    g.block_key = (uint32_t*)take(NB * 4);
    g.rs = (RoundState*)take(2 * sizeof(RoundState));
    g.visits = (unsigned long long*)take(8);
    g.key = (unsigned long long*)take(8);
    g.vcount = (unsigned long long*)take(8);
    g.walls = (uint32_t*)take(4);
    g.part_count = (unsigned long long*)take(kRelaxGrid * 8ull);
    g.part_walls = (uint32_t*)take(kRelaxGrid * 4ull);
    return g;
}

std::mutex g_scratch_mutex;
void* g_scratch_ptr[64] = {};

// The device address of this device's scratch array (looked up once per device, with the
// device current).
void* device_scratch(int device) {
    PMA_REQUIRE(device >= 0 && device < 64, PMA_ERR_INVALID_ARGUMENT, "pma_geodesic_noalloc: device out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        PMA_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        PMA_REQUIRE(p != nullptr, PMA_ERR_HIP, "pma_geodesic_noalloc: no device address for the scratch array");
        g_scratch_ptr[device] = p;
    }
    return g_scratch_ptr[device];
}

}  // namespace

void run_geodesic(pma_ctx* c) {
    PMA_REQUIRE(c->valid[PMA_VOL_SOLID], PMA_ERR_STATE, "pma_geodesic: the solid volume has not been loaded");
    const int X = c->dims.x, Y = c->dims.y, Z = c->dims.z;
    check_geodesic_dims(X, Y, Z);
    c->ensure_volume(PMA_VOL_SPD);
    c->ensure_volume(PMA_VOL_CAVITY);
    const size_t NB = geodesic_blocks(X, Y, Z);
    GeoScratch g;
    g.fa = c->buf<uint32_t>("geo_active_a", NB);
    g.fb = c->buf<uint32_t>("geo_active_b", NB);
    g.fp = c->buf<uint32_t>("geo_active_proc", NB);
    g.block_tag = c->buf<uint32_t>("geo_block_tag", NB);
    g.block_key = c->buf<uint32_t>("geo_block_key", NB);
    g.rs = c->buf<RoundState>("geo_round_state", 2);
    g.visits = c->buf<unsigned long long>("geo_visits", 1);
    g.key = c->buf<unsigned long long>("geo_key", 1);
    g.vcount = c->buf<unsigned long long>("geo_vcount", 1);
    g.walls = c->buf<uint32_t>("geo_walls", 1);
    g.part_count = c->buf<unsigned long long>("geo_part_count", kRelaxGrid);
    g.part_walls = c->buf<uint32_t>("geo_part_walls", kRelaxGrid);
    GeoHost h{c->stream, c};
    c->cavities.clear();
    const unsigned long long visits =
        geodesic_core(h, c->v<uint8_t>(PMA_VOL_SOLID), c->v<float>(PMA_VOL_SPD), c->v<uint32_t>(PMA_VOL_CAVITY), X, Y,
                      Z, c->stencil == PMA_STENCIL_D3Q7, g, [&](const CavityRow& r) {
                          c->cavities.push_back(r);
                          return true;
                      });
    c->have_cavities = true;
    c->valid[PMA_VOL_SPD] = c->valid[PMA_VOL_CAVITY] = true;
    c->stage[PMA_STAGE_GEODESIC].work_items = visits;
}

void run_geodesic_noalloc(int device, const uint8_t* solid, pma_dims dims, pma_stencil stencil, float* dist,
                          uint32_t* cavity, pma_cavity* rows, size_t cap, size_t* count, hipStream_t stream) {
    PMA_REQUIRE(solid && dist && cavity && count && (rows || cap == 0), PMA_ERR_INVALID_ARGUMENT,
                "pma_geodesic_noalloc: null pointer");
    PMA_REQUIRE(stencil == PMA_STENCIL_D3Q27 || stencil == PMA_STENCIL_D3Q7, PMA_ERR_INVALID_ARGUMENT,
                "pma_geodesic_noalloc: unknown stencil");
    const int X = dims.x, Y = dims.y, Z = dims.z;
    check_geodesic_dims(X, Y, Z);
    const size_t NB = geodesic_blocks(X, Y, Z);
    PMA_REQUIRE(geodesic_scratch_bytes(NB) <= kNoAllocScratch, PMA_ERR_CAPACITY,
                "pma_geodesic_noalloc: the volume needs " + std::to_string(geodesic_scratch_bytes(NB)) +
                    " bytes of scratch, more than the " + std::to_string(kNoAllocScratch) + " it has");
    PMA_HIP(hipSetDevice(device));
    GeoScratch g = carve(static_cast<unsigned char*>(device_scratch(device)), NB);
    GeoHost h{stream, nullptr};
    *count = 0;
    geodesic_core(h, solid, dist, cavity, X, Y, Z, stencil == PMA_STENCIL_D3Q7, g, [&](const CavityRow& r) {
        if (*count >= cap) return false;
        pma_cavity& o = rows[(*count)++];
        o.id = r.id;
        o.x = r.x;
        o.y = r.y;
        o.z = r.z;
        o.distance_to_center = r.distance_to_center;
        o.void_count = r.void_count;
        o.reached_walls = r.reached_walls;
        o.reaches_all = r.reaches_all ? 1 : 0;
        return true;
    });
}

}  // namespace pma
