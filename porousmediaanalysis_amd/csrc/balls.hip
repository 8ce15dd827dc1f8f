// Containing balls and the maximum-ball mask (ContainingBallsAssigner.cpp:101-250).
//
// Reference semantics with one thread: for every void centre c with r2 = e(c), every void
// voxel v with |v - c|^2 <= r2 gets CR(v) = max r2; among equal maxima the first centre in
// x, y, z scan order wins (strict '<' at :170).  Both rules are one 64-bit maximum of
//     key(c) = r2 << 32 | ~scan(c),   scan(c) = (x * Y + y) * Z + z,
// so the result does not depend on the order in which balls are applied.
//
// Work plan:
//   1. far tables: far_k(a) = max over lattice p with |p|^2 <= a of |p - d_k|^2 for the face,
//      edge and corner offsets d_k.  A centre c is skipped when a stencil neighbour q has
//      e(q) > e(c) and far_k(e(c)) <= e(q): then B(c) is a subset of B(q) and every voxel of
//      B(c) has CR >= e(q) > e(c), so c can never be a maximiser (exact pruning).
//   2. surviving centres are bucketed by the voxel tile that contains them.
//   3. one workgroup per tile keeps the tile's 64-bit keys in LDS, applies every candidate
//      ball that touches the tile with ds_max_u64 (one wave per ball, lanes along y, rows
//      packed several per instruction for small balls) and writes CR, CX, CY, CZ once.
//   4. every void voxel marks its centre 255 in the mask (SetMaximumBallsMaskInActiveArea).
#include "pma_internal.h"

namespace pma {
namespace {

constexpr int TZ = 8, TX = 16, TY = 64;
constexpr int kTileVoxels = TZ * TX * TY;
constexpr uint8_t kCandidate = 126;  // temporary mask value of a surviving centre

struct Candidate {
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
        int s = a + b + c;
        atomicMax(&shell[p2], p2 + 2u * hi + 1u);
        atomicMax(&shell[L + p2], p2 + 2u * (uint32_t)(s - lo) + 2u);
        atomicMax(&shell[2 * L + p2], p2 + 2u * (uint32_t)s + 3u);
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

__device__ __forceinline__ int tile_of(int x, int y, int z, int NTX, int NTY) {
    return ((z / TZ) * NTX + (x / TX)) * NTY + (y / TY);
}

// Block = 4 rows of 64 voxels along y (one wave per row, inside one tile).
__global__ void __launch_bounds__(256) prune_centres(const uint32_t* __restrict__ e, uint8_t* __restrict__ mask,
                                                     const uint32_t* __restrict__ far, uint32_t L,
                                                     uint32_t* __restrict__ tile_count, int X, int Y, int Z,
                                                     int NTX, int NTY, int d3q7) {
    int y = blockIdx.x * 64 + (threadIdx.x & 63);
    int x = blockIdx.y * 4 + (threadIdx.x >> 6);
    int z = blockIdx.z;
    bool keep = false;
    if (x < X && y < Y) {
        size_t i = ((size_t)z * X + x) * Y + y;
        uint32_t c = e[i];
        if (c == 0) {
            mask[i] = 0;
        } else {
            keep = true;
            const uint32_t* fc = far;
            for (int dz = -1; dz <= 1 && keep; ++dz)
                for (int dx = -1; dx <= 1 && keep; ++dx)
                    for (int dy = -1; dy <= 1; ++dy) {
                        int k = abs(dz) + abs(dx) + abs(dy);
                        if (k == 0) continue;
                        int nx = x + dx, ny = y + dy, nz = z + dz;
                        if (nx < 0 || nx >= X || ny < 0 || ny >= Y || nz < 0 || nz >= Z) continue;
                        uint32_t q = e[((size_t)nz * X + nx) * Y + ny];
                        if (q > c && fc[(size_t)(k - 1) * L + c] <= q) { keep = false; break; }
                    }
            mask[i] = keep ? kCandidate : 127;
        }
    }
    uint64_t bal = __ballot(keep);
    if ((threadIdx.x & 63) == 0 && bal && x < X)
        atomicAdd(&tile_count[tile_of(x, blockIdx.x * 64, z, NTX, NTY)], (uint32_t)__popcll(bal));
    (void)d3q7;
}

__global__ void __launch_bounds__(256) emit_centres(const uint32_t* __restrict__ e, uint8_t* __restrict__ mask,
                                                    const uint32_t* __restrict__ tile_off,
                                                    uint32_t* __restrict__ tile_cursor, Candidate* __restrict__ cand,
                                                    int X, int Y, int NTX, int NTY) {
    int lane = threadIdx.x & 63;
    int y = blockIdx.x * 64 + lane;
    int x = blockIdx.y * 4 + (threadIdx.x >> 6);
    int z = blockIdx.z;
    bool keep = false;
    size_t i = 0;
    if (x < X && y < Y) {
        i = ((size_t)z * X + x) * Y + y;
        keep = mask[i] == kCandidate;
    }
    uint64_t bal = __ballot(keep);
    if (!bal) return;
    int t = x < X ? tile_of(x, blockIdx.x * 64, z, NTX, NTY) : 0;
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(&tile_cursor[t], (uint32_t)__popcll(bal));
    base = __shfl(base, 0);
    if (keep) {
        uint32_t rank = (uint32_t)__popcll(bal & ((1ull << lane) - 1));
        cand[tile_off[t] + base + rank] = Candidate{x, y, z, e[i]};
        mask[i] = 127;
    }
}

__global__ void __launch_bounds__(256) scatter_tiles(const Candidate* __restrict__ cand,
                                                     const uint32_t* __restrict__ tile_off,
                                                     const uint32_t* __restrict__ tile_count,
                                                     const uint32_t* __restrict__ e, uint32_t* __restrict__ cr,
                                                     uint32_t* __restrict__ cx, uint32_t* __restrict__ cy,
                                                     uint32_t* __restrict__ cz, uint8_t* __restrict__ mask,
                                                     int X, int Y, int Z, int NTX, int NTY, int NTZ, int Rmax,
                                                     uint32_t* __restrict__ error_flag) {
    __shared__ unsigned long long key[kTileVoxels];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int t = blockIdx.x;
    const int ty = t % NTY, tx = (t / NTY) % NTX, tz = t / (NTY * NTX);
    const int z0 = tz * TZ, x0 = tx * TX, y0 = ty * TY;
    const int z1 = min(z0 + TZ, Z) - 1, x1 = min(x0 + TX, X) - 1, y1 = min(y0 + TY, Y) - 1;
    for (int k = threadIdx.x; k < kTileVoxels; k += blockDim.x) key[k] = 0ull;
    __syncthreads();

    const int rz = (Rmax + TZ - 1) / TZ, rx = (Rmax + TX - 1) / TX, ry = (Rmax + TY - 1) / TY;
    for (int ntz = max(tz - rz, 0); ntz <= min(tz + rz, NTZ - 1); ++ntz)
        for (int ntx = max(tx - rx, 0); ntx <= min(tx + rx, NTX - 1); ++ntx)
            for (int nty = max(ty - ry, 0); nty <= min(ty + ry, NTY - 1); ++nty) {
                const int nt = (ntz * NTX + ntx) * NTY + nty;
                const uint32_t start = tile_off[nt], cnt = tile_count[nt];
                for (uint32_t k0 = wave * 64; k0 < cnt; k0 += blockDim.x) {
                    Candidate cd{0, 0, 0, 0};
                    bool touch = false;
                    if (k0 + lane < cnt) {
                        cd = cand[start + k0 + lane];
                        int dz = max(max(z0 - cd.z, cd.z - z1), 0);
                        int dx = max(max(x0 - cd.x, cd.x - x1), 0);
                        int dy = max(max(y0 - cd.y, cd.y - y1), 0);
                        touch = (uint32_t)(dz * dz + dx * dx + dy * dy) <= cd.r2;
                    }
                    uint64_t todo = __ballot(touch);
                    while (todo) {
                        int src = __ffsll((long long)todo) - 1;
                        todo &= todo - 1;
                        const int bx = __shfl(cd.x, src), by = __shfl(cd.y, src), bz = __shfl(cd.z, src);
                        const uint32_t r2 = (uint32_t)__shfl((int)cd.r2, src);
                        const int R = (int)isqrt_u32(r2);
                        const uint32_t scan = (uint32_t)(((uint64_t)bx * Y + by) * (uint64_t)Z + bz);
                        const unsigned long long K = ((unsigned long long)r2 << 32) | (uint32_t)~scan;
                        const int zlo = max(bz - R, z0), zhi = min(bz + R, z1);
                        const int xlo = max(bx - R, x0), xhi = min(bx + R, x1);
                        const int nxr = xhi - xlo + 1;
                        const int rows = (zhi - zlo + 1) * nxr;
                        int width = min(2 * R + 1, TY);
                        int P = 1;
                        while (P < width) P <<= 1;
                        const int G = 64 / P, g = lane / P, off = lane % P;
                        for (int r = g; r < rows; r += G) {
                            const int zz = zlo + r / nxr, xx = xlo + r % nxr;
                            const int rem = (int)r2 - (zz - bz) * (zz - bz) - (xx - bx) * (xx - bx);
                            if (rem < 0) continue;
                            const int h = (int)isqrt_u32((uint32_t)rem);
                            const int ylo = max(by - h, y0), yhi = min(by + h, y1);
                            const int yy = ylo + off;
                            if (yy <= yhi)
                                atomicMax(&key[((zz - z0) * TX + (xx - x0)) * TY + (yy - y0)], K);
                        }
                    }
                }
            }
    __syncthreads();
    const uint64_t YZ = (uint64_t)Y * Z;
    for (int k = threadIdx.x; k < kTileVoxels; k += blockDim.x) {
        const int zz = z0 + k / (TX * TY), xx = x0 + (k / TY) % TX, yy = y0 + k % TY;
        if (zz > z1 || xx > x1 || yy > y1) continue;
        const size_t i = ((size_t)zz * X + xx) * Y + yy;
        const unsigned long long K = key[k];
        if (e[i] == 0) {
            cr[i] = 0u;
            cx[i] = cy[i] = cz[i] = 0xFFFFFFFFu;
            continue;
        }
        if (K == 0ull) { atomicOr(error_flag, 1u); continue; }
        const uint64_t scan = (uint32_t)~(uint32_t)K;
        const uint32_t ox = (uint32_t)(scan / YZ), oy = (uint32_t)((scan / Z) % Y), oz = (uint32_t)(scan % Z);
        cr[i] = (uint32_t)(K >> 32);
        cx[i] = ox; cy[i] = oy; cz[i] = oz;
        mask[((size_t)oz * X + ox) * Y + oy] = 255;
    }
}

__global__ void __launch_bounds__(256) count_equal_u8(const uint8_t* __restrict__ v, size_t n, uint8_t value,
                                                      unsigned long long* __restrict__ out) {
    unsigned long long local = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
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
    uint32_t* far = c->buf<uint32_t>("cb_far", 3 * L);
    PMA_HIP(hipMemsetAsync(far, 0, 3 * L * sizeof(uint32_t), c->stream));
    far_shells<<<dim3(grid_for(R + 1, 64), R + 1), 64, 0, c->stream>>>(far, max_e, R);
    PMA_LAUNCHED(c);
    far_prefix_max<<<3, 1024, 0, c->stream>>>(far, L);
    PMA_LAUNCHED(c);

    const int NTZ = (Z + TZ - 1) / TZ, NTX = (X + TX - 1) / TX, NTY = (Y + TY - 1) / TY;
    const size_t NT = (size_t)NTZ * NTX * NTY;
    uint32_t* tile_count = c->buf<uint32_t>("cb_tile_count", NT);
    uint32_t* tile_off = c->buf<uint32_t>("cb_tile_off", NT);
    uint32_t* tile_cursor = c->buf<uint32_t>("cb_tile_cursor", NT);
    PMA_HIP(hipMemsetAsync(tile_count, 0, NT * sizeof(uint32_t), c->stream));
    PMA_HIP(hipMemsetAsync(tile_cursor, 0, NT * sizeof(uint32_t), c->stream));
    const uint32_t* e = c->v<uint32_t>(PMA_VOL_EDT2);
    uint8_t* mask = c->v<uint8_t>(PMA_VOL_MASK);
    dim3 vgrid(grid_for(Y, 64), grid_for(X, 4), Z);
    {
        PMA_TIMED(c, "prune_centres");
        prune_centres<<<vgrid, 256, 0, c->stream>>>(e, mask, far, (uint32_t)L, tile_count, X, Y, Z, NTX, NTY,
                                                     c->stencil == PMA_STENCIL_D3Q7);
        PMA_LAUNCHED(c);
    }
    uint64_t n_cand = exclusive_scan_u32(c, tile_count, tile_off, NT);
    Candidate* cand = c->buf<Candidate>("cb_candidates", std::max<uint64_t>(n_cand, 1));
    {
        PMA_TIMED(c, "emit_centres");
        emit_centres<<<vgrid, 256, 0, c->stream>>>(e, mask, tile_off, tile_cursor, cand, X, Y, NTX, NTY);
        PMA_LAUNCHED(c);
    }
    uint32_t* err = c->buf<uint32_t>("cb_error", 1);
    PMA_HIP(hipMemsetAsync(err, 0, sizeof(uint32_t), c->stream));
    {
        PMA_TIMED(c, "scatter_tiles");
        scatter_tiles<<<(unsigned)NT, 256, 0, c->stream>>>(cand, tile_off, tile_count, e, c->v<uint32_t>(PMA_VOL_CR2),
                                                            c->v<uint32_t>(PMA_VOL_CX), c->v<uint32_t>(PMA_VOL_CY),
                                                            c->v<uint32_t>(PMA_VOL_CZ), mask, X, Y, Z, NTX, NTY, NTZ,
                                                            R, err);
        PMA_LAUNCHED(c);
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
    c->stage[PMA_STAGE_CONTAINING_BALLS].work_items = n_cand;
    for (pma_volume w : {PMA_VOL_CR2, PMA_VOL_CX, PMA_VOL_CY, PMA_VOL_CZ, PMA_VOL_MASK}) c->valid[w] = true;
}

}  // namespace pma
