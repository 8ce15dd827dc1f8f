// Squared Euclidean distance transform (EuclideanDistanceComputer.cpp:149-343).
//
// The reference computes, per void voxel, the exact squared distance to the nearest solid
// voxel or to the virtual solid planes x = -X and x = 2X-1 that its X-pass start value
// X-1 implies (:164, :180).  The same value is produced here by three separable passes:
//   pass X: one lane per (y, z) column, forward + backward scan of the 1-D distance, kept
//           unsquared in 16 bits (at most 2X - 1 < 2^15),
//   pass Y: one workgroup per (x, z) row staged in LDS, exact windowed minimum of
//           g(y') + (y-y')^2 over |y-y'| <= floor(sqrt(g(y))) (the reference's window), found
//           by scanning outward from y until dy^2 >= the best value so far,
//   pass Z: 64 (x, y) columns per workgroup, lanes along y so every load is coalesced, staged
//           in LDS per chunk of z with a halo; the same windowed minimum along z, fused with the
//           maximum and the histogram of squares whose ball does not cross the image boundary
//           (:310-327).
#include "pma_internal.h"

namespace pma {
namespace {

constexpr int kLocalBins = 2048;

__global__ void __launch_bounds__(256) edt_pass_x(const uint8_t* __restrict__ solid,
                                                  uint16_t* __restrict__ g, int X, int Y) {
    int y = blockIdx.x * blockDim.x + threadIdx.x;
    if (y >= Y) return;
    size_t base = (size_t)blockIdx.y * X * Y + y;
    uint32_t d = (uint32_t)(X - 1);
    for (int x = 0; x < X; ++x) {
        size_t i = base + (size_t)x * Y;
        d = solid[i] ? 0u : d + 1u;
        g[i] = (uint16_t)d;
    }
    d = (uint32_t)(X - 1);
    for (int x = X - 1; x >= 0; --x) {
        size_t i = base + (size_t)x * Y;
        d = solid[i] ? 0u : d + 1u;
        g[i] = (uint16_t)min((uint32_t)g[i], d);
    }
}

// Pass X for Y a multiple of 4: a thread takes four adjacent (y, z) columns (4-byte loads,
// 8-byte stores), the same scans as edt_pass_x.
__global__ void __launch_bounds__(256) edt_pass_x4(const uint8_t* __restrict__ solid,
                                                   uint16_t* __restrict__ g, int X, int Y) {
    const int y = (blockIdx.x * blockDim.x + threadIdx.x) * 4;
    if (y >= Y) return;
    const size_t base = (size_t)blockIdx.y * X * Y + y;
    uint4 d = make_uint4(X - 1, X - 1, X - 1, X - 1);
    for (int x = 0; x < X; ++x) {
        const size_t i = base + (size_t)x * Y;
        const uchar4 s = *reinterpret_cast<const uchar4*>(solid + i);
        d = make_uint4(s.x ? 0u : d.x + 1u, s.y ? 0u : d.y + 1u, s.z ? 0u : d.z + 1u, s.w ? 0u : d.w + 1u);
        *reinterpret_cast<ushort4*>(g + i) = make_ushort4(d.x, d.y, d.z, d.w);
    }
    d = make_uint4(X - 1, X - 1, X - 1, X - 1);
    for (int x = X - 1; x >= 0; --x) {
        const size_t i = base + (size_t)x * Y;
        const uchar4 s = *reinterpret_cast<const uchar4*>(solid + i);
        d = make_uint4(s.x ? 0u : d.x + 1u, s.y ? 0u : d.y + 1u, s.z ? 0u : d.z + 1u, s.w ? 0u : d.w + 1u);
        const ushort4 f = *reinterpret_cast<const ushort4*>(g + i);
        *reinterpret_cast<ushort4*>(g + i) = make_ushort4(min((uint32_t)f.x, d.x), min((uint32_t)f.y, d.y),
                                                          min((uint32_t)f.z, d.z), min((uint32_t)f.w, d.w));
    }
}

// Pass Y: one workgroup per (x, z) row, staged in LDS with kHaloY entries of kFarY either side.
// Each lane searches outward from its y over k = 1, 2, ..., taking min(g(y-k), g(y+k)) + k^2,
// until k^2 >= best on every lane of its wave: no farther y' can then do better, so the result
// is the exact minimum over the row, which equals the minimum over the reference's window
// |dy| <= floor(sqrt(g(y))).  k is the same on every lane (k^2 is a constant of the unrolled loop);
// lanes that still need k > kHaloY continue alone with bounds checks.
constexpr int kHaloY = 32;
// Inputs of pass Y are below (2X)^2 <= 2^30 for X <= 16383, and g + k^2 stays below 2^31.
constexpr uint32_t kFarY = 0x40000000u;

__global__ void __launch_bounds__(256) edt_pass_y(const uint16_t* __restrict__ dx,
                                                  uint32_t* __restrict__ h, int X, int Y) {
    extern __shared__ uint32_t line[];  // line[kHaloY + y] = g(y) = dx(y)^2 for 0 <= y < Y, kFarY outside
    const size_t base = ((size_t)blockIdx.y * X + blockIdx.x) * Y;
    for (int i = threadIdx.x; i < Y + 2 * kHaloY; i += blockDim.x) {
        const int y = i - kHaloY;
        const uint32_t d = (y >= 0 && y < Y) ? dx[base + y] : 0u;
        line[i] = (y >= 0 && y < Y) ? d * d : kFarY;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    for (int yw = (int)(threadIdx.x & ~63u); yw < Y; yw += blockDim.x) {  // the same trips on every lane of a wave
        const int y = yw + lane;
        const bool in_y = y < Y;
        const uint32_t* c = &line[kHaloY + (in_y ? y : Y - 1)];  // c[-kHaloY .. kHaloY] lie in line
        uint32_t best = in_y ? c[0] : 0u;
#pragma unroll
        for (int k0 = 1; k0 <= kHaloY; k0 += 4) {
            if (!__any((uint32_t)(k0 * k0) < best)) break;
#pragma unroll
            for (int k = k0; k < k0 + 4; ++k) best = min(best, min(c[-k], c[k]) + (uint32_t)(k * k));
        }
        if (best > (uint32_t)((kHaloY + 1) * (kHaloY + 1))) {
            for (uint32_t k = kHaloY + 1; k * k < best; ++k) {
                const uint32_t a = y >= (int)k ? line[kHaloY + y - (int)k] : kFarY;
                const uint32_t b = y + (int)k < Y ? line[kHaloY + y + (int)k] : kFarY;
                best = min(best, min(a, b) + k * k);
            }
        }
        if (in_y) h[base + y] = best;
    }
}

// Pass Z on chunks of kChunkZ planes: a block stages 64 columns (lanes along y) of its chunk and
// kHaloZ planes either side in LDS (planes outside the volume hold kFarZ, which never wins), and
// each of its 4 waves takes every 4th plane.  The search runs outward from z over k = 1, 2, ...,
// taking min(h(z-k), h(z+k)) + k^2, and stops once k^2 >= best on every lane of the wave: no
// plane at distance k or beyond can then improve best, so the result is the exact minimum over
// the column, which equals the minimum over the reference's window |dz| <= floor(sqrt(h(z))).
// k is the same on every lane (k^2 is a constant of the unrolled loop); lanes that still need
// k > kHaloZ continue alone from global memory.
constexpr int kChunkZ = 64, kHaloZ = 32, kStageZ = kChunkZ + 2 * kHaloZ;
// Inputs of pass Z are below (2X)^2 <= 2^30 for X <= 16383, and h + k^2 stays below 2^31.
constexpr uint32_t kFarZ = 0x40000000u;

__global__ void __launch_bounds__(256) edt_pass_z(const uint32_t* __restrict__ h,
                                                  uint32_t* __restrict__ e, int X, int Y, int Z,
                                                  unsigned long long* __restrict__ hist,
                                                  uint32_t hist_len, uint32_t* __restrict__ max_out) {
    __shared__ uint32_t col[kStageZ][64];
    __shared__ uint32_t lhist[kLocalBins];
    __shared__ uint32_t wave_max[4];
    const int lane = threadIdx.x & 63, sub = threadIdx.x >> 6;
    const int y = blockIdx.x * 64 + lane, x = blockIdx.y;
    const int z0 = blockIdx.z * kChunkZ, z1 = min(z0 + kChunkZ, Z);
    const int s0 = z0 - kHaloZ;  // LDS row r holds plane s0 + r
    const size_t stride = (size_t)X * Y;
    const size_t base = (size_t)x * Y + y;
    const bool in_y = y < Y;
    for (int i = threadIdx.x; i < kLocalBins; i += blockDim.x) lhist[i] = 0;
    for (int r = sub; r < kStageZ; r += 4) {
        const int z = s0 + r;
        col[r][lane] = (in_y && z >= 0 && z < Z) ? h[base + (size_t)z * stride] : kFarZ;
    }
    __syncthreads();
    uint32_t my_max = 0;
    const int mxy = min(min(x, y), min(X - 1 - x, Y - 1 - y));
    // two planes, z and z + 4, per step (their searches interleave); the trip count and `has_b`
    // are the same on every lane of a wave
    for (int z = z0 + sub; z < z1; z += 8) {
        const bool has_b = z + 4 < z1;
        const int ra = z - s0, rb = ra + (has_b ? 4 : 0);
        uint32_t best_a = in_y ? col[ra][lane] : 0u;
        uint32_t best_b = (in_y && has_b) ? col[rb][lane] : 0u;
        const uint32_t* below_a = &col[ra - kHaloZ][lane];  // below_a[(kHaloZ - k) * 64] is plane z - k
        const uint32_t* above_a = &col[ra][lane];           // above_a[k * 64] is plane z + k
        const uint32_t* below_b = &col[rb - kHaloZ][lane];
        const uint32_t* above_b = &col[rb][lane];
#pragma unroll
        for (int k0 = 1; k0 <= kHaloZ; k0 += 4) {
            if (!__any((uint32_t)(k0 * k0) < max(best_a, best_b))) break;
#pragma unroll
            for (int k = k0; k < k0 + 4; ++k) {
                best_a = min(best_a, min(below_a[(kHaloZ - k) * 64], above_a[k * 64]) + (uint32_t)(k * k));
                best_b = min(best_b, min(below_b[(kHaloZ - k) * 64], above_b[k * 64]) + (uint32_t)(k * k));
            }
        }
        auto finish = [&](int zz, uint32_t best) {
            if (best > (uint32_t)((kHaloZ + 1) * (kHaloZ + 1))) {
                for (uint32_t k = kHaloZ + 1; k * k < best; ++k) {
                    const uint32_t a = zz >= (int)k ? h[base + (size_t)(zz - (int)k) * stride] : kFarZ;
                    const uint32_t b = zz + (int)k < Z ? h[base + (size_t)(zz + (int)k) * stride] : kFarZ;
                    best = min(best, min(a, b) + k * k);
                }
            }
            if (best != 0) {
                my_max = max(my_max, best);
                // sqrt(best) > m  <=>  best > m*m for integer m >= 0
                const int m = min(mxy, min(zz, Z - 1 - zz));
                if ((uint64_t)best <= (uint64_t)m * (uint64_t)m) {
                    if (best < kLocalBins) atomicAdd(&lhist[best], 1u);
                    else if (best < hist_len) atomicAdd(&hist[best], 1ull);
                }
            }
            if (in_y) e[base + (size_t)zz * stride] = best;
        };
        finish(z, best_a);
        if (has_b) finish(z + 4, best_b);
    }
    for (int off = 32; off > 0; off >>= 1) my_max = max(my_max, (uint32_t)__shfl_xor((int)my_max, off));
    if (lane == 0) wave_max[sub] = my_max;
    __syncthreads();
    if (threadIdx.x == 0) atomicMax(max_out, max(max(wave_max[0], wave_max[1]), max(wave_max[2], wave_max[3])));
    const int lim = min((uint32_t)kLocalBins, hist_len);
    for (int i = threadIdx.x; i < lim; i += blockDim.x)
        if (lhist[i]) atomicAdd(&hist[i], (unsigned long long)lhist[i]);
}

}  // namespace

void run_edt(pma_ctx* c) {
    PMA_REQUIRE(c->valid[PMA_VOL_SOLID], PMA_ERR_STATE, "pma_edt: the solid volume has not been loaded");
    const int X = c->dims.x, Y = c->dims.y, Z = c->dims.z;
    PMA_REQUIRE(Z <= 65535 && X <= 16383, PMA_ERR_INVALID_ARGUMENT,
                "pma_edt: X must be <= 16383 (squared distances and their sums in 32 bits) and Z <= 65535");
    PMA_REQUIRE((size_t)(Y + 2 * kHaloY) * 4 <= 64 * 1024, PMA_ERR_INVALID_ARGUMENT, "pma_edt: Y must be <= 16320");
    c->ensure_volume(PMA_VOL_EDT2);
    uint32_t* e = c->v<uint32_t>(PMA_VOL_EDT2);
    uint32_t* tmp = c->buf<uint32_t>("edt_tmp", c->n);
    uint16_t* dx = reinterpret_cast<uint16_t*>(e);  // pass X's 16-bit distances, in the first half of e
    const uint8_t* solid = c->v<uint8_t>(PMA_VOL_SOLID);

    // Counted squares satisfy e <= m*m with m <= (min(X,Y,Z)-1)/2.
    int mind = std::min(X, std::min(Y, Z));
    uint32_t mmax = (uint32_t)std::max(0, (mind - 1) / 2);
    uint32_t hist_len = mmax * mmax + 1;
    unsigned long long* hist = c->buf<unsigned long long>("edt_hist", hist_len);
    uint32_t* dmax = c->buf<uint32_t>("edt_max", 1);
    PMA_HIP(hipMemsetAsync(hist, 0, hist_len * sizeof(unsigned long long), c->stream));
    PMA_HIP(hipMemsetAsync(dmax, 0, sizeof(uint32_t), c->stream));

    {

        PMA_TIMED(c, "edt_pass_x");

        if (Y % 4 == 0)
            edt_pass_x4<<<dim3(grid_for(Y / 4, 256), Z), 256, 0, c->stream>>>(solid, dx, X, Y);
        else
            edt_pass_x<<<dim3(grid_for(Y, 256), Z), 256, 0, c->stream>>>(solid, dx, X, Y);

        PMA_LAUNCHED(c);

    }
    {
        PMA_TIMED(c, "edt_pass_y");
        edt_pass_y<<<dim3(X, Z), 256, (size_t)(Y + 2 * kHaloY) * sizeof(uint32_t), c->stream>>>(dx, tmp, X, Y);
        PMA_LAUNCHED(c);
    }
    {
        PMA_TIMED(c, "edt_pass_z");
        edt_pass_z<<<dim3(grid_for(Y, 64), X, grid_for(Z, kChunkZ)), 256, 0, c->stream>>>(tmp, e, X, Y, Z, hist,
                                                                                        hist_len, dmax);
        PMA_LAUNCHED(c);
    }

    uint32_t m = read_scalar(c, dmax);
    c->max_e = m;
    c->n_max_balls = 0;
    c->histogram.assign((size_t)m + 1, 0);
    size_t copy = std::min<size_t>((size_t)m + 1, hist_len);
    PMA_HIP(hipMemcpyAsync(c->histogram.data(), hist, copy * sizeof(uint64_t), hipMemcpyDeviceToHost, c->stream));
    PMA_HIP(hipStreamSynchronize(c->stream));
    c->have_edt_stats = true;
    c->valid[PMA_VOL_EDT2] = true;
    c->stage[PMA_STAGE_EDT].work_items = c->n;
}

}  // namespace pma
