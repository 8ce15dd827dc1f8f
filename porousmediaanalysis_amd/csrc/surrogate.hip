// Learned screening of subvolumes: a small 3-D U-Net that predicts, per void voxel, the
// containing-ball radius and the ratio of the geodesic to the Euclidean distance from the voxel's
// cavity start.  No counterpart in the reference.  The exact pipeline makes the training labels
// and stays the judge: the network's fields are estimates for screening many subvolumes, and never
// replace a volume or table the pipeline writes.
//
// Network (channels in brackets, every convolution 3x3x3 with zero padding, ReLU):
//   input [2: void indicator, sqrt(EDT square) / 16, padded to 16]
//   e1: conv 16 -> 16, conv 16 -> 16                  (skip s1)   full resolution
//   avg-pool 2 ; e2: conv 16 -> 32, conv 32 -> 32     (skip s2)   1/2
//   avg-pool 2 ; b:  conv 32 -> 64, conv 64 -> 64                 1/4
//   upsample 2 ; concat s2 [96] ; d2: conv 96 -> 32, conv 32 -> 32
//   upsample 2 ; concat s1 [48] ; d1: conv 48 -> 16, conv 16 -> 16
//   out: conv 1x1x1 16 -> 2 (no ReLU): radius / 16 and ratio - 1.
// About 0.36 M parameters.
//
// Layout: activations channels-last, [z][x][y][c] (the volumes' [z][x][y] order with the channel
// innermost), bf16; weights [tap][out][in] bf16 with fp32 master copies; accumulation in fp32.
// A convolution is an implicit GEMM on v_mfma_f32_16x16x16bf16_1k: M = 16 voxels along y,
// N = 16 output channels, K = 16 input channels of one tap.
#include "pma_internal.h"

#include <algorithm>
#include <cmath>
#include <functional>
#include <random>
#include <vector>

namespace pma {
namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

struct Grid3 {
    int D, H, W;  // z, x, y sizes; y fastest
};

// Y = act(conv(X, Wt) + bias).  X [D][H][W][cin] bf16, Wt [taps][cout][cin] bf16, cin and cout
// multiples of 16 (cout <= 96), taps 27 (3x3x3) or 1.  One wave: 16 voxels of a y-row segment
// times every output channel; four waves per workgroup.  Output bf16 (ybf) or fp32 (yf).
__global__ void __launch_bounds__(256) conv3d(const uint16_t* __restrict__ X, const uint16_t* __restrict__ Wt,
                                              const float* __restrict__ bias, uint16_t* __restrict__ ybf,
                                              float* __restrict__ yf, Grid3 g, int cin, int cout, int taps, int relu) {
    const int lane = threadIdx.x & 63;
    const int segs = (g.W + 15) / 16;
    const long tile = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (tile >= (long)g.D * g.H * segs) return;  // the whole wave leaves together
    const int ys = (int)(tile % segs) * 16;
    const long r = tile / segs;
    const int x = (int)(r % g.H), z = (int)(r / g.H);
    const int m = lane & 15, kq = lane >> 4;
    const int nsub = cout / 16;
    f32x4 acc[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) acc[j] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int t = 0; t < taps; ++t) {
        const int dz = taps == 27 ? t / 9 - 1 : 0, dx = taps == 27 ? (t / 3) % 3 - 1 : 0,
                  dy = taps == 27 ? t % 3 - 1 : 0;
        const int zz = z + dz, xx = x + dx, yy = ys + m + dy;
        const bool in = zz >= 0 && zz < g.D && xx >= 0 && xx < g.H && yy >= 0 && yy < g.W;
        const uint16_t* xrow = X + (in ? (((long)zz * g.H + xx) * g.W + yy) * cin : 0);
        const uint16_t* wt = Wt + (long)t * cout * cin;
        for (int c0 = 0; c0 < cin; c0 += 16) {
            bf16x4 a = bf16x4{0, 0, 0, 0};
            if (in) a = *reinterpret_cast<const bf16x4*>(xrow + c0 + 4 * kq);
#pragma unroll
            for (int j = 0; j < 6; ++j) {
                if (j < nsub) {
                    const bf16x4 b = *reinterpret_cast<const bf16x4*>(wt + (long)(16 * j + m) * cin + c0 + 4 * kq);
                    acc[j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[j], 0, 0, 0);
                }
            }
        }
    }
    // lane holds rows 4 kq + i (voxels) of column m (output channel) of each 16x16 tile
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        if (j >= nsub) break;
        const int co = 16 * j + m;
        const float bv = bias ? bias[co] : 0.f;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int yv = ys + 4 * kq + i;
            if (yv >= g.W) continue;
            float v = acc[j][i] + bv;
            if (relu) v = fmaxf(v, 0.f);
            const long o = (((long)z * g.H + x) * g.W + yv) * cout + co;
            if (ybf) ybf[o] = to_bf16(v);
            else yf[o] = v;
        }
    }
}

// The 3x3x3 convolution with the input staged in LDS: a workgroup computes 4 x-rows of 16 voxels
// along y at one z (one row per wave, every output channel), and first copies the 3 x 6 x 18
// voxels those outputs read (zero outside the volume) into LDS, each voxel's channels padded by 4
// to spread the banks.  The 27 taps then read LDS instead of global memory.
constexpr int kTileX = 4, kTileY = 16;
constexpr int kHaloX = kTileX + 2, kHaloY = kTileY + 2;
__global__ void __launch_bounds__(256) conv3d_lds(const uint16_t* __restrict__ X, const uint16_t* __restrict__ Wt,
                                                  const float* __restrict__ bias, uint16_t* __restrict__ ybf,
                                                  float* __restrict__ yf, Grid3 g, int cin, int cout, int relu) {
    extern __shared__ uint16_t halo[];  // [3][kHaloX][kHaloY][cin + 4]
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int segs = (g.W + kTileY - 1) / kTileY, xrows = (g.H + kTileX - 1) / kTileX;
    const long tile = blockIdx.x;
    const int ys = (int)(tile % segs) * kTileY;
    const long r = tile / segs;
    const int x0 = (int)(r % xrows) * kTileX, z = (int)(r / xrows);
    const int cs = cin + 4;  // LDS stride of a voxel
    // stage: chunks of 4 channels (8 bytes), zero outside the volume
    const int cq = cin / 4, nchunks = 3 * kHaloX * kHaloY * cq;
    for (int q = threadIdx.x; q < nchunks; q += 256) {
        const int c4 = q % cq;
        const int vox = q / cq;
        const int hy = vox % kHaloY, hx = (vox / kHaloY) % kHaloX, hz = vox / (kHaloY * kHaloX);
        const int zz = z + hz - 1, xx = x0 + hx - 1, yy = ys + hy - 1;
        bf16x4 v = bf16x4{0, 0, 0, 0};
        if (zz >= 0 && zz < g.D && xx >= 0 && xx < g.H && yy >= 0 && yy < g.W)
            v = *reinterpret_cast<const bf16x4*>(X + (((long)zz * g.H + xx) * g.W + yy) * cin + 4 * c4);
        *reinterpret_cast<bf16x4*>(halo + vox * cs + 4 * c4) = v;
    }
    __syncthreads();
    const int x = x0 + wave;
    const int m = lane & 15, kq = lane >> 4;
    const int nsub = cout / 16;
    f32x4 acc[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) acc[j] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int t = 0; t < 27; ++t) {
        const int dz = t / 9, dx = (t / 3) % 3, dy = t % 3;  // halo coordinates of the tap (offset + 1)
        const uint16_t* arow = halo + ((dz * kHaloX + wave + dx) * kHaloY + m + dy) * cs + 4 * kq;
        const uint16_t* wt = Wt + (long)t * cout * cin;
        for (int c0 = 0; c0 < cin; c0 += 16) {
            const bf16x4 a = *reinterpret_cast<const bf16x4*>(arow + c0);
#pragma unroll
            for (int j = 0; j < 6; ++j) {
                if (j < nsub) {
                    const bf16x4 b = *reinterpret_cast<const bf16x4*>(wt + (long)(16 * j + m) * cin + c0 + 4 * kq);
                    acc[j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[j], 0, 0, 0);
                }
            }
        }
    }
    if (x >= g.H) return;  // after the barrier: a row past the volume computed zeros
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        if (j >= nsub) break;
        const int co = 16 * j + m;
        const float bv = bias ? bias[co] : 0.f;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int yv = ys + 4 * kq + i;
            if (yv >= g.W) continue;
            float v = acc[j][i] + bv;
            if (relu) v = fmaxf(v, 0.f);
            const long o = (((long)z * g.H + x) * g.W + yv) * cout + co;
            if (ybf) ybf[o] = to_bf16(v);
            else yf[o] = v;
        }
    }
}

// dW[t][co][ci] += sum over voxels v of dP[v][co] * X[v + offset(t)][ci]: an MFMA GEMM with
// M = 16 output channels, N = 16 input channels, K = 16 voxels, over a chunk of voxels per
// workgroup (one wave per 16 output channels), added to dW with atomics.
constexpr int kWgradChunk = 1024;
__global__ void __launch_bounds__(256) conv3d_wgrad(const uint16_t* __restrict__ X, const uint16_t* __restrict__ dP,
                                                    float* __restrict__ dW, Grid3 g, int cin, int cout, int taps) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int t = blockIdx.y;
    if (wave >= cout / 16) return;
    const long V = (long)g.D * g.H * g.W;
    const long v0 = (long)blockIdx.x * kWgradChunk;
    const int dz = taps == 27 ? t / 9 - 1 : 0, dx = taps == 27 ? (t / 3) % 3 - 1 : 0, dy = taps == 27 ? t % 3 - 1 : 0;
    const int m = lane & 15, kq = lane >> 4;
    const int co = 16 * wave + m;
    const int nci = cin / 16;
    f32x4 acc[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) acc[j] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (long k0 = v0; k0 < v0 + kWgradChunk && k0 < V; k0 += 16) {
        bf16x4 a;
        long src[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const long v = k0 + 4 * kq + i;
            a[i] = 0;
            src[i] = -1;
            if (v < V) {
                a[i] = (short)dP[v * cout + co];
                const int y = (int)(v % g.W);
                const long r = v / g.W;
                const int x = (int)(r % g.H), z = (int)(r / g.H);
                const int zz = z + dz, xx = x + dx, yy = y + dy;
                if (zz >= 0 && zz < g.D && xx >= 0 && xx < g.H && yy >= 0 && yy < g.W)
                    src[i] = ((long)zz * g.H + xx) * g.W + yy;
            }
        }
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            if (j < nci) {
                bf16x4 b;
#pragma unroll
                for (int i = 0; i < 4; ++i) b[i] = src[i] >= 0 ? (short)X[src[i] * cin + 16 * j + m] : (short)0;
                acc[j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[j], 0, 0, 0);
            }
        }
    }
    // lane holds rows 4 kq + i (output channels of the wave's 16) of column m (input channel)
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        if (j >= nci) break;
#pragma unroll
        for (int i = 0; i < 4; ++i)
            atomicAdd(&dW[((long)t * cout + 16 * wave + 4 * kq + i) * cin + 16 * j + m], acc[j][i]);
    }
}

// db[c] += sum over voxels of dP[v][c]: per-thread sums over a grid stride, summed over the
// workgroup's rows in LDS, one atomic per channel per workgroup.  blockDim.x = 256, a multiple of C.
__global__ void __launch_bounds__(256) bias_grad(const uint16_t* __restrict__ dP, long V, int C, float* __restrict__ db) {
    __shared__ float part[256];
    const int c = threadIdx.x % C;
    const int rows = blockDim.x / C;
    float s = 0.f;
    for (long v = (long)blockIdx.x * rows + threadIdx.x / C; v < V; v += (long)gridDim.x * rows)
        s += from_bf16(dP[v * C + c]);
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x < C) {
        float t = 0.f;
        for (int r = 0; r < rows; ++r) t += part[r * C + threadIdx.x];
        atomicAdd(&db[threadIdx.x], t);
    }
}

// Sum of a value over a 256-thread workgroup (every thread calls it); valid in thread 0.
__device__ __forceinline__ float block_sum(float v, float* lds4) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
    if ((threadIdx.x & 63) == 0) lds4[threadIdx.x >> 6] = v;
    __syncthreads();
    const float t = lds4[0] + lds4[1] + lds4[2] + lds4[3];
    __syncthreads();
    return t;
}

// Backward-data weights: Wf[26 - t][ci][co] = W[t][co][ci] (taps mirrored, in and out swapped).
__global__ void flip_weights(const uint16_t* __restrict__ W, uint16_t* __restrict__ Wf, int cin, int cout, int taps) {
    const long n = (long)taps * cout * cin;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const int ci = (int)(i % cin);
        const long r = i / cin;
        const int co = (int)(r % cout), t = (int)(r / cout);
        Wf[((long)(taps - 1 - t) * cin + ci) * cout + co] = W[i];
    }
}

// dP = bf16(dY * (y > 0)) (ReLU backward), or bf16(dY) when y is null.
__global__ void relu_back(const float* __restrict__ dY, const uint16_t* __restrict__ y, uint16_t* __restrict__ dP,
                          long n) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        dP[i] = to_bf16(y && y[i] == 0 ? 0.f : dY[i]);
}

// 2x2x2 average pooling, bf16; g = the fine grid (even sizes).
__global__ void pool_fwd(const uint16_t* __restrict__ X, uint16_t* __restrict__ Y, Grid3 g, int C) {
    const long n = (long)(g.D / 2) * (g.H / 2) * (g.W / 2) * C;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const int c = (int)(i % C);
        long r = i / C;
        const int y = (int)(r % (g.W / 2));
        r /= g.W / 2;
        const int x = (int)(r % (g.H / 2)), z = (int)(r / (g.H / 2));
        float s = 0.f;
        for (int k = 0; k < 8; ++k)
            s += from_bf16(X[(((long)(2 * z + (k >> 2)) * g.H + 2 * x + ((k >> 1) & 1)) * g.W + 2 * y + (k & 1)) * C + c]);
        Y[i] = to_bf16(s * 0.125f);
    }
}
// dX[fine] = dY[coarse] / 8, fp32; written, not added.
__global__ void pool_back(const float* __restrict__ dY, float* __restrict__ dX, Grid3 g, int C) {
    const long n = (long)g.D * g.H * g.W * C;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const int c = (int)(i % C);
        long r = i / C;
        const int y = (int)(r % g.W);
        r /= g.W;
        const int x = (int)(r % g.H), z = (int)(r / g.H);
        dX[i] = 0.125f * dY[(((long)(z / 2) * (g.H / 2) + x / 2) * (g.W / 2) + y / 2) * C + c];
    }
}
// Nearest 2x upsampling of the coarse X (g = fine grid) into channels [0, C) of the fine Y
// (Ct channels), and the skip S into channels [C, Ct).
__global__ void up_concat(const uint16_t* __restrict__ X, const uint16_t* __restrict__ S, uint16_t* __restrict__ Y,
                          Grid3 g, int C, int Cs) {
    const int Ct = C + Cs;
    const long n = (long)g.D * g.H * g.W * Ct;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const int c = (int)(i % Ct);
        const long v = i / Ct;
        if (c >= C) {
            Y[i] = S[v * Cs + (c - C)];
            continue;
        }
        const int y = (int)(v % g.W);
        const long r = v / g.W;
        const int x = (int)(r % g.H), z = (int)(r / g.H);
        Y[i] = X[(((long)(z / 2) * (g.H / 2) + x / 2) * (g.W / 2) + y / 2) * C + c];
    }
}
// Backward of up_concat: dX[coarse] = sum of the 8 children's first C channels; dS[fine] += the
// remaining channels (the skip also feeds the encoder path).
__global__ void up_concat_back(const float* __restrict__ dY, float* __restrict__ dX, float* __restrict__ dS, Grid3 g,
                               int C, int Cs) {
    const int Ct = C + Cs;
    const long nc = (long)(g.D / 2) * (g.H / 2) * (g.W / 2) * C;
    const long nf = (long)g.D * g.H * g.W * Cs;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < nc + nf; i += (long)gridDim.x * blockDim.x) {
        if (i < nc) {
            const int c = (int)(i % C);
            long r = i / C;
            const int y = (int)(r % (g.W / 2));
            r /= g.W / 2;
            const int x = (int)(r % (g.H / 2)), z = (int)(r / (g.H / 2));
            float s = 0.f;
            for (int k = 0; k < 8; ++k)
                s += dY[(((long)(2 * z + (k >> 2)) * g.H + 2 * x + ((k >> 1) & 1)) * g.W + 2 * y + (k & 1)) * Ct + c];
            dX[i] = s;
        } else {
            const long j = i - nc;
            const int c = (int)(j % Cs);
            const long v = j / Cs;
            dS[j] += dY[v * Ct + C + c];
        }
    }
}
// dA += dB (fp32).
__global__ void add_into(float* __restrict__ a, const float* __restrict__ b, long n) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) a[i] += b[i];
}

// Network input from the volumes: channel 0 = void, channel 1 = sqrt(EDT square) / 16.
__global__ void make_input(const uint8_t* __restrict__ solid, const uint32_t* __restrict__ e, uint16_t* __restrict__ X,
                           long n) {
    for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (long)gridDim.x * blockDim.x) {
#pragma unroll
        for (int c = 0; c < 16; ++c) X[v * 16 + c] = 0;
        if (solid[v] == 0) {
            X[v * 16] = to_bf16(1.f);
            X[v * 16 + 1] = to_bf16(sqrtf((float)e[v]) * 0.0625f);
        }
    }
}

// Targets: channel 0 = sqrt(CR square) / 16 on void voxels; channel 1 = geodesic / Euclidean
// distance from the voxel's cavity start, minus 1, on reached void voxels other than the start.
// Weights 1 where a target is defined, else 0.
__global__ void make_targets(const uint8_t* __restrict__ solid, const uint32_t* __restrict__ cr,
                             const float* __restrict__ dist, const uint32_t* __restrict__ cavity,
                             const int* __restrict__ starts, uint32_t ncav, Grid3 g, float* __restrict__ tgt,
                             float* __restrict__ wgt) {
    const long n = (long)g.D * g.H * g.W;
    for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (long)gridDim.x * blockDim.x) {
        tgt[2 * v] = tgt[2 * v + 1] = 0.f;
        wgt[2 * v] = wgt[2 * v + 1] = 0.f;
        if (solid[v] != 0) continue;
        tgt[2 * v] = sqrtf((float)cr[v]) * 0.0625f;
        wgt[2 * v] = 1.f;
        const uint32_t id = cavity[v];
        const float d = dist[v];
        if (id == 0 || id % 127u != 0 || id / 127u > ncav || !(d < 3.0e38f)) continue;
        const int* s = starts + 3 * (id / 127u - 1);
        const int y = (int)(v % g.W);
        const long r = v / g.W;
        const int x = (int)(r % g.H), z = (int)(r / g.H);
        const float ex = (float)(x - s[0]), ey = (float)(y - s[1]), ez = (float)(z - s[2]);
        const float eu = sqrtf(ex * ex + ey * ey + ez * ez);
        if (eu <= 0.f) continue;
        tgt[2 * v + 1] = d / eu - 1.f;
        wgt[2 * v + 1] = 1.f;
    }
}

// Mean squared error over the weighted entries of the network's two output channels (out has 16
// channels, 2 used); writes the gradient dOut (16 channels, 2 used) and adds to sums[0] (squared
// error) and sums[1] (weight).
__global__ void __launch_bounds__(256) mse_loss(const float* __restrict__ out, const float* __restrict__ tgt, const float* __restrict__ wgt,
                         long n, float inv_count, float* __restrict__ dout, float* __restrict__ sums) {
    float se = 0.f, sw = 0.f;
    for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (long)gridDim.x * blockDim.x) {
#pragma unroll
        for (int c = 0; c < 16; ++c) {
            float gval = 0.f;
            if (c < 2) {
                const float d = (out[v * 16 + c] - tgt[2 * v + c]) * wgt[2 * v + c];
                se += d * d;
                sw += wgt[2 * v + c];
                gval = 2.f * d * inv_count;
            }
            dout[v * 16 + c] = gval;
        }
    }
    __shared__ float lds4[4];
    se = block_sum(se, lds4);
    sw = block_sum(sw, lds4);
    if (threadIdx.x == 0) {
        atomicAdd(&sums[0], se);
        atomicAdd(&sums[1], sw);
    }
}

// Weighted count of the targets (for the mean).
__global__ void __launch_bounds__(256) weight_sum(const float* __restrict__ wgt, long n, float* __restrict__ s) {
    __shared__ float lds4[4];
    float a = 0.f;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) a += wgt[i];
    a = block_sum(a, lds4);
    if (threadIdx.x == 0) atomicAdd(s, a);
}

// Adam on the fp32 master weights, then their bf16 copies.
__global__ void adam(float* __restrict__ w, const float* __restrict__ gr, float* __restrict__ m1, float* __restrict__ m2,
                     uint16_t* __restrict__ wbf, long n, float lr, float b1, float b2, float c1, float c2, float eps) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const float gi = gr[i];
        const float a = b1 * m1[i] + (1.f - b1) * gi, b = b2 * m2[i] + (1.f - b2) * gi * gi;
        m1[i] = a;
        m2[i] = b;
        const float wi = w[i] - lr * (a * c1) / (sqrtf(b * c2) + eps);
        w[i] = wi;
        if (wbf) wbf[i] = to_bf16(wi);
    }
}
__global__ void to_bf16_kernel(const float* __restrict__ a, uint16_t* __restrict__ b, long n) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        b[i] = to_bf16(a[i]);
}

// Predicted fields from the output: radius (voxels) and ratio, 0 on solid voxels.
__global__ void write_prediction(const float* __restrict__ out, const uint8_t* __restrict__ solid, long n,
                                 float* __restrict__ radius, float* __restrict__ ratio) {
    for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (long)gridDim.x * blockDim.x) {
        const bool s = solid[v] != 0;
        radius[v] = s ? 0.f : fmaxf(out[v * 16], 0.f) * 16.f;
        ratio[v] = s ? 0.f : out[v * 16 + 1] + 1.f;
    }
}

unsigned blocks_for(long n) { return (unsigned)std::min<long>((n + 255) / 256, 16384); }

}  // namespace
}  // namespace pma

// ---- host side ---------------------------------------------------------------------------
struct pma_surrogate {
    struct Layer {
        int cin, cout, taps, cin_real, cout_real;
        size_t w_off, b_off;  // offsets in the parameter vector (weights then bias)
    };
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<Layer> layers;
    size_t nparam = 0;
    pma::DeviceBuffer w32, g32, m1, m2, wbf, wflip, sums;
    long step = 0;
    // activations and gradients, by name, sized for the last grid
    std::map<std::string, pma::DeviceBuffer> act;
    template <class T> T* buf(const std::string& name, size_t count) {
        pma::DeviceBuffer& b = act[name];
        b.reserve(count * sizeof(T), stream);
        return b.as<T>();
    }
};

namespace {
using pma::Grid3;

const int kLayerSpec[][4] = {  // cin, cout, taps, cin_real
    {16, 16, 27, 2},  {16, 16, 27, 16}, {16, 32, 27, 16}, {32, 32, 27, 32}, {32, 64, 27, 32}, {64, 64, 27, 64},
    {96, 32, 27, 96}, {32, 32, 27, 32}, {48, 16, 27, 48}, {16, 16, 27, 16}, {16, 16, 1, 16}};
constexpr int kLayers = 11;

// One convolution: 3x3x3 ones through the LDS-staged kernel, the 1x1x1 one directly.
void launch_conv(pma_surrogate* s, const uint16_t* X, const uint16_t* W, const float* b, uint16_t* ybf, float* yf,
                 Grid3 g, int cin, int cout, int taps, bool relu) {
    if (taps == 27) {
        const long tiles = (long)g.D * ((g.H + pma::kTileX - 1) / pma::kTileX) * ((g.W + pma::kTileY - 1) / pma::kTileY);
        const size_t lds = (size_t)3 * pma::kHaloX * pma::kHaloY * (cin + 4) * sizeof(uint16_t);
        hipLaunchKernelGGL(pma::conv3d_lds, dim3((unsigned)tiles), dim3(256), lds, s->stream, X, W, b, ybf, yf, g, cin,
                           cout, relu ? 1 : 0);
    } else {
        const long tiles = (long)g.D * g.H * ((g.W + 15) / 16);
        hipLaunchKernelGGL(pma::conv3d, dim3((unsigned)((tiles + 3) / 4)), dim3(256), 0, s->stream, X, W, b, ybf, yf,
                           g, cin, cout, taps, relu ? 1 : 0);
    }
    PMA_HIP(hipGetLastError());
}

void conv(pma_surrogate* s, int li, const uint16_t* X, Grid3 g, uint16_t* ybf, float* yf, bool relu) {
    const pma_surrogate::Layer& L = s->layers[li];
    launch_conv(s, X, s->wbf.as<uint16_t>() + L.w_off, s->w32.as<float>() + L.b_off, ybf, yf, g, L.cin, L.cout, L.taps,
                relu);
}

// Gradients of layer li: dW, db from (X, dP); dX = conv(dP, flipped W) when wanted (fp32).
void conv_back(pma_surrogate* s, int li, const uint16_t* X, const uint16_t* dP, Grid3 g, float* dX) {
    const pma_surrogate::Layer& L = s->layers[li];
    const long V = (long)g.D * g.H * g.W;
    float* dW = s->g32.as<float>() + L.w_off;
    float* db = s->g32.as<float>() + L.b_off;
    hipLaunchKernelGGL(pma::conv3d_wgrad, dim3((unsigned)((V + pma::kWgradChunk - 1) / pma::kWgradChunk), L.taps),
                       dim3(256), 0, s->stream, X, dP, dW, g, L.cin, L.cout, L.taps);
    PMA_HIP(hipGetLastError());
    const int rows = 256 / L.cout;  // cout is 16, 32 or 64
    hipLaunchKernelGGL(pma::bias_grad, dim3((unsigned)std::min<long>((V + rows - 1) / rows, 1024)), dim3(256), 0,
                       s->stream, dP, V, L.cout, db);
    PMA_HIP(hipGetLastError());
    if (!dX) return;
    uint16_t* Wf = s->wflip.as<uint16_t>();
    const long nw = (long)L.taps * L.cout * L.cin;
    hipLaunchKernelGGL(pma::flip_weights, dim3(pma::blocks_for(nw)), dim3(256), 0, s->stream,
                       s->wbf.as<uint16_t>() + L.w_off, Wf, L.cin, L.cout, L.taps);
    PMA_HIP(hipGetLastError());
    launch_conv(s, dP, Wf, nullptr, nullptr, dX, g, L.cout, L.cin, L.taps, false);
}

template <class K, class... A> void launch1(pma_surrogate* s, long n, K kernel, A... args) {
    hipLaunchKernelGGL(kernel, dim3(pma::blocks_for(n)), dim3(256), 0, s->stream, args...);
    PMA_HIP(hipGetLastError());
}

Grid3 grid_of(pma_ctx* c) { return Grid3{c->dims.z, c->dims.x, c->dims.y}; }

void check_grid(pma_ctx* c) {
    PMA_REQUIRE(c && c->n > 0, PMA_ERR_STATE, "the context has no volume");
    PMA_REQUIRE(c->dims.x % 4 == 0 && c->dims.y % 4 == 0 && c->dims.z % 4 == 0, PMA_ERR_INVALID_ARGUMENT,
                "the surrogate needs every side a multiple of 4 (two 2x poolings)");
    PMA_REQUIRE(c->valid[PMA_VOL_SOLID] && c->valid[PMA_VOL_EDT2], PMA_ERR_STATE,
                "the surrogate needs the solid volume and its EDT (run pma_edt first)");
}

// Forward pass; leaves the activations in s->act and the output in "out" (fp32, 16 channels).
void forward(pma_surrogate* s, pma_ctx* c) {
    const Grid3 g1 = grid_of(c), g2{g1.D / 2, g1.H / 2, g1.W / 2}, g3{g2.D / 2, g2.H / 2, g2.W / 2};
    const long V1 = (long)g1.D * g1.H * g1.W, V2 = (long)g2.D * g2.H * g2.W, V3 = (long)g3.D * g3.H * g3.W;
    uint16_t* x0 = s->buf<uint16_t>("x0", V1 * 16);
    launch1(s, V1, pma::make_input, (const uint8_t*)c->v<uint8_t>(PMA_VOL_SOLID),
            (const uint32_t*)c->v<uint32_t>(PMA_VOL_EDT2), x0, V1);
    uint16_t* e1a = s->buf<uint16_t>("e1a", V1 * 16);
    uint16_t* e1b = s->buf<uint16_t>("e1b", V1 * 16);
    conv(s, 0, x0, g1, e1a, nullptr, true);
    conv(s, 1, e1a, g1, e1b, nullptr, true);
    uint16_t* p1 = s->buf<uint16_t>("p1", V2 * 16);
    launch1(s, V2 * 16, pma::pool_fwd, (const uint16_t*)e1b, p1, g1, 16);
    uint16_t* e2a = s->buf<uint16_t>("e2a", V2 * 32);
    uint16_t* e2b = s->buf<uint16_t>("e2b", V2 * 32);
    conv(s, 2, p1, g2, e2a, nullptr, true);
    conv(s, 3, e2a, g2, e2b, nullptr, true);
    uint16_t* p2 = s->buf<uint16_t>("p2", V3 * 32);
    launch1(s, V3 * 32, pma::pool_fwd, (const uint16_t*)e2b, p2, g2, 32);
    uint16_t* ba = s->buf<uint16_t>("ba", V3 * 64);
    uint16_t* bb = s->buf<uint16_t>("bb", V3 * 64);
    conv(s, 4, p2, g3, ba, nullptr, true);
    conv(s, 5, ba, g3, bb, nullptr, true);
    uint16_t* u2 = s->buf<uint16_t>("u2", V2 * 96);
    launch1(s, V2 * 96, pma::up_concat, (const uint16_t*)bb, (const uint16_t*)e2b, u2, g2, 64, 32);
    uint16_t* d2a = s->buf<uint16_t>("d2a", V2 * 32);
    uint16_t* d2b = s->buf<uint16_t>("d2b", V2 * 32);
    conv(s, 6, u2, g2, d2a, nullptr, true);
    conv(s, 7, d2a, g2, d2b, nullptr, true);
    uint16_t* u1 = s->buf<uint16_t>("u1", V1 * 48);
    launch1(s, V1 * 48, pma::up_concat, (const uint16_t*)d2b, (const uint16_t*)e1b, u1, g1, 32, 16);
    uint16_t* d1a = s->buf<uint16_t>("d1a", V1 * 16);
    uint16_t* d1b = s->buf<uint16_t>("d1b", V1 * 16);
    conv(s, 8, u1, g1, d1a, nullptr, true);
    conv(s, 9, d1a, g1, d1b, nullptr, true);
    float* out = s->buf<float>("out", V1 * 16);
    conv(s, 10, d1b, g1, nullptr, out, false);
}

// Backward pass from dOut ("dout", fp32, 16 channels): every parameter gradient into g32.
void backward(pma_surrogate* s, pma_ctx* c) {
    const Grid3 g1 = grid_of(c), g2{g1.D / 2, g1.H / 2, g1.W / 2}, g3{g2.D / 2, g2.H / 2, g2.W / 2};
    const long V1 = (long)g1.D * g1.H * g1.W, V2 = (long)g2.D * g2.H * g2.W, V3 = (long)g3.D * g3.H * g3.W;
    auto A = [&](const char* n) { return s->act[n].as<uint16_t>(); };
    PMA_HIP(hipMemsetAsync(s->g32.ptr, 0, s->nparam * 4, s->stream));
    uint16_t* dp = s->buf<uint16_t>("dp", V1 * 16);  // bf16 gradient of a layer's pre-activation
    float* fa = s->buf<float>("fa", V1 * 48);        // fp32 gradient buffers
    float* fb = s->buf<float>("fb", V1 * 48);
    float* ds1 = s->buf<float>("ds1", V1 * 16);       // gradient reaching skip s1 (e1b) from the decoder
    float* ds2 = s->buf<float>("ds2", V2 * 32);
    // out (no ReLU): dP = dOut
    launch1(s, V1 * 16, pma::relu_back, (const float*)s->act["dout"].as<float>(), (const uint16_t*)nullptr, dp, V1 * 16);
    conv_back(s, 10, A("d1b"), dp, g1, fa);  // fa = d d1b
    launch1(s, V1 * 16, pma::relu_back, (const float*)fa, (const uint16_t*)A("d1b"), dp, V1 * 16);
    conv_back(s, 9, A("d1a"), dp, g1, fb);  // fb = d d1a
    launch1(s, V1 * 16, pma::relu_back, (const float*)fb, (const uint16_t*)A("d1a"), dp, V1 * 16);
    conv_back(s, 8, A("u1"), dp, g1, fa);  // fa = d u1 (48 channels)
    PMA_HIP(hipMemsetAsync(ds1, 0, V1 * 16 * 4, s->stream));
    float* dd2b = s->buf<float>("dd2b", V2 * 32);
    launch1(s, V2 * 32 + V1 * 16, pma::up_concat_back, (const float*)fa, dd2b, ds1, g1, 32, 16);
    launch1(s, V2 * 32, pma::relu_back, (const float*)dd2b, (const uint16_t*)A("d2b"), dp, V2 * 32);
    conv_back(s, 7, A("d2a"), dp, g2, fb);  // fb = d d2a
    launch1(s, V2 * 32, pma::relu_back, (const float*)fb, (const uint16_t*)A("d2a"), dp, V2 * 32);
    conv_back(s, 6, A("u2"), dp, g2, fa);  // fa = d u2 (96 channels)
    PMA_HIP(hipMemsetAsync(ds2, 0, V2 * 32 * 4, s->stream));
    float* dbb = s->buf<float>("dbb", V3 * 64);
    launch1(s, V3 * 64 + V2 * 32, pma::up_concat_back, (const float*)fa, dbb, ds2, g2, 64, 32);
    launch1(s, V3 * 64, pma::relu_back, (const float*)dbb, (const uint16_t*)A("bb"), dp, V3 * 64);
    conv_back(s, 5, A("ba"), dp, g3, fb);  // fb = d ba
    launch1(s, V3 * 64, pma::relu_back, (const float*)fb, (const uint16_t*)A("ba"), dp, V3 * 64);
    conv_back(s, 4, A("p2"), dp, g3, fa);  // fa = d p2 (32 channels, coarse)
    launch1(s, V2 * 32, pma::pool_back, (const float*)fa, fb, g2, 32);  // fb = d e2b from the pool
    launch1(s, V2 * 32, pma::add_into, fb, (const float*)ds2, V2 * 32);
    launch1(s, V2 * 32, pma::relu_back, (const float*)fb, (const uint16_t*)A("e2b"), dp, V2 * 32);
    conv_back(s, 3, A("e2a"), dp, g2, fa);  // fa = d e2a
    launch1(s, V2 * 32, pma::relu_back, (const float*)fa, (const uint16_t*)A("e2a"), dp, V2 * 32);
    conv_back(s, 2, A("p1"), dp, g2, fb);  // fb = d p1
    launch1(s, V1 * 16, pma::pool_back, (const float*)fb, fa, g1, 16);  // fa = d e1b from the pool
    launch1(s, V1 * 16, pma::add_into, fa, (const float*)ds1, V1 * 16);
    launch1(s, V1 * 16, pma::relu_back, (const float*)fa, (const uint16_t*)A("e1b"), dp, V1 * 16);
    conv_back(s, 1, A("e1a"), dp, g1, fb);  // fb = d e1a
    launch1(s, V1 * 16, pma::relu_back, (const float*)fb, (const uint16_t*)A("e1a"), dp, V1 * 16);
    conv_back(s, 0, A("x0"), dp, g1, nullptr);
    (void)V3;
}

// Loss and dOut against the exact fields of context c; returns the mean squared error.
float loss_and_grad(pma_surrogate* s, pma_ctx* c) {
    PMA_REQUIRE(c->valid[PMA_VOL_CR2] && c->valid[PMA_VOL_SPD] && c->valid[PMA_VOL_CAVITY] && c->have_cavities,
                PMA_ERR_STATE, "training needs the exact containing radii and geodesic distances of the volume");
    const Grid3 g = grid_of(c);
    const long V = (long)g.D * g.H * g.W;
    std::vector<int> starts;
    for (const pma::CavityRow& r : c->cavities) {
        starts.push_back(r.x);
        starts.push_back(r.y);
        starts.push_back(r.z);
    }
    starts.resize(std::max<size_t>(starts.size(), 3));
    int* d_starts = s->buf<int>("starts", starts.size());
    PMA_HIP(hipMemcpyAsync(d_starts, starts.data(), starts.size() * 4, hipMemcpyHostToDevice, s->stream));
    float* tgt = s->buf<float>("tgt", V * 2);
    float* wgt = s->buf<float>("wgt", V * 2);
    launch1(s, V, pma::make_targets, (const uint8_t*)c->v<uint8_t>(PMA_VOL_SOLID),
            (const uint32_t*)c->v<uint32_t>(PMA_VOL_CR2), (const float*)c->v<float>(PMA_VOL_SPD),
            (const uint32_t*)c->v<uint32_t>(PMA_VOL_CAVITY), (const int*)d_starts, (uint32_t)c->cavities.size(), g,
            tgt, wgt);
    float* sums = s->sums.as<float>();
    PMA_HIP(hipMemsetAsync(sums, 0, 16, s->stream));
    launch1(s, V * 2, pma::weight_sum, (const float*)wgt, V * 2, sums + 2);
    float wsum = 0.f;
    PMA_HIP(hipMemcpyAsync(&wsum, sums + 2, 4, hipMemcpyDeviceToHost, s->stream));
    PMA_HIP(hipStreamSynchronize(s->stream));
    float* dout = s->buf<float>("dout", V * 16);
    launch1(s, V, pma::mse_loss, (const float*)s->act["out"].as<float>(), (const float*)tgt, (const float*)wgt, V,
            wsum > 0 ? 1.f / wsum : 0.f, dout, sums);
    float h[2] = {0.f, 0.f};
    PMA_HIP(hipMemcpyAsync(h, sums, 8, hipMemcpyDeviceToHost, s->stream));
    PMA_HIP(hipStreamSynchronize(s->stream));
    return h[1] > 0 ? h[0] / h[1] : 0.f;
}

pma_status guarded(const std::function<void()>& f) {
    try {
        f();
        pma::set_last_error("");
        return PMA_OK;
    } catch (const pma::Error& e) {
        pma::set_last_error(e.what());
        return e.status;
    } catch (const std::exception& e) {
        pma::set_last_error(e.what());
        return PMA_ERR_INTERNAL;
    }
}

}  // namespace

extern "C" {

pma_status pma_surrogate_create(pma_surrogate** out, int device, uint64_t seed) {
    return guarded([&] {
        PMA_REQUIRE(out != nullptr, PMA_ERR_INVALID_ARGUMENT, "null output pointer");
        *out = nullptr;
        PMA_HIP(hipSetDevice(device));
        auto* s = new pma_surrogate();
        s->device = device;
        try {
            PMA_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
            size_t off = 0;
            for (int i = 0; i < kLayers; ++i) {
                pma_surrogate::Layer L{kLayerSpec[i][0], kLayerSpec[i][1], kLayerSpec[i][2], kLayerSpec[i][3],
                                       i == kLayers - 1 ? 2 : kLayerSpec[i][1], 0, 0};
                L.w_off = off;
                off += (size_t)L.taps * L.cout * L.cin;
                L.b_off = off;
                off += L.cout;
                s->layers.push_back(L);
            }
            s->nparam = off;
            // He-normal weights on the real channels, zero on the padded ones; zero biases.
            std::vector<float> w(off, 0.f);
            std::mt19937_64 rng(seed);
            for (const auto& L : s->layers) {
                std::normal_distribution<float> nd(0.f, std::sqrt(2.f / (float)(L.taps * L.cin_real)));
                for (int t = 0; t < L.taps; ++t)
                    for (int co = 0; co < L.cout_real; ++co)
                        for (int ci = 0; ci < L.cin_real; ++ci) w[L.w_off + ((size_t)t * L.cout + co) * L.cin + ci] = nd(rng);
            }
            s->w32.reserve(off * 4, s->stream);
            s->g32.reserve(off * 4, s->stream);
            s->m1.reserve(off * 4, s->stream);
            s->m2.reserve(off * 4, s->stream);
            s->wbf.reserve(off * 2, s->stream);
            s->wflip.reserve((size_t)27 * 96 * 64 * 2, s->stream);
            s->sums.reserve(64, s->stream);
            PMA_HIP(hipMemcpyAsync(s->w32.ptr, w.data(), off * 4, hipMemcpyHostToDevice, s->stream));
            PMA_HIP(hipMemsetAsync(s->m1.ptr, 0, off * 4, s->stream));
            PMA_HIP(hipMemsetAsync(s->m2.ptr, 0, off * 4, s->stream));
            launch1(s, (long)off, pma::to_bf16_kernel, (const float*)s->w32.as<float>(), s->wbf.as<uint16_t>(), (long)off);
            PMA_HIP(hipStreamSynchronize(s->stream));
        } catch (...) {
            pma_surrogate_destroy(s);
            throw;
        }
        *out = s;
    });
}

void pma_surrogate_destroy(pma_surrogate* s) {
    if (!s) return;
    (void)hipSetDevice(s->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (auto& kv : s->act) kv.second.release();
    for (pma::DeviceBuffer* b : {&s->w32, &s->g32, &s->m1, &s->m2, &s->wbf, &s->wflip, &s->sums}) b->release();
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

pma_status pma_surrogate_param_count(pma_surrogate* s, size_t* count) {
    return guarded([&] {
        PMA_REQUIRE(s && count, PMA_ERR_INVALID_ARGUMENT, "null argument");
        *count = s->nparam;
    });
}

pma_status pma_surrogate_params(pma_surrogate* s, int which, float* dst, size_t count) {
    return guarded([&] {
        PMA_REQUIRE(s && dst && count == s->nparam && (which == 0 || which == 1), PMA_ERR_INVALID_ARGUMENT,
                    "pma_surrogate_params: which is 0 (weights) or 1 (gradients), count the parameter count");
        PMA_HIP(hipSetDevice(s->device));
        PMA_HIP(hipMemcpyAsync(dst, which == 0 ? s->w32.ptr : s->g32.ptr, count * 4, hipMemcpyDeviceToHost, s->stream));
        PMA_HIP(hipStreamSynchronize(s->stream));
    });
}

pma_status pma_surrogate_set_params(pma_surrogate* s, const float* src, size_t count) {
    return guarded([&] {
        PMA_REQUIRE(s && src && count == s->nparam, PMA_ERR_INVALID_ARGUMENT, "count must be the parameter count");
        PMA_HIP(hipSetDevice(s->device));
        PMA_HIP(hipMemcpyAsync(s->w32.ptr, src, count * 4, hipMemcpyHostToDevice, s->stream));
        launch1(s, (long)count, pma::to_bf16_kernel, (const float*)s->w32.as<float>(), s->wbf.as<uint16_t>(), (long)count);
        PMA_HIP(hipStreamSynchronize(s->stream));
    });
}

pma_status pma_surrogate_train_step(pma_surrogate* s, pma_ctx* c, float learning_rate, int update, float* loss) {
    return guarded([&] {
        PMA_REQUIRE(s != nullptr, PMA_ERR_INVALID_ARGUMENT, "null surrogate");
        check_grid(c);
        PMA_REQUIRE(c->device == s->device, PMA_ERR_INVALID_ARGUMENT, "the context is on another device");
        PMA_HIP(hipSetDevice(s->device));
        PMA_HIP(hipStreamSynchronize(c->stream));  // the context's volumes are complete
        forward(s, c);
        const float l = loss_and_grad(s, c);
        backward(s, c);
        if (update) {
            ++s->step;
            const float b1 = 0.9f, b2 = 0.999f;
            const float c1 = 1.f / (1.f - std::pow(b1, (float)s->step)), c2 = 1.f / (1.f - std::pow(b2, (float)s->step));
            launch1(s, (long)s->nparam, pma::adam, s->w32.as<float>(), (const float*)s->g32.as<float>(),
                    s->m1.as<float>(), s->m2.as<float>(), s->wbf.as<uint16_t>(), (long)s->nparam, learning_rate, b1, b2,
                    c1, c2, 1e-8f);
        }
        PMA_HIP(hipStreamSynchronize(s->stream));
        if (loss) *loss = l;
    });
}

pma_status pma_surrogate_predict(pma_surrogate* s, pma_ctx* c, float* radius, float* ratio, int dst_on_device) {
    return guarded([&] {
        PMA_REQUIRE(s && radius && ratio, PMA_ERR_INVALID_ARGUMENT, "null argument");
        check_grid(c);
        PMA_REQUIRE(c->device == s->device, PMA_ERR_INVALID_ARGUMENT, "the context is on another device");
        PMA_HIP(hipSetDevice(s->device));
        PMA_HIP(hipStreamSynchronize(c->stream));
        forward(s, c);
        const long V = (long)c->n;
        float* r = dst_on_device ? radius : s->buf<float>("pred_radius", V);
        float* q = dst_on_device ? ratio : s->buf<float>("pred_ratio", V);
        launch1(s, V, pma::write_prediction, (const float*)s->act["out"].as<float>(),
                (const uint8_t*)c->v<uint8_t>(PMA_VOL_SOLID), V, r, q);
        if (!dst_on_device) {
            PMA_HIP(hipMemcpyAsync(radius, r, V * 4, hipMemcpyDeviceToHost, s->stream));
            PMA_HIP(hipMemcpyAsync(ratio, q, V * 4, hipMemcpyDeviceToHost, s->stream));
        }
        PMA_HIP(hipStreamSynchronize(s->stream));
    });
}

}  // extern "C"
