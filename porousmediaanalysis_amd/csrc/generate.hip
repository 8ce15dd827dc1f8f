// Seeded synthetic inputs for tests and benchmarks (0 = void, 255 = solid).
//   spheres : union of n balls, centres uniform in the box, radii uniform in [rmin, rmax];
//             a voxel (integer coordinates) is solid when it lies in any ball.
//   gaussian: i.i.d. N(0,1) noise from a counter-based hash (SplitMix64 + Box-Muller),
//             separable Gaussian blur (radius int(4 sigma + 0.5), reflected boundary
//             d c b a | a b c d | d c b a), void = the lowest `porosity` fraction of values
//             (exact order statistic by radix selection on the float bit patterns).
#include "pma_internal.h"

#include <cmath>

namespace pma {
namespace {

__host__ __device__ inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

struct Sphere {
    float x, y, z, r2;
};

__global__ void __launch_bounds__(256) spheres_kernel(uint8_t* __restrict__ out, const Sphere* __restrict__ s, int n,
                                                      int X, int Y, int Z) {
    extern __shared__ Sphere sh[];
    for (int k = threadIdx.x; k < n; k += blockDim.x) sh[k] = s[k];
    __syncthreads();
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t N = (size_t)X * Y * Z;
    if (i >= N) return;
    int y = (int)(i % Y);
    size_t r = i / Y;
    int x = (int)(r % X), z = (int)(r / X);
    bool solid = false;
    for (int k = 0; k < n && !solid; ++k) {
        float dx = x - sh[k].x, dy = y - sh[k].y, dz = z - sh[k].z;
        solid = dx * dx + dy * dy + dz * dz <= sh[k].r2;
    }
    out[i] = solid ? 255 : 0;
}

__global__ void __launch_bounds__(256) noise_kernel(float* __restrict__ f, size_t n, uint64_t seed) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t h = splitmix64(seed ^ splitmix64(i));
    float u1 = ((float)(h >> 40) + 1.0f) * (1.0f / 16777216.0f);  // (0, 1]
    float u2 = (float)((h >> 16) & 0xFFFFFF) * (1.0f / 16777216.0f);
    f[i] = sqrtf(-2.0f * logf(u1)) * cosf(6.28318530718f * u2);
}

__device__ __forceinline__ int reflect(int i, int n) {
    // half-sample symmetric extension, valid for any i
    int period = 2 * n;
    i %= period;
    if (i < 0) i += period;
    return i < n ? i : period - 1 - i;
}

// Blur along an axis with element stride `stride` and length `len`; lanes run along y.
__global__ void __launch_bounds__(256) blur_axis(const float* __restrict__ in, float* __restrict__ out,
                                                 const float* __restrict__ w, int radius, int X, int Y, int Z,
                                                 int axis) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t N = (size_t)X * Y * Z;
    if (i >= N) return;
    int y = (int)(i % Y);
    size_t r = i / Y;
    int x = (int)(r % X), z = (int)(r / X);
    float acc = 0.0f;
    for (int k = -radius; k <= radius; ++k) {
        size_t j;
        if (axis == 0) j = ((size_t)z * X + reflect(x + k, X)) * Y + y;
        else if (axis == 1) j = ((size_t)z * X + x) * Y + reflect(y + k, Y);
        else j = ((size_t)reflect(z + k, Z) * X + x) * Y + y;
        acc += w[k + radius] * in[j];
    }
    out[i] = acc;
}

__device__ __forceinline__ uint32_t float_key(float v) {
    uint32_t b = __float_as_uint(v);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__global__ void __launch_bounds__(256) radix_hist(const float* __restrict__ f, size_t n, uint32_t prefix,
                                                  uint32_t prefix_mask, int shift, unsigned long long* __restrict__ hist) {
    __shared__ unsigned int lh[256];
    lh[threadIdx.x] = 0;
    __syncthreads();
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        uint32_t k = float_key(f[i]);
        if ((k & prefix_mask) == prefix) atomicAdd(&lh[(k >> shift) & 0xFF], 1u);
    }
    __syncthreads();
    if (lh[threadIdx.x]) atomicAdd(&hist[threadIdx.x], (unsigned long long)lh[threadIdx.x]);
}

__global__ void __launch_bounds__(256) threshold_kernel(const float* __restrict__ f, uint8_t* __restrict__ out,
                                                        size_t n, uint32_t kth) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = float_key(f[i]) <= kth ? 0 : 255;
}

}  // namespace

void run_generate_spheres(pma_ctx* c, int32_t n, float rmin, float rmax, uint64_t seed) {
    PMA_REQUIRE(c->n > 0, PMA_ERR_STATE, "pma_generate_spheres: set the dimensions first");
    PMA_REQUIRE(n >= 0 && n <= 8192 && rmin >= 0 && rmax >= rmin, PMA_ERR_INVALID_ARGUMENT,
                "pma_generate_spheres: need 0 <= n <= 8192 and 0 <= rmin <= rmax");
    std::vector<Sphere> s(n);
    uint64_t st = seed;
    auto uni = [&]() {
        st = splitmix64(st);
        return (double)(st >> 11) * (1.0 / 9007199254740992.0);
    };
    for (int k = 0; k < n; ++k) {
        s[k].x = (float)(uni() * c->dims.x);
        s[k].y = (float)(uni() * c->dims.y);
        s[k].z = (float)(uni() * c->dims.z);
        float r = (float)(rmin + uni() * (rmax - rmin));
        s[k].r2 = r * r;
    }
    c->ensure_volume(PMA_VOL_SOLID);
    Sphere* d = c->buf<Sphere>("gen_spheres", std::max(n, 1));
    upload(c, d, s.data(), n * sizeof(Sphere));
    spheres_kernel<<<grid_for(c->n, 256), 256, std::max(n, 1) * sizeof(Sphere), c->stream>>>(
        c->v<uint8_t>(PMA_VOL_SOLID), d, n, c->dims.x, c->dims.y, c->dims.z);
    PMA_LAUNCHED(c);
    PMA_HIP(hipStreamSynchronize(c->stream));
    c->valid[PMA_VOL_SOLID] = true;
}

void run_generate_gaussian(pma_ctx* c, float sigma, float porosity, uint64_t seed) {
    PMA_REQUIRE(c->n > 0, PMA_ERR_STATE, "pma_generate_gaussian: set the dimensions first");
    PMA_REQUIRE(sigma > 0 && sigma < 256 && porosity >= 0 && porosity <= 1, PMA_ERR_INVALID_ARGUMENT,
                "pma_generate_gaussian: need 0 < sigma < 256 and 0 <= porosity <= 1");
    const size_t n = c->n;
    float* a = c->buf<float>("gen_field_a", n);
    float* b = c->buf<float>("gen_field_b", n);
    int radius = (int)(4.0 * sigma + 0.5);
    std::vector<float> w(2 * radius + 1);
    double sum = 0;
    for (int k = -radius; k <= radius; ++k) sum += (w[k + radius] = (float)std::exp(-0.5 * k * k / ((double)sigma * sigma)));
    for (float& v : w) v = (float)(v / sum);
    float* dw = c->buf<float>("gen_weights", w.size());
    upload(c, dw, w.data(), w.size() * sizeof(float));
    noise_kernel<<<grid_for(n, 256), 256, 0, c->stream>>>(a, n, seed);
    PMA_LAUNCHED(c);
    blur_axis<<<grid_for(n, 256), 256, 0, c->stream>>>(a, b, dw, radius, c->dims.x, c->dims.y, c->dims.z, 1);
    PMA_LAUNCHED(c);
    blur_axis<<<grid_for(n, 256), 256, 0, c->stream>>>(b, a, dw, radius, c->dims.x, c->dims.y, c->dims.z, 0);
    PMA_LAUNCHED(c);
    blur_axis<<<grid_for(n, 256), 256, 0, c->stream>>>(a, b, dw, radius, c->dims.x, c->dims.y, c->dims.z, 2);
    PMA_LAUNCHED(c);
    // k-th smallest value (0-based) with k = ceil(porosity * n) - 1: byte-wise radix select.
    uint64_t k = (uint64_t)std::ceil((double)porosity * (double)n);
    c->ensure_volume(PMA_VOL_SOLID);
    uint8_t* out = c->v<uint8_t>(PMA_VOL_SOLID);
    if (k == 0) {
        PMA_HIP(hipMemsetAsync(out, 255, n, c->stream));
    } else {
        k -= 1;
        unsigned long long* hist = c->buf<unsigned long long>("gen_hist", 256);
        uint32_t prefix = 0, pmask = 0;
        std::vector<unsigned long long> h(256);
        for (int shift = 24; shift >= 0; shift -= 8) {
            PMA_HIP(hipMemsetAsync(hist, 0, 256 * sizeof(unsigned long long), c->stream));
            radix_hist<<<1024, 256, 0, c->stream>>>(b, n, prefix, pmask, shift, hist);
            PMA_LAUNCHED(c);
            PMA_HIP(hipMemcpyAsync(h.data(), hist, 256 * sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream));
            PMA_HIP(hipStreamSynchronize(c->stream));
            uint32_t digit = 0;
            for (; digit < 256; ++digit) {
                if (k < h[digit]) break;
                k -= h[digit];
            }
            prefix |= digit << shift;
            pmask |= 0xFFu << shift;
        }
        threshold_kernel<<<grid_for(n, 256), 256, 0, c->stream>>>(b, out, n, prefix);
        PMA_LAUNCHED(c);
    }
    PMA_HIP(hipStreamSynchronize(c->stream));
    c->valid[PMA_VOL_SOLID] = true;
}

}  // namespace pma
