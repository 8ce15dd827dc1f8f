// Device-wide exclusive prefix sums (reduce, scan of block totals, downsweep).
#include "pma_internal.h"

namespace pma {
namespace {

constexpr int kBlock = 256;
constexpr int kItems = 4;
constexpr int kTile = kBlock * kItems;

template <class T>
__device__ T block_exclusive_scan(T v, T* total) {
    __shared__ T wave_sum[kBlock / 64];
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    T incl = v;
    for (int off = 1; off < 64; off <<= 1) {
        T o = __shfl_up(incl, off);
        if (lane >= off) incl += o;
    }
    if (lane == 63) wave_sum[wave] = incl;
    __syncthreads();
    T prefix = 0, all = 0;
    for (int w = 0; w < kBlock / 64; ++w) {
        if (w < wave) prefix += wave_sum[w];
        all += wave_sum[w];
    }
    __syncthreads();
    *total = all;
    return prefix + incl - v;
}

template <class T>
__global__ void __launch_bounds__(kBlock) scan_reduce(const T* __restrict__ in, T* __restrict__ sums, size_t n) {
    size_t base = (size_t)blockIdx.x * kTile + threadIdx.x * kItems;
    T s = 0;
    for (int k = 0; k < kItems; ++k)
        if (base + k < n) s += in[base + k];
    T total;
    block_exclusive_scan<T>(s, &total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

template <class T>
__global__ void __launch_bounds__(kBlock) scan_down(const T* in, T* out, const T* __restrict__ offsets, size_t n) {
    size_t base = (size_t)blockIdx.x * kTile + threadIdx.x * kItems;
    T v[kItems];
    T s = 0;
    for (int k = 0; k < kItems; ++k) {
        v[k] = base + k < n ? in[base + k] : T(0);
        s += v[k];
    }
    T total;
    T run = block_exclusive_scan<T>(s, &total) + (offsets ? offsets[blockIdx.x] : T(0));
    for (int k = 0; k < kItems; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
}

template <class T>
void scan_impl(pma_ctx* c, const T* in, T* out, size_t n, int level) {
    if (n == 0) return;
    size_t blocks = (n + kTile - 1) / kTile;
    if (blocks == 1) {
        scan_down<T><<<1, kBlock, 0, c->stream>>>(in, out, nullptr, n);
        PMA_LAUNCHED(c);
        return;
    }
    T* sums = c->buf<T>("scan_sums_" + std::to_string(level) + (sizeof(T) == 8 ? "_64" : "_32"), blocks);
    scan_reduce<T><<<(unsigned)blocks, kBlock, 0, c->stream>>>(in, sums, n);
    PMA_LAUNCHED(c);
    scan_impl<T>(c, sums, sums, blocks, level + 1);
    scan_down<T><<<(unsigned)blocks, kBlock, 0, c->stream>>>(in, out, sums, n);
    PMA_LAUNCHED(c);
}

}  // namespace

uint64_t exclusive_scan_u32(pma_ctx* c, const uint32_t* in, uint32_t* out, size_t n) {
    if (n == 0) return 0;
    uint32_t last_in = read_scalar(c, in + n - 1);
    scan_impl<uint32_t>(c, in, out, n, 0);
    uint32_t last_out = read_scalar(c, out + n - 1);
    return (uint64_t)last_out + last_in;
}

uint64_t exclusive_scan_u64(pma_ctx* c, const uint64_t* in, uint64_t* out, size_t n) {
    if (n == 0) return 0;
    uint64_t last_in = read_scalar(c, in + n - 1);
    scan_impl<unsigned long long>(c, reinterpret_cast<const unsigned long long*>(in),
                                  reinterpret_cast<unsigned long long*>(out), n, 0);
    return read_scalar(c, out + n - 1) + last_in;
}

}  // namespace pma
