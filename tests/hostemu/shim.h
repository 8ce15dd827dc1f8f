// Host emulation of the HIP kernels of this library (test tooling only, never shipped).
// The library's .hip sources are translated (translate.py) into one host C++ file over this
// header.  Ordinary kernels run thread by thread (or, with HOSTEMU_PAR set, their blocks on 8
// OS threads); wave kernels (those using __syncthreads, __shfl or __ballot) run one fiber per
// GPU thread, switched round-robin at every barrier, so the cross-lane operations see exactly
// the values of the other lanes.  Device helpers that the library builds on rocPRIM are
// replaced by std::stable_sort; scans and the libstdc++-sort emulation run their real kernels.
#pragma once
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <atomic>
#include <barrier>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <set>
#include <mutex>
#include <numeric>
#include <stdexcept>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(x)
#define __shared__ static
struct Dim3 {
    unsigned x = 0, y = 0, z = 0;
    Dim3() {}
    Dim3(unsigned a, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
typedef Dim3 dim3;
inline thread_local Dim3 threadIdx, blockIdx;
inline Dim3 blockDim, gridDim;
using hipStream_t = void*;
using std::isfinite;
struct uint2 { uint32_t x, y; };
struct alignas(16) uint4 { uint32_t x, y, z, w; };
struct alignas(16) float4 { float x, y, z, w; };
struct alignas(4) uchar4 { uint8_t x, y, z, w; };
struct alignas(8) ushort4 { uint16_t x, y, z, w; };
inline ushort4 make_ushort4(uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
    return ushort4{(uint16_t)a, (uint16_t)b, (uint16_t)c, (uint16_t)d};
}
inline uint4 make_uint4(uint32_t a, uint32_t b, uint32_t c, uint32_t d) { return uint4{a, b, c, d}; }
inline float4 make_float4(float a, float b, float c, float d) { return float4{a, b, c, d}; }
inline uint2 make_uint2(uint32_t a, uint32_t b) { return uint2{a, b}; }
#define hipSetDevice(d) 0

// ---- fibers: one per GPU thread of a wave-kernel block, switched round-robin at barriers ----
extern "C" void shim_fiber_switch(void** from_sp, void* to_sp);  // shim.cpp
#include <sys/mman.h>
#include <functional>
#if defined(__SANITIZE_ADDRESS__)
#include <sanitizer/common_interface_defs.h>
#define SHIM_ASAN 1
#else
#define SHIM_ASAN 0
#endif
struct Fiber {
    void* sp = nullptr;
    char* stack = nullptr;
    size_t size = 0;
    bool done = false;
    unsigned tid = 0;
};
struct WaveCtx {
    std::vector<uint64_t> slot;
    std::vector<Fiber> fibers;
    void* sched_sp = nullptr;
    int cur = -1;
    std::function<void()>* body = nullptr;
    Dim3 block;
    unsigned arrived = 0, generation = 0;  // block barrier
};
inline WaveCtx* g_wave = nullptr;
inline void shim_yield() {
    WaveCtx* w = g_wave;
    Fiber& f = w->fibers[w->cur];
#if SHIM_ASAN
    void* fake = nullptr;
    __sanitizer_start_switch_fiber(&fake, nullptr, 0);
#endif
    shim_fiber_switch(&f.sp, w->sched_sp);
#if SHIM_ASAN
    __sanitizer_finish_switch_fiber(fake, nullptr, nullptr);
#endif
    threadIdx = Dim3(w->fibers[w->cur].tid, 0, 0);
    blockIdx = w->block;
}
extern "C" inline void shim_fiber_entry() {
#if SHIM_ASAN
    __sanitizer_finish_switch_fiber(nullptr, nullptr, nullptr);
#endif
    WaveCtx* w = g_wave;
    threadIdx = Dim3(w->fibers[w->cur].tid, 0, 0);
    blockIdx = w->block;
    (*w->body)();
    w->fibers[w->cur].done = true;
    for (;;) shim_yield();
}
// A block barrier: a fiber waits (yielding) until every fiber of the block that has not ended
// has arrived.  The fiber that completes the barrier also yields once, so after a barrier the
// fibers of a wave that trail by one switch are always lower-numbered than those ahead of them,
// which keeps the slot exchange of the cross-lane operations in order.
inline void __syncthreads() {
    WaveCtx* w = g_wave;
    if (!w) return;
    const unsigned gen = w->generation;
    ++w->arrived;
    for (;;) {
        if (w->generation != gen) return;
        unsigned live = 0;
        for (const Fiber& f : w->fibers) live += !f.done;
        if (w->arrived >= live) {
            w->arrived = 0;
            ++w->generation;
        }
        shim_yield();
    }
}
template <class T> T wave_read(T v, int from_lane_abs) {
    uint64_t bits = 0;
    std::memcpy(&bits, &v, sizeof(T));
    g_wave->slot[threadIdx.x] = bits;
    shim_yield();
    uint64_t r = g_wave->slot[from_lane_abs];
    shim_yield();
    T out;
    std::memcpy(&out, &r, sizeof(T));
    return out;
}
template <class T> T __shfl(T v, int src) {
    if (!g_wave) throw std::runtime_error("__shfl outside a wave launch");
    return wave_read(v, (int)(threadIdx.x & ~63u) + (src & 63));
}
inline int __builtin_amdgcn_readlane(int v, int lane) {
    if (!g_wave) throw std::runtime_error("readlane outside a wave launch");
    return wave_read(v, (int)(threadIdx.x & ~63u) + (lane & 63));
}
template <class T> T __shfl_xor(T v, int mask) {
    if (!g_wave) throw std::runtime_error("__shfl_xor outside a wave launch");
    return wave_read(v, (int)(threadIdx.x & ~63u) + (((int)threadIdx.x & 63) ^ (mask & 63)));
}
template <class T> T __shfl_up(T v, unsigned d) {
    if (!g_wave) throw std::runtime_error("__shfl_up outside a wave launch");
    const int lane = threadIdx.x & 63;
    const int src = lane >= (int)d ? (int)threadIdx.x - (int)d : (int)threadIdx.x;
    return wave_read(v, src);
}
// Lanes whose thread has ended take no part in the wave operations below (on the GPU they are
// masked off), so their last slot values are skipped.
inline unsigned long long __ballot(int pred) {
    if (!g_wave) throw std::runtime_error("__ballot outside a wave launch");
    g_wave->slot[threadIdx.x] = pred ? 1 : 0;
    shim_yield();
    unsigned long long r = 0;
    const unsigned base = threadIdx.x & ~63u;
    for (unsigned l = 0; l < 64 && base + l < g_wave->slot.size(); ++l)
        if (!g_wave->fibers[base + l].done && g_wave->slot[base + l]) r |= 1ull << l;
    shim_yield();
    return r;
}
inline int __any(int pred) { return __ballot(pred) != 0; }
inline unsigned __ockl_wfred_or_u32(unsigned v) {
    if (!g_wave) throw std::runtime_error("__ockl_wfred_or_u32 outside a wave launch");
    g_wave->slot[threadIdx.x] = v;
    shim_yield();
    unsigned r = 0;
    const unsigned base = threadIdx.x & ~63u;
    for (unsigned l = 0; l < 64 && base + l < g_wave->slot.size(); ++l)
        if (!g_wave->fibers[base + l].done) r |= (unsigned)g_wave->slot[base + l];
    shim_yield();
    return r;
}
inline unsigned __ockl_wfred_add_u32(unsigned v) {
    if (!g_wave) throw std::runtime_error("__ockl_wfred_add_u32 outside a wave launch");
    g_wave->slot[threadIdx.x] = v;
    shim_yield();
    unsigned r = 0;
    const unsigned base = threadIdx.x & ~63u;
    for (unsigned l = 0; l < 64 && base + l < g_wave->slot.size(); ++l)
        if (!g_wave->fibers[base + l].done) r += (unsigned)g_wave->slot[base + l];
    shim_yield();
    return r;
}
inline unsigned __ockl_wfred_min_u32(unsigned v) {
    if (!g_wave) throw std::runtime_error("__ockl_wfred_min_u32 outside a wave launch");
    g_wave->slot[threadIdx.x] = v;
    shim_yield();
    unsigned r = ~0u;
    const unsigned base = threadIdx.x & ~63u;
    for (unsigned l = 0; l < 64 && base + l < g_wave->slot.size(); ++l)
        if (!g_wave->fibers[base + l].done) r = std::min(r, (unsigned)g_wave->slot[base + l]);
    shim_yield();
    return r;
}
// Lanes of a wave interleave at a wave barrier (on the GPU they run in lockstep).
inline void __builtin_amdgcn_wave_barrier() {
    if (g_wave) shim_yield();
}
#define __builtin_amdgcn_fence(order, scope) ((void)0)
// The hardware square root is only guaranteed within 1 ulp: return the correctly rounded root
// moved by -1, 0 or +1 ulp (pseudo-randomly), so that callers' corrections are exercised.
inline float __builtin_amdgcn_sqrtf(float x) {
    static thread_local uint32_t state = 12345u;
    state = state * 1664525u + 1013904223u;
    const float r = std::sqrt(x);
    if (r == 0.0f) return r;
    const int k = (int)(state >> 30) % 3;
    return k == 0 ? r : (k == 1 ? std::nextafter(r, 0.0f) : std::nextafter(r, INFINITY));
}
inline int __syncthreads_or(int pred) {
    if (!g_wave) throw std::runtime_error("__syncthreads_or outside a wave launch");
    g_wave->slot[threadIdx.x] = pred ? 1 : 0;
    __syncthreads();
    int r = 0;
    for (size_t t = 0; t < g_wave->slot.size(); ++t) r |= !g_wave->fibers[t].done && g_wave->slot[t] != 0;
    __syncthreads();
    return r;
}
// Signed 24-bit multiply: operands are the sign-extended low 24 bits, the result the low 32 bits.
inline int __mul24(int a, int b) {
    const int64_t x = (int32_t)((uint32_t)a << 8) >> 8, y = (int32_t)((uint32_t)b << 8) >> 8;
    return (int)(uint32_t)(uint64_t)(x * y);
}
// The library passes only wave-uniform values to readfirstlane, so the lane's own value is it.
inline int __builtin_amdgcn_readfirstlane(int v) { return v; }
// Unsigned 24-bit multiply: operands are the low 24 bits, the result the low 32 bits.
inline unsigned __umul24(unsigned a, unsigned b) {
    return (unsigned)((uint64_t)(a & 0xFFFFFFu) * (uint64_t)(b & 0xFFFFFFu));
}
inline unsigned __float_as_uint(float f) { unsigned u; std::memcpy(&u, &f, 4); return u; }
inline float __uint_as_float(unsigned u) { float f; std::memcpy(&f, &u, 4); return f; }
inline float __fadd_rn(float a, float b) { volatile float s = a + b; return s; }
inline int __popc(unsigned x) { return __builtin_popcount(x); }
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
inline int __ffsll(long long x) { return __builtin_ffsll(x); }
inline int __ffs(unsigned x) { return __builtin_ffs((int)x); }
inline int __clzll(long long x) { return x ? __builtin_clzll((unsigned long long)x) : 64; }

template <class T, class U> T atomicAdd(T* p, U v) { return __atomic_fetch_add(p, (T)v, __ATOMIC_SEQ_CST); }
template <class T, class U> T atomicExch(T* p, U v) { return __atomic_exchange_n(p, (T)v, __ATOMIC_SEQ_CST); }
template <class T, class U> T atomicOr(T* p, U v) { return __atomic_fetch_or(p, (T)v, __ATOMIC_SEQ_CST); }
template <class T, class U> T atomicAnd(T* p, U v) { return __atomic_fetch_and(p, (T)v, __ATOMIC_SEQ_CST); }
template <class T, class U> T atomicCAS(T* p, U cmp, T val) {
    T e = (T)cmp;
    __atomic_compare_exchange_n(p, &e, val, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST);
    return e;
}
template <class T, class U> T atomicMin(T* p, U v) {
    T o = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while ((T)v < o && !__atomic_compare_exchange_n(p, &o, (T)v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
    }
    return o;
}
template <class T, class U> T atomicMax(T* p, U v) {
    T o = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while ((T)v > o && !__atomic_compare_exchange_n(p, &o, (T)v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
    }
    return o;
}
#define __ATOMIC_RELAXED_ 0
#define __HIP_MEMORY_SCOPE_AGENT 0
#define __HIP_MEMORY_SCOPE_WORKGROUP 1
template <class T> T __hip_atomic_load(T* p, int, int) { return __atomic_load_n(p, __ATOMIC_SEQ_CST); }
template <class T, class U> void __hip_atomic_store(T* p, U v, int, int) { __atomic_store_n(p, (T)v, __ATOMIC_SEQ_CST); }
inline void fail_check(uint32_t* err, uint32_t code) { atomicCAS(err, 0u, code); }
inline uint32_t min(uint32_t a, uint32_t b) { return a < b ? a : b; }
inline uint32_t max(uint32_t a, uint32_t b) { return a > b ? a : b; }
inline int min(int a, int b) { return a < b ? a : b; }
inline size_t min(size_t a, size_t b) { return a < b ? a : b; }
inline unsigned long long min(unsigned long long a, unsigned long long b) { return a < b ? a : b; }
inline unsigned long long max(unsigned long long a, unsigned long long b) { return a > b ? a : b; }
inline size_t max(size_t a, size_t b) { return a > b ? a : b; }
inline int max(int a, int b) { return a > b ? a : b; }

inline std::vector<char*>& fiber_stacks() {
    static std::vector<char*> stacks;  // reused by every wave launch
    return stacks;
}

// Dynamic shared memory of the current launch, allocated at its exact size so that the
// sanitizer sees any access past its end.
inline std::vector<uint8_t> g_dyn_shared;
inline void* hostemu_dyn_shared() { return g_dyn_shared.data(); }

template <class F> void launch_cpu(bool wave, Dim3 g, Dim3 b, F f, size_t shmem = 0) {
    gridDim = g;
    blockDim = b;
    std::vector<uint8_t>(shmem, 0xCD).swap(g_dyn_shared);
    for (unsigned bz = 0; bz < g.z; ++bz)
        for (unsigned by = 0; by < g.y; ++by)
            for (unsigned bx = 0; bx < g.x; ++bx) {
                if (!wave) {
                    if (getenv("HOSTEMU_PAR") && bz == 0 && by == 0 && bx == 0 && g.y == 1 && g.z == 1) {
                        // all blocks of this launch on 8 OS threads, threads of a block interleaved
                        std::atomic<unsigned> nextb{0};
                        auto run = [&] {
                            for (;;) {
                                unsigned blk = nextb.fetch_add(1);
                                if (blk >= g.x) break;
                                for (unsigned t = 0; t < b.x; ++t) {
                                    blockIdx = Dim3(blk, 0, 0);
                                    threadIdx = Dim3(t, 0, 0);
                                    f();
                                }
                            }
                        };
                        std::vector<std::thread> th;
                        for (int k = 0; k < 8; ++k) th.emplace_back(run);
                        for (auto& x : th) x.join();
                        goto done_launch;
                    }
                    for (unsigned t = 0; t < b.x; ++t) {
                        blockIdx = Dim3(bx, by, bz);
                        threadIdx = Dim3(t, 0, 0);
                        f();
                    }
                    continue;
                }
                std::function<void()> body = f;
                WaveCtx w;
                w.slot.assign(b.x, 0);
                w.fibers.resize(b.x);
                w.body = &body;
                w.block = Dim3(bx, by, bz);
                const size_t SZ = 1 << 18;
                std::vector<char*>& pool = fiber_stacks();
                while (pool.size() < b.x)
                    pool.push_back((char*)mmap(nullptr, SZ, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS, -1, 0));
                for (unsigned t = 0; t < b.x; ++t) {
                    Fiber& fb = w.fibers[t];
                    fb.tid = t;
                    fb.size = SZ;
                    fb.stack = pool[t];
                    uintptr_t top = ((uintptr_t)(fb.stack + SZ)) & ~(uintptr_t)15;
                    void** sp = (void**)top;
                    *--sp = nullptr;                          // fake return address of the entry
                    *--sp = (void*)&shim_fiber_entry;         // popped by ret
                    for (int r = 0; r < 6; ++r) *--sp = nullptr;  // rbp rbx r12 r13 r14 r15
                    fb.sp = sp;
                }
                g_wave = &w;
                for (;;) {
                    bool any = false;
                    for (unsigned t = 0; t < b.x; ++t) {
                        if (w.fibers[t].done) continue;
                        any = true;
                        w.cur = (int)t;
#if SHIM_ASAN
                        void* fake = nullptr;
                        __sanitizer_start_switch_fiber(&fake, w.fibers[t].stack, w.fibers[t].size);
#endif
                        shim_fiber_switch(&w.sched_sp, w.fibers[t].sp);
#if SHIM_ASAN
                        __sanitizer_finish_switch_fiber(fake, nullptr, nullptr);
#endif
                    }
                    if (!any) break;
                }
                g_wave = nullptr;
            }
done_launch:;
}

// ---- host side of the library ----
enum pma_status { PMA_OK = 0, PMA_ERR_INVALID_ARGUMENT = 1, PMA_ERR_HIP = 2, PMA_ERR_OUT_OF_MEMORY = 3,
                  PMA_ERR_STATE = 4, PMA_ERR_CAPACITY = 5, PMA_ERR_INTERNAL = 6 };
enum pma_volume { PMA_VOL_SOLID, PMA_VOL_EDT2, PMA_VOL_CR2, PMA_VOL_CX, PMA_VOL_CY, PMA_VOL_CZ, PMA_VOL_MASK,
                  PMA_VOL_PT, PMA_VOL_PTW, PMA_VOL_SPD, PMA_VOL_CAVITY, PMA_VOL_COUNT };
enum pma_stencil { PMA_STENCIL_D3Q27 = 0, PMA_STENCIL_D3Q7 = 1 };
enum pma_stage { PMA_STAGE_EDT, PMA_STAGE_CONTAINING_BALLS, PMA_STAGE_PORES, PMA_STAGE_REMOVE_EMPTY_PORES,
                 PMA_STAGE_WATERSHED, PMA_STAGE_GEODESIC, PMA_STAGE_COUNT };
struct pma_dims { int x, y, z; };
struct pma_cavity {
    uint32_t id;
    int32_t x, y, z;
    double distance_to_center;
    uint64_t void_count;
    int32_t reached_walls;
    int32_t reaches_all;
};
#define PMA_GEODESIC_NOALLOC_SCRATCH_BYTES (48u << 20)
#define __align__(n) alignas(n)
using hipError_t = int;
constexpr hipError_t hipSuccess = 0;
inline hipError_t hipGetLastError() { return hipSuccess; }
inline const char* hipGetErrorString(hipError_t) { return "no error"; }
#define HIP_SYMBOL(x) ((void*)&(x))
inline hipError_t hipGetSymbolAddress(void** p, void* sym) { *p = sym; return hipSuccess; }
inline hipError_t hipMemGetInfo(size_t* free_b, size_t* total_b) {
    *free_b = (size_t)1 << 30;
    *total_b = (size_t)1 << 31;
    return hipSuccess;
}
struct StageInfo { uint64_t work_items = 0; };
namespace pma {
struct Error : std::runtime_error {
    int status;
    Error(int s, const std::string& m) : std::runtime_error(m), status(s) {}
};
struct PoreRow { uint32_t id = 0; int32_t x = 0, y = 0, z = 0; uint32_t r2 = 0; bool boundary = false; uint64_t volume = 0; };
struct ThroatRow { uint32_t id = 0; uint64_t volume = 0; std::vector<uint32_t> pores; };
struct CavityRow {
    uint32_t id = 0;
    int32_t x = 0, y = 0, z = 0;
    double distance_to_center = 0;
    uint64_t void_count = 0;
    int32_t reached_walls = 0;
    bool reaches_all = false;
};
inline unsigned grid_for(size_t n, unsigned block) { return (unsigned)((n + block - 1) / block); }
}  // namespace pma
#define PMA_REQUIRE(cond, status, msg) do { if (!(cond)) throw ::pma::Error((status), (msg)); } while (0)
#define PMA_LAUNCHED(c) do {} while (0)
// Kernels run to completion at their launch here, so a failed in-kernel check is raised at once,
// naming the kernel.
#define PMA_CHECKED(c, name, err)                                                                   \
    do {                                                                                            \
        const uint32_t* e_ = (err);                                                                 \
        if (e_ && *e_)                                                                              \
            throw ::pma::Error(PMA_ERR_INTERNAL, std::string(name) + ": in-kernel check " +         \
                                                     std::to_string(*e_) + " failed");               \
    } while (0)
#define PMA_TIMED(c, n) do {} while (0)
#define PMA_HIP(x) (void)(x)
#define hipMemsetAsync(p, v, bytes, s) (std::memset((p), (v), (bytes)), 0)
enum { hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice };
#define hipMemcpyAsync(d, s_, bytes, kind, st) (std::memcpy((d), (s_), (bytes)), 0)
#define hipStreamSynchronize(s) 0

struct pma_ctx {
    int device = 0;
    pma_dims dims{0, 0, 0};
    size_t n = 0;
    pma_stencil stencil = PMA_STENCIL_D3Q27;
    void* stream = nullptr;
    std::vector<uint8_t> vol[PMA_VOL_COUNT];
    bool valid[PMA_VOL_COUNT] = {};
    std::map<std::string, std::pair<void*, size_t>> bufs;
    uint32_t max_e = 0;
    std::vector<uint64_t> histogram;
    uint64_t n_max_balls = 0;
    bool have_edt_stats = false, have_pores = false, have_pores_ws = false;
    std::vector<pma::PoreRow> pores, pores_ws;
    std::vector<pma::ThroatRow> throats, throats_ws;
    std::vector<pma::CavityRow> cavities;
    bool have_cavities = false;
    StageInfo stage[PMA_STAGE_COUNT];
    std::vector<uint32_t> testing_ball_order;
    bool debug_sync = false;
    template <class T> T* v(int w) { return (T*)vol[w].data(); }
    void ensure_volume(int w) {
        const size_t bytes = n * ((w == PMA_VOL_SOLID || w == PMA_VOL_MASK) ? 1 : 4);
        if (vol[w].size() < bytes) {
            std::vector<uint8_t>().swap(vol[w]);
            vol[w].resize(bytes);  // exact size, so the sanitizer sees any access past the end
        }
    }
    void* buf(const std::string& name, size_t bytes) {
        auto& b = bufs[name];
        if (bytes > b.second) {
            std::free(b.first);
            b.first = std::malloc(std::max<size_t>(bytes, 1));
            b.second = bytes;
            { static int fillv = getenv("HOSTEMU_FILL") ? atoi(getenv("HOSTEMU_FILL")) : 0xCD;
              if (fillv >= 0) std::memset(b.first, fillv, std::max<size_t>(bytes, 1));
              else { uint8_t* q = (uint8_t*)b.first; for (size_t k = 0; k < bytes; ++k) q[k] = (uint8_t)(rand() % 7); } }
        }
        return b.first;
    }
    template <class T> T* buf(const std::string& name, size_t count) { return (T*)buf(name, count * sizeof(T)); }
    void drop(const std::string& name) {
        auto it = bufs.find(name);
        if (it == bufs.end()) return;
        std::free(it->second.first);
        bufs.erase(it);
    }
    void* host_staging(size_t bytes) { return buf("host_staging", bytes); }
    void sync() {}
    size_t buffer_bytes(const std::string& name) const {
        auto it = bufs.find(name);
        return it == bufs.end() ? 0 : it->second.second;
    }
    float geo_delta = 16.0f;
    uint32_t flood_grid = 4096;
    unsigned long long replay_mem = 128ull << 30;
    std::set<std::string> evictable;
    std::set<std::string> persistent;
    size_t evictable_bytes() const {
        size_t total = 0;
        for (const std::string& name : evictable) {
            auto it = bufs.find(name);
            if (it != bufs.end()) total += it->second.second;
        }
        return total;
    }
};

namespace pma {
template <class T> T read_scalar(pma_ctx*, const T* p) { return *p; }
inline void upload(pma_ctx*, void* d, const void* s, size_t b) { if (b) std::memcpy(d, s, b); }
uint64_t exclusive_scan_u32(pma_ctx* c, const uint32_t* in, uint32_t* out, size_t n);
uint64_t exclusive_scan_u64(pma_ctx* c, const uint64_t* in, uint64_t* out, size_t n);

template <class K, class V> void rs(const K* ki, K* ko, const V* vi, V* vo, size_t n, int b0, int b1) {
    std::vector<size_t> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    auto key = [&](size_t i) { K k = ki[i]; if (b1 < (int)(8 * sizeof(K))) k &= (((K)1 << b1) - 1); return k >> b0; };
    std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return key(a) < key(b); });
    std::vector<K> k2(n); std::vector<V> v2(n);
    for (size_t i = 0; i < n; ++i) { k2[i] = ki[idx[i]]; v2[i] = vi[idx[i]]; }
    for (size_t i = 0; i < n; ++i) { ko[i] = k2[i]; vo[i] = v2[i]; }
}
inline void radix_sort_pairs_u64(pma_ctx*, const uint64_t* a, uint64_t* b, const uint32_t* c, uint32_t* d, size_t n, int b0 = 0, int b1 = 64) { rs(a, b, c, d, n, b0, b1); }
inline void radix_sort_pairs_u64_u64(pma_ctx*, const uint64_t* a, uint64_t* b, const uint64_t* c, uint64_t* d, size_t n, int b0 = 0, int b1 = 64) { rs(a, b, c, d, n, b0, b1); }
inline void radix_sort_pairs_u32_u64(pma_ctx*, const uint32_t* a, uint32_t* b, const uint64_t* c, uint64_t* d, size_t n, int b0 = 0, int b1 = 32) { rs(a, b, c, d, n, b0, b1); }
inline void radix_sort_pairs_u32(pma_ctx*, const uint32_t* a, uint32_t* b, const uint32_t* c, uint32_t* d, size_t n, int b0 = 0, int b1 = 32) { rs(a, b, c, d, n, b0, b1); }
inline void radix_sort_keys_u64(pma_ctx*, const uint64_t* a, uint64_t* b, size_t n, int b0 = 0, int b1 = 64) {
    std::vector<uint8_t> none(n);
    std::vector<uint8_t> none2(n);
    rs(a, b, none.data(), none2.data(), n, b0, b1);
}
void sort_like_libstdcxx_desc(pma_ctx* c, uint64_t* r, size_t n);

constexpr uint32_t kTrieNode = 0x80000000u;
struct PrefixTrie {
    uint32_t distinct = 0;
    std::vector<uint64_t> key, min_time, finals;
    std::vector<uint32_t> members(uint32_t d) const;
};
PrefixTrie build_prefix_trie(pma_ctx* c, const uint32_t* elems, uint64_t elem_cap, const uint64_t* seq_off,
                             const uint32_t* seq_len, const uint64_t* rbase, const uint64_t* rec_time,
                             const uint32_t* seq_weight, size_t S, uint64_t R, uint32_t* rec_node);
}  // namespace pma
