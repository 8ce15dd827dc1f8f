// Internal declarations shared by the host orchestration and the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <map>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "pma/pma.h"

namespace pma {

// Raised inside the library, turned into a pma_status at the C boundary.
struct Error : std::runtime_error {
    pma_status status;
    Error(pma_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};

// The text pma_last_error returns on this thread (api.cpp).
void set_last_error(const std::string& message);

#define PMA_HIP(expr)                                                                         \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            throw ::pma::Error(e_ == hipErrorOutOfMemory ? PMA_ERR_OUT_OF_MEMORY : PMA_ERR_HIP, \
                               std::string(#expr) + ": " + hipGetErrorString(e_));            \
    } while (0)

#define PMA_LAUNCHED(ctx)                                                                     \
    do {                                                                                      \
        hipError_t e_ = hipGetLastError();                                                    \
        if (e_ != hipSuccess)                                                                 \
            throw ::pma::Error(PMA_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e_)); \
        (ctx)->launches++;                                                                    \
    } while (0)

#define PMA_REQUIRE(cond, status, msg)                                                        \
    do {                                                                                      \
        if (!(cond)) throw ::pma::Error((status), (msg));                                     \
    } while (0)

// A failed in-kernel check: the first code stored wins (kernel id << 8 | check id).  Kernels
// that index buffers with values read from memory test each such index against the buffer's
// length and report here instead of accessing it.
__device__ __forceinline__ void fail_check(uint32_t* err, uint32_t code) { atomicCAS(err, 0u, code); }

// Host tables (Types.h:157-180 of the reference).
struct PoreRow {
    uint32_t id = 0;
    int32_t x = 0, y = 0, z = 0;
    uint32_t r2 = 0;
    bool boundary = false;
    uint64_t volume = 0;
};
struct ThroatRow {
    uint32_t id = 0;
    uint64_t volume = 0;
    std::vector<uint32_t> pores;
};
struct CavityRow {
    uint32_t id = 0;
    int32_t x = 0, y = 0, z = 0;
    double distance_to_center = 0;
    uint64_t void_count = 0;
    int32_t reached_walls = 0;
    bool reaches_all = false;
};

// A device allocation that only grows.  Its work runs on one stream, which reserve() waits for
// before it frees a buffer too small.
struct DeviceBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
    void reserve(size_t n, hipStream_t stream);
    void release();
    template <class T> T* as() const { return static_cast<T*>(ptr); }
};

}  // namespace pma

struct pma_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    pma_dims dims{0, 0, 0};
    size_t n = 0;
    pma_stencil stencil = PMA_STENCIL_D3Q27;

    pma::DeviceBuffer vol[PMA_VOL_COUNT];
    bool valid[PMA_VOL_COUNT] = {};

    // named scratch buffers, kept between calls
    std::map<std::string, pma::DeviceBuffer> scratch;
    // small pinned host staging buffer for scalar read-backs
    void* pinned = nullptr;
    size_t pinned_bytes = 0;

    // statistics of the EDT and containing-ball stages (intermediateStatistics.txt)
    bool have_edt_stats = false;
    uint32_t max_e = 0;
    uint64_t n_max_balls = 0;
    std::vector<uint64_t> histogram;

    // tables
    bool have_pores = false, have_pores_ws = false, have_cavities = false;
    std::vector<pma::PoreRow> pores, pores_ws;
    std::vector<pma::ThroatRow> throats, throats_ws;
    std::vector<pma::CavityRow> cavities;

    pma_stage_info stage[PMA_STAGE_COUNT] = {};
    uint64_t launches = 0;
    // test hook: the next pma_pores processes the maximum balls in this order (voxel indices)
    // instead of the sorted one, then clears it
    std::vector<uint32_t> testing_ball_order;
    // PMA_DEBUG_SYNC=1 at pma_create: every checked launch is named on stderr, waited for and
    // its stage's error flag read, so that a failing kernel is identified.
    bool debug_sync = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;

    // per-kernel event timing (pma_set_kernel_timing)
    bool kernel_timing = false;
    struct KernelTotal {
        double ms = 0;
        uint64_t launches = 0;
    };
    std::map<std::string, KernelTotal> kernel_totals;
    std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> kernel_pending;
    std::vector<hipEvent_t> event_pool;
    hipEvent_t kernel_mark();                      // records an event on the stream
    void kernel_done(const char* name, hipEvent_t start);  // records the end event
    void resolve_kernel_times();                   // after a synchronisation

    // knobs of pma_set_tuning (none changes a result)
    float geo_delta = 16.0f;                   // distance width of a geodesic round
    uint32_t flood_grid = 4096;                // workgroups of ws_bid and ws_resolve
    unsigned long long replay_mem = 128ull << 30;  // device bytes the DFS replays may hold

    // Stage scratch: every stage's buffers come from one arena that the next stage reuses (no
    // buffer of a stage is read by another), so the device holds the largest stage's scratch, not
    // the sum.  Within a stage a name keeps its region; regions are bumped from chunks that never
    // move, and a stage that outgrows the arena adds a chunk, merged into one at the next stage.
    struct ArenaRegion {
        void* ptr;
        size_t bytes;
    };
    std::vector<pma::DeviceBuffer> arena_chunks;
    size_t arena_used = 0;   // bytes bumped in the last chunk
    size_t arena_stage = 0;  // bytes taken by the running stage
    size_t arena_peak = 0;   // most bytes any stage took
    std::map<std::string, ArenaRegion> arena_names;
    void arena_begin_stage();
    // buffers kept between calls outside the arena (the DFS replay's, sized from the free memory)
    std::set<std::string> persistent;
    // scratch buffers that a failing allocation may free (large ones sized from the free memory,
    // kept between calls only to avoid allocating them again)
    std::set<std::string> evictable;

    template <class T> T* v(pma_volume w) { return vol[w].as<T>(); }
    void* buf(const std::string& name, size_t bytes);
    template <class T> T* buf(const std::string& name, size_t count) {
        return static_cast<T*>(buf(name, count * sizeof(T)));
    }
    // Frees a named scratch buffer once the context's stream is idle (for the large, short-lived ones).
    void drop(const std::string& name);
    // Frees the evictable buffers other than `keep`; false if there was none.
    bool evict(const std::string& keep);
    // Bytes held by the evictable buffers.
    size_t evictable_bytes() const;
    // Bytes held by a named scratch buffer (0 if none).
    size_t buffer_bytes(const std::string& name) const {
        auto it = scratch.find(name);
        return it == scratch.end() ? 0 : it->second.bytes;
    }
    void* host_staging(size_t bytes);
    void ensure_volume(pma_volume w);
    void sync();
};

namespace pma {

size_t volume_element_size(pma_volume w);

// Stage implementations (one .hip file each).
void run_edt(pma_ctx* c);
void run_containing_balls(pma_ctx* c);
void run_pores(pma_ctx* c);
void run_remove_empty_pores(pma_ctx* c);
void run_watershed(pma_ctx* c);
void run_geodesic(pma_ctx* c);
// pma_geodesic_noalloc: the geodesic on caller buffers with the library's static scratch.
void run_geodesic_noalloc(int device, const uint8_t* solid, pma_dims dims, pma_stencil stencil, float* dist,
                          uint32_t* cavity, pma_cavity* rows, size_t cap, size_t* count, hipStream_t stream);
void run_generate_spheres(pma_ctx* c, int32_t n, float rmin, float rmax, uint64_t seed);
void run_generate_gaussian(pma_ctx* c, float sigma, float porosity, uint64_t seed);

// Device-wide helpers (scan.hip).
// Exclusive scan of n u32 values into out (out may alias in); returns the total.
uint64_t exclusive_scan_u32(pma_ctx* c, const uint32_t* in, uint32_t* out, size_t n);
uint64_t exclusive_scan_u64(pma_ctx* c, const uint64_t* in, uint64_t* out, size_t n);
// Stable LSD radix sorts over bits [begin_bit, end_bit) of the keys (rsort.hip).
void radix_sort_pairs_u64(pma_ctx* c, const uint64_t* keys_in, uint64_t* keys_out, const uint32_t* vals_in,
                          uint32_t* vals_out, size_t n, int begin_bit = 0, int end_bit = 64);
void radix_sort_pairs_u64_u64(pma_ctx* c, const uint64_t* keys_in, uint64_t* keys_out, const uint64_t* vals_in,
                              uint64_t* vals_out, size_t n, int begin_bit = 0, int end_bit = 64);
void radix_sort_pairs_u32_u64(pma_ctx* c, const uint32_t* keys_in, uint32_t* keys_out, const uint64_t* vals_in,
                              uint64_t* vals_out, size_t n, int begin_bit = 0, int end_bit = 32);
void radix_sort_pairs_u32(pma_ctx* c, const uint32_t* keys_in, uint32_t* keys_out, const uint32_t* vals_in,
                          uint32_t* vals_out, size_t n, int begin_bit = 0, int end_bit = 32);
void radix_sort_keys_u64(pma_ctx* c, const uint64_t* keys_in, uint64_t* keys_out, size_t n, int begin_bit = 0,
                         int end_bit = 64);
// Exact emulation of libstdc++ std::sort with comp(a, b) = (a >> 32) > (b >> 32) (sort.hip).
void sort_like_libstdcxx_desc(pma_ctx* c, uint64_t* records, size_t n);

// Ordered pore lists of any length (lists.hip).  Sequence s holds the elements
// elems[seq_off[s] .. seq_off[s] + seq_len[s]) (each < 2^31); its prefixes of length 2 .. len
// are its formation records rbase[s] + (length - 2).  Distinct lists are the nodes of a prefix
// trie: key[d] = (parent << 32) | last element, where parent is the first element itself
// for a list of length 2 and kTrieNode | node otherwise; nodes are numbered by length, then
// by key.
constexpr uint32_t kTrieNode = 0x80000000u;
struct PrefixTrie {
    uint32_t distinct = 0;           // number of distinct lists
    std::vector<uint64_t> key;       // per node: parent << 32 | last element
    std::vector<uint64_t> min_time;  // per node: earliest rec_time of its records (~0 without times)
    std::vector<uint64_t> finals;    // per node: summed seq_weight of sequences whose full list it is
    std::vector<uint32_t> members(uint32_t d) const;  // the list, in sequence order
};
// rec_time and seq_weight may be null.  rec_node[r] receives the node of record r (device).
// elems holds elem_cap entries; every sequence must lie inside it.
PrefixTrie build_prefix_trie(pma_ctx* c, const uint32_t* elems, uint64_t elem_cap, const uint64_t* seq_off,
                             const uint32_t* seq_len, const uint64_t* rbase, const uint64_t* rec_time,
                             const uint32_t* seq_weight, size_t S, uint64_t R, uint32_t* rec_node);

// Host -> device copy that returns only when the source may be released.
inline void upload(pma_ctx* c, void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return;
    PMA_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    PMA_HIP(hipStreamSynchronize(c->stream));
}

// Copies one value from device memory through pinned staging.
template <class T> T read_scalar(pma_ctx* c, const T* dptr) {
    T* h = static_cast<T*>(c->host_staging(sizeof(T)));
    PMA_HIP(hipMemcpyAsync(h, dptr, sizeof(T), hipMemcpyDeviceToHost, c->stream));
    PMA_HIP(hipStreamSynchronize(c->stream));
    return *h;
}

inline unsigned grid_for(size_t n, unsigned block) { return (unsigned)((n + block - 1) / block); }

// After a launch of `kernel`: checks the launch and, in debug-sync mode, names the kernel on
// stderr, waits for it and reads the stage's error flag `err` (may be null).
inline void launched(pma_ctx* c, const char* kernel, const uint32_t* err) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw Error(PMA_ERR_HIP, std::string("launch of ") + kernel + ": " + hipGetErrorString(e));
    c->launches++;
    if (!c->debug_sync) return;
    std::fprintf(stderr, "[pma] %s\n", kernel);
    std::fflush(stderr);
    const hipError_t s = hipStreamSynchronize(c->stream);
    if (s != hipSuccess) throw Error(PMA_ERR_HIP, std::string(kernel) + ": " + hipGetErrorString(s));
    if (err) {
        const uint32_t v = read_scalar(c, err);
        if (v) throw Error(PMA_ERR_INTERNAL, std::string(kernel) + ": in-kernel check " + std::to_string(v) + " failed");
    }
}
#define PMA_CHECKED(ctx, name, err) ::pma::launched((ctx), (name), (err))

// Host-side guard of a copy or fill against the buffer it targets.
inline void require_fits(size_t bytes, size_t capacity_bytes, const char* what) {
    PMA_REQUIRE(bytes <= capacity_bytes, PMA_ERR_INTERNAL, std::string(what) + ": copy larger than its buffer");
}

// Brackets one kernel launch with events when kernel timing is enabled.
struct KernelScope {
    pma_ctx* c;
    const char* name;
    hipEvent_t start = nullptr;
    KernelScope(pma_ctx* ctx, const char* n) : c(ctx), name(n) {
        if (c && c->kernel_timing) start = c->kernel_mark();
    }
    ~KernelScope() {
        if (start) c->kernel_done(name, start);
    }
};
#define PMA_TIMED(ctx, name) ::pma::KernelScope pma_scope_##__LINE__((ctx), (name))

}  // namespace pma
