// C ABI entry points: context, volumes, stage dispatch with timing, and the reference's
// text formats (Serializer.cpp:139-221).  No exception crosses the boundary.
#include "pma_internal.h"

#include <algorithm>
#include <cinttypes>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>

namespace pma {

void DeviceBuffer::reserve(size_t n, hipStream_t stream) {
    if (n <= bytes) return;
    if (ptr) {
        PMA_HIP(hipStreamSynchronize(stream));  // work queued on the old buffer has finished
        PMA_HIP(hipFree(ptr));
        ptr = nullptr;
        bytes = 0;
    }
    size_t want = std::max<size_t>(n, 256);
    PMA_HIP(hipMalloc(&ptr, want));
    bytes = want;
}

void DeviceBuffer::release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

size_t volume_element_size(pma_volume w) {
    switch (w) {
        case PMA_VOL_SOLID:
        case PMA_VOL_MASK:
            return 1;
        default:
            return 4;
    }
}

namespace {

thread_local std::string g_last_error;

pma_status guard(const std::function<void()>& f) {
    try {
        f();
        g_last_error.clear();
        return PMA_OK;
    } catch (const Error& e) {
        g_last_error = e.what();
        return e.status;
    } catch (const std::bad_alloc&) {
        g_last_error = "host allocation failed";
        return PMA_ERR_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return PMA_ERR_INTERNAL;
    }
}

}  // namespace

void set_last_error(const std::string& m) { g_last_error = m; }

namespace {

void check_ctx(pma_ctx* c) {
    PMA_REQUIRE(c != nullptr, PMA_ERR_INVALID_ARGUMENT, "null context");
    PMA_HIP(hipSetDevice(c->device));
}

void check_volume(pma_volume w) {
    PMA_REQUIRE((int)w >= 0 && (int)w < PMA_VOL_COUNT, PMA_ERR_INVALID_ARGUMENT, "unknown volume");
}

// Runs one stage with HIP events around it on the context's stream.
void timed(pma_ctx* c, pma_stage s, void (*fn)(pma_ctx*)) {
    check_ctx(c);
    PMA_REQUIRE(c->n > 0, PMA_ERR_STATE, "set the dimensions first");
    auto t0 = std::chrono::steady_clock::now();
    uint64_t l0 = c->launches;
    c->arena_begin_stage();
    PMA_HIP(hipEventRecord(c->ev0, c->stream));
    fn(c);
    PMA_HIP(hipEventRecord(c->ev1, c->stream));
    PMA_HIP(hipEventSynchronize(c->ev1));
    float ms = 0;
    PMA_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->stage[s].device_ms = ms;
    c->stage[s].host_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    c->stage[s].launches = c->launches - l0;
    c->resolve_kernel_times();
    if (c->debug_sync) {  // the scratch held after the stage
        size_t arena = 0;
        for (const auto& b : c->arena_chunks) arena += b.bytes;
        std::fprintf(stderr, "[pma] after stage %d: arena %.2f GiB in %zu chunk(s), stage took %.2f GiB\n", (int)s,
                     arena / double(1ull << 30), c->arena_chunks.size(), c->arena_stage / double(1ull << 30));
        for (const auto& kv : c->scratch)
            if (kv.second.bytes >= (256ull << 20))
                std::fprintf(stderr, "[pma] after stage %d: kept %-24s %8.2f GiB\n", (int)s, kv.first.c_str(),
                             kv.second.bytes / double(1ull << 30));
    }
}

std::string format_pores(const std::vector<PoreRow>& rows) {
    std::string s = "PoreId SeedX SeedY SeedZ SeedRadiusSquare IsBoundaryPore Volume\n";
    char b[160];
    for (const PoreRow& p : rows) {
        std::snprintf(b, sizeof b, "%u %d %d %d %d %d %zu\n", p.id, p.x, p.y, p.z, (int)p.r2, p.boundary ? 1 : 0,
                      (size_t)p.volume);
        s += b;
    }
    return s;
}

std::string format_throats(const std::vector<ThroatRow>& rows) {
    std::string s = "SharedPixelTypeId; Volume; PoreId1 PoreId2 ... \n";
    char b[64];
    for (const ThroatRow& t : rows) {
        std::snprintf(b, sizeof b, "%u; %zu;", t.id, (size_t)t.volume);
        s += b;
        for (uint32_t p : t.pores) {
            std::snprintf(b, sizeof b, " %u", p);
            s += b;
        }
        s += "\n";
    }
    return s;
}

std::string format_text(pma_ctx* c, pma_text which) {
    char b[256];
    switch (which) {
        case PMA_TEXT_INTERMEDIATE_STATISTICS:
            PMA_REQUIRE(c->have_edt_stats, PMA_ERR_STATE, "no statistics yet");
            std::snprintf(b, sizeof b, "Max Euclidean distance square: %u\nMaximum balls count: %zu\n", c->max_e,
                          (size_t)c->n_max_balls);
            return b;
        case PMA_TEXT_EDT_COUNTS: {
            // EuclideanDistanceComputer.cpp:83-110 with Serializer.cpp:148-159
            PMA_REQUIRE(c->have_edt_stats, PMA_ERR_STATE, "no EDT histogram yet");
            std::vector<uint64_t> h = c->histogram;
            long last = -1;
            for (long i = (long)h.size() - 1; i >= 0; --i)
                if (h[i]) { last = i; break; }
            h.resize(last == -1 ? 1 : (size_t)last + 2, 0);
            std::string s = "EuclideanDistanceSquaresCount (line index is EuclideanDistanceSquare, the next line has index one)\n";
            for (size_t i = 1; i < h.size(); ++i) {
                std::snprintf(b, sizeof b, "%zu\n", (size_t)h[i]);
                s += b;
            }
            return s;
        }
        case PMA_TEXT_PORES:
            PMA_REQUIRE(c->have_pores, PMA_ERR_STATE, "no pores yet");
            return format_pores(c->pores);
        case PMA_TEXT_THROATS:
            PMA_REQUIRE(c->have_pores, PMA_ERR_STATE, "no throats yet");
            return format_throats(c->throats);
        case PMA_TEXT_PORES_WATERSHED:
            PMA_REQUIRE(c->have_pores_ws, PMA_ERR_STATE, "no watershed pores yet");
            return format_pores(c->pores_ws);
        case PMA_TEXT_THROATS_WATERSHED:
            PMA_REQUIRE(c->have_pores_ws, PMA_ERR_STATE, "no watershed throats yet");
            return format_throats(c->throats_ws);
        case PMA_TEXT_CAVITIES: {
            PMA_REQUIRE(c->have_cavities, PMA_ERR_STATE, "no cavities yet");
            std::string s = "CavityId SeedX SeedY SeedZ DistanceToCenter VoidPixelsCount ReachedWallsCount IsReachingAllWalls\n";
            for (const CavityRow& r : c->cavities) {
                std::snprintf(b, sizeof b, "%u %d %d %d %f %zu %d %d\n", r.id, r.x, r.y, r.z, r.distance_to_center,
                              (size_t)r.void_count, r.reached_walls, r.reaches_all ? 1 : 0);
                s += b;
            }
            return s;
        }
        default:
            throw Error(PMA_ERR_INVALID_ARGUMENT, "unknown text");
    }
}

// Serializer::ReadPores (Serializer.cpp:103-128).
std::vector<PoreRow> parse_pores(const char* text) {
    std::vector<PoreRow> out;
    const char* nl = std::strchr(text, '\n');
    const char* p = nl ? nl + 1 : text + std::strlen(text);
    while (*p) {
        unsigned id;
        int x, y, z, r2, b, used = 0;
        size_t vol;
        if (std::sscanf(p, "%u %d %d %d %d %d %zu%n", &id, &x, &y, &z, &r2, &b, &vol, &used) != 7) break;
        PoreRow r;
        r.id = id; r.x = x; r.y = y; r.z = z; r.r2 = (uint32_t)r2; r.boundary = b != 0; r.volume = vol;
        out.push_back(r);
        p += used;
        while (*p == '\n' || *p == '\r' || *p == ' ') ++p;
    }
    std::sort(out.begin(), out.end(), [](const PoreRow& a, const PoreRow& b) { return a.id < b.id; });
    out.erase(std::unique(out.begin(), out.end(), [](const PoreRow& a, const PoreRow& b) { return a.id == b.id; }),
              out.end());
    return out;
}

// Serializer::ReadSharedPixelTypes (Serializer.cpp:66-101), one type per line.
std::vector<ThroatRow> parse_throats(const char* text) {
    std::vector<ThroatRow> out;
    const char* nl = std::strchr(text, '\n');
    const char* p = nl ? nl + 1 : text + std::strlen(text);
    while (*p) {
        const char* end = std::strchr(p, '\n');
        std::string line = end ? std::string(p, end) : std::string(p);
        p = end ? end + 1 : p + line.size();
        if (line.empty() || line == "\r") continue;
        ThroatRow t;
        unsigned id;
        size_t vol;
        int used = 0;
        if (std::sscanf(line.c_str(), "%u; %zu;%n", &id, &vol, &used) != 2) break;
        t.id = id;
        t.volume = vol;
        const char* q = line.c_str() + used;
        unsigned pid;
        int k = 0;
        while (std::sscanf(q, " %u%n", &pid, &k) == 1) {
            t.pores.push_back(pid);
            q += k;
        }
        out.push_back(t);
    }
    std::sort(out.begin(), out.end(), [](const ThroatRow& a, const ThroatRow& b) { return a.id < b.id; });
    return out;
}

}  // namespace
}  // namespace pma

using namespace pma;

void* pma_ctx::host_staging(size_t bytes) {
    if (bytes > pinned_bytes) {
        if (pinned) PMA_HIP(hipHostFree(pinned));
        pinned = nullptr;
        pinned_bytes = 0;
        PMA_HIP(hipHostMalloc(&pinned, std::max<size_t>(bytes, 4096), hipHostMallocDefault));
        pinned_bytes = std::max<size_t>(bytes, 4096);
    }
    return pinned;
}

namespace {
// Runs an allocation; when the device is out of memory, frees the evictable buffers (other than
// `keep`) and tries once more.
template <class F> void with_eviction(pma_ctx* c, const std::string& keep, F&& alloc) {
    try {
        alloc();
    } catch (const pma::Error& e) {
        if (e.status != PMA_ERR_OUT_OF_MEMORY) throw;
        (void)hipGetLastError();  // the failed hipMalloc must not surface at the next launch check
        if (!c->evict(keep)) throw;
        alloc();
    }
}
}  // namespace

void pma_ctx::ensure_volume(pma_volume w) {
    with_eviction(this, std::string(), [&] { vol[w].reserve(n * volume_element_size(w), stream); });
}

void* pma_ctx::buf(const std::string& name, size_t bytes) {
    if (persistent.count(name)) {
        pma::DeviceBuffer& b = scratch[name];
        with_eviction(this, name, [&] { b.reserve(bytes, stream); });
        return b.ptr;
    }
    bytes = std::max<size_t>(bytes, 256);
    auto it = arena_names.find(name);
    if (it != arena_names.end() && bytes <= it->second.bytes) return it->second.ptr;
    const size_t want = (bytes + 255) & ~size_t(255);
    if (arena_chunks.empty() || arena_used + want > arena_chunks.back().bytes) {
        // a new chunk: earlier regions of this stage stay where they are
        const size_t chunk = std::max<size_t>(want, arena_chunks.empty() ? (size_t(1) << 30) : arena_chunks.back().bytes);
        arena_chunks.emplace_back();
        pma::DeviceBuffer& b = arena_chunks.back();
        try {
            with_eviction(this, std::string(), [&] { b.reserve(chunk, stream); });
        } catch (...) {
            arena_chunks.pop_back();
            throw;
        }
        arena_used = 0;
    }
    void* p = static_cast<unsigned char*>(arena_chunks.back().ptr) + arena_used;
    arena_used += want;
    arena_stage += want;
    arena_peak = std::max(arena_peak, arena_stage);
    arena_names[name] = ArenaRegion{p, want};
    return p;
}

void pma_ctx::arena_begin_stage() {
    arena_names.clear();
    arena_stage = 0;
    arena_used = 0;
    if (arena_chunks.size() > 1) {
        // merge into one chunk of the largest stage's size
        PMA_HIP(hipStreamSynchronize(stream));
        for (pma::DeviceBuffer& b : arena_chunks) b.release();
        arena_chunks.clear();
        arena_chunks.emplace_back();
        pma::DeviceBuffer& b = arena_chunks.back();
        try {
            with_eviction(this, std::string(), [&] { b.reserve(arena_peak, stream); });
        } catch (...) {
            arena_chunks.clear();
            throw;
        }
    }
}

bool pma_ctx::evict(const std::string& keep) {
    bool any = false;
    for (const std::string& name : evictable) {
        if (name == keep) continue;
        auto it = scratch.find(name);
        if (it == scratch.end() || !it->second.ptr) continue;
        if (!any) PMA_HIP(hipStreamSynchronize(stream));
        it->second.release();
        scratch.erase(it);
        any = true;
    }
    return any;
}

size_t pma_ctx::evictable_bytes() const {
    size_t total = 0;
    for (const std::string& name : evictable) {
        auto it = scratch.find(name);
        if (it != scratch.end()) total += it->second.bytes;
    }
    return total;
}

void pma_ctx::drop(const std::string& name) {
    arena_names.erase(name);  // an arena region is taken back when the next stage starts
    auto it = scratch.find(name);
    if (it == scratch.end()) return;
    PMA_HIP(hipStreamSynchronize(stream));
    it->second.release();
    scratch.erase(it);
}

void pma_ctx::sync() { PMA_HIP(hipStreamSynchronize(stream)); }

hipEvent_t pma_ctx::kernel_mark() {
    hipEvent_t e;
    if (!event_pool.empty()) {
        e = event_pool.back();
        event_pool.pop_back();
    } else {
        PMA_HIP(hipEventCreate(&e));
    }
    PMA_HIP(hipEventRecord(e, stream));
    return e;
}

void pma_ctx::kernel_done(const char* name, hipEvent_t start) {
    hipEvent_t end = kernel_mark();
    kernel_pending.push_back({name, {start, end}});
}

void pma_ctx::resolve_kernel_times() {
    if (kernel_pending.empty()) return;
    PMA_HIP(hipStreamSynchronize(stream));
    for (auto& k : kernel_pending) {
        float ms = 0;
        PMA_HIP(hipEventElapsedTime(&ms, k.second.first, k.second.second));
        KernelTotal& t = kernel_totals[k.first];
        t.ms += ms;
        t.launches++;
        event_pool.push_back(k.second.first);
        event_pool.push_back(k.second.second);
    }
    kernel_pending.clear();
}

extern "C" {

const char* pma_last_error(void) { return g_last_error.c_str(); }

pma_status pma_create(pma_ctx** out, int device) {
    return guard([&] {
        PMA_REQUIRE(out != nullptr, PMA_ERR_INVALID_ARGUMENT, "null output pointer");
        *out = nullptr;
        int count = 0;
        PMA_HIP(hipGetDeviceCount(&count));
        PMA_REQUIRE(device >= 0 && device < count, PMA_ERR_INVALID_ARGUMENT, "no such HIP device");
        PMA_HIP(hipSetDevice(device));
        pma_ctx* c = new pma_ctx();
        c->device = device;
        const char* dbg = std::getenv("PMA_DEBUG_SYNC");
        c->debug_sync = dbg && dbg[0] == '1';
        try {
            PMA_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
            PMA_HIP(hipEventCreate(&c->ev0));
            PMA_HIP(hipEventCreate(&c->ev1));
        } catch (...) {
            delete c;
            throw;
        }
        *out = c;
    });
}

void pma_destroy(pma_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto& b : c->vol) b.release();
    for (auto& kv : c->scratch) kv.second.release();
    for (auto& b : c->arena_chunks) b.release();
    if (c->pinned) (void)hipHostFree(c->pinned);
    for (auto& k : c->kernel_pending) {
        (void)hipEventDestroy(k.second.first);
        (void)hipEventDestroy(k.second.second);
    }
    for (hipEvent_t e : c->event_pool) (void)hipEventDestroy(e);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

pma_status pma_set_stencil(pma_ctx* c, pma_stencil s) {
    return guard([&] {
        check_ctx(c);
        PMA_REQUIRE(s == PMA_STENCIL_D3Q27 || s == PMA_STENCIL_D3Q7, PMA_ERR_INVALID_ARGUMENT, "unknown stencil");
        c->stencil = s;
    });
}

pma_status pma_set_dims(pma_ctx* c, pma_dims d) {
    return guard([&] {
        check_ctx(c);
        PMA_REQUIRE(d.x > 0 && d.y > 0 && d.z > 0, PMA_ERR_INVALID_ARGUMENT, "dimensions must be positive");
        c->dims = d;
        c->n = (size_t)d.x * d.y * d.z;
        for (bool& v : c->valid) v = false;
        c->have_edt_stats = c->have_pores = c->have_pores_ws = c->have_cavities = false;
    });
}

pma_status pma_get_dims(pma_ctx* c, pma_dims* d) {
    return guard([&] {
        check_ctx(c);
        PMA_REQUIRE(d != nullptr, PMA_ERR_INVALID_ARGUMENT, "null output pointer");
        *d = c->dims;
    });
}

pma_status pma_load_volume(pma_ctx* c, pma_volume w, const void* src, int on_device) {
    return guard([&] {
        check_ctx(c);
        check_volume(w);
        PMA_REQUIRE(c->n > 0, PMA_ERR_STATE, "set the dimensions first");
        PMA_REQUIRE(src != nullptr, PMA_ERR_INVALID_ARGUMENT, "null source");
        c->ensure_volume(w);
        size_t bytes = c->n * volume_element_size(w);
        PMA_HIP(hipMemcpyAsync(c->vol[w].ptr, src, bytes, on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                               c->stream));
        PMA_HIP(hipStreamSynchronize(c->stream));
        c->valid[w] = true;
    });
}

pma_status pma_read_volume(pma_ctx* c, pma_volume w, void* dst, size_t bytes, int on_device) {
    return guard([&] {
        check_ctx(c);
        check_volume(w);
        PMA_REQUIRE(c->valid[w], PMA_ERR_STATE, "volume has not been produced or loaded");
        PMA_REQUIRE(dst != nullptr, PMA_ERR_INVALID_ARGUMENT, "null destination");
        PMA_REQUIRE(bytes == c->n * volume_element_size(w), PMA_ERR_INVALID_ARGUMENT, "size does not match the volume");
        PMA_HIP(hipMemcpyAsync(dst, c->vol[w].ptr, bytes, on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                               c->stream));
        PMA_HIP(hipStreamSynchronize(c->stream));
    });
}

pma_status pma_volume_device_ptr(pma_ctx* c, pma_volume w, void** dptr) {
    return guard([&] {
        check_ctx(c);
        check_volume(w);
        PMA_REQUIRE(dptr != nullptr, PMA_ERR_INVALID_ARGUMENT, "null output pointer");
        PMA_REQUIRE(c->n > 0, PMA_ERR_STATE, "set the dimensions first");
        c->ensure_volume(w);
        *dptr = c->vol[w].ptr;
    });
}

pma_status pma_edt(pma_ctx* c) { return guard([&] { timed(c, PMA_STAGE_EDT, run_edt); }); }
pma_status pma_containing_balls(pma_ctx* c) {
    return guard([&] { timed(c, PMA_STAGE_CONTAINING_BALLS, run_containing_balls); });
}
pma_status pma_pores(pma_ctx* c) { return guard([&] { timed(c, PMA_STAGE_PORES, run_pores); }); }
// Test hook (not part of the public header): the next pma_pores takes the maximum balls in this
// order (flat [z][x][y] voxel indices, every maximum ball once, r2 non-increasing) instead of
// sorting them, so that a test can replay the tie order another STL's std::sort produced.
pma_status pma_testing_set_ball_order(pma_ctx* c, const uint32_t* order, size_t m) {
    return guard([&] {
        PMA_REQUIRE(c && (order || m == 0), PMA_ERR_INVALID_ARGUMENT, "pma_testing_set_ball_order: null argument");
        c->testing_ball_order.assign(order, order + m);
    });
}
pma_status pma_remove_empty_pores(pma_ctx* c) {
    return guard([&] { timed(c, PMA_STAGE_REMOVE_EMPTY_PORES, run_remove_empty_pores); });
}
pma_status pma_watershed(pma_ctx* c) { return guard([&] { timed(c, PMA_STAGE_WATERSHED, run_watershed); }); }
pma_status pma_geodesic(pma_ctx* c) { return guard([&] { timed(c, PMA_STAGE_GEODESIC, run_geodesic); }); }

pma_status pma_run_all(pma_ctx* c) {
    return guard([&] {
        timed(c, PMA_STAGE_EDT, run_edt);
        timed(c, PMA_STAGE_CONTAINING_BALLS, run_containing_balls);
        timed(c, PMA_STAGE_PORES, run_pores);
        timed(c, PMA_STAGE_REMOVE_EMPTY_PORES, run_remove_empty_pores);
        timed(c, PMA_STAGE_WATERSHED, run_watershed);
        timed(c, PMA_STAGE_GEODESIC, run_geodesic);
    });
}

pma_status pma_geodesic_noalloc(int device, const uint8_t* solid, pma_dims dims, pma_stencil stencil, float* dist,
                                uint32_t* cavity, pma_cavity* cavities, size_t cap, size_t* count, void* stream) {
    return guard([&] {
        run_geodesic_noalloc(device, solid, dims, stencil, dist, cavity, cavities, cap, count,
                             static_cast<hipStream_t>(stream));
    });
}

pma_status pma_get_statistics(pma_ctx* c, uint32_t* max_e, uint64_t* n_max) {
    return guard([&] {
        check_ctx(c);
        PMA_REQUIRE(c->have_edt_stats, PMA_ERR_STATE, "no statistics yet");
        if (max_e) *max_e = c->max_e;
        if (n_max) *n_max = c->n_max_balls;
    });
}

pma_status pma_get_histogram(pma_ctx* c, uint64_t* dst, size_t cap, size_t* len) {
    return guard([&] {
        check_ctx(c);
        PMA_REQUIRE(c->have_edt_stats, PMA_ERR_STATE, "no EDT histogram yet");
        PMA_REQUIRE(len != nullptr, PMA_ERR_INVALID_ARGUMENT, "null length pointer");
        *len = c->histogram.size();
        PMA_REQUIRE(dst != nullptr && cap >= c->histogram.size(), PMA_ERR_CAPACITY, "histogram buffer too small");
        std::copy(c->histogram.begin(), c->histogram.end(), dst);
    });
}

pma_status pma_get_text(pma_ctx* c, pma_text which, char* dst, size_t cap, size_t* len) {
    return guard([&] {
        check_ctx(c);
        PMA_REQUIRE(len != nullptr, PMA_ERR_INVALID_ARGUMENT, "null length pointer");
        std::string s = format_text(c, which);
        *len = s.size() + 1;
        PMA_REQUIRE(dst != nullptr && cap >= s.size() + 1, PMA_ERR_CAPACITY, "text buffer too small");
        std::memcpy(dst, s.c_str(), s.size() + 1);
    });
}

pma_status pma_set_text(pma_ctx* c, pma_text which, const char* text) {
    return guard([&] {
        check_ctx(c);
        PMA_REQUIRE(text != nullptr, PMA_ERR_INVALID_ARGUMENT, "null text");
        switch (which) {
            case PMA_TEXT_INTERMEDIATE_STATISTICS: {
                // Serializer::ReadIntermediateStatistics (Serializer.cpp:130-137)
                unsigned m = 0;
                size_t k = 0;
                PMA_REQUIRE(std::sscanf(text, "Max Euclidean distance square: %u\nMaximum balls count: %zu", &m, &k) == 2,
                            PMA_ERR_INVALID_ARGUMENT, "cannot parse intermediateStatistics.txt");
                c->max_e = m;
                c->n_max_balls = k;
                if (c->histogram.size() != (size_t)m + 1) c->histogram.assign((size_t)m + 1, 0);
                c->have_edt_stats = true;
                break;
            }
            case PMA_TEXT_PORES:
                c->pores = parse_pores(text);
                if (!c->have_pores) c->throats.clear();
                c->have_pores = true;
                break;
            case PMA_TEXT_THROATS:
                c->throats = parse_throats(text);
                break;
            default:
                throw Error(PMA_ERR_INVALID_ARGUMENT, "this text cannot be set");
        }
    });
}

pma_status pma_get_stage_info(pma_ctx* c, pma_stage s, pma_stage_info* info) {
    return guard([&] {
        check_ctx(c);
        PMA_REQUIRE((int)s >= 0 && (int)s < PMA_STAGE_COUNT && info, PMA_ERR_INVALID_ARGUMENT, "bad stage or pointer");
        *info = c->stage[s];
    });
}

pma_status pma_set_kernel_timing(pma_ctx* c, int enabled) {
    return guard([&] {
        check_ctx(c);
        c->resolve_kernel_times();
        c->kernel_timing = enabled != 0;
        c->kernel_totals.clear();
    });
}

pma_status pma_get_kernel_time(pma_ctx* c, const char* kernel, double* total_ms, uint64_t* launches) {
    return guard([&] {
        check_ctx(c);
        PMA_REQUIRE(kernel != nullptr, PMA_ERR_INVALID_ARGUMENT, "null kernel name");
        c->resolve_kernel_times();
        auto it = c->kernel_totals.find(kernel);
        if (total_ms) *total_ms = it == c->kernel_totals.end() ? 0.0 : it->second.ms;
        if (launches) *launches = it == c->kernel_totals.end() ? 0 : it->second.launches;
    });
}

pma_status pma_list_kernel_times(pma_ctx* c, char* dst, size_t cap, size_t* len) {
    return guard([&] {
        check_ctx(c);
        PMA_REQUIRE(len != nullptr, PMA_ERR_INVALID_ARGUMENT, "null length pointer");
        c->resolve_kernel_times();
        std::string s;
        for (const auto& kv : c->kernel_totals) s += kv.first + "\n";
        *len = s.size() + 1;
        PMA_REQUIRE(dst != nullptr && cap >= s.size() + 1, PMA_ERR_CAPACITY, "buffer too small");
        std::memcpy(dst, s.c_str(), s.size() + 1);
    });
}

pma_status pma_generate_spheres(pma_ctx* c, int32_t n, float rmin, float rmax, uint64_t seed) {
    return guard([&] {
        check_ctx(c);
        run_generate_spheres(c, n, rmin, rmax, seed);
    });
}

pma_status pma_generate_gaussian(pma_ctx* c, float sigma, float porosity, uint64_t seed) {
    return guard([&] {
        check_ctx(c);
        run_generate_gaussian(c, sigma, porosity, seed);
    });
}

}  // extern "C"
