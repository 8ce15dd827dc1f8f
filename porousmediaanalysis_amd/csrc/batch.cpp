// Batches of subvolumes and the knobs an autotuner picks for them.  No counterpart in the
// reference, which analyses one volume per console pass (ImageProcessingManager.cpp:44-108).
//
// A batch runs on a pool of host threads, each with its own context (so its own stream and
// scratch): small subvolumes leave most of the GPU idle between the dependent rounds of the
// flood stages, and several contexts fill it.  The pipeline cannot be one captured graph: every
// stage reads counts back to size its next launches (levels, rounds, plateaus, keys).
#include "pma_internal.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <mutex>
#include <numeric>
#include <thread>
#include <vector>

namespace {

struct Knob {
    double lo, hi;
};
constexpr Knob kKnobs[] = {{1.0, 1024.0}, {64.0, 65536.0}, {256.0, 1048576.0}};

void set_knob(pma_ctx* c, pma_tunable which, double value) {
    PMA_REQUIRE(c != nullptr, PMA_ERR_INVALID_ARGUMENT, "null context");
    PMA_REQUIRE((int)which >= 0 && (int)which < 3, PMA_ERR_INVALID_ARGUMENT, "no such tuning knob");
    const Knob k = kKnobs[(int)which];
    PMA_REQUIRE(value >= k.lo && value <= k.hi, PMA_ERR_INVALID_ARGUMENT,
                "tuning value out of range [" + std::to_string(k.lo) + ", " + std::to_string(k.hi) + "]");
    switch (which) {
        case PMA_TUNE_GEODESIC_BUCKET:
            c->geo_delta = (float)value;
            break;
        case PMA_TUNE_FLOOD_GRID:
            c->flood_grid = (uint32_t)value;
            break;
        case PMA_TUNE_REPLAY_MEMORY_MB:
            c->replay_mem = (unsigned long long)value << 20;
            break;
    }
}

struct Outcome {
    std::mutex m;
    pma_status status = PMA_OK;
    std::string message;
    void fail(size_t index, pma_status s, const std::string& what) {
        std::lock_guard<std::mutex> lock(m);
        if (status != PMA_OK) return;
        status = s;
        message = "subvolume " + std::to_string(index) + ": " + what;
    }
};

// The configuration with defaults filled in.
pma_batch_config resolved(int device, const pma_batch_config* in, size_t count) {
    pma_batch_config c = in ? *in : pma_batch_config{0, 0.0, 0, 0};
    if (c.workers == 0) c.workers = 8;
    c.workers = (uint32_t)std::max<size_t>(1, std::min<size_t>(c.workers, std::max<size_t>(count, 1)));
    if (c.geodesic_bucket == 0.0) c.geodesic_bucket = 16.0;
    if (c.flood_grid == 0) c.flood_grid = 4096;
    if (c.replay_memory_mb == 0) {
        PMA_HIP(hipSetDevice(device));
        size_t free_b = 0, total_b = 0;
        PMA_HIP(hipMemGetInfo(&free_b, &total_b));
        const unsigned long long mb = free_b / (2ull * c.workers) >> 20;
        c.replay_memory_mb = (uint32_t)std::max<unsigned long long>(256, std::min<unsigned long long>(mb, 128ull << 10));
    }
    return c;
}

// Runs volumes[order[...]] on the given contexts (one thread each), largest first.
void run_on(std::vector<pma_ctx*>& ctxs, const pma_batch_volume* volumes, const std::vector<size_t>& order,
            pma_stencil stencil, const pma_batch_config& cfg, pma_batch_done done, void* user, Outcome& out) {
    std::atomic<size_t> next{0};
    auto work = [&](pma_ctx* c) {
        for (pma_tunable k : {PMA_TUNE_GEODESIC_BUCKET, PMA_TUNE_FLOOD_GRID, PMA_TUNE_REPLAY_MEMORY_MB}) {
            const double v = k == PMA_TUNE_GEODESIC_BUCKET ? cfg.geodesic_bucket
                             : k == PMA_TUNE_FLOOD_GRID    ? (double)cfg.flood_grid
                                                           : (double)cfg.replay_memory_mb;
            if (pma_set_tuning(c, k, v) != PMA_OK) {
                out.fail(0, PMA_ERR_INVALID_ARGUMENT, pma_last_error());
                return;
            }
        }
        for (size_t i; (i = next.fetch_add(1)) < order.size();) {
            const size_t index = order[i];
            const pma_batch_volume& v = volumes[index];
            pma_status s = v.solid ? PMA_OK : PMA_ERR_INVALID_ARGUMENT;
            std::string what = v.solid ? "" : "null solid volume";
            if (s == PMA_OK) s = pma_set_dims(c, v.dims);
            if (s == PMA_OK) s = pma_set_stencil(c, stencil);
            if (s == PMA_OK) s = pma_load_volume(c, PMA_VOL_SOLID, v.solid, 0);
            if (s == PMA_OK) s = pma_run_all(c);
            if (s != PMA_OK) {
                if (what.empty()) what = pma_last_error();
                out.fail(index, s, what);
            }
            if (done) done(user, index, c, s);
        }
    };
    std::vector<std::thread> threads;
    for (size_t w = 1; w < ctxs.size(); ++w) threads.emplace_back(work, ctxs[w]);
    work(ctxs[0]);
    for (std::thread& t : threads) t.join();
}

std::vector<size_t> largest_first(const pma_batch_volume* volumes, size_t count) {
    std::vector<size_t> order(count);
    std::iota(order.begin(), order.end(), size_t(0));
    auto voxels = [&](size_t i) {
        const pma_dims d = volumes[i].dims;
        return d.x > 0 && d.y > 0 && d.z > 0 ? (uint64_t)d.x * (uint64_t)d.y * (uint64_t)d.z : 0ull;
    };
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return voxels(a) > voxels(b); });
    return order;
}

struct Contexts {
    std::vector<pma_ctx*> v;
    Contexts(int device, size_t n) {
        for (size_t i = 0; i < n; ++i) {
            pma_ctx* c = nullptr;
            if (pma_create(&c, device) != PMA_OK) throw pma::Error(PMA_ERR_HIP, pma_last_error());
            v.push_back(c);
        }
    }
    ~Contexts() {
        for (pma_ctx* c : v) pma_destroy(c);
    }
};

}  // namespace

extern "C" {

pma_status pma_set_tuning(pma_ctx* c, pma_tunable which, double value) {
    try {
        set_knob(c, which, value);
        pma::set_last_error("");
        return PMA_OK;
    } catch (const pma::Error& e) {
        pma::set_last_error(e.what());
        return e.status;
    }
}

pma_status pma_get_tuning(pma_ctx* c, pma_tunable which, double* value) {
    if (!c || !value || (int)which < 0 || (int)which >= 3) {
        pma::set_last_error("null argument or no such tuning knob");
        return PMA_ERR_INVALID_ARGUMENT;
    }
    *value = which == PMA_TUNE_GEODESIC_BUCKET ? (double)c->geo_delta
             : which == PMA_TUNE_FLOOD_GRID    ? (double)c->flood_grid
                                               : (double)(c->replay_mem >> 20);
    pma::set_last_error("");
    return PMA_OK;
}

pma_status pma_run_batch(int device, const pma_batch_volume* volumes, size_t count, pma_stencil stencil,
                         const pma_batch_config* config, pma_batch_done done, void* user) {
    try {
        PMA_REQUIRE(volumes != nullptr || count == 0, PMA_ERR_INVALID_ARGUMENT, "null volume list");
        PMA_REQUIRE(stencil == PMA_STENCIL_D3Q27 || stencil == PMA_STENCIL_D3Q7, PMA_ERR_INVALID_ARGUMENT,
                    "no such stencil");
        if (count == 0) {
            pma::set_last_error("");
            return PMA_OK;
        }
        const pma_batch_config cfg = resolved(device, config, count);
        Contexts ctxs(device, cfg.workers);
        Outcome out;
        run_on(ctxs.v, volumes, largest_first(volumes, count), stencil, cfg, done, user, out);
        pma::set_last_error(out.message);
        return out.status;
    } catch (const pma::Error& e) {
        pma::set_last_error(e.what());
        return e.status;
    } catch (const std::exception& e) {
        pma::set_last_error(e.what());
        return PMA_ERR_INTERNAL;
    }
}

pma_status pma_tune_batch(int device, const pma_batch_volume* sample, size_t count, pma_stencil stencil,
                          pma_batch_config* best, double* best_voxels_per_second) {
    try {
        PMA_REQUIRE(sample != nullptr && count > 0 && best != nullptr, PMA_ERR_INVALID_ARGUMENT,
                    "pma_tune_batch needs a sample and an output");
        PMA_REQUIRE(stencil == PMA_STENCIL_D3Q27 || stencil == PMA_STENCIL_D3Q7, PMA_ERR_INVALID_ARGUMENT,
                    "no such stencil");
        const uint32_t kWorkers[] = {1, 2, 4, 8, 12, 16};
        const double kBuckets[] = {8.0, 16.0, 32.0};
        const uint32_t kGrids[] = {1024, 4096, 16384};
        const size_t max_workers = std::min<size_t>(16, count);
        pma_batch_config base{(uint32_t)max_workers, 0.0, 0, 0};
        base = resolved(device, &base, count);
        // the replay memory of the largest pool, for every candidate (buffers are sized once)
        const uint32_t replay_mb = base.replay_memory_mb;
        Contexts ctxs(device, max_workers);
        const std::vector<size_t> order = largest_first(sample, count);
        uint64_t voxels = 0;
        for (size_t i = 0; i < count; ++i)
            voxels += (uint64_t)std::max(sample[i].dims.x, 0) * (uint64_t)std::max(sample[i].dims.y, 0) *
                      (uint64_t)std::max(sample[i].dims.z, 0);
        auto timed = [&](const pma_batch_config& cfg) -> double {
            std::vector<pma_ctx*> use(ctxs.v.begin(), ctxs.v.begin() + cfg.workers);
            Outcome out;
            const auto t0 = std::chrono::steady_clock::now();
            run_on(use, sample, order, stencil, cfg, nullptr, nullptr, out);
            const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (out.status != PMA_OK) throw pma::Error(out.status, out.message);
            return s;
        };
        (void)timed(base);  // untimed: every context allocates its buffers
        pma_batch_config cur{1, 16.0, 4096, replay_mb};
        double best_s = -1.0;
        for (uint32_t w : kWorkers) {
            if (w > max_workers) continue;
            pma_batch_config c = cur;
            c.workers = w;
            const double s = timed(c);
            if (best_s < 0 || s < best_s) best_s = s, cur.workers = w;
        }
        for (double b : kBuckets) {
            if (b == cur.geodesic_bucket) continue;
            pma_batch_config c = cur;
            c.geodesic_bucket = b;
            const double s = timed(c);
            if (s < best_s) best_s = s, cur.geodesic_bucket = b;
        }
        for (uint32_t g : kGrids) {
            if (g == cur.flood_grid) continue;
            pma_batch_config c = cur;
            c.flood_grid = g;
            const double s = timed(c);
            if (s < best_s) best_s = s, cur.flood_grid = g;
        }
        *best = cur;
        if (best_voxels_per_second) *best_voxels_per_second = best_s > 0 ? (double)voxels / best_s : 0.0;
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

}  // extern "C"
