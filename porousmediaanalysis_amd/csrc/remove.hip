// Empty-pore removal (EmptyPoresRemover.cpp:51-244): table rules on the host (the tables
// hold one row per pore and per throat), relabelling of the PT volume on the device.
//   1. pores with volume 0 are dropped and removed from every throat's list,
//   2. every throat's list is sorted ascending,
//   3. throats left with one pore are folded into it, throats with none are dropped,
//   4. throats with equal lists are merged into the largest ID,
//   5. voxels carrying a folded or merged throat ID take its new ID.
#include "pma_internal.h"

#include <algorithm>
#include <iterator>
#include <map>

namespace pma {
namespace {

__global__ void remap_labels(uint32_t* __restrict__ pt, size_t n, const uint32_t* __restrict__ from,
                             const uint32_t* __restrict__ to, uint32_t m) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t v = pt[i];
    if (v == 0 || v < from[0]) return;
    uint32_t lo = 0, hi = m;
    while (lo < hi) {
        uint32_t mid = (lo + hi) / 2;
        if (from[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    if (lo < m && from[lo] == v) pt[i] = to[lo];
}

}  // namespace

void run_remove_empty_pores(pma_ctx* c) {
    PMA_REQUIRE(c->have_pores, PMA_ERR_STATE, "pma_remove_empty_pores: no pores (run pma_pores or set pores.txt)");
    PMA_REQUIRE(c->valid[PMA_VOL_PT], PMA_ERR_STATE, "pma_remove_empty_pores: the PoreThroatIds volume is missing");
    std::map<uint32_t, uint32_t> remap;
    std::vector<uint32_t> empty;
    std::vector<PoreRow> kept;
    for (const PoreRow& p : c->pores) {
        if (p.volume == 0) empty.push_back(p.id);
        else kept.push_back(p);
    }
    std::sort(empty.begin(), empty.end());
    std::vector<ThroatRow> throats;
    for (ThroatRow t : c->throats) {
        std::sort(t.pores.begin(), t.pores.end());
        std::vector<uint32_t> rest;
        std::set_difference(t.pores.begin(), t.pores.end(), empty.begin(), empty.end(), std::back_inserter(rest));
        t.pores = rest;
        if (t.pores.size() < 2) {
            if (t.pores.size() == 1) remap[t.id] = t.pores[0];
            continue;
        }
        throats.push_back(t);
    }
    std::map<std::vector<uint32_t>, std::vector<uint32_t>> ids_by_list;
    for (const ThroatRow& t : throats) ids_by_list[t.pores].push_back(t.id);
    std::vector<ThroatRow> merged;
    for (const ThroatRow& t : throats) {
        const std::vector<uint32_t>& ids = ids_by_list[t.pores];
        uint32_t keep = *std::max_element(ids.begin(), ids.end());
        if (t.id == keep) merged.push_back(t);
        else remap[t.id] = keep;
    }
    c->pores = kept;
    c->throats = merged;
    if (!remap.empty()) {
        std::vector<uint32_t> from, to;
        for (const auto& kv : remap) {
            from.push_back(kv.first);
            to.push_back(kv.second);
        }
        uint32_t m = (uint32_t)from.size();
        uint32_t* d_from = c->buf<uint32_t>("rm_from", m);
        uint32_t* d_to = c->buf<uint32_t>("rm_to", m);
        upload(c, d_from, from.data(), m * 4ull);
        upload(c, d_to, to.data(), m * 4ull);
        remap_labels<<<grid_for(c->n, 256), 256, 0, c->stream>>>(c->v<uint32_t>(PMA_VOL_PT), c->n, d_from, d_to, m);
        PMA_LAUNCHED(c);
        PMA_HIP(hipStreamSynchronize(c->stream));
    }
    c->stage[PMA_STAGE_REMOVE_EMPTY_PORES].work_items = remap.size();
}

}  // namespace pma
