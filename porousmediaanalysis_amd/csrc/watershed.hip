// Watershed flood (WatershedComputer.cpp:30-219), reproduced exactly.
//
// Reference: one FIFO per depth d (= CR square); seeds are pushed in ascending pore ID with
// depth = seed r2; depths run from maxE down to 0; a popped node (v, p, d) visits its 26
// (or 6) neighbours in x, y, z loop order; a neighbour n qualifies when it is in the image,
// void and CR(n) <= d; an unlabelled n is claimed by p and pushed to FIFO CR(n); a labelled
// n owned by another pore gets that pore appended to its ordered throat list (without a
// push).  Keys are never erased; their IDs count down from UINT_MAX - 1 in order of first
// formation.
//
// GPU plan (level-synchronous, exact FIFO order):
//   * an event's global time is T = (maxE - d + 1) << 40 | pos << 5 | k, where pos is the
//     popped node's position in FIFO d and k its neighbour index; T increases exactly in the
//     reference's event order (pushes into FIFO d all happen before d is popped, and FIFO d
//     is popped in position order).
//   * FIFO d holds its seeds (ascending k) and then the pushes from higher depths in push
//     time order; its pops are split into rounds (round r+1 = the nodes pushed into d by
//     round r).
//   * in a round, every qualifying contact to an unlabelled voxel bids its time with a 64-bit
//     atomicMin (round-stamped so older bids lose); the minimum wins the claim.  All other
//     contacts whose pore differs from the owner are logged (voxel, time, pore).
//   * claims into the same depth are appended to FIFO d in (pos, k) order with a prefix sum;
//     claims into lower depths are staged the same way, so the level's staged list is in push
//     time order; when the level ends, a stable sort by target depth moves each depth's run to
//     the end of its FIFO (no queue atomics, and no sort by time).
//   * after the flood the log is sorted by (voxel, time): the distinct foreign pores in order
//     of first contact, c2, c3, ..., give the list [c2, owner, c3, ...] and its prefixes, each
//     formed at the first contact of its last pore; keys are numbered by formation time.
#include "pma_internal.h"

#include <algorithm>

namespace pma {
namespace {

constexpr uint64_t kPosMask = (1ull << 35) - 1;

struct LogEntry {
    unsigned long long t;
    uint32_t voxel;
    uint32_t pore;
};

__device__ __forceinline__ void decode(uint32_t i, int X, int Y, int& x, int& y, int& z) {
    y = (int)(i % (uint32_t)Y);
    uint32_t r = i / (uint32_t)Y;
    x = (int)(r % (uint32_t)X);
    z = (int)(r / (uint32_t)X);
}

__global__ void cr_histogram(const uint32_t* __restrict__ cr, size_t n, uint32_t* __restrict__ hist, uint32_t len) {
    __shared__ uint32_t lh[4096];
    for (int i = threadIdx.x; i < 4096; i += blockDim.x) lh[i] = 0;
    __syncthreads();
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        uint32_t v = cr[i];
        if (v == 0 || v >= len) continue;
        if (v < 4096) atomicAdd(&lh[v], 1u);
        else atomicAdd(&hist[v], 1u);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < min(4096u, len); i += blockDim.x)
        if (lh[i]) atomicAdd(&hist[i], lh[i]);
}

// The initial pushes: seed k at the slot the host gave it (its level's seeds in ascending k, the
// first entries of their FIFOs).
__global__ void push_seeds(const uint32_t* __restrict__ seed_voxel, const uint64_t* __restrict__ seed_slot,
                           uint32_t np, uint32_t* __restrict__ label, unsigned long long* __restrict__ qitem) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= np) return;
    uint32_t v = seed_voxel[k];
    label[v] = k + 1;
    qitem[seed_slot[k]] = (unsigned long long)v | ((unsigned long long)k << 32);
}

// The bounds [a, b) of a round's FIFO wave and the number of lower-level pushes the level has
// staged before the round, kept on the device: the kernels of a round read them, and
// ws_scan_parts writes the next round's (rounds alternate between two states).
struct WsRound {
    uint32_t a, b, staged;
};

// A level starts with its FIFO as the earlier levels left it.
__global__ void ws_level_start(const uint32_t* __restrict__ qcount, WsRound* __restrict__ state) {
    if (threadIdx.x == 0) *state = WsRound{0u, *qcount, 0u};
}
constexpr uint32_t kWsGrid = 512;       // workgroups of the round's scan kernels (grid-stride over the wave)

// A round's items are spread over lanes, one lane per (item, neighbour): the item's lanes form
// an aligned group of 32 (the 26 neighbours of D3Q27) or 8 (the 6 of D3Q7), so a group never
// straddles a wave and its loads run side by side instead of one neighbour after another.
__device__ __forceinline__ uint32_t ws_group_shift(int d3q7) { return d3q7 ? 3u : 5u; }

// Offset index k = 9 (dx + 1) + 3 (dy + 1) + (dz + 1) of lane kk of a group, or -1 for a spare
// lane; k increases with kk.
__device__ __forceinline__ int ws_offset(uint32_t kk, int d3q7) {
    if (d3q7) return kk < 6 ? (int)((0x16100E0C0A04ull >> (8 * kk)) & 0xFFu) : -1;  // 4 10 12 14 16 22
    return kk < 26 ? (int)(kk + (kk >= 13 ? 1u : 0u)) : -1;
}

// Neighbour k of voxel v, when it lies in the image.
__device__ __forceinline__ bool ws_neighbour(uint32_t v, int k, int X, int Y, int Z, uint32_t& j) {
    int x, y, z;
    decode(v, X, Y, x, y, z);
    const int nx = x + k / 9 - 1, ny = y + (k / 3) % 3 - 1, nz = z + k % 3 - 1;
    if (nx < 0 || nx >= X || ny < 0 || ny >= Y || nz < 0 || nz >= Z) return false;
    j = ((uint32_t)nz * (uint32_t)X + (uint32_t)nx) * (uint32_t)Y + (uint32_t)ny;
    return true;
}

__global__ void __launch_bounds__(256) ws_bid(const unsigned long long* __restrict__ items,
                                              const WsRound* __restrict__ round_state,
                                              const uint32_t* __restrict__ cr, const uint32_t* __restrict__ label,
                                              unsigned long long* __restrict__ claim, unsigned long long hi,
                                              uint32_t d, int X, int Y, int Z, int d3q7) {
    const uint32_t a = round_state->a, b = round_state->b;
    const uint32_t sh = ws_group_shift(d3q7);
    const size_t n = (size_t)(b - a) << sh, stride = (size_t)gridDim.x * blockDim.x;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) {
        const int k = ws_offset((uint32_t)t & ((1u << sh) - 1u), d3q7);
        if (k < 0) continue;
        const uint32_t pos = a + (uint32_t)(t >> sh);
        uint32_t j;
        if (!ws_neighbour((uint32_t)items[pos], k, X, Y, Z, j)) continue;
        const uint32_t c = cr[j];
        if (c == 0 || c > d || label[j] != 0) continue;
        atomicMin(&claim[j], hi | ((unsigned long long)pos << 5) | (unsigned long long)k);
    }
}

// Each contact of a round: a won claim labels the neighbour and marks it in one of its item's
// two masks over the 27 offsets (same level: appended to this level's FIFO by ws_append; lower
// level: staged by ws_append); a contact with another pore's voxel is logged.  A group's claims
// become its item's masks with one ballot each, and a wave's log entries take one atomic.
// same_cnt holds an item's same-level claims in its low 16 bits and its lower-level ones above.
__global__ void __launch_bounds__(256) ws_resolve(const unsigned long long* __restrict__ items,
                                                  const WsRound* __restrict__ round_state,
                                                  const uint32_t* __restrict__ cr, uint32_t* __restrict__ label,
                                                  const unsigned long long* __restrict__ claim, unsigned long long hi,
                                                  unsigned long long tlevel, uint32_t d, int X, int Y, int Z, int d3q7,
                                                  uint32_t* __restrict__ same_mask, uint32_t* __restrict__ lower_mask,
                                                  uint32_t* __restrict__ same_cnt, LogEntry* __restrict__ log,
                                                  unsigned long long* __restrict__ log_count, unsigned long long log_cap) {
    const uint32_t a = round_state->a, b = round_state->b;
    const uint32_t sh = ws_group_shift(d3q7), span = 1u << sh;
    const int lane = (int)(threadIdx.x & 63);
    const size_t n = (size_t)(b - a) << sh, stride = (size_t)gridDim.x * blockDim.x;
    // n and the stride are multiples of the group size, so a group is wholly in or out of the loop
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) {
        const uint32_t kk = (uint32_t)t & (span - 1u);
        const int k = ws_offset(kk, d3q7);
        const uint32_t pos = a + (uint32_t)(t >> sh);
        const unsigned long long it = items[pos];
        const uint32_t p = (uint32_t)(it >> 32);
        const unsigned long long when = (unsigned long long)(pos) << 5 | (unsigned long long)(k < 0 ? 0 : k);
        bool same = false, logged = false, lower = false;
        uint32_t j = 0;
        if (k >= 0 && ws_neighbour((uint32_t)it, k, X, Y, Z, j)) {
            const uint32_t c = cr[j];
            const uint32_t lab = label[j];  // loaded beside c: the voxel is in the image either way
            if (c != 0 && c <= d) {
                uint32_t owner = p;
                if (lab == 0) {
                    const unsigned long long w = claim[j];
                    if (w == (hi | when)) {
                        label[j] = p + 1;
                        same = c == d;
                        lower = c != d;
                    } else {
                        owner = (uint32_t)(items[(w >> 5) & kPosMask] >> 32);
                    }
                } else {
                    owner = lab - 1;
                }
                logged = owner != p;
            }
        }
        const unsigned long long logs = __ballot(logged);
        if (logs) {
            const int leader = __ffsll((long long)logs) - 1;
            unsigned long long base = 0;
            if (lane == leader) base = atomicAdd(log_count, (unsigned long long)__popcll(logs));
            base = __shfl(base, leader);
            if (logged) {
                const unsigned long long s = base + (unsigned long long)__popcll(logs & ((1ull << lane) - 1ull));
                if (s < log_cap) log[s] = LogEntry{tlevel | when, j, p};
            }
        }
        const unsigned long long sames = __ballot(same), lowers = __ballot(lower);
        if (kk == 0) {
            const int first = lane & ~(int)(span - 1u);
            const unsigned long long gm = (1ull << span) - 1ull;
            const uint32_t gs = (uint32_t)((sames >> first) & gm), gl = (uint32_t)((lowers >> first) & gm);
            uint32_t ms, ml;
            if (d3q7) {
                ms = ml = 0;
                for (uint32_t i = 0; i < 6; ++i) {
                    if (gs >> i & 1u) ms |= 1u << ws_offset(i, 1);
                    if (gl >> i & 1u) ml |= 1u << ws_offset(i, 1);
                }
            } else {
                ms = (gs & 0x1FFFu) | ((gs >> 13) << 14);
                ml = (gl & 0x1FFFu) | ((gl >> 13) << 14);
            }
            same_mask[pos - a] = ms;
            lower_mask[pos - a] = ml;
            same_cnt[pos - a] = (uint32_t)__popc(ms) | (uint32_t)__popc(ml) << 16;
        }
    }
}

// Workgroup w of the round kernels below takes the slice [w * chunk, min(n, (w + 1) * chunk)) of
// the wave's n items, chunk = ceil(n / gridDim.x).
__device__ __forceinline__ void ws_slice(uint32_t n, uint32_t& lo, uint32_t& hi) {
    const uint32_t chunk = (n + gridDim.x - 1) / gridDim.x;
    lo = min(n, blockIdx.x * chunk);
    hi = min(n, lo + chunk);
}

// Exclusive scan of one value per thread over a 256-thread workgroup; returns the total.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t& excl, uint32_t* __restrict__ lds) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;  // inclusive scan within the wave
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)x, (unsigned)off);
        if (lane >= off) x += y;
    }
    if (lane == 63) lds[wave] = x;
    __syncthreads();
    uint32_t before = 0, total = 0;
    for (int w = 0; w < 4; ++w) {
        if (w < wave) before += lds[w];
        total += lds[w];
    }
    excl = before + x - v;
    __syncthreads();  // lds is reused by the next call
    return total;
}

// Appended and staged items per workgroup slice (part, part_low).
__global__ void __launch_bounds__(256) ws_count_parts(const WsRound* __restrict__ round_state,
                                                      const uint32_t* __restrict__ same_cnt,
                                                      uint32_t* __restrict__ part, uint32_t* __restrict__ part_low) {
    __shared__ uint32_t lds[4];
    uint32_t lo, hi;
    ws_slice(round_state->b - round_state->a, lo, hi);
    uint32_t sum = 0, sum_low = 0;
    for (uint32_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
        const uint32_t v = same_cnt[i];
        sum += v & 0xFFFFu;
        sum_low += v >> 16;
    }
    uint32_t excl;
    const uint32_t total = block_exclusive_scan(sum, excl, lds);
    const uint32_t total_low = block_exclusive_scan(sum_low, excl, lds);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = total;
        part_low[blockIdx.x] = total_low;
    }
}

// One workgroup: the slices' offsets (exclusive scans of part and part_low, in place), the next
// round's bounds [b, b + appended) and the level's staged count after this round.  A round that
// would overflow its level's FIFO (capacity cap) or the staging list (capacity stage_cap) is
// flagged, and ws_append skips what does not fit.
__global__ void __launch_bounds__(256) ws_scan_parts(const WsRound* __restrict__ round_state,
                                                     WsRound* __restrict__ next_state, uint32_t* __restrict__ part,
                                                     uint32_t* __restrict__ part_low, uint32_t nparts, uint32_t cap,
                                                     uint32_t stage_cap, uint32_t* __restrict__ error) {
    __shared__ uint32_t lds[4];
    uint32_t carry = 0, carry_low = 0;
    for (uint32_t base = 0; base < nparts; base += blockDim.x) {  // the same trips on every thread
        const uint32_t i = base + threadIdx.x;
        const uint32_t v = i < nparts ? part[i] : 0u, w = i < nparts ? part_low[i] : 0u;
        uint32_t excl, excl_low;
        const uint32_t total = block_exclusive_scan(v, excl, lds);
        const uint32_t total_low = block_exclusive_scan(w, excl_low, lds);
        if (i < nparts) {
            part[i] = carry + excl;
            part_low[i] = carry_low + excl_low;
        }
        carry += total;
        carry_low += total_low;
    }
    if (threadIdx.x == 0) {
        const uint32_t b = round_state->b, staged = round_state->staged;
        uint64_t nb = (uint64_t)b + carry, ns = (uint64_t)staged + carry_low;
        if (nb > cap || ns > stage_cap) {
            atomicOr(error, 1u);
            nb = b;  // stop the level; the stage fails at its end
            ns = staged;
        }
        next_state->a = b;
        next_state->b = (uint32_t)nb;
        next_state->staged = (uint32_t)ns;
    }
}

// Appends each item's same-level claims after the wave and its lower-level claims after the
// level's staged list (target depth CR(j) beside each), both in item order, then offset order
// (slice offsets from ws_scan_parts plus scans within the slice), skipping places past the
// capacities.
__global__ void __launch_bounds__(256) ws_append(unsigned long long* __restrict__ items,
                                                 const WsRound* __restrict__ round_state,
                                                 const uint32_t* __restrict__ part, const uint32_t* __restrict__ part_low,
                                                 const uint32_t* __restrict__ same_mask,
                                                 const uint32_t* __restrict__ lower_mask,
                                                 const uint32_t* __restrict__ same_cnt, uint32_t cap,
                                                 const uint32_t* __restrict__ cr, unsigned long long* __restrict__ stage_item,
                                                 uint32_t* __restrict__ stage_level, uint32_t stage_cap, int X, int Y) {
    __shared__ uint32_t lds[4];
    const uint32_t a = round_state->a, b = round_state->b;
    uint32_t lo, hi;
    ws_slice(b - a, lo, hi);
    uint32_t carry = b + part[blockIdx.x];
    uint32_t carry_low = round_state->staged + part_low[blockIdx.x];
    for (uint32_t base = lo; base < hi; base += blockDim.x) {  // the same trips on every thread
        const uint32_t t = base + threadIdx.x;
        const uint32_t cnt = t < hi ? same_cnt[t] : 0u;
        uint32_t excl, excl_low;
        const uint32_t total = block_exclusive_scan(cnt & 0xFFFFu, excl, lds);
        const uint32_t total_low = block_exclusive_scan(cnt >> 16, excl_low, lds);
        if (cnt) {
            const unsigned long long it = items[a + t];
            const uint32_t v = (uint32_t)it;
            const unsigned long long p = it >> 32;
            int x, y, z;
            decode(v, X, Y, x, y, z);
            uint32_t mask = same_mask[t], o = carry + excl;
            while (mask) {
                const int k = __ffs(mask) - 1;
                mask &= mask - 1;
                const int dx = k / 9 - 1, dy = (k / 3) % 3 - 1, dz = k % 3 - 1;
                const uint32_t j = ((uint32_t)(z + dz) * (uint32_t)X + (uint32_t)(x + dx)) * (uint32_t)Y +
                                   (uint32_t)(y + dy);
                if (o < cap) items[o] = (unsigned long long)j | (p << 32);
                ++o;
            }
            mask = lower_mask[t];
            o = carry_low + excl_low;
            while (mask) {
                const int k = __ffs(mask) - 1;
                mask &= mask - 1;
                const int dx = k / 9 - 1, dy = (k / 3) % 3 - 1, dz = k % 3 - 1;
                const uint32_t j = ((uint32_t)(z + dz) * (uint32_t)X + (uint32_t)(x + dx)) * (uint32_t)Y +
                                   (uint32_t)(y + dy);
                if (o < stage_cap) {
                    stage_item[o] = (unsigned long long)j | (p << 32);
                    stage_level[o] = cr[j];
                }
                ++o;
            }
        }
        carry += total;
        carry_low += total_low;
    }
}

// The level's staged pushes, stably sorted by target depth, go to the ends of their FIFOs: run
// heads record where each depth's run starts, every push takes FIFO slot count + (i - start),
// and run tails then advance the counts.
__global__ void ws_run_first(const uint32_t* __restrict__ lvl, uint32_t n, uint32_t* __restrict__ first) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && (i == 0 || lvl[i - 1] != lvl[i])) first[lvl[i]] = i;
}

__global__ void ws_distribute(const uint32_t* __restrict__ lvl, const unsigned long long* __restrict__ item, uint32_t n,
                              const uint32_t* __restrict__ first, const uint64_t* __restrict__ qoff,
                              const uint32_t* __restrict__ qcount, unsigned long long* __restrict__ qitem,
                              uint32_t* __restrict__ error) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t c = lvl[i];
    const uint64_t slot = qoff[c] + qcount[c] + (i - first[c]);
    if (slot < qoff[c + 1]) qitem[slot] = item[i];
    else atomicOr(error, 1u);
}

// A count never passes its FIFO's capacity (ws_distribute has flagged what did not fit), so the
// rounds of that depth read only its own slots.
__global__ void ws_run_commit(const uint32_t* __restrict__ lvl, uint32_t n, const uint32_t* __restrict__ first,
                              const uint64_t* __restrict__ qoff, uint32_t* __restrict__ qcount) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || (i != n - 1 && lvl[i + 1] == lvl[i])) return;
    const uint32_t c = lvl[i];
    const uint64_t grown = (uint64_t)qcount[c] + (i - first[c] + 1), cap = qoff[c + 1] - qoff[c];
    qcount[c] = (uint32_t)(grown < cap ? grown : cap);
}

__global__ void log_split(const LogEntry* __restrict__ log, size_t n, unsigned long long* __restrict__ t,
                          uint32_t* __restrict__ idx) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    t[i] = log[i].t;
    idx[i] = (uint32_t)i;
}

__global__ void gather_voxel(const LogEntry* __restrict__ log, const uint32_t* __restrict__ perm,
                             uint32_t* __restrict__ key, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) key[i] = log[perm[i]].voxel;
}

__global__ void run_heads(const LogEntry* __restrict__ log, const uint32_t* __restrict__ perm,
                          uint32_t* __restrict__ head, size_t n) {
    size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    head[s] = (s == 0 || log[perm[s]].voxel != log[perm[s - 1]].voxel) ? 1u : 0u;
}

__global__ void run_starts(const uint32_t* __restrict__ head, const uint32_t* __restrict__ hscan,
                           uint32_t* __restrict__ start, size_t n) {
    size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n && head[s]) start[hscan[s]] = (uint32_t)s;
}

// Distinct foreign pores of one voxel in order of first contact.  A voxel is contacted only
// by the popped nodes among its neighbours, each popped once, so it has at most 26 foreign
// pores.
constexpr int kMaxContacts = 26;

__device__ int foreign_list(const LogEntry* __restrict__ log, const uint32_t* __restrict__ perm, uint32_t s0,
                            uint32_t s1, uint32_t owner, uint32_t* pores, unsigned long long* times, int* overflow) {
    int m = 0;
    for (uint32_t s = s0; s < s1; ++s) {
        const LogEntry e = log[perm[s]];
        if (e.pore == owner) continue;
        bool seen = false;
        for (int q = 0; q < m; ++q) seen |= pores[q] == e.pore;
        if (seen) continue;
        if (m == kMaxContacts) {
            *overflow = 1;
            break;
        }
        pores[m] = e.pore;
        times[m] = e.t;
        ++m;
    }
    return m;
}

// Per voxel with contacts: the length of its list [f0, owner, f1, ...] and its formations.
__global__ void count_contacts(const LogEntry* __restrict__ log, const uint32_t* __restrict__ perm,
                               const uint32_t* __restrict__ start, uint32_t runs, size_t n,
                               const uint32_t* __restrict__ label, uint32_t* __restrict__ seq_len,
                               unsigned long long* __restrict__ nrec, uint32_t* __restrict__ overflow) {
    uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= runs) return;
    uint32_t s0 = start[r], s1 = r + 1 < runs ? start[r + 1] : (uint32_t)n;
    uint32_t owner = label[log[perm[s0]].voxel] - 1;
    uint32_t pores[kMaxContacts];
    unsigned long long times[kMaxContacts];
    int of = 0;
    const int m = foreign_list(log, perm, s0, s1, owner, pores, times, &of);
    seq_len[r] = m ? (uint32_t)m + 1 : 0;
    nrec[r] = (unsigned long long)m;
    if (of) atomicOr(overflow, 1u);
}

// Writes the list [f0, owner, f1, f2, ...] of every voxel with contacts; the prefix with j
// foreign pores (length j + 1) is formed at the first contact of f(j-1).
__global__ void write_sequences(const LogEntry* __restrict__ log, const uint32_t* __restrict__ perm,
                                const uint32_t* __restrict__ start, uint32_t runs, size_t n,
                                const uint32_t* __restrict__ label, const unsigned long long* __restrict__ seq_off,
                                const unsigned long long* __restrict__ rbase, uint32_t* __restrict__ elems,
                                unsigned long long* __restrict__ rec_time, uint32_t* __restrict__ weight,
                                uint32_t* __restrict__ voxel, uint32_t* __restrict__ shared_owned) {
    uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= runs) return;
    uint32_t s0 = start[r], s1 = r + 1 < runs ? start[r + 1] : (uint32_t)n;
    uint32_t vox = log[perm[s0]].voxel;
    uint32_t owner = label[vox] - 1;
    uint32_t pores[kMaxContacts];
    unsigned long long times[kMaxContacts];
    int of = 0;
    int m = foreign_list(log, perm, s0, s1, owner, pores, times, &of);
    voxel[r] = vox;
    weight[r] = m ? 1u : 0u;
    if (m == 0) return;
    atomicAdd(&shared_owned[owner], 1u);
    uint32_t* e = elems + seq_off[r];
    e[0] = pores[0];
    e[1] = owner;
    for (int j = 2; j <= m; ++j) e[j] = pores[j - 1];
    for (int j = 1; j <= m; ++j) rec_time[rbase[r] + j - 1] = times[j - 1];
}

__global__ void widen_u32(const uint32_t* __restrict__ in, uint64_t* __restrict__ out, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[i];
}

// Voxels per label.  Neighbouring voxels mostly share a label, so a wave adds its lanes of one
// label with one atomic (a few distinct labels per wave) rather than one atomic per voxel on a
// handful of counters.
__global__ void owned_counts(const uint32_t* __restrict__ label, size_t n, unsigned long long* __restrict__ counts) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t lab = i < n ? label[i] : 0u;  // every lane stays for the wave operations
    const int lane = (int)(threadIdx.x & 63);
    while (true) {
        const unsigned long long pending = __ballot(lab != 0);
        if (!pending) break;
        const int leader = __ffsll((long long)pending) - 1;
        const uint32_t l = (uint32_t)__shfl((int)lab, leader);
        const unsigned long long same = __ballot(lab == l);
        if (lane == leader) atomicAdd(&counts[l - 1], (unsigned long long)__popcll(same));
        if (lab == l) lab = 0;
    }
}

__global__ void boundary_owners(const uint32_t* __restrict__ label, int X, int Y, int Z, uint32_t* __restrict__ flag) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)X * Y * Z) return;
    uint32_t lab = label[i];
    if (!lab) return;
    int x, y, z;
    decode((uint32_t)i, X, Y, x, y, z);
    if (x == 0 || y == 0 || z == 0 || x == X - 1 || y == Y - 1 || z == Z - 1) flag[lab - 1] = 1;
}

__global__ void labels_to_ids(uint32_t* __restrict__ label, size_t n, const uint32_t* __restrict__ pore_id) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && label[i]) label[i] = pore_id[label[i] - 1];
}

__global__ void shared_ids(uint32_t* __restrict__ label, const uint32_t* __restrict__ voxel,
                           const uint32_t* __restrict__ seq_len, const unsigned long long* __restrict__ rbase,
                           const uint32_t* __restrict__ rec_node, const uint32_t* __restrict__ key_id, size_t runs) {
    size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < runs && seq_len[r] >= 2) label[voxel[r]] = key_id[rec_node[rbase[r] + seq_len[r] - 2]];
}

}  // namespace

namespace {
bool watershed_once(pma_ctx* c, unsigned long long log_cap, unsigned long long* needed);
}

void run_watershed(pma_ctx* c) {
    // The contact log is sized from the volume; a run that overflows it is repeated once
    // with the exact size it reported.
    unsigned long long needed = 0;
    if (watershed_once(c, std::max<unsigned long long>(c->n / 2, 1 << 16), &needed)) return;
    PMA_REQUIRE(watershed_once(c, needed + needed / 8 + 1024, &needed), PMA_ERR_INTERNAL,
                "pma_watershed: contact log overflow on the second attempt");
}

namespace {
bool watershed_once(pma_ctx* c, unsigned long long log_cap, unsigned long long* needed) {
    PMA_REQUIRE(c->valid[PMA_VOL_CR2], PMA_ERR_STATE, "pma_watershed: containing radii have not been computed or loaded");
    PMA_REQUIRE(c->have_pores, PMA_ERR_STATE, "pma_watershed: no pores (run pma_pores or set pores.txt)");
    PMA_REQUIRE(c->have_edt_stats, PMA_ERR_STATE, "pma_watershed: no maximum EDT square (intermediate statistics)");
    const int X = c->dims.x, Y = c->dims.y, Z = c->dims.z;
    const size_t n = c->n;
    PMA_REQUIRE(n < 0xFFFFFFFFull, PMA_ERR_INVALID_ARGUMENT, "pma_watershed: volume too large");
    const uint32_t maxE = c->max_e;
    const uint32_t np = (uint32_t)c->pores.size();
    PMA_REQUIRE(np < (1u << 21) - 1, PMA_ERR_INVALID_ARGUMENT, "pma_watershed: too many pores");
    c->ensure_volume(PMA_VOL_PTW);
    uint32_t* label = c->v<uint32_t>(PMA_VOL_PTW);
    const uint32_t* cr = c->v<uint32_t>(PMA_VOL_CR2);
    const int d3q7 = c->stencil == PMA_STENCIL_D3Q7;

    // FIFO capacities: voxels with CR == d plus seeds with r2 == d.
    const size_t L = (size_t)maxE + 1;
    uint32_t* hist = c->buf<uint32_t>("ws_hist", L);
    PMA_HIP(hipMemsetAsync(hist, 0, L * sizeof(uint32_t), c->stream));
    cr_histogram<<<1024, 256, 0, c->stream>>>(cr, n, hist, (uint32_t)L);
    PMA_LAUNCHED(c);
    std::vector<uint32_t> h(L);
    PMA_HIP(hipMemcpyAsync(h.data(), hist, L * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    PMA_HIP(hipStreamSynchronize(c->stream));
    std::vector<uint32_t> seed_voxel(np), pore_id(np), seeds_at(L, 0);
    for (uint32_t k = 0; k < np; ++k) {
        const PoreRow& p = c->pores[k];
        PMA_REQUIRE(p.x >= 0 && p.x < X && p.y >= 0 && p.y < Y && p.z >= 0 && p.z < Z, PMA_ERR_INVALID_ARGUMENT,
                    "pma_watershed: a pore seed lies outside the image");
        PMA_REQUIRE(p.r2 <= maxE, PMA_ERR_INVALID_ARGUMENT, "pma_watershed: a seed radius exceeds the maximum EDT square");
        seed_voxel[k] = (uint32_t)(((size_t)p.z * X + p.x) * Y + p.y);
        pore_id[k] = p.id;
        h[p.r2] += 1;
    }
    std::vector<uint64_t> qoff_h(L + 1, 0);
    for (size_t d = 0; d < L; ++d) qoff_h[d + 1] = qoff_h[d] + h[d];
    const uint64_t qtotal = qoff_h[L];
    PMA_REQUIRE(qtotal < 0xFFFFFFFFull, PMA_ERR_CAPACITY, "pma_watershed: 2^32 or more FIFO entries (void voxels and seeds)");
    // the initial pushes: each level's seeds first in its FIFO, in ascending k
    std::vector<uint64_t> seed_slot(np);
    for (uint32_t k = 0; k < np; ++k) seed_slot[k] = qoff_h[c->pores[k].r2] + seeds_at[c->pores[k].r2]++;
    uint64_t* qoff = c->buf<uint64_t>("ws_qoff", L + 1);
    uint32_t* qcount = c->buf<uint32_t>("ws_qcount", L);
    unsigned long long* qitem = c->buf<unsigned long long>("ws_qitem", std::max<uint64_t>(qtotal, 1));
    upload(c, qoff, qoff_h.data(), (L + 1) * sizeof(uint64_t));
    upload(c, qcount, seeds_at.data(), L * sizeof(uint32_t));
    // a level's lower-level pushes, staged in push time order and then sorted by target depth
    const uint32_t stage_cap = (uint32_t)std::max<uint64_t>(qtotal, 1);
    unsigned long long* st_item = c->buf<unsigned long long>("ws_stage_item", stage_cap);
    uint32_t* st_lvl = c->buf<uint32_t>("ws_stage_level", stage_cap);
    unsigned long long* st_item2 = c->buf<unsigned long long>("ws_stage_item2", stage_cap);
    uint32_t* st_lvl2 = c->buf<uint32_t>("ws_stage_level2", stage_cap);
    uint32_t* run_first = c->buf<uint32_t>("ws_run_first", L);
    PMA_HIP(hipMemsetAsync(label, 0, n * sizeof(uint32_t), c->stream));
    unsigned long long* claim = c->buf<unsigned long long>("ws_claim", n);
    PMA_HIP(hipMemsetAsync(claim, 0xFF, n * sizeof(unsigned long long), c->stream));
    uint32_t* d_seed_voxel = c->buf<uint32_t>("ws_seed_voxel", std::max(np, 1u));
    uint64_t* d_seed_slot = c->buf<uint64_t>("ws_seed_slot", std::max(np, 1u));
    if (np) {
        upload(c, d_seed_voxel, seed_voxel.data(), np * 4ull);
        upload(c, d_seed_slot, seed_slot.data(), np * 8ull);
        push_seeds<<<grid_for(np, 256), 256, 0, c->stream>>>(d_seed_voxel, d_seed_slot, np, label, qitem);
        PMA_LAUNCHED(c);
    }
    LogEntry* log = c->buf<LogEntry>("ws_log", log_cap);
    unsigned long long* log_count = c->buf<unsigned long long>("ws_log_count", 1);
    PMA_HIP(hipMemsetAsync(log_count, 0, sizeof(unsigned long long), c->stream));
    uint32_t* ws_err = c->buf<uint32_t>("ws_error", 1);
    PMA_HIP(hipMemsetAsync(ws_err, 0, 4, c->stream));
    uint32_t* same_mask = c->buf<uint32_t>("ws_same_mask", std::max<uint64_t>(qtotal, 1));
    uint32_t* lower_mask = c->buf<uint32_t>("ws_lower_mask", std::max<uint64_t>(qtotal, 1));
    uint32_t* same_cnt = c->buf<uint32_t>("ws_same_cnt", std::max<uint64_t>(qtotal, 1));
    uint32_t* part = c->buf<uint32_t>("ws_parts", kWsGrid);
    uint32_t* part_low = c->buf<uint32_t>("ws_parts_low", kWsGrid);
    WsRound* rstate = c->buf<WsRound>("ws_round_state", 2);

    uint64_t rounds = 0;
    uint32_t gr = 0;  // global round counter for claim stamps
    for (long d = (long)maxE; d >= 1; --d) {
        if (h[d] == 0) continue;  // no voxel and no seed at this depth: its FIFO stays empty
        unsigned long long* items = qitem + qoff_h[d];
        const unsigned long long tlevel = (unsigned long long)(maxE - d + 1) << 40;
        // the level's rounds run back to back on the device, each reading its bounds from the
        // previous one's; the host looks for an empty round only after rounds 1, 2, 4, 8 and
        // then every 8 (the rounds after the last are empty and change nothing)
        ws_level_start<<<1, 64, 0, c->stream>>>(qcount + d, rstate);
        PMA_LAUNCHED(c);
        const uint32_t g = std::min(grid_for(h[d], 256), kWsGrid);
        // ws_bid and ws_resolve: grid-stride over the lanes on at most c->flood_grid workgroups
        const uint32_t gi = std::min(grid_for((size_t)h[d] << (d3q7 ? 3 : 5), 256), c->flood_grid);
        uint32_t r = 0, check_at = 1, step = 1, staged = 0;
        while (true) {
            PMA_REQUIRE(gr < (1u << 24) - 1, PMA_ERR_INTERNAL, "pma_watershed: too many rounds");
            const unsigned long long hi = (unsigned long long)((1u << 24) - 1 - gr) << 40;
            ++gr;
            ++rounds;
            const WsRound* cur = rstate + (r & 1);
            WsRound* nxt = rstate + ((r + 1) & 1);
            {
                PMA_TIMED(c, "ws_bid");
                ws_bid<<<gi, 256, 0, c->stream>>>(items, cur, cr, label, claim, hi, (uint32_t)d, X, Y, Z, d3q7);
                PMA_LAUNCHED(c);
            }
            {
                PMA_TIMED(c, "ws_resolve");
                ws_resolve<<<gi, 256, 0, c->stream>>>(items, cur, cr, label, claim, hi, tlevel, (uint32_t)d, X, Y, Z,
                                                     d3q7, same_mask, lower_mask, same_cnt, log, log_count, log_cap);
                PMA_LAUNCHED(c);
            }
            ws_count_parts<<<g, 256, 0, c->stream>>>(cur, same_cnt, part, part_low);
            PMA_LAUNCHED(c);
            ws_scan_parts<<<1, 256, 0, c->stream>>>(cur, nxt, part, part_low, g, h[d], stage_cap, ws_err);
            PMA_LAUNCHED(c);
            ws_append<<<g, 256, 0, c->stream>>>(items, cur, part, part_low, same_mask, lower_mask, same_cnt, h[d], cr,
                                                st_item, st_lvl, stage_cap, X, Y);
            PMA_LAUNCHED(c);
            ++r;
            if (r == check_at) {
                const WsRound next = read_scalar(c, nxt);
                if (next.a == next.b) {
                    staged = next.staged;
                    break;
                }
                step = std::min(step * 2, 8u);
                check_at += step;
            }
        }
        if (staged == 0) continue;
        // the level's pushes into lower depths join their FIFOs in push time order
        PMA_TIMED(c, "ws_distribute");
        const uint32_t* lv = st_lvl;
        const unsigned long long* iv = st_item;
        if (staged > 1) {
            int end_bit = 1;
            while (end_bit < 32 && ((unsigned long)d - 1) >> end_bit) ++end_bit;
            radix_sort_pairs_u32_u64(c, st_lvl, st_lvl2, reinterpret_cast<const uint64_t*>(st_item),
                                     reinterpret_cast<uint64_t*>(st_item2), staged, 0, end_bit);
            lv = st_lvl2;
            iv = st_item2;
        }
        const uint32_t gs = grid_for(staged, 256);
        ws_run_first<<<gs, 256, 0, c->stream>>>(lv, staged, run_first);
        PMA_LAUNCHED(c);
        ws_distribute<<<gs, 256, 0, c->stream>>>(lv, iv, staged, run_first, qoff, qcount, qitem, ws_err);
        PMA_LAUNCHED(c);
        ws_run_commit<<<gs, 256, 0, c->stream>>>(lv, staged, run_first, qoff, qcount);
        PMA_LAUNCHED(c);
    }
    // a FIFO or the staging list that filled up was flagged by ws_scan_parts or ws_distribute
    // (neither they nor ws_append write past it)
    PMA_REQUIRE(read_scalar(c, ws_err) == 0, PMA_ERR_INTERNAL, "pma_watershed: FIFO capacity exceeded");
    const unsigned long long nlog = read_scalar(c, log_count);
    *needed = nlog;
    if (nlog > log_cap) return false;

    // Order the log by (voxel, time): stable sort by time, then by voxel.
    std::vector<ThroatRow> throats;
    uint32_t* shared_owned = c->buf<uint32_t>("ws_shared_owned", std::max(np, 1u));
    PMA_HIP(hipMemsetAsync(shared_owned, 0, std::max(np, 1u) * sizeof(uint32_t), c->stream));
    if (nlog) {
        unsigned long long* tk = c->buf<unsigned long long>("ws_log_t", nlog);
        unsigned long long* tk2 = c->buf<unsigned long long>("ws_log_t2", nlog);
        uint32_t* perm = c->buf<uint32_t>("ws_log_perm", nlog);
        uint32_t* perm2 = c->buf<uint32_t>("ws_log_perm2", nlog);
        uint32_t* vk = c->buf<uint32_t>("ws_log_vk", nlog);
        uint32_t* vk2 = c->buf<uint32_t>("ws_log_vk2", nlog);
        log_split<<<grid_for(nlog, 256), 256, 0, c->stream>>>(log, nlog, tk, perm);
        PMA_LAUNCHED(c);
        radix_sort_pairs_u64(c, reinterpret_cast<uint64_t*>(tk), reinterpret_cast<uint64_t*>(tk2), perm, perm2, nlog);
        gather_voxel<<<grid_for(nlog, 256), 256, 0, c->stream>>>(log, perm2, vk, nlog);
        PMA_LAUNCHED(c);
        radix_sort_pairs_u32(c, vk, vk2, perm2, perm, nlog);
        uint32_t* head = c->buf<uint32_t>("ws_head", nlog);
        uint32_t* hscan = c->buf<uint32_t>("ws_hscan", nlog);
        run_heads<<<grid_for(nlog, 256), 256, 0, c->stream>>>(log, perm, head, nlog);
        PMA_LAUNCHED(c);
        uint32_t runs = (uint32_t)exclusive_scan_u32(c, head, hscan, nlog);
        uint32_t* start = c->buf<uint32_t>("ws_run_start", runs);
        run_starts<<<grid_for(nlog, 256), 256, 0, c->stream>>>(head, hscan, start, nlog);
        PMA_LAUNCHED(c);
        uint32_t* slen = c->buf<uint32_t>("ws_seq_len", runs);
        unsigned long long* nrec = c->buf<unsigned long long>("ws_nrec", runs);
        unsigned long long* soff = c->buf<unsigned long long>("ws_seq_off", runs);
        unsigned long long* rbase = c->buf<unsigned long long>("ws_rbase", runs);
        uint32_t* of = c->buf<uint32_t>("ws_overflow", 1);
        PMA_HIP(hipMemsetAsync(of, 0, 4, c->stream));
        count_contacts<<<grid_for(runs, 256), 256, 0, c->stream>>>(log, perm, start, runs, nlog, label, slen, nrec, of);
        PMA_LAUNCHED(c);
        PMA_REQUIRE(read_scalar(c, of) == 0, PMA_ERR_INTERNAL, "pma_watershed: a voxel has more than 26 contacts");
        uint64_t* slen64 = c->buf<uint64_t>("ws_seq_len64", runs);
        widen_u32<<<grid_for(runs, 256), 256, 0, c->stream>>>(slen, slen64, runs);
        PMA_LAUNCHED(c);
        const uint64_t nelem = exclusive_scan_u64(c, slen64, reinterpret_cast<uint64_t*>(soff), runs);
        const uint64_t R = exclusive_scan_u64(c, reinterpret_cast<const uint64_t*>(nrec),
                                              reinterpret_cast<uint64_t*>(rbase), runs);
        if (R) {
            uint32_t* elems = c->buf<uint32_t>("ws_elems", nelem);
            unsigned long long* tm = c->buf<unsigned long long>("ws_tm", R);
            uint32_t* wt = c->buf<uint32_t>("ws_weight", runs);
            uint32_t* vox = c->buf<uint32_t>("ws_vox", runs);
            uint32_t* rd = c->buf<uint32_t>("ws_rd", R);
            write_sequences<<<grid_for(runs, 256), 256, 0, c->stream>>>(log, perm, start, runs, nlog, label, soff, rbase,
                                                                        elems, tm, wt, vox, shared_owned);
            PMA_LAUNCHED(c);
            PrefixTrie lt = build_prefix_trie(c, elems, nelem, reinterpret_cast<const uint64_t*>(soff), slen,
                                              reinterpret_cast<const uint64_t*>(rbase),
                                              reinterpret_cast<const uint64_t*>(tm), wt, runs, R, rd);
            std::vector<uint32_t> order(lt.distinct);
            for (uint32_t i = 0; i < lt.distinct; ++i) order[i] = i;
            std::sort(order.begin(), order.end(), [&](uint32_t a2, uint32_t b2) { return lt.min_time[a2] < lt.min_time[b2]; });
            std::vector<uint32_t> key_id(lt.distinct);
            throats.resize(lt.distinct);
            for (uint32_t r = 0; r < lt.distinct; ++r) {
                uint32_t dix = order[r];
                key_id[dix] = 0xFFFFFFFFu - 1u - r;
                ThroatRow& t = throats[lt.distinct - 1 - r];
                t.id = key_id[dix];
                t.volume = lt.finals[dix];
                for (uint32_t q : lt.members(dix)) t.pores.push_back(pore_id[q]);
            }
            // owners first, then the shared voxels
            uint32_t* d_pid = c->buf<uint32_t>("ws_pore_id", std::max(np, 1u));
            uint32_t* d_kid = c->buf<uint32_t>("ws_key_id", lt.distinct);
            upload(c, d_pid, pore_id.data(), np * 4ull);
            upload(c, d_kid, key_id.data(), lt.distinct * 4ull);
            // volumes and boundary flags need the pore-index labels: compute them before relabelling
            c->pores_ws = c->pores;
            {
                unsigned long long* owned = c->buf<unsigned long long>("ws_owned", std::max(np, 1u));
                uint32_t* bflag = c->buf<uint32_t>("ws_bflag", std::max(np, 1u));
                PMA_HIP(hipMemsetAsync(owned, 0, std::max(np, 1u) * 8ull, c->stream));
                PMA_HIP(hipMemsetAsync(bflag, 0, std::max(np, 1u) * 4ull, c->stream));
                owned_counts<<<grid_for(n, 256), 256, 0, c->stream>>>(label, n, owned);
                PMA_LAUNCHED(c);
                boundary_owners<<<grid_for(n, 256), 256, 0, c->stream>>>(label, X, Y, Z, bflag);
                PMA_LAUNCHED(c);
                std::vector<unsigned long long> ow(np);
                std::vector<uint32_t> bf(np), so(np);
                PMA_HIP(hipMemcpyAsync(ow.data(), owned, np * 8ull, hipMemcpyDeviceToHost, c->stream));
                PMA_HIP(hipMemcpyAsync(bf.data(), bflag, np * 4ull, hipMemcpyDeviceToHost, c->stream));
                PMA_HIP(hipMemcpyAsync(so.data(), shared_owned, np * 4ull, hipMemcpyDeviceToHost, c->stream));
                PMA_HIP(hipStreamSynchronize(c->stream));
                for (uint32_t k = 0; k < np; ++k) {
                    c->pores_ws[k].volume = ow[k] - so[k];
                    c->pores_ws[k].boundary = bf[k] != 0;
                }
            }
            labels_to_ids<<<grid_for(n, 256), 256, 0, c->stream>>>(label, n, d_pid);
            PMA_LAUNCHED(c);
            shared_ids<<<grid_for(runs, 256), 256, 0, c->stream>>>(label, vox, slen, rbase, rd, d_kid, runs);
            PMA_LAUNCHED(c);
            PMA_HIP(hipStreamSynchronize(c->stream));
            c->throats_ws = throats;
            c->have_pores_ws = true;
            c->valid[PMA_VOL_PTW] = true;
            c->stage[PMA_STAGE_WATERSHED].work_items = rounds;
            return true;
        }
    }
    // no shared voxels
    c->pores_ws = c->pores;
    {
        unsigned long long* owned = c->buf<unsigned long long>("ws_owned", std::max(np, 1u));
        uint32_t* bflag = c->buf<uint32_t>("ws_bflag", std::max(np, 1u));
        PMA_HIP(hipMemsetAsync(owned, 0, std::max(np, 1u) * 8ull, c->stream));
        PMA_HIP(hipMemsetAsync(bflag, 0, std::max(np, 1u) * 4ull, c->stream));
        owned_counts<<<grid_for(n, 256), 256, 0, c->stream>>>(label, n, owned);
        PMA_LAUNCHED(c);
        boundary_owners<<<grid_for(n, 256), 256, 0, c->stream>>>(label, X, Y, Z, bflag);
        PMA_LAUNCHED(c);
        std::vector<unsigned long long> ow(std::max(np, 1u));
        std::vector<uint32_t> bf(std::max(np, 1u));
        PMA_HIP(hipMemcpyAsync(ow.data(), owned, np * 8ull, hipMemcpyDeviceToHost, c->stream));
        PMA_HIP(hipMemcpyAsync(bf.data(), bflag, np * 4ull, hipMemcpyDeviceToHost, c->stream));
        PMA_HIP(hipStreamSynchronize(c->stream));
        for (uint32_t k = 0; k < np; ++k) {
            c->pores_ws[k].volume = ow[k];
            c->pores_ws[k].boundary = bf[k] != 0;
        }
        uint32_t* d_pid = c->buf<uint32_t>("ws_pore_id", std::max(np, 1u));
        upload(c, d_pid, pore_id.data(), np * 4ull);
        labels_to_ids<<<grid_for(n, 256), 256, 0, c->stream>>>(label, n, d_pid);
        PMA_LAUNCHED(c);
        PMA_HIP(hipStreamSynchronize(c->stream));
    }
    c->throats_ws.clear();
    c->have_pores_ws = true;
    c->valid[PMA_VOL_PTW] = true;
    c->stage[PMA_STAGE_WATERSHED].work_items = rounds;
    return true;
}
}  // namespace

}  // namespace pma
