// pma_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of the reference (VasiliBaranov/PorousMediaAnalysis) algorithms on
// the pore-throat + geodesic path, with the reference's single-thread semantics.  Only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this library,
// and only as the checker or the timed CPU baseline -- never as part of the product.
//
// Layout everywhere: voxel (x, y, z) lives at ((z * X) + x) * Y + y, i.e. the reference's
// [z][x][y] (ImageProcessing/Model/Headers/DataArray.h:85-94), X = TIFF rows,
// Y = TIFF columns, Z = slice files.
//
// Tables cross this API as the reference's own text formats
// (ImageProcessing/Services/Source/Serializer.cpp:139-221), so a stage can be fed the
// previous stage's text exactly like the reference feeds it the previous stage's files.
//
// Parity pinning (what checks each function; tests/test_oracle_example.py and
// tests/test_oracle_bruteforce.py):
//   oracle_edt              : Docs/Example EuclideanDistanceSquares bit-exact, histogram and
//                             max file bit-exact; brute force on tiny volumes; scipy EDT of
//                             an x-padded array.
//   oracle_containing_balls : Docs/Example ContainingBallRadiiSquares bit-exact; brute force
//                             (max over covering balls, lexicographic centre) on tiny volumes.
//   oracle_pores            : Docs/Example PoreThroatIds partition (up to relabelling, fed the
//                             fixture's own mask/EDT/CR); on tiny volumes, both stencils: seeds =
//                             the first libstdc++-sorted maximum ball of each regional-maximum CR
//                             plateau (scipy labels) in pore-ID order; every voxel's list = the
//                             pores whose seed reaches it by BFS, in [second, first, ...] order;
//                             volumes and boundary flags by the same reachability; the DFS order
//                             of key numbering on a row worked out by hand.  The sort itself is
//                             libstdc++'s std::sort.
//   oracle_pores_in_order   : the same loop in a given ball order.  Docs/Example PoreThroatIds,
//                             pores.txt and sharedPixelTypes.txt bit-exact, throat-key IDs
//                             included, in the fixture's seed order (the fixture was made with
//                             MSVC's std::sort, whose tie order differs from libstdc++'s).
//   oracle_remove_empty_pores : a hand-made table with known answers (drop, fold into the one
//                             remaining pore, drop without remap, merge duplicates into the
//                             largest ID); the Docs/Example check above runs through it.
//   oracle_watershed        : Docs/Example PoreThroatIdsWatershed, poresWatershed.txt and
//                             sharedPixelTypesWatershed.txt bit-exact from the fixture's
//                             pores.txt; the FIFO order and throat-key order on a row worked
//                             out by hand, both stencils.
//   oracle_geodesic         : Docs/Example ShortestPathDistances + CavityIds bit-exact,
//                             cavityDescriptions.txt; float32 Bellman-Ford from every cavity's
//                             seed on tiny volumes with up to 4 cavities, both stencils; box-shell
//                             restarts worked out by hand (shell and split volumes).
#include <algorithm>
#include <array>
#include <iterator>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <queue>
#include <stack>
#include <string>
#include <vector>
#include <omp.h>

namespace {

struct Grid {
    int X, Y, Z;
    size_t idx(int x, int y, int z) const { return ((size_t)z * X + x) * (size_t)Y + y; }
    size_t count() const { return (size_t)X * Y * Z; }
    bool outer(int x, int y, int z) const {
        return x < 0 || x >= X || y < 0 || y >= Y || z < 0 || z >= Z;
    }
};

// ModelUtilities.h:22-29: a stencil vector is diagonal when two or more axes change.
inline bool diagonal(int dx, int dy, int dz) {
    return (dx != 0 && dy != 0) || (dx != 0 && dz != 0) || (dy != 0 && dz != 0);
}

char* dup_string(const std::string& s) {
    char* p = (char*)std::malloc(s.size() + 1);
    std::memcpy(p, s.c_str(), s.size() + 1);
    return p;
}

// ---------------------------------------------------------------------------------------
// Tables (Types.h:157-180) and their text forms (Serializer.cpp:80-137, 187-221).
struct Pore {
    uint32_t id = 0;
    int x = 0, y = 0, z = 0;
    uint32_t r2 = 0;
    size_t volume = 0;
    bool boundary = false;
};
struct Key {
    uint32_t id = 0;
    std::vector<uint32_t> pores;
    size_t volume = 0;
};

std::string pores_text(const std::map<uint32_t, Pore>& pores) {
    std::string s = "PoreId SeedX SeedY SeedZ SeedRadiusSquare IsBoundaryPore Volume\n";
    char buf[160];
    for (const auto& kv : pores) {
        const Pore& p = kv.second;
        std::snprintf(buf, sizeof buf, "%u %d %d %d %d %d %zu\n", p.id, p.x, p.y, p.z, (int)p.r2,
                      p.boundary ? 1 : 0, p.volume);
        s += buf;
    }
    return s;
}

std::string keys_text(const std::map<uint32_t, Key>& keys) {
    std::string s = "SharedPixelTypeId; Volume; PoreId1 PoreId2 ... \n";
    char buf[64];
    for (const auto& kv : keys) {
        std::snprintf(buf, sizeof buf, "%u; %zu;", kv.second.id, kv.second.volume);
        s += buf;
        for (uint32_t p : kv.second.pores) {
            std::snprintf(buf, sizeof buf, " %u", p);
            s += buf;
        }
        s += "\n";
    }
    return s;
}

// Serializer::ReadPores (Serializer.cpp:103-128): fscanf until a line does not parse.
std::map<uint32_t, Pore> parse_pores(const char* text) {
    std::map<uint32_t, Pore> out;
    const char* nl = std::strchr(text, '\n');
    const char* p = nl ? nl + 1 : text + std::strlen(text);
    while (*p) {
        unsigned id; int x, y, z, r2, b; size_t vol; int used = 0;
        if (std::sscanf(p, "%u %d %d %d %d %d %zu%n", &id, &x, &y, &z, &r2, &b, &vol, &used) != 7)
            break;
        Pore q; q.id = id; q.x = x; q.y = y; q.z = z; q.r2 = (uint32_t)r2; q.boundary = b != 0;
        q.volume = vol;
        out[id] = q;
        p += used;
        while (*p == '\n' || *p == '\r' || *p == ' ') ++p;
    }
    return out;
}

// Serializer::ReadSharedPixelTypes (Serializer.cpp:66-101).
std::map<uint32_t, Key> parse_keys(const char* text) {
    std::map<uint32_t, Key> out;
    const char* nl = std::strchr(text, '\n');
    const char* p = nl ? nl + 1 : text + std::strlen(text);
    while (*p) {
        const char* end = std::strchr(p, '\n');
        std::string line = end ? std::string(p, end) : std::string(p);
        p = end ? end + 1 : p + line.size();
        if (line.empty() || line == "\r") continue;
        Key k;
        unsigned id; size_t vol; int used = 0;
        if (std::sscanf(line.c_str(), "%u; %zu;%n", &id, &vol, &used) != 2) break;
        k.id = id; k.volume = vol;
        const char* q = line.c_str() + used;
        unsigned pid; int n = 0;
        while (std::sscanf(q, " %u%n", &pid, &n) == 1) { k.pores.push_back(pid); q += n; }
        out[id] = k;
    }
    return out;
}

// BasePoreAndThroatComputer state: pores, keys (shared pixel types), the list->id index and
// the counter of key IDs that counts down from UINT_MAX (BasePoreAndThroatComputer.cpp:28-34).
struct Labeller {
    std::map<uint32_t, Pore> pores;
    std::map<uint32_t, Key> keys;
    std::map<std::vector<uint32_t>, uint32_t> key_by_list;
    uint32_t first_shared = UINT32_MAX;

    // GetOrAddSharedPixelType (BasePoreAndThroatComputer.cpp:112-133)
    Key& get_or_add(const std::vector<uint32_t>& list) {
        auto it = key_by_list.find(list);
        if (it != key_by_list.end()) return keys[it->second];
        --first_shared;
        Key& k = keys[first_shared];
        k.id = first_shared; k.volume = 0; k.pores = list;
        key_by_list[list] = first_shared;
        return k;
    }

    // UpdatePixelAndSharedBalls (BasePoreAndThroatComputer.cpp:37-110); returns "continue".
    bool update(uint32_t* labels, size_t i, Pore& pore, bool continue_after_shared) {
        uint32_t cur = labels[i];
        if (cur == pore.id) return false;
        if (cur == 0) { labels[i] = pore.id; pore.volume++; return true; }
        if (cur < first_shared) {
            std::vector<uint32_t> list{pore.id, cur};
            Key& k = get_or_add(list);
            labels[i] = k.id; k.volume++;
            pores[cur].volume--;
            return continue_after_shared;
        }
        Key& old = keys[cur];
        if (std::find(old.pores.begin(), old.pores.end(), pore.id) != old.pores.end()) return false;
        std::vector<uint32_t> list = old.pores;
        list.push_back(pore.id);
        Key& k = get_or_add(list);
        labels[i] = k.id;
        old.volume--;
        k.volume++;
        return continue_after_shared;
    }
};

}  // namespace

extern "C" {

void oracle_free(void* p) { std::free(p); }

// std::sort of u64 records by their high 32 bits, descending: the comparator shape of
// MaximumBallComparer (Types.h:189-195) applied through StlUtilities::Sort (StlUtilities.h:95-98).
void oracle_sort_desc(uint64_t* a, size_t n) {
    std::sort(a, a + n, [](uint64_t x, uint64_t y) { return (x >> 32) > (y >> 32); });
}

// ---------------------------------------------------------------------------------------
// EuclideanDistanceComputer (EuclideanDistanceComputer.cpp:149-343): X two-scan with the
// start value X-1 (:164,:180), Y and Z brute-force windows of floor(sqrt(g))+1 (:234,:295),
// max and the histogram of squares whose ball does not cross the image boundary (:310-327).
// hist receives counts[0..maxE]; *hist_len = maxE + 1 (returns 1 if hist_cap is too small).
int oracle_edt(const uint8_t* solid, int X, int Y, int Z, uint32_t* e, uint32_t* max_e,
               uint64_t* hist, size_t hist_cap, size_t* hist_len, int threads) {
    Grid g{X, Y, Z};
    if (threads < 1) threads = 1;
#pragma omp parallel for schedule(static) num_threads(threads)
    for (int y = 0; y < Y; ++y) {
        for (int z = 0; z < Z; ++z) {
            uint32_t d = (uint32_t)(X - 1);
            for (int x = 0; x < X; ++x) {
                size_t i = g.idx(x, y, z);
                d = solid[i] == 0 ? d + 1 : 0;
                e[i] = d * d;
            }
            d = (uint32_t)(X - 1);
            for (int x = X - 1; x >= 0; --x) {
                size_t i = g.idx(x, y, z);
                d = solid[i] == 0 ? d + 1 : 0;
                uint32_t b = d * d;
                if (b < e[i]) e[i] = b;
            }
        }
    }
#pragma omp parallel for schedule(static) num_threads(threads)
    for (int x = 0; x < X; ++x) {
        std::vector<uint32_t> line(Y);
        for (int z = 0; z < Z; ++z) {
            for (int y = 0; y < Y; ++y) line[y] = e[g.idx(x, y, z)];
            for (int y = 0; y < Y; ++y) {
                uint32_t best = line[y];
                if (best == 0) continue;
                int w = (int)(std::sqrt((double)best) + 1);
                int lo = std::min(w, y), hi = std::min(w, Y - y);
                for (int dy = -lo; dy < hi; ++dy) {
                    uint32_t c = line[y + dy] + (uint32_t)(dy * dy);
                    if (c < best) best = c;
                }
                e[g.idx(x, y, z)] = best;
            }
        }
    }
    std::vector<std::vector<uint64_t>> hists(threads);
    std::vector<uint32_t> maxes(threads, 0);
#pragma omp parallel for schedule(static) num_threads(threads)
    for (int x = 0; x < X; ++x) {
        int t = omp_get_thread_num();
        std::vector<uint64_t>& h = hists[t];
        uint32_t& m = maxes[t];
        std::vector<uint32_t> line(Z);
        for (int y = 0; y < Y; ++y) {
            for (int z = 0; z < Z; ++z) line[z] = e[g.idx(x, y, z)];
            for (int z = 0; z < Z; ++z) {
                uint32_t best = line[z];
                if (best == 0) continue;
                int w = (int)(std::sqrt((double)best) + 1);
                int lo = std::min(w, z), hi = std::min(w, Z - z);
                for (int dz = -lo; dz < hi; ++dz) {
                    uint32_t c = line[z + dz] + (uint32_t)(dz * dz);
                    if (c < best) best = c;
                }
                e[g.idx(x, y, z)] = best;
                if (m < best) { m = best; h.resize((size_t)m + 1, 0); }
                double r = std::sqrt((double)best);
                bool cross = r > x || r > y || r > z || x + r > X - 1 || y + r > Y - 1 || z + r > Z - 1;
                if (!cross) h[best]++;
            }
        }
    }
    uint32_t m = 0;
    for (uint32_t v : maxes) m = std::max(m, v);
    *max_e = m;
    *hist_len = (size_t)m + 1;
    if (hist_cap < (size_t)m + 1) return 1;
    std::fill(hist, hist + (size_t)m + 1, 0);
    for (const auto& h : hists)
        for (size_t i = 0; i < h.size(); ++i) hist[i] += h[i];
    return 0;
}

// EuclideanDistanceComputer::WriteEuclideanDistanceSquareCounts (:83-110) and
// Serializer::WriteEuclideanDistanceSquareCounts (Serializer.cpp:148-159).
char* oracle_histogram_text(const uint64_t* hist, size_t len) {
    std::vector<uint64_t> c(hist, hist + len);
    long last = -1;
    for (long i = (long)c.size() - 1; i >= 0; --i)
        if (c[i] != 0) { last = i; break; }
    if (last == -1) c.resize(1, 0);
    else c.resize((size_t)last + 2, 0);
    std::string s = "EuclideanDistanceSquaresCount (line index is EuclideanDistanceSquare, the next line has index one)\n";
    char buf[32];
    for (size_t i = 1; i < c.size(); ++i) {
        std::snprintf(buf, sizeof buf, "%zu\n", (size_t)c[i]);
        s += buf;
    }
    return dup_string(s);
}

// ---------------------------------------------------------------------------------------
// ContainingBallsAssigner (ContainingBallsAssigner.cpp:101-250) with one thread: centres in
// x, y, z order, the cube of half-size (int)sqrtf(r2)+1 (:142-144, the float table of
// EuclideanDistanceDerivativesComputer.cpp:33-36), strict '<' so the first centre in scan order
// wins among equal radii (:170).  With threads > 1 the same result is produced by a 64-bit
// compare-and-swap max on (r2, ~scan index), which is the single-thread rule restated.
int oracle_containing_balls(const uint32_t* e, int X, int Y, int Z, uint32_t* cr, uint32_t* cx,
                            uint32_t* cy, uint32_t* cz, uint8_t* mask, uint64_t* n_max,
                            int threads) {
    Grid g{X, Y, Z};
    size_t n = g.count();
    if (threads < 1) threads = 1;
    std::vector<uint64_t> key(n, 0);
    auto scan_index = [&](int x, int y, int z) -> uint64_t {
        return ((uint64_t)x * Y + y) * (uint64_t)Z + z;
    };
#pragma omp parallel for schedule(static) num_threads(threads)
    for (int x = 0; x < X; ++x)
        for (int y = 0; y < Y; ++y)
            for (int z = 0; z < Z; ++z) {
                uint32_t r2 = e[g.idx(x, y, z)];
                if (r2 == 0) continue;
                int R = (int)(std::sqrt((float)r2) + 1.0f);
                int x0 = std::max(0, x - R), x1 = std::min(X - 1, x + R);
                int y0 = std::max(0, y - R), y1 = std::min(Y - 1, y + R);
                int z0 = std::max(0, z - R), z1 = std::min(Z - 1, z + R);
                uint64_t k = ((uint64_t)r2 << 32) | (uint32_t)~(uint32_t)scan_index(x, y, z);
                for (int ix = x0; ix <= x1; ++ix)
                    for (int iy = y0; iy <= y1; ++iy)
                        for (int iz = z0; iz <= z1; ++iz) {
                            uint32_t d2 = (uint32_t)((ix - x) * (ix - x) + (iy - y) * (iy - y) +
                                                     (iz - z) * (iz - z));
                            if (d2 > r2) continue;
                            size_t i = g.idx(ix, iy, iz);
                            if (e[i] == 0) continue;
                            if (threads == 1) {
                                if ((uint32_t)(key[i] >> 32) < r2) key[i] = k;
                            } else {
                                uint64_t cur = __atomic_load_n(&key[i], __ATOMIC_RELAXED);
                                while (cur < k && !__atomic_compare_exchange_n(&key[i], &cur, k, true,
                                                                                __ATOMIC_RELAXED,
                                                                                __ATOMIC_RELAXED)) {
                                }
                            }
                        }
            }
    size_t count = 0;
    for (size_t i = 0; i < n; ++i) {
        if (key[i] == 0) {
            cr[i] = 0; cx[i] = cy[i] = cz[i] = 0xFFFFFFFFu; mask[i] = 0;
        } else {
            cr[i] = (uint32_t)(key[i] >> 32);
            uint64_t s = (uint32_t)~(uint32_t)key[i];
            cz[i] = (uint32_t)(s % Z); s /= Z;
            cy[i] = (uint32_t)(s % Y); s /= Y;
            cx[i] = (uint32_t)s;
            mask[i] = 127;
        }
    }
    // SetMaximumBallsMaskInActiveArea (:190-250): every void voxel marks its centre.
    for (size_t i = 0; i < n; ++i) {
        if (key[i] == 0) continue;
        size_t c = g.idx((int)cx[i], (int)cy[i], (int)cz[i]);
        if (mask[c] == 127) { mask[c] = 255; ++count; }
    }
    *n_max = count;
    return 0;
}

// ---------------------------------------------------------------------------------------
// PoreAndThroatComputer (PoreAndThroatComputer.cpp:81-269) with one active area:
// maximum balls in z, x, y order sorted by r2 descending with std::sort (:93-123), pores
// opened at still-unlabelled centres with IDs 127, +1 (+127 after a multiple of 255)
// (:125-163), DFS over non-increasing containing radii (:192-269), empty keys erased after
// every pore (:168-189).
namespace {
struct Ball { int x, y, z; uint32_t r2; };

// The pore loop (:125-269) over maximum balls already in processing order.
void pores_from_balls(const std::vector<Ball>& balls, const uint32_t* cr, const Grid& g, int d3q7,
                      uint32_t* pt, char** pores_out, char** keys_out) {
    std::fill(pt, pt + g.count(), 0u);
    Labeller L;
    uint32_t pore_id = 127;
    for (const Ball& b : balls) {
        if (pt[g.idx(b.x, b.y, b.z)] != 0) continue;
        Pore& pore = L.pores[pore_id];
        pore.id = pore_id; pore.x = b.x; pore.y = b.y; pore.z = b.z; pore.r2 = b.r2;
        pore.volume = 0; pore.boundary = false;
        pore_id += (pore_id % 255 == 0) ? 127 : 1;

        std::stack<std::array<int, 3>> st;
        st.push({b.x, b.y, b.z});
        while (!st.empty()) {
            std::array<int, 3> v = st.top();
            st.pop();
            if (!L.update(pt, g.idx(v[0], v[1], v[2]), pore, true)) continue;
            uint32_t here = cr[g.idx(v[0], v[1], v[2])];
            for (int nx = v[0] - 1; nx <= v[0] + 1; ++nx)
                for (int ny = v[1] - 1; ny <= v[1] + 1; ++ny)
                    for (int nz = v[2] - 1; nz <= v[2] + 1; ++nz) {
                        int dx = nx - v[0], dy = ny - v[1], dz = nz - v[2];
                        if (!dx && !dy && !dz) continue;
                        if (d3q7 && diagonal(dx, dy, dz)) continue;
                        if (g.outer(nx, ny, nz)) { pore.boundary = true; continue; }
                        uint32_t c = cr[g.idx(nx, ny, nz)];
                        if (c > 0 && c <= here) st.push({nx, ny, nz});
                    }
        }
        std::vector<uint32_t> empty;
        for (const auto& kv : L.keys)
            if (kv.second.volume == 0) empty.push_back(kv.first);
        for (uint32_t id : empty) {
            L.key_by_list.erase(L.keys[id].pores);
            L.keys.erase(id);
        }
    }
    *pores_out = dup_string(pores_text(L.pores));
    *keys_out = dup_string(keys_text(L.keys));
}
}  // namespace

int oracle_pores(const uint8_t* mask, const uint32_t* e, const uint32_t* cr, int X, int Y, int Z,
                 int d3q7, uint32_t* pt, char** pores_out, char** keys_out) {
    Grid g{X, Y, Z};
    std::vector<Ball> balls;
    for (int z = 0; z < Z; ++z)
        for (int x = 0; x < X; ++x)
            for (int y = 0; y < Y; ++y) {
                size_t i = g.idx(x, y, z);
                if (mask[i] != 255) continue;
                balls.push_back(Ball{x, y, z, e[i]});
            }
    std::sort(balls.begin(), balls.end(), [](const Ball& a, const Ball& b) { return a.r2 > b.r2; });
    pores_from_balls(balls, cr, g, d3q7, pt, pores_out, keys_out);
    return 0;
}

// The same pore loop over the maximum balls in a caller-given order (voxel indices), in place of
// the std::sort permutation: lets a test replay the tie order another STL's sort produced. The
// loop itself takes any order; the reference's is non-increasing in r2.
int oracle_pores_in_order(const uint32_t* e, const uint32_t* cr, int X, int Y, int Z, int d3q7,
                          const uint32_t* order, size_t m, uint32_t* pt, char** pores_out, char** keys_out) {
    Grid g{X, Y, Z};
    std::vector<Ball> balls(m);
    for (size_t k = 0; k < m; ++k) {
        size_t i = order[k];
        if (i >= g.count()) return 1;
        int y = (int)(i % (size_t)Y), x = (int)(i / (size_t)Y % (size_t)X), z = (int)(i / ((size_t)X * Y));
        balls[k] = Ball{x, y, z, e[i]};
    }
    pores_from_balls(balls, cr, g, d3q7, pt, pores_out, keys_out);
    return 0;
}

// ---------------------------------------------------------------------------------------
// EmptyPoresRemover (EmptyPoresRemover.cpp:51-244): drop volume-0 pores, sort each key's
// pore list and remove dropped pores from it, fold keys with one pore into that pore and
// drop keys with none, merge keys with equal lists into the largest ID, then remap labels.
int oracle_remove_empty_pores(uint32_t* pt, int X, int Y, int Z, const char* pores_in,
                              const char* keys_in, char** pores_out, char** keys_out) {
    size_t n = (size_t)X * Y * Z;
    std::map<uint32_t, Pore> pores = parse_pores(pores_in);
    std::map<uint32_t, Key> keys = parse_keys(keys_in);
    std::map<uint32_t, uint32_t> remap;

    std::vector<uint32_t> empty;
    for (const auto& kv : pores)
        if (kv.second.volume == 0) empty.push_back(kv.first);
    for (uint32_t id : empty) pores.erase(id);
    std::sort(empty.begin(), empty.end());
    for (auto& kv : keys) {
        std::vector<uint32_t>& l = kv.second.pores;
        std::sort(l.begin(), l.end());
        std::vector<uint32_t> kept;
        std::set_difference(l.begin(), l.end(), empty.begin(), empty.end(), std::back_inserter(kept));
        l = kept;
    }
    std::vector<uint32_t> few;
    for (const auto& kv : keys)
        if (kv.second.pores.size() < 2) few.push_back(kv.first);
    for (uint32_t id : few) {
        if (keys[id].pores.size() == 1) remap[id] = keys[id].pores[0];
        keys.erase(id);
    }
    std::map<std::vector<uint32_t>, std::vector<uint32_t>> ids_by_list;
    for (const auto& kv : keys) ids_by_list[kv.second.pores].push_back(kv.first);
    for (const auto& kv : ids_by_list) {
        if (kv.second.size() < 2) continue;
        uint32_t keep = *std::max_element(kv.second.begin(), kv.second.end());
        for (uint32_t id : kv.second)
            if (id != keep) { keys.erase(id); remap[id] = keep; }
    }
    *pores_out = dup_string(pores_text(pores));
    *keys_out = dup_string(keys_text(keys));
    if (!remap.empty())
        for (size_t i = 0; i < n; ++i) {
            if (pt[i] == 0) continue;
            auto it = remap.find(pt[i]);
            if (it != remap.end()) pt[i] = it->second;
        }
    return 0;
}

// ---------------------------------------------------------------------------------------
// WatershedComputer (WatershedComputer.cpp:30-219): one FIFO per depth (= containing radius
// square), seeds pushed in ascending pore ID with depth = seed r2 (:183-219), depths
// processed from maxE down to 0 (:80-102), neighbours in x, y, z order; a neighbour qualifies
// if in the image, void and CR <= node depth (:130-176); the first claim wins and is pushed,
// later contacts only extend the throat key (continue-after-shared = false).
int oracle_watershed(const uint32_t* cr, int X, int Y, int Z, int d3q7, uint32_t max_e,
                     const char* pores_in, uint32_t* ptw, char** pores_out, char** keys_out) {
    Grid g{X, Y, Z};
    size_t n = g.count();
    std::fill(ptw, ptw + n, 0u);
    Labeller L;
    L.pores = parse_pores(pores_in);
    struct Node { int x, y, z; uint32_t depth; uint32_t pore; };
    std::vector<std::queue<Node>> q((size_t)max_e + 1);
    for (const auto& kv : L.pores)
        if (kv.second.r2 > max_e) return 1;
    for (auto& kv : L.pores) {
        Pore& p = kv.second;
        p.boundary = false;
        p.volume = 0;
        q[p.r2].push(Node{p.x, p.y, p.z, p.r2, p.id});
        L.update(ptw, g.idx(p.x, p.y, p.z), p, false);
    }
    for (long depth = (long)max_e; depth >= 0; --depth) {
        std::queue<Node>& fifo = q[(size_t)depth];
        while (!fifo.empty()) {
            Node v = fifo.front();
            fifo.pop();
            Pore& pore = L.pores[v.pore];
            for (int nx = v.x - 1; nx <= v.x + 1; ++nx)
                for (int ny = v.y - 1; ny <= v.y + 1; ++ny)
                    for (int nz = v.z - 1; nz <= v.z + 1; ++nz) {
                        int dx = nx - v.x, dy = ny - v.y, dz = nz - v.z;
                        if (!dx && !dy && !dz) continue;
                        if (d3q7 && diagonal(dx, dy, dz)) continue;
                        if (g.outer(nx, ny, nz)) { pore.boundary = true; continue; }
                        uint32_t c = cr[g.idx(nx, ny, nz)];
                        if (c == 0 || c > v.depth) continue;
                        if (L.update(ptw, g.idx(nx, ny, nz), pore, false))
                            q[c].push(Node{nx, ny, nz, c, v.pore});
                    }
        }
    }
    *pores_out = dup_string(pores_text(L.pores));
    *keys_out = dup_string(keys_text(L.keys));
    return 0;
}

// ---------------------------------------------------------------------------------------
// ShortestPathComputer (ShortestPathComputer.cpp:39-511): distances are the minimum over
// paths of the left-fold float32 sum of sqrtf(1|2|3) steps (1.0f with D3Q7), so any
// label-setting order gives the same values; a binary heap with lazy deletion is used here.
// Start = centre voxel, else the first void unprocessed voxel on box shells of growing half
// size; cavities are restarted until one reaches every wall.
char* oracle_geodesic(const uint8_t* solid, int X, int Y, int Z, int d3q7, float* dist,
                      uint32_t* cavity) {
    Grid g{X, Y, Z};
    size_t n = g.count();
    const int size[3] = {X, Y, Z};
    for (size_t i = 0; i < n; ++i) {
        bool s = solid[i] != 0;
        dist[i] = s ? -1.0f : FLT_MAX;
        cavity[i] = s ? 0u : 0xFFFFFFFFu;
    }
    const float w[4] = {std::sqrt(0.0f), std::sqrt(1.0f), std::sqrt(2.0f), std::sqrt(3.0f)};
    int c[3];
    for (int i = 0; i < 3; ++i) c[i] = std::min((int)(0.5 * size[i] + 0.5), size[i] - 1);
    int max_half = 0;
    for (int i = 0; i < 3; ++i) max_half = std::max(max_half, std::max(size[i] - 1 - c[i], c[i]));
    std::vector<std::pair<int, int>> walls;
    for (int i = 0; i < 3; ++i)
        for (int v : {0, size[i] - 1})
            if (std::find(walls.begin(), walls.end(), std::make_pair(i, v)) == walls.end())
                walls.push_back({i, v});
    auto usable = [&](const int p[3]) {
        size_t i = g.idx(p[0], p[1], p[2]);
        return solid[i] == 0 && dist[i] == FLT_MAX;
    };
    auto on_box = [&](int h, int out[3]) {
        for (int k = 0; k < 3; ++k) {
            int a0 = k == 0 ? 1 : 0, a1 = k == 2 ? 1 : 2;
            int f0 = std::max(c[a0] - h, 0), f1 = std::min(c[a0] + h, size[a0] - 1);
            int s0 = std::max(c[a1] - h, 0), s1 = std::min(c[a1] + h, size[a1] - 1);
            std::vector<int> fixed{std::max(c[k] - h, 0), std::min(c[k] + h, size[k] - 1)};
            if (fixed[0] == fixed[1]) fixed.pop_back();
            for (int fv : fixed)
                for (int a = f0; a <= f1; ++a)
                    for (int b = s0; b <= s1; ++b) {
                        out[a0] = a; out[a1] = b; out[k] = fv;
                        if (usable(out)) return true;
                    }
        }
        return false;
    };
    std::string text = "CavityId SeedX SeedY SeedZ DistanceToCenter VoidPixelsCount ReachedWallsCount IsReachingAllWalls\n";
    int half = 0;
    uint32_t id = 0;
    typedef std::pair<float, size_t> Item;
    while (true) {
        ++half;
        int start[3];
        bool found = false;
        if (usable(c)) { start[0] = c[0]; start[1] = c[1]; start[2] = c[2]; found = true; }
        else {
            for (; half < max_half; ++half)
                if (on_box(half, start)) { found = true; break; }
        }
        if (!found) break;
        id += 127;
        std::vector<std::pair<int, int>> reached;
        size_t settled = 0;
        std::priority_queue<Item, std::vector<Item>, std::greater<Item>> heap;
        size_t s = g.idx(start[0], start[1], start[2]);
        dist[s] = 0.0f;
        heap.push({0.0f, s});
        while (!heap.empty()) {
            Item it = heap.top();
            heap.pop();
            size_t i = it.second;
            if (it.first != dist[i] || cavity[i] == id) continue;
            cavity[i] = id;
            ++settled;
            int p[3];
            p[1] = (int)(i % Y);
            size_t r = i / Y;
            p[0] = (int)(r % X);
            p[2] = (int)(r / X);
            for (int k = 0; k < 3; ++k)
                if (p[k] == 0 || p[k] == size[k] - 1)
                    if (std::find(reached.begin(), reached.end(), std::make_pair(k, p[k])) == reached.end())
                        reached.push_back({k, p[k]});
            float d = dist[i];
            for (int dx = -1; dx <= 1; ++dx)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dz = -1; dz <= 1; ++dz) {
                        if (!dx && !dy && !dz) continue;
                        if (d3q7 && diagonal(dx, dy, dz)) continue;
                        int nx = p[0] + dx, ny = p[1] + dy, nz = p[2] + dz;
                        if (g.outer(nx, ny, nz)) continue;
                        size_t j = g.idx(nx, ny, nz);
                        if (solid[j] != 0) continue;
                        float nd = d + (d3q7 ? 1.0f : w[std::abs(dx) + std::abs(dy) + std::abs(dz)]);
                        if (nd < dist[j]) { dist[j] = nd; heap.push({nd, j}); }
                    }
        }
        int diff2 = 0;
        for (int k = 0; k < 3; ++k) diff2 += (start[k] - c[k]) * (start[k] - c[k]);
        char buf[200];
        std::snprintf(buf, sizeof buf, "%u %d %d %d %f %zu %d %d\n", id, start[0], start[1], start[2],
                      std::sqrt((double)diff2), settled, (int)reached.size(),
                      reached.size() == walls.size() ? 1 : 0);
        text += buf;
        if (reached.size() == walls.size()) break;
    }
    return dup_string(text);
}

}  // extern "C"
