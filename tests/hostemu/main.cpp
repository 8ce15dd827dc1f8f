// Driver of the host emulation: runs one stage of the library (translated kernels) on inputs
// written by tests/test_hostemu.py and writes its volume and text outputs.
//   hostemu pores     <in> <out_volume> <out_text>   in: X Y Z d3q7 | EDT2 u32 | CR2 u32 | mask u8
//                                                    [| u64 m | m u32 maximum-ball order]
//   hostemu watershed <in> <out_volume> <out_text>   in: X Y Z d3q7 maxE | CR2 u32 | np | np x 7 i32
//   hostemu remove    <in> <out_volume> <out_text>   in: X Y Z | PT u32 | text (pores, "@@", throats)
//   hostemu edt       <in> <out_volume> <out_text>   in: X Y Z | solid u8; out: EDT2 u32;
//                     text: maximum square, then the histogram counts one a line
//   hostemu balls     <in> <out_volumes> <out_text>  in: X Y Z maxE | EDT2 u32;
//                     out: CR2, CX, CY, CZ u32 | mask u8; text: maximum-ball count
//   hostemu geodesic  <in> <out_volumes> <out_text>  in: X Y Z d3q7 | solid u8;
//                     out: distance f32 | cavity u32; text: cavityDescriptions.txt
//   hostemu pores_seq <list> - -                     list: lines "<in> <out_volume> <out_text>" of
//                     pore-stage cases run one after another on ONE context, whose scratch
//                     buffers keep the previous case's contents (as in a process that runs many)
#include <fstream>
#include <sstream>

#include "shim.h"

namespace pma {
void run_edt(pma_ctx* c);
void run_pores(pma_ctx* c);
void run_watershed(pma_ctx* c);
void run_remove_empty_pores(pma_ctx* c);
void run_containing_balls(pma_ctx* c);
void run_geodesic(pma_ctx* c);
}  // namespace pma

namespace {

std::string format_pores(const std::vector<pma::PoreRow>& rows) {
    std::string s = "PoreId SeedX SeedY SeedZ SeedRadiusSquare IsBoundaryPore Volume\n";
    char b[160];
    for (const auto& p : rows) {
        std::snprintf(b, sizeof b, "%u %d %d %d %d %d %zu\n", p.id, p.x, p.y, p.z, (int)p.r2, p.boundary ? 1 : 0,
                      (size_t)p.volume);
        s += b;
    }
    return s;
}

std::string format_throats(const std::vector<pma::ThroatRow>& rows) {
    std::string s = "SharedPixelTypeId; Volume; PoreId1 PoreId2 ... \n";
    char b[64];
    for (const auto& t : rows) {
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

template <class T> void read_into(FILE* f, T* dst, size_t n) {
    if (std::fread(dst, sizeof(T), n, f) != n) throw std::runtime_error("short input file");
}

void write_outputs(pma_ctx& c, pma_volume w, const std::vector<pma::PoreRow>& pores,
                   const std::vector<pma::ThroatRow>& throats, const char* vol_path, const char* text_path) {
    FILE* o = std::fopen(vol_path, "wb");
    std::fwrite(c.v<uint32_t>(w), 4, c.n, o);
    std::fclose(o);
    std::string t = format_pores(pores) + "@@\n" + format_throats(throats);
    o = std::fopen(text_path, "wb");
    std::fwrite(t.data(), 1, t.size(), o);
    std::fclose(o);
}

}  // namespace

void read_pores_input(pma_ctx& c, FILE* f) {
    int32_t h[4] = {0, 0, 0, 0};
    read_into(f, h, 4);
    c.dims = {h[0], h[1], h[2]};
    c.n = (size_t)h[0] * h[1] * h[2];
    c.stencil = h[3] ? PMA_STENCIL_D3Q7 : PMA_STENCIL_D3Q27;
    c.ensure_volume(PMA_VOL_EDT2);
    c.ensure_volume(PMA_VOL_CR2);
    c.ensure_volume(PMA_VOL_MASK);
    read_into(f, c.v<uint32_t>(PMA_VOL_EDT2), c.n);
    read_into(f, c.v<uint32_t>(PMA_VOL_CR2), c.n);
    read_into(f, c.v<uint8_t>(PMA_VOL_MASK), c.n);
    c.valid[PMA_VOL_EDT2] = c.valid[PMA_VOL_CR2] = c.valid[PMA_VOL_MASK] = true;
    // optional: u64 count | u32 voxel indices, the maximum balls' processing order
    uint64_t m = 0;
    if (std::fread(&m, 8, 1, f) == 1) {
        c.testing_ball_order.resize(m);
        read_into(f, c.testing_ball_order.data(), m);
    }
}

int main(int argc, char** argv) {
    if (argc != 5) return 2;
    const std::string stage = argv[1];
    if (stage == "pores_seq") {
        std::ifstream list(argv[2]);
        pma_ctx c;
        std::string in, vol, txt;
        while (list >> in >> vol >> txt) {
            FILE* f = std::fopen(in.c_str(), "rb");
            if (!f) return 2;
            try {
                read_pores_input(c, f);
                pma::run_pores(&c);
                write_outputs(c, PMA_VOL_PT, c.pores, c.throats, vol.c_str(), txt.c_str());
            } catch (const std::exception& ex) {
                std::fprintf(stderr, "error: %s (%s)\n", ex.what(), in.c_str());
                return 1;
            }
            std::fclose(f);
        }
        return 0;
    }
    FILE* f = std::fopen(argv[2], "rb");
    if (!f) return 2;
    pma_ctx c;
    try {
        if (stage == "pores") {
            read_pores_input(c, f);
            pma::run_pores(&c);
            write_outputs(c, PMA_VOL_PT, c.pores, c.throats, argv[3], argv[4]);
        } else if (stage == "watershed") {
            int32_t h[5] = {0, 0, 0, 0, 0};
            read_into(f, h, 5);
            c.dims = {h[0], h[1], h[2]};
            c.n = (size_t)h[0] * h[1] * h[2];
            c.stencil = h[3] ? PMA_STENCIL_D3Q7 : PMA_STENCIL_D3Q27;
            {
                c.max_e = (uint32_t)h[4];
                c.have_edt_stats = true;
                c.ensure_volume(PMA_VOL_CR2);
                read_into(f, c.v<uint32_t>(PMA_VOL_CR2), c.n);
                c.valid[PMA_VOL_CR2] = true;
                int32_t np = 0;
                read_into(f, &np, 1);
                for (int32_t k = 0; k < np; ++k) {
                    int32_t r[7];
                    read_into(f, r, 7);
                    pma::PoreRow p;
                    p.id = (uint32_t)r[0];
                    p.x = r[1];
                    p.y = r[2];
                    p.z = r[3];
                    p.r2 = (uint32_t)r[4];
                    p.boundary = r[5] != 0;
                    p.volume = (uint64_t)r[6];
                    c.pores.push_back(p);
                }
                c.have_pores = true;
                pma::run_watershed(&c);
                write_outputs(c, PMA_VOL_PTW, c.pores_ws, c.throats_ws, argv[3], argv[4]);
            }
        } else if (stage == "edt") {
            int32_t h[3];
            read_into(f, h, 3);
            c.dims = {h[0], h[1], h[2]};
            c.n = (size_t)h[0] * h[1] * h[2];
            c.ensure_volume(PMA_VOL_SOLID);
            read_into(f, c.v<uint8_t>(PMA_VOL_SOLID), c.n);
            c.valid[PMA_VOL_SOLID] = true;
            pma::run_edt(&c);
            FILE* o = std::fopen(argv[3], "wb");
            std::fwrite(c.v<uint32_t>(PMA_VOL_EDT2), 4, c.n, o);
            std::fclose(o);
            o = std::fopen(argv[4], "wb");
            std::fprintf(o, "%u\n", c.max_e);
            for (uint64_t v : c.histogram) std::fprintf(o, "%llu\n", (unsigned long long)v);
            std::fclose(o);
        } else if (stage == "balls") {
            int32_t h[4];
            read_into(f, h, 4);
            c.dims = {h[0], h[1], h[2]};
            c.n = (size_t)h[0] * h[1] * h[2];
            c.max_e = (uint32_t)h[3];
            c.have_edt_stats = true;
            c.ensure_volume(PMA_VOL_EDT2);
            read_into(f, c.v<uint32_t>(PMA_VOL_EDT2), c.n);
            c.valid[PMA_VOL_EDT2] = true;
            pma::run_containing_balls(&c);
            FILE* o = std::fopen(argv[3], "wb");
            for (pma_volume w : {PMA_VOL_CR2, PMA_VOL_CX, PMA_VOL_CY, PMA_VOL_CZ}) std::fwrite(c.v<uint32_t>(w), 4, c.n, o);
            std::fwrite(c.v<uint8_t>(PMA_VOL_MASK), 1, c.n, o);
            std::fclose(o);
            o = std::fopen(argv[4], "wb");
            std::fprintf(o, "%llu\n", (unsigned long long)c.n_max_balls);
            std::fclose(o);
        } else if (stage == "geodesic") {
            int32_t h[4];
            read_into(f, h, 4);
            c.dims = {h[0], h[1], h[2]};
            c.n = (size_t)h[0] * h[1] * h[2];
            c.stencil = h[3] ? PMA_STENCIL_D3Q7 : PMA_STENCIL_D3Q27;
            c.ensure_volume(PMA_VOL_SOLID);
            read_into(f, c.v<uint8_t>(PMA_VOL_SOLID), c.n);
            c.valid[PMA_VOL_SOLID] = true;
            pma::run_geodesic(&c);
            FILE* o = std::fopen(argv[3], "wb");
            std::fwrite(c.v<float>(PMA_VOL_SPD), 4, c.n, o);
            std::fwrite(c.v<uint32_t>(PMA_VOL_CAVITY), 4, c.n, o);
            std::fclose(o);
            std::string s = "CavityId SeedX SeedY SeedZ DistanceToCenter VoidPixelsCount ReachedWallsCount IsReachingAllWalls\n";
            char b[256];
            for (const auto& r : c.cavities) {
                std::snprintf(b, sizeof b, "%u %d %d %d %f %zu %d %d\n", r.id, r.x, r.y, r.z, r.distance_to_center,
                              (size_t)r.void_count, r.reached_walls, r.reaches_all ? 1 : 0);
                s += b;
            }
            o = std::fopen(argv[4], "wb");
            std::fwrite(s.data(), 1, s.size(), o);
            std::fclose(o);
        } else if (stage == "remove") {
            int32_t h[3];
            read_into(f, h, 3);
            c.dims = {h[0], h[1], h[2]};
            c.n = (size_t)h[0] * h[1] * h[2];
            c.ensure_volume(PMA_VOL_PT);
            read_into(f, c.v<uint32_t>(PMA_VOL_PT), c.n);
            std::stringstream rest;
            char buf[4096];
            size_t got;
            while ((got = std::fread(buf, 1, sizeof buf, f)) > 0) rest.write(buf, (std::streamsize)got);
            std::string line;
            int part = 0;
            bool header = true;
            while (std::getline(rest, line)) {
                if (line == "@@") {
                    part = 1;
                    header = true;
                    continue;
                }
                if (header) {
                    header = false;
                    continue;
                }
                if (line.empty()) continue;
                if (part == 0) {
                    std::istringstream is(line);
                    pma::PoreRow r;
                    int b;
                    is >> r.id >> r.x >> r.y >> r.z >> r.r2 >> b >> r.volume;
                    r.boundary = b != 0;
                    c.pores.push_back(r);
                } else {
                    for (char& ch : line)
                        if (ch == ';') ch = ' ';
                    std::istringstream is(line);
                    pma::ThroatRow t;
                    is >> t.id >> t.volume;
                    uint32_t p;
                    while (is >> p) t.pores.push_back(p);
                    c.throats.push_back(t);
                }
            }
            c.have_pores = true;
            c.valid[PMA_VOL_PT] = true;
            pma::run_remove_empty_pores(&c);
            write_outputs(c, PMA_VOL_PT, c.pores, c.throats, argv[3], argv[4]);
        } else {
            return 2;
        }
    } catch (const std::exception& e) {
        std::printf("%s threw: %s\n", stage.c_str(), e.what());
        return 3;
    }
    std::fclose(f);
    std::printf("%s ok\n", stage.c_str());
    return 0;
}
