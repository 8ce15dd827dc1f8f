// The reference's folder contract (host code above the C ABI): one -all / -edt / -rib / -apt /
// -spd pass over a folder holding config.txt and Images/, reading and writing the same files and
// sub-folders as the reference's console program.
//
//   config.txt          Serializer::ReadConfig (Serializer.cpp:31-64): resolution of the
//                       initial images, available memory, maximum pore-to-seed radius ratio,
//                       D3Q7 stencil flag, number of threads.  The stencil flag and the
//                       resolution are used; memory, ratio and threads have no effect here (one
//                       active area, results as with one thread).
//   Images/             the binary stack; files ordered by their names without extension
//                       (Serializer.cpp:253-280); size X = rows, Y = columns, Z = files
//                       (Serializer.cpp:223-251).
//   resampling          ImageResampler (ImageResampler.cpp:28-206): when the normalised
//                       resolution is not 1 1 1 and Original/ does not exist, Images/ becomes
//                       Original/ and Images/ receives round(size x resolution) voxels by
//                       nearest-neighbour duplication (slice i <- slice floor(i / r_z), names of
//                       repeated slices get "_1", "_2", ...; voxel (x, y) <- (floor(x / r_x),
//                       floor(y / r_y))).
//   outputs             one folder per volume (Constants.h:26-40), one file per slice named like
//                       the slice's file in Images/, carrying its descriptive tags
//                       (DataArray.h:226-252); the text files of each stage (Serializer.cpp).
// Stage order and inputs follow ImageProcessingManager.cpp:63-108: -edt = (resample +) EDT;
// -rib = containing balls from EuclideanDistanceSquares/; -apt = pores, empty-pore removal and
// watershed from IsMaximumBall/, EuclideanDistanceSquares/ and ContainingBallRadiiSquares/;
// -all = the three; -spd = (resample +) geodesic distances.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "pma/pma.h"
#include "pma_internal.h"
#include "tiff_io.h"

namespace fs = std::filesystem;

namespace pma {
namespace {

struct FolderConfig {
    double resolution[3] = {1, 1, 1};  // x, y, z as written
    double normalized[3] = {1, 1, 1};  // divided by the smallest
    int memory_mb = 0, threads = 1;
    double max_ratio = 10.0;
    bool d3q7 = false;
};

void chk(pma_status s) {
    if (s != PMA_OK) throw Error(s, pma_last_error());
}

FolderConfig read_config(const fs::path& folder) {
    std::ifstream f(folder / "config.txt");
    PMA_REQUIRE(f.good(), PMA_ERR_INVALID_ARGUMENT, "cannot open " + (folder / "config.txt").string());
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string text = ss.str();
    FolderConfig c;
    int d3q7 = 0;
    const int got = std::sscanf(text.c_str(),
                                "Resolution of initial images: %lf %lf %lf "
                                "Available memory in megabytes: %d "
                                "Max ratio of pore radius to seed radius: %lf "
                                "Use D3Q7 stencil: %d "
                                "Number of threads: %d",
                                &c.resolution[0], &c.resolution[1], &c.resolution[2], &c.memory_mb, &c.max_ratio, &d3q7,
                                &c.threads);
    PMA_REQUIRE(got == 7, PMA_ERR_INVALID_ARGUMENT,
                (folder / "config.txt").string() + ": expected the five lines of the reference's config.txt");
    const double mn = std::min({c.resolution[0], c.resolution[1], c.resolution[2]});
    PMA_REQUIRE(mn > 0, PMA_ERR_INVALID_ARGUMENT, "config.txt: resolutions must be positive");
    for (int i = 0; i < 3; ++i) c.normalized[i] = c.resolution[i] / mn;
    c.d3q7 = d3q7 != 0;
    return c;
}

// .tif / .tiff files of a folder, ordered by their names without extension.
std::vector<fs::path> image_paths(const fs::path& folder) {
    std::vector<fs::path> out;
    if (!fs::is_directory(folder)) return out;
    for (const auto& e : fs::directory_iterator(folder)) {
        const std::string ext = e.path().extension().string();
        if (e.is_regular_file() && (ext == ".tif" || ext == ".tiff")) out.push_back(e.path());
    }
    std::stable_sort(out.begin(), out.end(),
                     [](const fs::path& a, const fs::path& b) { return a.stem().string() < b.stem().string(); });
    return out;
}

long round_half_up(double v) { return (long)(v + 0.5); }

// ImageResampler::ResampleImages: a no-op when Original/ exists or the size is unchanged.
void resample(const fs::path& folder, const FolderConfig& cfg) {
    const fs::path images = folder / "Images", original = folder / "Original";
    if (fs::exists(original)) return;
    std::vector<fs::path> in = image_paths(images);
    PMA_REQUIRE(!in.empty(), PMA_ERR_INVALID_ARGUMENT, "Images/ holds no .tif or .tiff files");
    const tiff::Image first = tiff::read(in[0].string());
    const long X0 = first.height, Y0 = first.width, Z0 = (long)in.size();
    const long X = round_half_up(X0 * cfg.normalized[0]), Y = round_half_up(Y0 * cfg.normalized[1]),
               Z = round_half_up(Z0 * cfg.normalized[2]);
    if (X == X0 && Y == Y0 && Z == Z0) return;
    fs::rename(images, original);
    fs::create_directory(images);
    in = image_paths(original);
    std::string prev;
    int dup = 0;
    for (long i = 0; i < Z; ++i) {
        const long src = std::min<long>((long)(i / cfg.normalized[2]), Z0 - 1);
        std::string name = in[src].filename().string();
        if (name == prev) {
            ++dup;
            name = in[src].stem().string() + "_" + std::to_string(dup) + in[src].extension().string();
        } else {
            dup = 0;
            prev = name;
        }
        tiff::Image im = tiff::read(in[src].string());
        PMA_REQUIRE(im.samples == 1 && (long)im.height == X0 && (long)im.width == Y0, PMA_ERR_INVALID_ARGUMENT,
                    in[src].string() + ": slices must be 8-bit and of one size");
        if (X != X0 || Y != Y0) {
            std::vector<uint8_t> px((size_t)X * Y);
            for (long x = 0; x < X; ++x)
                for (long y = 0; y < Y; ++y) {
                    const long sx = std::min<long>((long)(x / cfg.normalized[0]), X0 - 1);
                    const long sy = std::min<long>((long)(y / cfg.normalized[1]), Y0 - 1);
                    px[(size_t)x * Y + y] = im.pixels[(size_t)sx * Y0 + sy];
                }
            im.pixels.swap(px);
            im.width = (uint32_t)Y;
            im.height = (uint32_t)X;
        }
        tiff::write((images / name).string(), im);
    }
}

struct Stack {
    std::vector<fs::path> files;   // the slices of Images/ (names and tags of every output)
    std::vector<tiff::Image> meta;  // per slice: tags and photometric only
    pma_dims dims{0, 0, 0};
};

// The slices of an 8-bit stack folder (sizes, names, tags), and optionally their samples.
Stack read_stack(const fs::path& dir, std::vector<uint8_t>* solid) {
    Stack s;
    s.files = image_paths(dir);
    PMA_REQUIRE(!s.files.empty(), PMA_ERR_INVALID_ARGUMENT,
                "Initial images folder, \"Images\", which shall be near the config.txt file, is empty");
    for (size_t z = 0; z < s.files.size(); ++z) {
        tiff::Image im = tiff::read(s.files[z].string());
        if (z == 0) {
            s.dims.x = (int32_t)im.height;
            s.dims.y = (int32_t)im.width;
            s.dims.z = (int32_t)s.files.size();
            if (solid) solid->resize((size_t)s.dims.x * s.dims.y * s.dims.z);
        }
        PMA_REQUIRE((int32_t)im.height == s.dims.x && (int32_t)im.width == s.dims.y && im.samples == 1,
                    PMA_ERR_INVALID_ARGUMENT, s.files[z].string() + ": slices must be 8-bit and of one size");
        if (solid) std::memcpy(solid->data() + z * (size_t)s.dims.x * s.dims.y, im.pixels.data(), im.pixels.size());
        im.pixels.clear();
        im.pixels.shrink_to_fit();
        s.meta.push_back(std::move(im));
    }
    return s;
}

// Writes a host volume ([z][x][y], es = 1 or 4 bytes per voxel) as a folder of slices named and
// tagged like the stack's files.
void write_stack(const uint8_t* host, size_t es, const Stack& s, const fs::path& dir) {
    const size_t plane = (size_t)s.dims.x * s.dims.y;
    fs::create_directories(dir);
    for (int z = 0; z < s.dims.z; ++z) {
        tiff::Image im;
        im.width = (uint32_t)s.dims.y;
        im.height = (uint32_t)s.dims.x;
        im.samples = (uint32_t)es;
        im.photometric = s.meta[z].photometric;
        im.keep = s.meta[z].keep;
        // 32-bit values: the little-endian bytes are the four samples
        im.pixels.assign(host + z * plane * es, host + (z + 1) * plane * es);
        tiff::write((dir / s.files[z].filename()).string(), im);
    }
}

// Writes one volume of the context as a folder of slices named and tagged like Images/.
void write_volume(pma_ctx* c, pma_volume w, const Stack& s, const fs::path& dir) {
    const size_t n = (size_t)s.dims.x * s.dims.y * s.dims.z;
    const size_t es = (w == PMA_VOL_SOLID || w == PMA_VOL_MASK) ? 1 : 4;
    std::vector<uint8_t> host(n * es);
    chk(pma_read_volume(c, w, host.data(), host.size(), 0));
    write_stack(host.data(), es, s, dir);
}

// Reads a folder written by write_volume (or by the reference) into a context volume.
void read_volume(pma_ctx* c, pma_volume w, const Stack& s, const fs::path& dir) {
    const size_t plane = (size_t)s.dims.x * s.dims.y, n = plane * s.dims.z;
    const size_t es = (w == PMA_VOL_SOLID || w == PMA_VOL_MASK) ? 1 : 4;
    const std::vector<fs::path> files = image_paths(dir);
    PMA_REQUIRE(files.size() == (size_t)s.dims.z, PMA_ERR_STATE,
                dir.string() + ": expected " + std::to_string(s.dims.z) + " images (run the earlier stage first)");
    std::vector<uint8_t> host(n * es);
    for (int z = 0; z < s.dims.z; ++z) {
        const tiff::Image im = tiff::read(files[z].string());
        PMA_REQUIRE(im.samples == es && (int32_t)im.height == s.dims.x && (int32_t)im.width == s.dims.y,
                    PMA_ERR_INVALID_ARGUMENT, files[z].string() + ": unexpected image size or sample count");
        std::memcpy(host.data() + z * plane * es, im.pixels.data(), plane * es);
    }
    chk(pma_load_volume(c, w, host.data(), 0));
}

std::string text_of(pma_ctx* c, pma_text t) {
    size_t len = 0;
    const pma_status q = pma_get_text(c, t, nullptr, 0, &len);  // the size query reports CAPACITY
    if (q != PMA_ERR_CAPACITY) chk(q);
    std::string s(len, '\0');
    chk(pma_get_text(c, t, s.data(), len, &len));
    s.resize(std::strlen(s.c_str()));
    return s;
}

void write_text(const fs::path& p, const std::string& s) {
    std::ofstream f(p, std::ios::binary);
    f << s;
    PMA_REQUIRE(f.good(), PMA_ERR_INVALID_ARGUMENT, "cannot write " + p.string());
}

std::string read_text(const fs::path& p) {
    std::ifstream f(p, std::ios::binary);
    PMA_REQUIRE(f.good(), PMA_ERR_STATE, "cannot read " + p.string() + " (run the earlier stage first)");
    std::stringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

void stage_edt(pma_ctx* c, const fs::path& folder, const Stack& s) {
    chk(pma_edt(c));
    write_volume(c, PMA_VOL_EDT2, s, folder / "EuclideanDistanceSquares");
    uint32_t max_e = 0;
    chk(pma_get_statistics(c, &max_e, nullptr));
    // EuclideanDistanceComputer.cpp:74-81: the maximum-ball count is 0 until -rib
    char b[128];
    std::snprintf(b, sizeof b, "Max Euclidean distance square: %u\nMaximum balls count: 0\n", max_e);
    write_text(folder / "intermediateStatistics.txt", b);
    write_text(folder / "euclideanDistanceSquareCounts.txt", text_of(c, PMA_TEXT_EDT_COUNTS));
}

void stage_rib(pma_ctx* c, const fs::path& folder, const Stack& s) {
    chk(pma_containing_balls(c));
    write_volume(c, PMA_VOL_CR2, s, folder / "ContainingBallRadiiSquares");
    write_volume(c, PMA_VOL_CX, s, folder / "ContainingBallCoordinatesX");
    write_volume(c, PMA_VOL_CY, s, folder / "ContainingBallCoordinatesY");
    write_volume(c, PMA_VOL_CZ, s, folder / "ContainingBallCoordinatesZ");
    write_volume(c, PMA_VOL_MASK, s, folder / "IsMaximumBall");
    write_text(folder / "intermediateStatistics.txt", text_of(c, PMA_TEXT_INTERMEDIATE_STATISTICS));
}

void stage_apt(pma_ctx* c, const fs::path& folder, const Stack& s) {
    chk(pma_pores(c));
    chk(pma_remove_empty_pores(c));
    write_volume(c, PMA_VOL_PT, s, folder / "PoreThroatIds");
    write_text(folder / "pores.txt", text_of(c, PMA_TEXT_PORES));
    write_text(folder / "sharedPixelTypes.txt", text_of(c, PMA_TEXT_THROATS));
    chk(pma_watershed(c));
    write_volume(c, PMA_VOL_PTW, s, folder / "PoreThroatIdsWatershed");
    write_text(folder / "poresWatershed.txt", text_of(c, PMA_TEXT_PORES_WATERSHED));
    write_text(folder / "sharedPixelTypesWatershed.txt", text_of(c, PMA_TEXT_THROATS_WATERSHED));
}

void run_folder(pma_ctx* c, const fs::path& folder, pma_mode mode) {
    const FolderConfig cfg = read_config(folder);
    chk(pma_set_stencil(c, cfg.d3q7 ? PMA_STENCIL_D3Q7 : PMA_STENCIL_D3Q27));
    if (mode == PMA_MODE_EDT || mode == PMA_MODE_ALL || mode == PMA_MODE_SPD) resample(folder, cfg);
    std::vector<uint8_t> solid;
    const bool need_solid = mode == PMA_MODE_EDT || mode == PMA_MODE_ALL || mode == PMA_MODE_SPD;
    const Stack s = read_stack(folder / "Images", need_solid ? &solid : nullptr);
    chk(pma_set_dims(c, s.dims));
    if (need_solid) chk(pma_load_volume(c, PMA_VOL_SOLID, solid.data(), 0));
    switch (mode) {
        case PMA_MODE_EDT:
            stage_edt(c, folder, s);
            break;
        case PMA_MODE_RIB:
            read_volume(c, PMA_VOL_EDT2, s, folder / "EuclideanDistanceSquares");
            chk(pma_set_text(c, PMA_TEXT_INTERMEDIATE_STATISTICS, read_text(folder / "intermediateStatistics.txt").c_str()));
            stage_rib(c, folder, s);
            break;
        case PMA_MODE_APT:
            read_volume(c, PMA_VOL_EDT2, s, folder / "EuclideanDistanceSquares");
            read_volume(c, PMA_VOL_CR2, s, folder / "ContainingBallRadiiSquares");
            read_volume(c, PMA_VOL_MASK, s, folder / "IsMaximumBall");
            chk(pma_set_text(c, PMA_TEXT_INTERMEDIATE_STATISTICS, read_text(folder / "intermediateStatistics.txt").c_str()));
            stage_apt(c, folder, s);
            break;
        case PMA_MODE_ALL:
            stage_edt(c, folder, s);
            stage_rib(c, folder, s);
            stage_apt(c, folder, s);
            break;
        case PMA_MODE_SPD:
            chk(pma_geodesic(c));
            write_volume(c, PMA_VOL_SPD, s, folder / "ShortestPathDistances");
            write_volume(c, PMA_VOL_CAVITY, s, folder / "CavityIds");
            write_text(folder / "cavityDescriptions.txt", text_of(c, PMA_TEXT_CAVITIES));
            break;
        default:
            throw Error(PMA_ERR_INVALID_ARGUMENT, "unknown mode");
    }
}

}  // namespace
}  // namespace pma

extern "C" {

pma_status pma_run_folder(pma_ctx* c, const char* folder, pma_mode mode) {
    try {
        if (!c || !folder) throw pma::Error(PMA_ERR_INVALID_ARGUMENT, "null context or folder");
        pma::run_folder(c, fs::path(folder), mode);
        pma::set_last_error("");
        return PMA_OK;
    } catch (const pma::Error& e) {
        pma::set_last_error(e.what());
        return e.status;
    } catch (const std::exception& e) {
        pma::set_last_error(std::string("pma_run_folder: ") + e.what());
        return PMA_ERR_INVALID_ARGUMENT;
    }
}

pma_status pma_write_tiff_stack(const char* folder, const char* like, const void* src, pma_dims dims,
                                int bytes_per_voxel) {
    try {
        if (!folder || !like || !src) throw pma::Error(PMA_ERR_INVALID_ARGUMENT, "null argument");
        if (bytes_per_voxel != 1 && bytes_per_voxel != 4)
            throw pma::Error(PMA_ERR_INVALID_ARGUMENT, "pma_write_tiff_stack: 1 or 4 bytes per voxel");
        const pma::Stack s = pma::read_stack(fs::path(like), nullptr);
        if (s.dims.x != dims.x || s.dims.y != dims.y || s.dims.z != dims.z)
            throw pma::Error(PMA_ERR_INVALID_ARGUMENT, "pma_write_tiff_stack: size differs from the stack it is named after");
        pma::write_stack(static_cast<const uint8_t*>(src), (size_t)bytes_per_voxel, s, fs::path(folder));
        pma::set_last_error("");
        return PMA_OK;
    } catch (const pma::Error& e) {
        pma::set_last_error(e.what());
        return e.status;
    } catch (const std::exception& e) {
        pma::set_last_error(std::string("pma_write_tiff_stack: ") + e.what());
        return PMA_ERR_INVALID_ARGUMENT;
    }
}

pma_status pma_read_tiff_stack(const char* folder, void* dst, size_t bytes, pma_dims* dims, int* bytes_per_voxel) {
    try {
        if (!folder) throw pma::Error(PMA_ERR_INVALID_ARGUMENT, "null folder");
        const std::vector<fs::path> files = pma::image_paths(fs::path(folder));
        if (files.empty()) throw pma::Error(PMA_ERR_INVALID_ARGUMENT, std::string(folder) + ": no .tif or .tiff files");
        size_t plane = 0, es = 0;
        for (size_t z = 0; z < files.size(); ++z) {
            const pma::tiff::Image im = pma::tiff::read(files[z].string());
            if (z == 0) {
                plane = (size_t)im.width * im.height;
                es = im.samples;
                if (dims) *dims = pma_dims{(int32_t)im.height, (int32_t)im.width, (int32_t)files.size()};
                if (bytes_per_voxel) *bytes_per_voxel = (int)es;
                if (!dst) return PMA_OK;
                if (bytes != plane * es * files.size())
                    throw pma::Error(PMA_ERR_INVALID_ARGUMENT, "pma_read_tiff_stack: buffer size differs from the stack's");
            }
            if ((size_t)im.width * im.height != plane || im.samples != es)
                throw pma::Error(PMA_ERR_INVALID_ARGUMENT, files[z].string() + ": slices differ in size or samples");
            std::memcpy(static_cast<uint8_t*>(dst) + z * plane * es, im.pixels.data(), plane * es);
        }
        pma::set_last_error("");
        return PMA_OK;
    } catch (const pma::Error& e) {
        pma::set_last_error(e.what());
        return e.status;
    } catch (const std::exception& e) {
        pma::set_last_error(std::string("pma_read_tiff_stack: ") + e.what());
        return PMA_ERR_INVALID_ARGUMENT;
    }
}

}  // extern "C"
