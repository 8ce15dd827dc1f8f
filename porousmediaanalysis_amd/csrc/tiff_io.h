// TIFF stacks in the reference's on-disk form: codec (tiff.cpp) and folder runner (folder.cpp).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace pma {
namespace tiff {

// A descriptive tag carried from an input image to the outputs made from it (ImageDescription,
// XResolution, YResolution, ResolutionUnit), values little-endian.
struct Tag {
    uint16_t tag = 0, type = 0;
    uint32_t count = 0;
    std::vector<uint8_t> bytes;
};

struct Image {
    uint32_t width = 0, height = 0;  // TIFF columns (y) and rows (x)
    uint32_t samples = 1;            // 1 (8-bit) or 4 (32-bit values as CMYK bytes)
    uint32_t photometric = 1;        // kept for 8-bit outputs
    std::vector<uint8_t> pixels;     // rows top to bottom, samples interleaved
    std::vector<Tag> keep;
};

Image read(const std::string& path);
void write(const std::string& path, const Image& image);

}  // namespace tiff
}  // namespace pma
