// TIFF stacks in the reference's on-disk form (host code).
//
// The reference reads and writes its volumes as folders of TIFF files through FreeImage
// (Image.h:45-214; DataArray.h:185-252, 339-375): one file per z slice, TIFF row = x, column = y.
//   * input: 8-bit images; a voxel is solid when its raw sample is non-zero (no photometric
//     inversion, so MinIsWhite images are read as stored);
//   * 8-bit outputs (the maximum-ball mask) keep the input's photometric interpretation and
//     descriptive tags and are written LZW-compressed with the horizontal predictor;
//   * 32-bit outputs are 4-sample "CMYK" images whose samples are the little-endian bytes of the
//     value (C = lowest byte), LZW-compressed with the horizontal predictor (Image.h:72-83).
// This is an independent codec: baseline strips, compression none / LZW / PackBits, predictor
// 1 or 2, 8 bits per sample, 1 or 4 samples per pixel, either byte order.
#include "tiff_io.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <stdexcept>

namespace pma {
namespace tiff {
namespace {

struct Reader {
    const std::vector<uint8_t>& d;
    bool be;
    uint16_t u16(size_t o) const {
        if (o + 2 > d.size()) throw std::runtime_error("truncated TIFF");
        return be ? (uint16_t)(d[o] << 8 | d[o + 1]) : (uint16_t)(d[o] | d[o + 1] << 8);
    }
    uint32_t u32(size_t o) const {
        if (o + 4 > d.size()) throw std::runtime_error("truncated TIFF");
        return be ? ((uint32_t)d[o] << 24 | (uint32_t)d[o + 1] << 16 | (uint32_t)d[o + 2] << 8 | d[o + 3])
                  : ((uint32_t)d[o] | (uint32_t)d[o + 1] << 8 | (uint32_t)d[o + 2] << 16 | (uint32_t)d[o + 3] << 24);
    }
};

size_t type_size(uint16_t t) {
    switch (t) {
        case 1: case 2: case 6: case 7: return 1;
        case 3: case 8: return 2;
        case 4: case 9: case 11: return 4;
        case 5: case 10: case 12: return 8;
        default: return 0;
    }
}

// Values of an integer-typed entry (BYTE, SHORT or LONG).
std::vector<uint32_t> int_values(const Reader& r, size_t entry) {
    const uint16_t type = r.u16(entry + 2);
    const uint32_t count = r.u32(entry + 4);
    const size_t sz = type_size(type);
    if (sz == 0 || (type != 1 && type != 3 && type != 4)) throw std::runtime_error("unexpected TIFF tag type");
    const size_t off = (size_t)sz * count <= 4 ? entry + 8 : r.u32(entry + 8);
    if (off + (size_t)sz * count > r.d.size()) throw std::runtime_error("truncated TIFF");
    std::vector<uint32_t> v(count);
    for (uint32_t i = 0; i < count; ++i)
        v[i] = type == 1 ? r.d[off + i] : type == 3 ? r.u16(off + 2 * i) : r.u32(off + 4 * i);
    return v;
}

std::vector<uint8_t> raw_bytes(const Reader& r, size_t entry) {
    const uint16_t type = r.u16(entry + 2);
    const uint32_t count = r.u32(entry + 4);
    const size_t n = type_size(type) * (size_t)count;
    const size_t off = n <= 4 ? entry + 8 : r.u32(entry + 8);
    if (off + n > r.d.size()) throw std::runtime_error("truncated TIFF");
    return std::vector<uint8_t>(r.d.begin() + off, r.d.begin() + off + n);
}

// TIFF LZW (codes most significant bit first, 9 to 12 bits, width grows one code early).
void lzw_decode(const uint8_t* in, size_t n, std::vector<uint8_t>& out, size_t want) {
    struct Entry {
        int prefix;
        uint8_t last, first;
        uint32_t len;
    };
    std::vector<Entry> tab(4096);
    for (int i = 0; i < 256; ++i) tab[i] = Entry{-1, (uint8_t)i, (uint8_t)i, 1};
    int next = 258, width = 9, prev = -1;
    uint64_t acc = 0;
    int bits = 0;
    size_t pos = 0;
    std::vector<uint8_t> tmp;
    const size_t start = out.size();
    auto emit = [&](int code) {
        tmp.resize(tab[code].len);
        for (int c = code, i = (int)tab[code].len - 1; i >= 0; c = tab[c].prefix, --i) tmp[i] = tab[c].last;
        out.insert(out.end(), tmp.begin(), tmp.end());
    };
    while (out.size() - start < want) {
        while (bits < width && pos < n) {
            acc = acc << 8 | in[pos++];
            bits += 8;
        }
        if (bits < width) break;
        const int code = (int)((acc >> (bits - width)) & ((1u << width) - 1));
        bits -= width;
        if (code == 257) break;
        if (code == 256) {
            next = 258;
            width = 9;
            prev = -1;
            continue;
        }
        if (prev < 0) {
            if (code > 255) throw std::runtime_error("bad LZW stream");
            emit(code);
            prev = code;
            continue;
        }
        if (code < next) {
            emit(code);
            if (next < 4096) tab[next++] = Entry{prev, tab[code].first, tab[prev].first, tab[prev].len + 1};
        } else if (code == next && next < 4096) {
            tab[next] = Entry{prev, tab[prev].first, tab[prev].first, tab[prev].len + 1};
            ++next;
            emit(code);
        } else {
            throw std::runtime_error("bad LZW code");
        }
        prev = code;
        if (next + 1 >= (1 << width) && width < 12) ++width;
    }
}

void packbits_decode(const uint8_t* in, size_t n, std::vector<uint8_t>& out, size_t want) {
    size_t pos = 0, produced = 0;
    while (pos < n && produced < want) {
        const int8_t h = (int8_t)in[pos++];
        if (h >= 0) {
            const size_t k = std::min<size_t>((size_t)h + 1, n - pos);
            out.insert(out.end(), in + pos, in + pos + k);
            pos += k;
            produced += k;
        } else if (h != -128) {
            if (pos >= n) break;
            out.insert(out.end(), (size_t)(1 - h), in[pos++]);
            produced += (size_t)(1 - h);
        }
    }
}

class BitWriter {
   public:
    std::vector<uint8_t> out;
    void put(uint32_t code, int width) {
        acc = acc << width | code;
        bits += width;
        while (bits >= 8) {
            out.push_back((uint8_t)(acc >> (bits - 8)));
            bits -= 8;
        }
    }
    void flush() {
        if (bits > 0) out.push_back((uint8_t)(acc << (8 - bits)));
        bits = 0;
    }

   private:
    uint64_t acc = 0;
    int bits = 0;
};

std::vector<uint8_t> lzw_encode(const uint8_t* in, size_t n) {
    // string table: open addressing on (prefix code << 8 | byte), cleared by bumping a stamp
    constexpr uint32_t kSlots = 1u << 14;
    thread_local std::vector<uint32_t> key(kSlots), stamp(kSlots, 0);
    thread_local std::vector<uint16_t> val(kSlots);
    thread_local uint32_t gen = 0;
    auto reset = [&] {
        if (++gen == 0) {
            std::fill(stamp.begin(), stamp.end(), 0u);
            gen = 1;
        }
    };
    auto slot = [&](uint32_t k) {
        uint32_t h = (k * 2654435761u) >> 18;
        while (stamp[h] == gen && key[h] != k) h = (h + 1) & (kSlots - 1);
        return h;
    };
    reset();
    BitWriter w;
    int next = 258, width = 9;
    w.put(256, width);
    if (n == 0) {
        w.put(257, width);
        w.flush();
        return w.out;
    }
    uint32_t cur = in[0];
    for (size_t i = 1; i < n; ++i) {
        const uint32_t k = cur << 8 | in[i];
        const uint32_t h = slot(k);
        if (stamp[h] == gen) {
            cur = val[h];
            continue;
        }
        w.put(cur, width);
        stamp[h] = gen;
        key[h] = k;
        val[h] = (uint16_t)next++;
        // the encoder widens when the next free code needs another bit; the decoder, one code
        // behind, widens one code early
        if (next == 4094) {  // table full: clear (the decoder resets to 9 bits)
            w.put(256, width);
            reset();
            next = 258;
            width = 9;
        } else if (next >= (1 << width) && width < 12) {
            ++width;
        }
        cur = in[i];
    }
    w.put(cur, width);
    ++next;  // the decoder adds an entry after this code
    if (next < 4094 && next >= (1 << width) && width < 12) ++width;
    w.put(257, width);
    w.flush();
    return w.out;
}

void put16(std::vector<uint8_t>& b, uint16_t v) {
    b.push_back((uint8_t)v);
    b.push_back((uint8_t)(v >> 8));
}
void put32(std::vector<uint8_t>& b, uint32_t v) {
    for (int i = 0; i < 4; ++i) b.push_back((uint8_t)(v >> (8 * i)));
}

}  // namespace

Image read(const std::string& path) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) throw std::runtime_error("cannot open " + path);
    std::vector<uint8_t> d;
    uint8_t buf[1 << 16];
    size_t k;
    while ((k = std::fread(buf, 1, sizeof buf, f)) > 0) d.insert(d.end(), buf, buf + k);
    std::fclose(f);
    if (d.size() < 8 || !((d[0] == 'I' && d[1] == 'I') || (d[0] == 'M' && d[1] == 'M')))
        throw std::runtime_error(path + ": not a TIFF file");
    Reader r{d, d[0] == 'M'};
    if (r.u16(2) != 42) throw std::runtime_error(path + ": not a classic TIFF file");
    const size_t ifd = r.u32(4);
    const uint16_t nent = r.u16(ifd);
    Image im;
    std::vector<uint32_t> offsets, counts, bps{1};
    uint32_t compression = 1, predictor = 1, rows_per_strip = 0xFFFFFFFFu, planar = 1;
    for (uint16_t e = 0; e < nent; ++e) {
        const size_t ent = ifd + 2 + 12 * (size_t)e;
        const uint16_t tag = r.u16(ent);
        switch (tag) {
            case 256: im.width = int_values(r, ent)[0]; break;
            case 257: im.height = int_values(r, ent)[0]; break;
            case 258: bps = int_values(r, ent); break;
            case 259: compression = int_values(r, ent)[0]; break;
            case 262: im.photometric = int_values(r, ent)[0]; break;
            case 273: offsets = int_values(r, ent); break;
            case 277: im.samples = int_values(r, ent)[0]; break;
            case 278: rows_per_strip = int_values(r, ent)[0]; break;
            case 279: counts = int_values(r, ent); break;
            case 284: planar = int_values(r, ent)[0]; break;
            case 317: predictor = int_values(r, ent)[0]; break;
            case 270: case 282: case 283: case 296: {
                Tag t;
                t.tag = tag;
                t.type = r.u16(ent + 2);
                t.count = r.u32(ent + 4);
                t.bytes = raw_bytes(r, ent);
                if (r.be && t.type != 2) {  // keep the values little-endian
                    const size_t el = t.type == 5 || t.type == 10 ? 4 : type_size(t.type);
                    for (size_t i = 0; el > 1 && i + el <= t.bytes.size(); i += el)
                        std::reverse(t.bytes.begin() + i, t.bytes.begin() + i + el);
                }
                im.keep.push_back(t);
                break;
            }
            default: break;
        }
    }
    for (uint32_t b : bps)
        if (b != 8) throw std::runtime_error(path + ": only 8 bits per sample are supported");
    if (im.samples != 1 && im.samples != 4) throw std::runtime_error(path + ": 1 or 4 samples per pixel expected");
    if (planar != 1 && im.samples > 1) throw std::runtime_error(path + ": planar images are not supported");
    if (compression != 1 && compression != 5 && compression != 32773)
        throw std::runtime_error(path + ": compression " + std::to_string(compression) + " is not supported");
    if (offsets.size() != counts.size() || offsets.empty()) throw std::runtime_error(path + ": bad strips");
    const size_t row = (size_t)im.width * im.samples;
    const size_t total = row * im.height;
    rows_per_strip = std::min<uint32_t>(rows_per_strip, im.height);
    im.pixels.reserve(total);
    for (size_t s = 0; s < offsets.size() && im.pixels.size() < total; ++s) {
        if ((size_t)offsets[s] + counts[s] > d.size()) throw std::runtime_error(path + ": strip past the end");
        const uint8_t* p = d.data() + offsets[s];
        const size_t want = std::min(total - im.pixels.size(), row * rows_per_strip);
        const size_t before = im.pixels.size();
        if (compression == 1) im.pixels.insert(im.pixels.end(), p, p + std::min<size_t>(counts[s], want));
        else if (compression == 5) lzw_decode(p, counts[s], im.pixels, want);
        else packbits_decode(p, counts[s], im.pixels, want);
        im.pixels.resize(before + want, 0);  // a short strip decodes as zeros
    }
    if (im.pixels.size() != total) throw std::runtime_error(path + ": image data missing");
    if (predictor == 2)
        for (uint32_t y = 0; y < im.height; ++y) {
            uint8_t* q = im.pixels.data() + y * row;
            for (size_t i = im.samples; i < row; ++i) q[i] = (uint8_t)(q[i] + q[i - im.samples]);
        }
    return im;
}

void write(const std::string& path, const Image& im) {
    if (im.samples != 1 && im.samples != 4) throw std::runtime_error("TIFF writer: 1 or 4 samples per pixel");
    const size_t row = (size_t)im.width * im.samples;
    if (im.pixels.size() != row * im.height) throw std::runtime_error("TIFF writer: pixel count");
    const uint32_t rps = im.samples == 1 ? 8 : 2;
    std::vector<uint8_t> body;  // strips after the 8-byte header
    std::vector<uint32_t> offsets, counts;
    std::vector<uint8_t> diff;
    for (uint32_t y0 = 0; y0 < im.height; y0 += rps) {
        const uint32_t rows = std::min(rps, im.height - y0);
        diff.assign(im.pixels.begin() + y0 * row, im.pixels.begin() + (y0 + rows) * row);
        for (uint32_t r = 0; r < rows; ++r) {  // horizontal predictor
            uint8_t* q = diff.data() + r * row;
            for (size_t i = row; i-- > (size_t)im.samples;) q[i] = (uint8_t)(q[i] - q[i - im.samples]);
        }
        std::vector<uint8_t> c = lzw_encode(diff.data(), diff.size());
        offsets.push_back((uint32_t)(8 + body.size()));
        counts.push_back((uint32_t)c.size());
        body.insert(body.end(), c.begin(), c.end());
        if (body.size() & 1) body.push_back(0);
    }
    struct E {
        uint16_t tag, type;
        uint32_t count;
        std::vector<uint8_t> bytes;
    };
    std::vector<E> es;
    auto shorts = [](std::initializer_list<uint32_t> v) {
        std::vector<uint8_t> b;
        for (uint32_t x : v) put16(b, (uint16_t)x);
        return b;
    };
    auto longs = [](const std::vector<uint32_t>& v) {
        std::vector<uint8_t> b;
        for (uint32_t x : v) put32(b, x);
        return b;
    };
    es.push_back({256, 4, 1, longs({im.width})});
    es.push_back({257, 4, 1, longs({im.height})});
    if (im.samples == 1) es.push_back({258, 3, 1, shorts({8})});
    else es.push_back({258, 3, 4, shorts({8, 8, 8, 8})});
    es.push_back({259, 3, 1, shorts({5})});
    es.push_back({262, 3, 1, shorts({im.samples == 4 ? 5u : im.photometric})});
    es.push_back({273, 4, (uint32_t)offsets.size(), longs(offsets)});
    es.push_back({277, 3, 1, shorts({im.samples})});
    es.push_back({278, 4, 1, longs({rps})});
    es.push_back({279, 4, (uint32_t)counts.size(), longs(counts)});
    es.push_back({284, 3, 1, shorts({1})});
    es.push_back({317, 3, 1, shorts({2})});
    if (im.samples == 4) {
        es.push_back({332, 3, 1, shorts({1})});  // InkSet: CMYK
        es.push_back({334, 3, 1, shorts({4})});  // NumberOfInks
        es.push_back({339, 3, 4, shorts({1, 1, 1, 1})});
    } else {
        es.push_back({339, 3, 1, shorts({1})});
    }
    for (const Tag& t : im.keep) es.push_back({t.tag, t.type, t.count, t.bytes});
    std::sort(es.begin(), es.end(), [](const E& a, const E& b) { return a.tag < b.tag; });
    std::vector<uint8_t> out{'I', 'I'};
    put16(out, 42);
    const uint32_t ifd_off = (uint32_t)(8 + body.size());
    put32(out, ifd_off);
    out.insert(out.end(), body.begin(), body.end());
    // IFD, then the values that do not fit in an entry
    const size_t ifd_size = 2 + 12 * es.size() + 4;
    std::vector<uint8_t> extra;
    put16(out, (uint16_t)es.size());
    for (const E& e : es) {
        put16(out, e.tag);
        put16(out, e.type);
        put32(out, e.count);
        if (e.bytes.size() <= 4) {
            std::vector<uint8_t> v = e.bytes;
            v.resize(4, 0);
            out.insert(out.end(), v.begin(), v.end());
        } else {
            put32(out, (uint32_t)(ifd_off + ifd_size + extra.size()));
            extra.insert(extra.end(), e.bytes.begin(), e.bytes.end());
            if (extra.size() & 1) extra.push_back(0);
        }
    }
    put32(out, 0);
    out.insert(out.end(), extra.begin(), extra.end());
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error("cannot write " + path);
    const bool ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
    if (std::fclose(f) != 0 || !ok) throw std::runtime_error("cannot write " + path);
}

}  // namespace tiff
}  // namespace pma
