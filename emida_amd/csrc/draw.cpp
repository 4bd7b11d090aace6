// Visualisation of offset vectors (the reference's draw_tiff, src/draw_tiff.hpp:13-59):
// the pattern as grey RGBA (high byte of each pixel), each vector drawn as 20 points
// from its position, written as an uncompressed 8-bit RGBA TIFF.  The reference
// writes through libtiff; this writer emits the raster as one strip with the same
// tags plus ExtraSamples (unassociated alpha), so readers decode the same pixels.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "offset.hpp"

namespace {

void put16(std::vector<unsigned char>& b, uint16_t v) {
    b.push_back(v & 0xff);
    b.push_back(v >> 8);
}
void put32(std::vector<unsigned char>& b, uint32_t v) {
    for (int i = 0; i < 4; ++i) b.push_back((v >> (8 * i)) & 0xff);
}

// Little-endian baseline TIFF of a width x height RGBA8 raster in one strip.
void write_rgba_tiff(const std::string& path, const std::vector<uint32_t>& rgba, uint32_t w, uint32_t h) {
    const uint32_t data_bytes = w * h * 4;
    const uint32_t bps_at = 8 + data_bytes;  // BitsPerSample array (4 SHORTs)
    const uint32_t ifd_at = bps_at + 8;
    std::vector<unsigned char> b;
    b.reserve(ifd_at + 2 + 13 * 12 + 4);
    b.push_back('I');
    b.push_back('I');
    put16(b, 42);
    put32(b, ifd_at);
    for (uint32_t i = 0; i < w * h; ++i) {  // R, G, B, A bytes of 0xAABBGGRR
        const uint32_t v = rgba[i];
        b.push_back(v & 0xff);
        b.push_back((v >> 8) & 0xff);
        b.push_back((v >> 16) & 0xff);
        b.push_back((v >> 24) & 0xff);
    }
    for (int i = 0; i < 4; ++i) put16(b, 8);
    struct entry {
        uint16_t tag, type;
        uint32_t count, value;
    };
    const entry e[] = {
        {256, 4, 1, w},        {257, 4, 1, h},   {258, 3, 4, bps_at}, {259, 3, 1, 1},
        {262, 3, 1, 2},        {273, 4, 1, 8},   {274, 3, 1, 1},      {277, 3, 1, 4},
        {278, 4, 1, h},        {279, 4, 1, data_bytes}, {284, 3, 1, 1}, {338, 3, 1, 2},
    };
    put16(b, (uint16_t)(sizeof(e) / sizeof(e[0])));
    for (const auto& x : e) {
        put16(b, x.tag);
        put16(b, x.type);
        put32(b, x.count);
        if (x.type == 3 && x.count == 1) {
            put16(b, (uint16_t)x.value);
            put16(b, 0);
        } else {
            put32(b, x.value);
        }
    }
    put32(b, 0);
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) throw arg_error("cannot create " + path);
    const size_t n = std::fwrite(b.data(), 1, b.size(), f);
    const bool ok = std::fclose(f) == 0 && n == b.size();
    if (!ok) throw arg_error("cannot write " + path);
}

}  // namespace

// draw_tiff(pic, size, file_name, vectors, vectors_pos), src/draw_tiff.hpp:26-59
extern "C" int emida_draw_tiff(const uint16_t* pic, emida_size2 size, const char* path, const double* vectors,
                               const emida_size2* positions, uint32_t n_vectors) {
    return guarded([&] {
        if (!pic || !path || (n_vectors && (!vectors || !positions))) throw arg_error("null argument");
        if (size.x == 0 || size.y == 0) throw arg_error("empty picture");
        const size_t area = (size_t)size.x * size.y;
        std::vector<uint32_t> rgba(area);
        for (size_t i = 0; i < area; ++i) {
            const uint32_t v = pic[i] >> 8;
            rgba[i] = 0xFF000000u | v | (v << 8) | (v << 16);
        }
        // draw_vector (src/draw_tiff.hpp:13-23): 20 points at d = 0, 0.05, ... (the
        // same double accumulation), rounded half away from zero, added to the
        // unsigned position; points outside the picture are skipped
        for (uint32_t k = 0; k < n_vectors; ++k) {
            if (!std::isfinite(vectors[2 * k]) || !std::isfinite(vectors[2 * k + 1])) continue;  // NaN offsets
            for (double d = 0; d < 1; d += 0.05) {
                const uint32_t x = positions[k].x + (uint32_t)(int)std::round(d * vectors[2 * k]);
                const uint32_t y = positions[k].y + (uint32_t)(int)std::round(d * vectors[2 * k + 1]);
                if (x < size.x && y < size.y) rgba[(size_t)y * size.x + x] = 0xFF0000FFu;
            }
        }
        write_rgba_tiff(path, rgba, size.x, size.y);
    });
}
