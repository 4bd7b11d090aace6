// Raw-strip TIFF input (the reference's src/load_tiff.hpp:14-69, which calls libtiff's
// TIFFReadRawStrip).  A baseline TIFF reader of its own: the IFD of the first image
// gives width, height, orientation, compression and the strip table; the strips'
// bytes are copied in file order into the raster, exactly what raw strip reads of an
// uncompressed image produce (no byte swapping, as the reference).  Deliberate
// differences: compressed images, other than 16-bit single-sample pixels and strip
// data that would overflow or not fill the raster are rejected instead of copied as
// they are.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "io.hpp"
#include "offset.hpp"

namespace {

struct tiff_file {
    FILE* f = nullptr;
    explicit tiff_file(const char* path) : f(std::fopen(path, "rb")) {}
    ~tiff_file() {
        if (f) std::fclose(f);
    }
};

struct tiff_image {
    uint32_t width = 0, height = 0;
    uint32_t bits = 1, samples = 1, compression = 1, orientation = 1, planar = 1;
    std::vector<uint64_t> offsets, counts;
};

struct reader {
    FILE* f;
    bool big;  // "MM" byte order
    std::string name;

    void read_at(uint64_t pos, void* dst, size_t n) {
        if (std::fseek(f, (long)pos, SEEK_SET) != 0 || std::fread(dst, 1, n, f) != n)
            throw arg_error(name + ": truncated TIFF file");
    }
    uint32_t u(const unsigned char* p, int n) const {
        uint32_t v = 0;
        for (int i = 0; i < n; ++i) v |= (uint32_t)p[big ? n - 1 - i : i] << (8 * i);
        return v;
    }
    // Values of one IFD entry (SHORT or LONG), inline or at their offset.
    std::vector<uint64_t> values(const unsigned char* e) {
        const uint32_t type = u(e + 2, 2), count = u(e + 4, 4);
        const int sz = type == 3 ? 2 : (type == 4 ? 4 : 0);
        if (sz == 0) throw arg_error(name + ": unsupported TIFF field type " + std::to_string(type));
        if (count == 0) throw arg_error(name + ": empty TIFF field");
        if (count > (1u << 26)) throw arg_error(name + ": TIFF field too long");
        std::vector<unsigned char> raw((size_t)count * sz);
        if (raw.size() <= 4)
            std::memcpy(raw.data(), e + 8, raw.size());
        else
            read_at(u(e + 8, 4), raw.data(), raw.size());
        std::vector<uint64_t> out(count);
        for (uint32_t i = 0; i < count; ++i) out[i] = u(raw.data() + (size_t)i * sz, sz);
        return out;
    }
};

tiff_image parse(const char* path, tiff_file& tf) {
    if (!tf.f) throw arg_error(std::string(path) + ": cannot open");
    reader rd{tf.f, false, path};
    unsigned char hdr[8];
    rd.read_at(0, hdr, 8);
    if (hdr[0] == 'I' && hdr[1] == 'I') rd.big = false;
    else if (hdr[0] == 'M' && hdr[1] == 'M') rd.big = true;
    else throw arg_error(std::string(path) + ": not a TIFF file");
    if (rd.u(hdr + 2, 2) != 42) throw arg_error(std::string(path) + ": not a classic TIFF file");
    const uint32_t ifd = rd.u(hdr + 4, 4);
    unsigned char nbuf[2];
    rd.read_at(ifd, nbuf, 2);
    const uint32_t n = rd.u(nbuf, 2);
    std::vector<unsigned char> entries((size_t)n * 12);
    rd.read_at(ifd + 2, entries.data(), entries.size());
    tiff_image im;
    for (uint32_t i = 0; i < n; ++i) {
        const unsigned char* e = entries.data() + (size_t)i * 12;
        const uint32_t tag = rd.u(e, 2);
        switch (tag) {
            case 256: im.width = (uint32_t)rd.values(e)[0]; break;
            case 257: im.height = (uint32_t)rd.values(e)[0]; break;
            case 258: im.bits = (uint32_t)rd.values(e)[0]; break;
            case 259: im.compression = (uint32_t)rd.values(e)[0]; break;
            case 273: im.offsets = rd.values(e); break;
            case 274: im.orientation = (uint32_t)rd.values(e)[0]; break;
            case 277: im.samples = (uint32_t)rd.values(e)[0]; break;
            case 279: im.counts = rd.values(e); break;
            case 284: im.planar = (uint32_t)rd.values(e)[0]; break;
            default: break;
        }
    }
    return im;
}

}  // namespace

namespace emida_amd {

size2 tiff_size(const char* path) {
    tiff_file tf(path);
    const tiff_image im = parse(path, tf);
    return {im.width, im.height};
}

void tiff_read(const char* path, uint16_t* raster, size2 size) {
    tiff_file tf(path);
    const tiff_image im = parse(path, tf);
    const std::string name(path);
    if (im.width != size.x || im.height != size.y)
        throw arg_error("The size of " + name + " is different than specified size " + std::to_string(size.x) + "," +
                        std::to_string(size.y));
    if (im.orientation != 1)
        throw arg_error("The orientation of '" + name +
                        "' is wrong, only able to process row 0 TOP, col 0 LEFT orientation.");
    if (im.compression != 1) throw unsupported_error(name + ": compressed TIFF images are not supported");
    if (im.bits != 16 || im.samples != 1)
        throw unsupported_error(name + ": only 16-bit single-sample TIFF images are supported");
    if (im.offsets.empty() || im.offsets.size() != im.counts.size())
        throw arg_error(name + ": missing or inconsistent strip table");
    const uint64_t cap = (uint64_t)size.x * size.y * sizeof(uint16_t);
    uint64_t done = 0;
    for (size_t i = 0; i < im.offsets.size(); ++i) {
        if (done + im.counts[i] > cap) throw arg_error(name + ": strip data exceed the image size");
        if (im.counts[i] == 0) continue;
        if (std::fseek(tf.f, (long)im.offsets[i], SEEK_SET) != 0 ||
            std::fread((unsigned char*)raster + done, 1, im.counts[i], tf.f) != im.counts[i])
            throw arg_error(name + ": truncated strip " + std::to_string(i));
        done += im.counts[i];
    }
    // an uncompressed image fills the raster exactly; fewer bytes would leave the
    // rest of the caller's buffer as it was
    if (done != cap) throw arg_error(name + ": strip data are shorter than the image");
}

}  // namespace emida_amd

extern "C" {

int emida_tiff_size(const char* path, emida_size2* size) {
    return guarded([&] {
        if (!path || !size) throw arg_error("null argument");
        const size2 s = tiff_size(path);
        *size = {s.x, s.y};
    });
}

int emida_tiff_load(const char* path, uint16_t* raster, emida_size2 size) {
    return guarded([&] {
        if (!path || !raster) throw arg_error("null argument");
        tiff_read(path, raster, {size.x, size.y});
    });
}

}  // extern "C"
