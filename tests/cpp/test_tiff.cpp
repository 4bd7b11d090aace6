// C++ drop-in check of emida/load_tiff.hpp (the reference's load_tiff / get_size /
// draw_tiff signatures) on the CPU: prints the size and a checksum of the reference's
// TIFF fixture for the Python test to compare with an independent decoder, and
// exercises the error returns.  Usage: test_tiff <tiff> <out.tif>
#include <cstdio>
#include <vector>

#include "emida/load_tiff.hpp"

int main(int argc, char** argv) {
    using namespace emida;
    if (argc < 3) return 2;
    size2_t size{};
    if (!get_size(argv[1], size)) return 3;
    std::vector<uint16_t> raster(size.area());
    if (!load_tiff(argv[1], raster.data(), size)) return 4;
    unsigned long long sum = 0, weighted = 0;
    for (size_t i = 0; i < raster.size(); ++i) {
        sum += raster[i];
        weighted += (unsigned long long)raster[i] * (i % 9973);
    }
    std::printf("size %u %u sum %llu weighted %llu\n", size.x, size.y, sum, weighted);
    // a wrong size is refused with false, as the reference
    if (load_tiff(argv[1], raster.data(), {size.x + 1, size.y})) return 5;
    size2_t none{};
    if (get_size("/nonexistent/file.tif", none)) return 6;
    draw_tiff(raster.data(), size, argv[2], {{10.0, -5.0}}, {{100, 100}});
    std::printf("ok\n");
    return 0;
}
