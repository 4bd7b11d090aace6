// File input of the pattern-offset pipeline: raw-strip TIFF reading (tiff.cpp),
// used by the C ABI and by the multi-threaded file pipeline (files.cpp).
#pragma once

#include <cstdint>

#include "common.hpp"

namespace emida_amd {

// Width and height of the first image of a TIFF file (src/load_tiff.hpp:57-69).
size2 tiff_size(const char* path);

// Strips of an uncompressed 16-bit single-sample TIFF, copied in file order into a
// size.x * size.y raster (src/load_tiff.hpp:14-53); throws arg_error or
// unsupported_error with a message naming the file.
void tiff_read(const char* path, uint16_t* raster, size2 size);

}  // namespace emida_amd
