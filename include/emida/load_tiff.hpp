// emida/load_tiff.hpp - the reference's TIFF helpers (src/load_tiff.hpp:14-69,
// src/draw_tiff.hpp:26-59) with the same signatures, over the C ABI in emida_amd.h.
// load_tiff and get_size print the error to stderr and return false, as the reference
// does after libtiff's message.
#pragma once

#include <iostream>
#include <string>
#include <vector>

#include "emida/gpu_offset.hpp"

namespace emida {

// src/load_tiff.hpp:14-53
inline bool load_tiff(const std::string& file_name, uint16_t* raster, size2_t size) {
    if (emida_tiff_load(file_name.c_str(), raster, detail::c(size)) != EMIDA_OK) {
        std::cerr << emida_last_error() << "\n";
        return false;
    }
    return true;
}

// src/load_tiff.hpp:55-69
inline bool get_size(const std::string& file_name, size2_t& size) {
    emida_size2 s{};
    if (emida_tiff_size(file_name.c_str(), &s) != EMIDA_OK) {
        std::cerr << emida_last_error() << "\n";
        return false;
    }
    size = {s.x, s.y};
    return true;
}

// src/draw_tiff.hpp:26-59 (uint16 patterns)
inline void draw_tiff(const uint16_t* pic, size2_t size, std::string file_name, std::vector<vec2<double>> vectors,
                      std::vector<size2_t> vectors_pos) {
    std::vector<double> v(2 * vectors.size());
    std::vector<emida_size2> pos(vectors_pos.size());
    for (size_t i = 0; i < vectors.size(); ++i) {
        v[2 * i] = vectors[i].x;
        v[2 * i + 1] = vectors[i].y;
    }
    for (size_t i = 0; i < pos.size(); ++i) pos[i] = detail::c(vectors_pos[i]);
    if (vectors.size() != vectors_pos.size()) throw std::runtime_error("draw_tiff: one position per vector");
    detail::check(emida_draw_tiff(pic, detail::c(size), file_name.c_str(), v.data(), pos.data(),
                                  (uint32_t)v.size() / 2),
                  "draw_tiff");
}

}  // namespace emida
