// Text inputs of the file driver: the slicepos file (ROI radius and middles) and
// the deformed list (pattern positions and file names).  Same line formats and
// the same treatment of bad lines as the reference (README.md:47-69).
#include <cerrno>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "offset.hpp"

namespace {

// std::stoul: optional leading white space and sign, then digits; trailing
// characters are ignored; a '-' negates modulo 2^64.  False where stoul throws.
bool parse_ulong(const std::string& s, unsigned long long& out) {
    const char* b = s.c_str();
    char* end = nullptr;
    errno = 0;
    out = std::strtoull(b, &end, 10);
    return end != b && errno != ERANGE;
}

}  // namespace

extern "C" {

// load_slice_begins, src/args_parser.cpp:62-108
int emida_read_slicepos(const char* path, uint64_t* radius, int* radius_ok, emida_size2* mids, uint32_t capacity,
                        uint32_t* count) {
    return guarded([&] {
        if (!path || !count || (capacity && !mids)) throw arg_error("null argument");
        std::ifstream fin(path);
        if (!fin) throw arg_error(std::string("cannot open ") + path);
        std::string line;
        std::getline(fin, line);
        unsigned long long size = 0;
        const bool ok = parse_ulong(line, size);
        const bool warn = mids != nullptr;  // a counting call stays quiet
        if (!ok && warn)
            std::cerr << "Warning: could not parse the first line with size of file " << path << ". Skipping.\n";
        if (radius) *radius = size;
        if (radius_ok) *radius_ok = ok ? 1 : 0;
        uint32_t n = 0, i = 0;
        while (std::getline(fin, line)) {
            ++i;
            const size_t split = line.find(' ');
            if (split == std::string::npos) {
                if (warn) std::cerr << "Warning: could not parse line " << i << " of file " << path << ". Skipping.\n";
                continue;
            }
            unsigned long long x, y;
            if (!parse_ulong(line.substr(0, split), x) || !parse_ulong(line.substr(split + 1), y)) {
                if (warn) std::cerr << "Warning: could not parse line " << i << " of file " << path << ". Skipping.\n";
                continue;
            }
            if (n < capacity) mids[n] = {(uint32_t)x, (uint32_t)y};
            ++n;
        }
        *count = n;
    });
}

// file_processor::load_work, src/process_files.hpp:54-73: "<x> <y> <name>" per line;
// the name is the rest of the line after the one character that follows y.  A line
// whose x or y does not parse keeps x = y = 0 (the reference leaves y unset) and
// an empty name, so its batch fails to load, as in the reference.
int emida_read_deformed_list(const char* path, double* xy, uint32_t capacity, uint32_t* count, char* names,
                             size_t names_capacity, size_t* names_bytes) {
    return guarded([&] {
        if (!path || !count || (capacity && !xy) || (names_capacity && !names)) throw arg_error("null argument");
        std::ifstream fin(path);
        if (!fin) throw arg_error(std::string("cannot open ") + path);
        std::string line;
        uint32_t n = 0;
        size_t nb = 0;
        while (std::getline(fin, line)) {
            std::istringstream iss(line);
            double x = 0, y = 0;
            std::string name;
            if (iss >> x && iss >> y) {
                iss.ignore();
                std::getline(iss, name);
            } else {
                x = 0;
                y = 0;
            }
            if (n < capacity) {
                xy[2 * n] = x;
                xy[2 * n + 1] = y;
            }
            if (nb + name.size() + 1 <= names_capacity) std::memcpy(names + nb, name.c_str(), name.size() + 1);
            nb += name.size() + 1;
            ++n;
        }
        *count = n;
        if (names_bytes) *names_bytes = nb;
    });
}

}  // extern "C"
