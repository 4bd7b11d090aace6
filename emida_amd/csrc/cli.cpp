// emida: command-line driver of the pattern-offset pipeline, the counterpart of
// the reference's executable (src/main.cpp, src/args_parser.cpp:110-283,
// src/process_files.hpp:135-264).  Built on the C ABI only, as any user program.
//
// Same options, input files and standard output:
//   per pattern   "%f %f %llu %s\n"  (x, y, number of ROIs, file name)
//   per ROI       "%lu %lu %f %f\n"  (ROI middle x, y, offset x, y)
//                 plus six "%f" fit coefficients with -q/--writecoefs
// Deliberate differences: patterns are printed in list order (the reference prints
// batches in the order they finish); --crosspolicy also accepts "mfma" (matrix-core
// brute correlation with exact refinement, results equal to "brute"); invalid
// --batchsize/--loadworkers values and an unparsable slicepos size line are errors
// (the reference goes on with defaults or an unset size); a batch that fails to
// load is named on stderr; -a prints the pipeline's own timings.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "emida_amd.h"

namespace {

struct option_spec {
    const char* short_name;  // "" for none
    const char* long_name;
    bool takes_value;
    bool mandatory;
    const char* value_name;
    const char* help;
};

const option_spec kOptions[] = {
    {"i", "initial", true, true, "TIFF_FILE", "Path to picture of initial undeformed picture."},
    {"d", "deformed", true, true, "TXT_FILE", "Path to file with list of paths to deformed pictures."},
    {"o", "outpics", true, false, "FOLDER",
     "Accepted for compatibility; offsets are not drawn into pictures."},
    {"c", "crosssize", true, false, "X_SIZE,Y_SIZE",
     "Size of neighbourhood that is analyzed in each slice of picture. Must be odd numbers."},
    {"s", "slicesize", true, false, "X_RADIUS,Y_RADIUS",
     "Half of the slice size; overrides the size in the slicepos file."},
    {"", "slicestep", true, false, "X_SIZE,Y_SIZE", "Cannot be combined with slicepos."},
    {"b", "slicepos", true, true, "FILE_PATH", "Path to file with positions of slice middles in each picture"},
    {"a", "analysis", false, false, "", "Write time measurements of the pipeline to standard error output."},
    {"q", "writecoefs", false, false, "",
     "In addition to offsets, output also coefficients of parabola fitting for each region of interest."},
    {"", "precision", true, false, "double|float", "Specifies the floating type to be used."},
    {"f", "fitsize", true, false, "SIZE",
     "Size of neighbourhood used to find the subpixel maximum. Allowed values: 3, 5, 7, 9"},
    {"", "crosspolicy", true, false, "brute|fft|mfma", "How the cross correlation is computed."},
    {"", "batchsize", true, false, "NUMBER", "Specifies how many files are processed in one batch"},
    {"", "loadworkers", true, false, "NUMBER", "Specifies number of workers that load input patterns simultaneously"},
    {"h", "help", false, false, "", "Print a usage message on standard error output and exit."},
};

std::string help_text() {
    std::ostringstream o;
    o << "Usage: emida [options]\n";
    for (const auto& s : kOptions) {
        o << "  ";
        if (s.short_name[0]) o << "-" << s.short_name << ",";
        o << "--" << s.long_name;
        if (s.takes_value) o << " " << s.value_name;
        if (s.mandatory) o << " (mandatory)";
        o << "\n      " << s.help << "\n";
    }
    return o.str();
}

const option_spec* find_option(const std::string& name) {
    for (const auto& s : kOptions)
        if (name == s.long_name || (s.short_name[0] && name == s.short_name)) return &s;
    return nullptr;
}

// "-x value", "-xvalue", grouped flags "-aq", "--name value", "--name=value".
bool parse_args(int argc, char** argv, std::map<std::string, std::string>& out) {
    bool ok = true;
    for (int i = 1; i < argc; ++i) {
        const std::string arg = argv[i];
        if (arg.size() < 2 || arg[0] != '-') {
            ok = false;  // no plain arguments are used
            continue;
        }
        if (arg[1] == '-') {
            const size_t eq = arg.find('=');
            const std::string name = arg.substr(2, eq == std::string::npos ? std::string::npos : eq - 2);
            const option_spec* s = find_option(name);
            if (!s || name.size() == 1) {
                ok = false;
                continue;
            }
            if (!s->takes_value) {
                if (eq != std::string::npos) ok = false;
                out[s->long_name] = "";
            } else if (eq != std::string::npos) {
                out[s->long_name] = arg.substr(eq + 1);
            } else if (i + 1 < argc) {
                out[s->long_name] = argv[++i];
            } else {
                ok = false;
            }
            continue;
        }
        for (size_t k = 1; k < arg.size(); ++k) {
            const option_spec* s = find_option(std::string(1, arg[k]));
            if (!s) {
                ok = false;
                break;
            }
            if (!s->takes_value) {
                out[s->long_name] = "";
                continue;
            }
            if (k + 1 < arg.size()) out[s->long_name] = arg.substr(k + 1);
            else if (i + 1 < argc) out[s->long_name] = argv[++i];
            else ok = false;
            break;
        }
    }
    for (const auto& s : kOptions)
        if (s.mandatory && !out.count(s.long_name)) ok = false;
    return ok;
}

// get_ints, src/args_parser.cpp:14-33: exactly two comma-separated unsigned values.
bool parse_pair(const std::string& v, emida_size2& r) {
    std::istringstream ss(v);
    std::string tok;
    uint32_t vals[2];
    int n = 0;
    while (n < 2 && std::getline(ss, tok, ',')) {
        char* end = nullptr;
        const char* b = tok.c_str();
        const unsigned long x = std::strtoul(b, &end, 10);
        if (end == b) return false;
        vals[n++] = (uint32_t)x;
    }
    if (n != 2) return false;
    r = {vals[0], vals[1]};
    return true;
}

bool parse_positive(const std::string& v, uint32_t& r) {
    char* end = nullptr;
    const long x = std::strtol(v.c_str(), &end, 10);
    if (end == v.c_str() || x <= 0) return false;
    r = (uint32_t)x;
    return true;
}

struct run_config {
    std::string initial, deformed;
    emida_size2 pic{}, slice{64, 64}, cross{};
    std::vector<emida_size2> mids, begins;
    bool write_coefs = false, analysis = false;
    emida_type T = EMIDA_F64;
    int fit = 3;
    emida_cross_policy policy = EMIDA_CROSS_POLICY_FFT;
    uint32_t batch = 1, load_workers = 1;
};

// params::parse, src/args_parser.cpp:110-283.  0: run; 1: error; 2: help shown.
int configure(int argc, char** argv, run_config& c) {
    std::map<std::string, std::string> a;
    const bool ok = parse_args(argc, argv, a);
    if (a.count("help")) {
        std::cerr << help_text() << "\n";
        return 2;
    }
    if (!ok) {
        std::cerr << "Argument parsing error.\n" << help_text() << "\n";
        return 1;
    }
    c.initial = a["initial"];
    c.deformed = a["deformed"];
    if (a.count("slicesize")) {
        if (!parse_pair(a["slicesize"], c.slice)) {
            std::cerr << "Argument parsing error.\n" << help_text() << "\n";
            return 1;
        }
        c.slice = {c.slice.x * 2, c.slice.y * 2};
    }
    if (a.count("slicestep")) {
        std::cerr << "Error: cannot use slicestep when slicepos specified\n";
        return 1;
    }
    {
        const std::string& path = a["slicepos"];
        uint64_t radius = 0;
        int radius_ok = 0;
        uint32_t n = 0;
        if (emida_read_slicepos(path.c_str(), &radius, &radius_ok, nullptr, 0, &n) != EMIDA_OK) {
            std::cerr << "Error: " << emida_last_error() << "\n";
            return 1;
        }
        c.mids.resize(n);
        if (emida_read_slicepos(path.c_str(), &radius, &radius_ok, c.mids.data(), n, &n) != EMIDA_OK) {
            std::cerr << "Error: " << emida_last_error() << "\n";
            return 1;
        }
        if (a.count("slicesize")) {
            std::cerr << "Warning: size written in --slicepos file will be overriden by --slicesize.\n";
        } else {
            if (!radius_ok) {
                std::cerr << "Error: the slicepos file " << path << " does not give the slice size.\n";
                return 1;
            }
            c.slice = {(uint32_t)radius * 2U, (uint32_t)radius * 2U};
        }
        for (const auto& m : c.mids) {
            if (m.x < c.slice.x / 2 || m.y < c.slice.y / 2) {
                std::cerr << "Error: A slice out of bounds\n";
                return 1;
            }
            c.begins.push_back({m.x - c.slice.x / 2, m.y - c.slice.y / 2});
        }
    }
    const bool cross_given = a.count("crosssize") != 0;
    if (cross_given) {
        if (!parse_pair(a["crosssize"], c.cross)) {
            std::cerr << "Argument parsing error.\n" << help_text() << "\n";
            return 1;
        }
    } else {
        c.cross = {c.slice.x * 2 - 1, c.slice.y * 2 - 1};
    }
    if (c.cross.x > c.slice.x * 2 - 1 || c.cross.y > c.slice.y * 2 - 1) {
        std::cerr << "Error: cross size must be lesser or equal to slice_size * 2 - 1.\n";
        return 1;
    }
    if (a.count("precision")) {
        const std::string& p = a["precision"];
        if (p == "float") c.T = EMIDA_F32;
        else if (p == "double") c.T = EMIDA_F64;
        else {
            std::cerr << "Invalid precision argument.\n";
            return 1;
        }
    }
    if (a.count("fitsize")) {
        char* end = nullptr;
        c.fit = (int)std::strtol(a["fitsize"].c_str(), &end, 10);
        if (end == a["fitsize"].c_str() || (c.fit != 3 && c.fit != 5 && c.fit != 7 && c.fit != 9)) {
            std::cerr << "Error: the fitsize " << a["fitsize"] << " is not supported.\n";
            return 1;
        }
    }
    if (a.count("crosspolicy")) {
        const std::string& p = a["crosspolicy"];
        if (p == "fft") {
            c.policy = EMIDA_CROSS_POLICY_FFT;
            if (cross_given) std::cerr << "Warning: --crosssize (-c) is ignored with FFT policy.\n";
        } else if (p == "brute") {
            c.policy = EMIDA_CROSS_POLICY_BRUTE;
        } else if (p == "mfma") {
            c.policy = EMIDA_CROSS_POLICY_BRUTE_MFMA;
        } else {
            std::cerr << "Invalid crosspolicy argument.\n";
            return 1;
        }
    }
    if (a.count("batchsize") && !parse_positive(a["batchsize"], c.batch)) {
        std::cerr << "Error: batchsize must be greater than 0.\n";
        return 1;
    }
    if (a.count("loadworkers") && (!parse_positive(a["loadworkers"], c.load_workers) || c.load_workers > 64)) {
        std::cerr << "Error: loadworkers must be in 1..64.\n";
        return 1;
    }
    c.analysis = a.count("analysis") != 0;
    c.write_coefs = a.count("writecoefs") != 0;
    if (emida_tiff_size(c.initial.c_str(), &c.pic) != EMIDA_OK) {
        std::cerr << emida_last_error() << "\n";
        return 1;
    }
    return 0;
}

struct printer {
    const run_config* c;
    const std::vector<double>* xy;
    const std::vector<std::string>* names;
};

double norm_nan(double v) { return std::isnan(v) ? std::nan("") : v; }

// finalize_offsets, src/process_files.hpp:235-264
void print_batch(void* user, uint32_t first, uint32_t count, int status, const char* error, const double* offs,
                 const double* coefs) {
    const printer& p = *(const printer*)user;
    if (status != EMIDA_OK) {
        std::fprintf(stderr, "%s\n", error ? error : "batch failed");
        return;
    }
    const size_t n = p.c->mids.size();
    for (uint32_t j = 0; j < count; ++j) {
        const uint32_t f = first + j;
        std::printf("%f %f %llu %s\n", (*p.xy)[2 * f], (*p.xy)[2 * f + 1], (unsigned long long)n,
                    (*p.names)[f].c_str());
        for (size_t r = 0; r < n; ++r) {
            const size_t i = (size_t)j * n + r;
            const auto& m = p.c->mids[r];
            const double ox = norm_nan(offs[2 * i]), oy = norm_nan(offs[2 * i + 1]);
            if (p.c->write_coefs) {
                const double* k = coefs + 6 * i;
                std::printf("%lu %lu %f %f %f %f %f %f %f %f\n", (unsigned long)m.x, (unsigned long)m.y, ox, oy, k[0],
                            k[1], k[2], k[3], k[4], k[5]);
            } else {
                std::printf("%lu %lu %f %f\n", (unsigned long)m.x, (unsigned long)m.y, ox, oy);
            }
        }
    }
}

struct offset_handle {
    emida_offset* h = nullptr;
    ~offset_handle() {
        if (h) emida_offset_destroy(h);
    }
};

void check(int rc) {
    if (rc != EMIDA_OK) throw std::runtime_error(emida_last_error());
}

// file_processor::process_files, src/process_files.hpp:135-181
int run(const run_config& c) {
    std::vector<uint16_t> initial((size_t)c.pic.x * c.pic.y);
    if (emida_tiff_load(c.initial.c_str(), initial.data(), c.pic) != EMIDA_OK) {
        std::cerr << emida_last_error() << "\n";
        return 0;  // the reference returns quietly after load_tiff's message
    }
    offset_handle off;
    check(emida_offset_create(&off.h, c.T, EMIDA_U16, c.pic, c.begins.data(), (uint32_t)c.begins.size(), c.slice,
                              c.cross, c.batch, c.policy, c.fit, c.load_workers * 2));
    check(emida_offset_allocate_memory(off.h, initial.data()));

    uint32_t n = 0;
    size_t nb = 0;
    check(emida_read_deformed_list(c.deformed.c_str(), nullptr, 0, &n, nullptr, 0, &nb));
    std::vector<double> xy(2 * (size_t)n);
    std::vector<char> packed(nb);
    check(emida_read_deformed_list(c.deformed.c_str(), xy.data(), n, &n, packed.data(), packed.size(), &nb));
    if (n == 0) {
        std::cerr << "The deformed list is empty.\n";
        return 0;
    }
    std::vector<std::string> names;
    std::vector<const char*> paths;
    names.reserve(n);
    for (size_t pos = 0; pos < nb; pos += names.back().size() + 1) names.emplace_back(packed.data() + pos);
    for (const auto& s : names) paths.push_back(s.c_str());

    printer p{&c, &xy, &names};
    emida_files_stats st{};
    check(emida_process_files(off.h, paths.data(), n, c.load_workers, 1, print_batch, &p, &st));
    std::fflush(stdout);
    if (c.analysis)
        std::fprintf(stderr,
                     "files %u batches %u failed %u wall %.3f ms load %.3f ms (summed over %u loaders) "
                     "upload wait %.3f ms\n",
                     n, st.batches, st.failed_batches, st.wall_ms, st.load_ms, c.load_workers, st.upload_wait_ms);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    run_config c;
    const int rc = configure(argc, argv, c);
    if (rc == 1) return 1;
    if (rc == 2) return 0;
    try {
        return run(c);
    } catch (const std::exception& e) {
        std::cout << "Error: " << e.what() << "\n";
        return 1;
    }
}
