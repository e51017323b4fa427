// fieldTools.cpp -- CPU-only fieldConvert and fieldCalculate on the C ABI of
// lbm_case.h (case files, .LBMBin reader) and lbm_post.h (writers, schemes).
//
// Command lines as the reference's tools (run in a case directory):
//   fieldConvert  -fileType <vtu|vts|tecplot> [-fieldName <k|kMean|S|SMean|momentsMean>]
//                 [-cutPlane <x|y|z>=<coordinate>] [-latestTime]
//                 (applications/fieldConvert/fieldConvert.cu:55-113)
//   fieldCalculate -calculationType <containsNaN|spatialMean|vorticity|div[U]>
//                 [-fileType <vtu|vts|tecplot>] [-latestTime] [-scheme <reference|central>]
//                 (applications/fieldCalculate/fieldCalculate.cu:60-200)
// Output files go to postProcess/ (src/postProcess/postProcess.cuh:61).
// -scheme is an addition of this tool: the reference's derivative stencils are
// reproduced by default; "central" selects the consistent central differences.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <iostream>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "lbm_case.h"
#include "lbm_post.h"

namespace {

[[noreturn]] void die(const std::string &msg) {
    std::cerr << msg << std::endl;
    std::exit(1);
}

struct Args {
    int argc;
    char **argv;
    bool present(const char *name) const {
        for (int i = 1; i < argc; i++)
            if (std::strcmp(argv[i], name) == 0) return true;
        return false;
    }
    std::string value(const char *name) const {
        for (int i = 1; i < argc; i++)
            if (std::strcmp(argv[i], name) == 0) {
                if (i + 1 >= argc) die(std::string("Input argument ") + name + " has not been supplied with a value");
                return argv[i + 1];
            }
        die(std::string("Input argument ") + name + " has not been supplied");
    }
};

const std::vector<std::string> kSolution = {"rho", "u", "v", "w", "m_xx", "m_xy", "m_xz", "m_yy", "m_yz", "m_zz"};

// functionObjects.cuh:60-70
const std::vector<std::string> &field_names(const std::string &prefix, bool custom) {
    static const std::map<std::string, std::vector<std::string>> comp = {
        {"momentsMean", {"rhoMean", "uMean", "vMean", "wMean", "m_xxMean", "m_xyMean", "m_xzMean", "m_yyMean",
                         "m_yzMean", "m_zzMean"}},
        {"S", {"S_xx", "S_xy", "S_xz", "S_yy", "S_yz", "S_zz"}},
        {"SMean", {"S_xxMean", "S_xyMean", "S_xzMean", "S_yyMean", "S_yzMean", "S_zzMean"}},
        {"k", {"k"}},
        {"kMean", {"kMean"}}};
    if (!custom) return kSolution;
    const auto it = comp.find(prefix);
    if (it == comp.end()) throw std::runtime_error("Invalid argument passed to -fieldName");
    return it->second;
}

int format_of(const std::string &t) {
    if (t == "vtu") return LBM_POST_VTU;
    if (t == "vts") return LBM_POST_VTS;
    if (t == "tecplot") return LBM_POST_TECPLOT;
    throw std::runtime_error("Invalid writer \"" + t + "\": valid types are vtu, vts, tecplot");
}

std::vector<int64_t> times(const std::string &prefix) {
    const int n = lbm_post_time_indices(".", prefix.c_str(), nullptr, 0);
    if (n < 0) die(lbm_case_last_error());
    std::vector<int64_t> t((size_t)n);
    if (n) lbm_post_time_indices(".", prefix.c_str(), t.data(), n);
    if (t.empty()) die("No files " + prefix + "_<time>.LBMBin in this directory");
    return t;
}

std::vector<std::vector<float>> read_fields(const std::string &file, const std::vector<std::string> &names, int64_t n) {
    std::vector<std::vector<float>> f(names.size(), std::vector<float>((size_t)n));
    for (size_t i = 0; i < names.size(); i++)
        if (lbm_lbmbin_read_field(file.c_str(), names[i].c_str(), f[i].data()) != 0) die(lbm_case_last_error());
    return f;
}

void write(int format, const std::string &name, int nx, int ny, int nz, const lbm_case_info &ci,
           const std::vector<std::string> &names, const std::vector<std::vector<float>> &f) {
    std::filesystem::create_directories("postProcess");
    std::vector<const char *> np;
    std::vector<const float *> fp;
    for (const auto &s : names) np.push_back(s.c_str());
    for (const auto &v : f) fp.push_back(v.data());
    const std::string path = "postProcess/" + name;
    static const char *ext[] = {".vtu", ".vts", ".dat"};
    std::cout << "writer:" << std::endl << "{" << std::endl;
    std::cout << "    fileName: " << path << ext[format] << ";" << std::endl;
    if (lbm_post_write(format, path.c_str(), nx, ny, nz, ci.Lx, ci.Ly, ci.Lz, (int)names.size(), np.data(),
                       fp.data()) != 0)
        die(lbm_case_last_error());
    std::cout << "    writeStatus: success;" << std::endl << "};" << std::endl;
}

int field_convert(const Args &a, const lbm_case_info &ci) {
    const bool custom = a.present("-fieldName");
    const std::string prefix = custom ? a.value("-fieldName") : std::string(ci.case_name);
    const auto &names = field_names(prefix, custom);
    const int format = format_of(a.value("-fileType"));
    const bool cut = a.present("-cutPlane");
    int axis = -1;
    double coord = 0.0;
    std::string cut_arg;
    if (cut) {
        cut_arg = a.value("-cutPlane");
        if (cut_arg.size() < 3 || cut_arg[1] != '=' || (cut_arg[0] != 'x' && cut_arg[0] != 'y' && cut_arg[0] != 'z'))
            die("Invalid argument passed to -cutPlane: the syntax is -cutPlane x=0.5 (or y=, z=)");
        axis = cut_arg[0] - 'x';
        coord = std::stold(cut_arg.substr(2));
    }
    const auto t = times(prefix);
    const int64_t n = (int64_t)ci.nx * ci.ny * ci.nz;
    for (size_t k = a.present("-latestTime") ? t.size() - 1 : 0; k < t.size(); k++) {
        const std::string file = prefix + "_" + std::to_string(t[k]) + ".LBMBin";
        auto f = read_fields(file, names, n);
        int nx = ci.nx, ny = ci.ny, nz = ci.nz;
        std::string name = prefix + "_" + std::to_string(t[k]);
        if (cut) {
            const double L = axis == 0 ? ci.Lx : axis == 1 ? ci.Ly : ci.Lz;
            (axis == 0 ? nx : axis == 1 ? ny : nz) = 1;
            std::vector<std::vector<float>> plane(f.size(), std::vector<float>((size_t)nx * ny * nz));
            for (size_t i = 0; i < f.size(); i++)
                if (lbm_post_cut_plane(axis, coord, ci.nx, ci.ny, ci.nz, L, f[i].data(), plane[i].data()) != 0)
                    die(lbm_case_last_error());
            f.swap(plane);
            name = prefix + "CutPlane_" + cut_arg + "_" + std::to_string(t[k]);
        }
        write(format, name, nx, ny, nz, ci, names, f);
    }
    return 0;
}

int field_calculate(const Args &a, const lbm_case_info &ci) {
    const std::string what = a.present("-calculationType") ? a.value("-calculationType") : "";
    const std::string prefix = ci.case_name;
    const auto t = times(prefix);
    const int64_t n = (int64_t)ci.nx * ci.ny * ci.nz;
    int scheme = LBM_POST_REFERENCE;
    if (a.present("-scheme")) {
        const std::string s = a.value("-scheme");
        if (s == "central") scheme = LBM_POST_CENTRAL;
        else if (s != "reference") die("Invalid argument passed to -scheme: reference or central");
    }
    constexpr int kOrder = 8;  // fieldCalculate.cuh:67
    for (size_t k = a.present("-latestTime") ? t.size() - 1 : 0; k < t.size(); k++) {
        const std::string file = prefix + "_" + std::to_string(t[k]) + ".LBMBin";
        const auto f = read_fields(file, kSolution, n);
        if (what == "containsNaN") {
            // fieldCalculate.cuh:70-118
            std::cout << "Time: " << t[k] << std::endl << "{" << std::endl;
            size_t bad = 0;
            for (size_t i = 0; i < f.size(); i++) {
                bool nan = false;
                for (float v : f[i]) nan = nan || std::isnan(v);
                if (nan) {
                    std::cout << "    NaN detected in field " << kSolution[i] << std::endl;
                    bad++;
                }
            }
            if (bad == 0)
                std::cout << "    No NaN values detected in any field." << std::endl;
            else
                std::cout << "    Total number of fields with NaN values: " << bad << std::endl;
            std::cout << "};" << std::endl << std::endl;
        } else if (what == "spatialMean") {
            // fieldCalculate.cuh:120-150: FP32 sum in index order
            std::cout << "Time: " << t[k] << std::endl << "{" << std::endl;
            for (size_t i = 0; i < f.size(); i++) {
                float s = 0.f;
                for (float v : f[i]) s += v;
                std::cout << "    mean(" << kSolution[i] << "): " << s / (float)f[i].size() << ";" << std::endl;
            }
            std::cout << "};" << std::endl << std::endl;
        } else if (what == "vorticity" || what == "div[U]") {
            const int format = format_of(a.value("-fileType"));
            if (what == "vorticity") {
                std::vector<std::vector<float>> o(4, std::vector<float>((size_t)n));
                if (lbm_post_curl(scheme, kOrder, ci.nx, ci.ny, ci.nz, f[1].data(), f[2].data(), f[3].data(),
                                  o[0].data(), o[1].data(), o[2].data()) != 0)
                    die(lbm_case_last_error());
                for (int64_t i = 0; i < n; i++)
                    o[3][(size_t)i] = std::sqrt(o[0][(size_t)i] * o[0][(size_t)i] + o[1][(size_t)i] * o[1][(size_t)i] +
                                                o[2][(size_t)i] * o[2][(size_t)i]);
                write(format, "vorticity_" + std::to_string(t[k]), ci.nx, ci.ny, ci.nz, ci,
                      {"omega_x", "omega_y", "omega_z", "mag[omega]"}, o);
            } else {
                std::vector<std::vector<float>> d(1, std::vector<float>((size_t)n));
                if (lbm_post_div(scheme, kOrder, ci.nx, ci.ny, ci.nz, f[1].data(), f[2].data(), f[3].data(),
                                 d[0].data()) != 0)
                    die(lbm_case_last_error());
                write(format, "div[U]_" + std::to_string(t[k]), ci.nx, ci.ny, ci.nz, ci, {"div[U]"}, d);
            }
        }
    }
    std::cout << "End" << std::endl << std::endl;
    return 0;
}

}  // namespace

int run_field_tool(bool calculate, int argc, char **argv) {
    try {
        lbm_case_info ci;
        if (lbm_case_read(".", &ci) != 0) die(lbm_case_last_error());
        const Args a{argc, argv};
        return calculate ? field_calculate(a, ci) : field_convert(a, ci);
    } catch (const std::exception &e) {
        die(e.what());
    }
}
