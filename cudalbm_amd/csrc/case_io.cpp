// case_io.cpp -- the reference's case interface and .LBMBin format on the host.
//
// Parsing follows the reference's rules so that its case directories run
// unchanged:
//   - a parameter is read from the FIRST line that contains its name as a
//     substring; that line is split on spaces (empty tokens dropped) and the
//     second token, minus its last character (the ';'), is the value
//     (src/strings.cuh:561-595);
//   - a block "name { ... };" is found by findBlockLine / extractBlock
//     (src/strings.cuh:262-409): '//' comments are ignored for matching, braces
//     are counted, the block is returned with its original text;
//   - a number is an optional sign, digits and at most one '.'
//     (src/strings.cuh:454-505); anything else in initialConditions must be
//     "equilibrium".
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "lbm_amd.h"
#include "lbm_case.h"

namespace fs = std::filesystem;

namespace {

thread_local std::string g_err;

using Lines = std::vector<std::string>;

Lines read_lines(const std::string &path) {
    if (!fs::exists(path)) throw std::runtime_error(path + " file not opened");
    std::ifstream in(path);
    Lines S;
    std::string s;
    while (std::getline(in, s)) S.push_back(s);
    return S;
}

std::string remove_comments(const std::string &s) {
    const auto p = s.find("//");
    return p == std::string::npos ? s : s.substr(0, p);
}

std::string trim(const std::string &s, bool semicolon) {
    const auto a = s.find_first_not_of(" \t\n\r\f\v");
    if (a == std::string::npos) return "";
    const auto b = s.find_last_not_of(semicolon ? " \t\n\r\f\v;" : " \t\n\r\f\v");
    return s.substr(a, b - a + 1);
}

bool only_space(const std::string &s) {
    for (unsigned char c : s)
        if (!std::isspace(c)) return false;
    return true;
}

Lines split_spaces(const std::string &s) {
    Lines r;
    std::string cur;
    bool any = false;
    for (char c : s) {
        if (c == ' ') {
            if (any || !cur.empty()) r.push_back(cur);
            cur.clear();
            any = false;
        } else {
            cur.push_back(c);
            any = true;
        }
    }
    if (!cur.empty()) r.push_back(cur);
    r.erase(std::remove(r.begin(), r.end(), std::string()), r.end());
    return r;
}

std::string parameter_line(const Lines &S, const std::string &name) {
    for (const auto &line : S)
        if (line.find(name) != std::string::npos) {
            const Lines t = split_spaces(line);
            if (t.size() < 2 || t[1].empty()) throw std::runtime_error("Parameter " + name + " malformed");
            return t[1].substr(0, t[1].size() - 1);
        }
    throw std::runtime_error("Parameter " + name + " not found");
}

bool is_number(const std::string &s) {
    if (s.empty()) return false;
    size_t i = 0;
    if (s[0] == '+' || s[0] == '-') {
        i = 1;
        if (s.size() == 1) return false;
    }
    bool digits = false, dot = false;
    for (; i < s.size(); i++) {
        if (std::isdigit((unsigned char)s[i]))
            digits = true;
        else if (s[i] == '.' && !dot)
            dot = true;
        else
            return false;
    }
    return digits;
}

size_t find_block_line(const Lines &L, const std::string &name) {
    for (size_t i = 0; i < L.size(); i++) {
        const std::string t = trim(remove_comments(L[i]), false);
        if (t.compare(0, name.size(), name) != 0) continue;
        const std::string rest = trim(t.substr(name.size()), false);
        if (rest.empty() || rest[0] == '{' || rest[0] == ';') return i;
        std::istringstream iss(rest);
        std::string tok;
        iss >> tok;
        if (tok.empty() || tok == "{" || tok == ";") return i;
    }
    throw std::runtime_error("Field block \"" + name + "\" not defined");
}

Lines extract_block(const Lines &L, const std::string &name) {
    Lines out;
    const size_t bl = find_block_line(L, name);
    size_t open = 0;
    bool found = false;
    if (remove_comments(L[bl]).find('{') != std::string::npos) {
        out.push_back(L[bl]);
        open = bl;
        found = true;
    } else {
        for (size_t i = bl + 1; i < L.size(); i++) {
            const std::string p = remove_comments(L[i]);
            if (p.find('}') != std::string::npos)
                throw std::runtime_error("Found closing brace before opening brace for block '" + name + "'");
            const std::string t = trim(p, false);
            if (!only_space(t) && t.find('{') == std::string::npos)
                throw std::runtime_error("Non-whitespace content found between block declaration and opening brace for block '" + name + "'");
            if (p.find('{') != std::string::npos) {
                out.push_back(L[i]);
                open = i;
                found = true;
                break;
            }
        }
    }
    if (!found) throw std::runtime_error("Opening brace not found for block '" + name + "'");
    int depth = 1;
    for (size_t i = open + 1; i < L.size(); i++) {
        for (char c : remove_comments(L[i])) {
            if (c == '{') depth++;
            if (c == '}') depth--;
        }
        out.push_back(L[i]);
        if (depth == 0) return out;
    }
    throw std::runtime_error("Unbalanced braces for block '" + name + "'");
}

const char *kMomentNames[10] = {"rho", "u", "v", "w", "m_xx", "m_xy", "m_xz", "m_yy", "m_yz", "m_zz"};

// Region value with the reference scaling (boundaryValue.cuh:147-254):
// rho as given; u,v,w x3; diagonal moments x4.5; off-diagonal x9;
// "equilibrium" second moments from the region's (unscaled) velocities, all
// scaled by 4.5 (the off-diagonal factor is 4.5 there, not 9).
double region_value(const Lines &ic, const std::string &field, const std::string &region) {
    const Lines fb = extract_block(ic, "field " + field);
    const Lines rb = extract_block(fb, region);
    const std::string v = parameter_line(rb, "value");
    auto raw = [&](const std::string &f) {
        const Lines rb2 = extract_block(extract_block(ic, "field " + f), region);
        const std::string s = parameter_line(rb2, "value");
        if (!is_number(s)) throw std::runtime_error("Invalid boundary value " + s + ". Value must be a number.");
        return std::stod(s);
    };
    if (is_number(v)) {
        const double x = std::stod(v);
        if (field == "rho") return x;
        if (field == "u" || field == "v" || field == "w") return 3.0 * x;
        if (field == "m_xx" || field == "m_yy" || field == "m_zz") return 4.5 * x;
        if (field == "m_xy" || field == "m_xz" || field == "m_yz") return 9.0 * x;
        throw std::runtime_error("Invalid field name \" " + field + "\" for equilibrium distribution");
    }
    if (v == "equilibrium") {
        if (field == "m_xx") { const double u = raw("u"); return 4.5 * u * u; }
        if (field == "m_xy") return 4.5 * raw("u") * raw("v");
        if (field == "m_xz") return 4.5 * raw("u") * raw("w");
        if (field == "m_yy") { const double v2 = raw("v"); return 4.5 * v2 * v2; }
        if (field == "m_yz") return 4.5 * raw("v") * raw("w");
        if (field == "m_zz") { const double w = raw("w"); return 4.5 * w * w; }
    }
    throw std::runtime_error("Entry for " + field + " in region " + region +
                             " not a valid numerical value and not an equilibrium moment");
}

std::string join(const std::string &dir, const std::string &f) {
    if (!dir.empty() && dir != ".") return (fs::path(dir) / f).string();
    return f;
}

bool all_digits(const std::string &s) {
    if (s.empty()) return false;
    for (unsigned char c : s)
        if (!std::isdigit(c)) return false;
    return true;
}

int64_t latest_time(const std::string &dir, const std::string &caseName) {
    int64_t best = -1;
    const std::string prefix = caseName + "_";
    const fs::path d = dir.empty() ? fs::current_path() : fs::path(dir);
    if (!fs::exists(d)) return 0;
    for (const auto &e : fs::directory_iterator(d)) {
        if (!e.is_regular_file() || e.path().extension() != ".LBMBin") continue;
        const std::string stem = e.path().stem().string();
        if (stem.size() <= prefix.size() || stem.compare(0, prefix.size(), prefix) != 0) continue;
        const std::string num = stem.substr(prefix.size());
        if (!all_digits(num)) continue;
        best = std::max<int64_t>(best, std::stoll(num));
    }
    return best < 0 ? 0 : best;
}

// reference 8^3 block order of a global (x,y,z) (src/globalFunctions.cuh:347-370)
inline size_t tiled8(int x, int y, int z, int nx, int ny) {
    const size_t NBX = nx / 8, NBY = ny / 8;
    return (size_t)(x % 8) + 8 * ((size_t)(y % 8) + 8 * ((size_t)(z % 8) + 8 * ((size_t)(x / 8) + NBX * ((size_t)(y / 8) + NBY * (size_t)(z / 8)))));
}

struct Header {
    int nx = 0, ny = 0, nz = 0, nvars = 0, scalar = 0;
    int64_t time_step = 0;
    bool little = false;
    std::vector<std::string> names;
    size_t data_pos = 0;
};

std::string strip(const std::string &s) { return trim(s, false); }

Header parse_header(const std::string &path) {
    if (!fs::exists(path)) throw std::runtime_error("File does not exist: " + path);
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("Cannot open file: " + path);
    in.seekg(0, std::ios::end);
    const size_t fsize = (size_t)in.tellg();
    in.seekg(0, std::ios::beg);
    Header h;
    std::string line;
    bool sys = false, info = false, names = false, data = false, found_data = false;
    bool seen_sys = false, seen_info = false;
    size_t declared = 0;
    while (std::getline(in, line)) {
        const std::string t = strip(line);
        if (t.empty()) continue;
        if (t == "systemInformation") { sys = true; seen_sys = true; continue; }
        if (t == "fieldInformation") { info = true; seen_info = true; continue; }
        if (t == "fieldData") { data = true; continue; }
        if (sys) {
            if (t == "};") sys = false;
            else if (t.find("binaryType") != std::string::npos) h.little = t.find("littleEndian") != std::string::npos;
            else if (t.find("scalarType") != std::string::npos) {
                if (t.find("32 bit") != std::string::npos) h.scalar = 4;
                else if (t.find("64 bit") != std::string::npos) h.scalar = 8;
                else throw std::runtime_error("Invalid scalarType");
            }
            continue;
        }
        if (info) {
            if (names) {
                if (t == "};") names = false;
                else if (t != "{") h.names.push_back(trim(t, true));
                continue;
            }
            if (t == "};") { info = false; continue; }
            if (t.find("timeStep") != std::string::npos) {
                const Lines tok = split_spaces(t);
                std::string v = t.substr(t.find("timeStep") + 8);
                v = trim(v, true);
                h.time_step = std::stoll(v);
            } else if (t.find("fieldNames[") != std::string::npos) {
                declared = (size_t)std::stoull(t.substr(t.find('[') + 1));
                names = true;
            }
            continue;
        }
        if (data && t.find("field[") != std::string::npos) {
            std::vector<size_t> dims;
            size_t p = 0;
            while ((p = t.find('[', p)) != std::string::npos) {
                const size_t q = t.find(']', p);
                if (q == std::string::npos) throw std::runtime_error("Malformed field dimensions");
                dims.push_back((size_t)std::stoull(t.substr(p + 1, q - p - 1)));
                p = q + 1;
            }
            if (dims.size() != 5) throw std::runtime_error("Expected 5 dimensions in field[...]");
            h.nvars = (int)dims[1];
            h.nz = (int)dims[2];
            h.ny = (int)dims[3];
            h.nx = (int)dims[4];
            if (!h.nx || !h.ny || !h.nz || !h.nvars) throw std::runtime_error("Invalid grid dimensions");
            if (dims[0] != dims[1] * dims[2] * dims[3] * dims[4]) throw std::runtime_error("Dimension mismatch");
            if (!std::getline(in, line)) throw std::runtime_error("Unexpected end of file after field declaration");
            h.data_pos = (size_t)in.tellg();
            found_data = true;
            break;
        }
    }
    if (!seen_sys) throw std::runtime_error("Missing systemInformation section");
    if (!seen_info) throw std::runtime_error("Missing fieldInformation section");
    if (!found_data) throw std::runtime_error("Missing fieldData section");
    if (h.scalar == 0) throw std::runtime_error("Missing or invalid scalarType in systemInformation");
    if (h.names.size() != (size_t)h.nvars) throw std::runtime_error("Field names count does not match nVars");
    if (h.names.size() != declared) throw std::runtime_error("Field names count does not match declared count");
    const size_t need = (size_t)h.nvars * h.nx * h.ny * h.nz * h.scalar;
    if (fsize - h.data_pos < need) throw std::runtime_error("Insufficient data in file");
    return h;
}

template <typename F>
int guard(F &&f) {
    try {
        f();
        return 0;
    } catch (const std::exception &e) {
        g_err = e.what();
        return -1;
    }
}

}  // namespace

// Shared with the post-processing unit (same thread-local message).
namespace lbm {
void case_set_error(const std::string &msg) { g_err = msg; }
}  // namespace lbm

extern "C" {

const char *lbm_case_last_error(void) { return g_err.c_str(); }

int lbm_case_parameter(const char *file, const char *name, char *out, int cap) {
    return guard([&] {
        if (!file || !name || !out || cap <= 0) throw std::runtime_error("null argument");
        const std::string v = parameter_line(read_lines(file), name);
        std::snprintf(out, (size_t)cap, "%s", v.c_str());
    });
}

int lbm_case_read(const char *dir, lbm_case_info *o) {
    return guard([&] {
        if (!o) throw std::runtime_error("null argument");
        const std::string d = dir ? dir : "";
        const Lines pc = read_lines(join(d, "programControl"));
        const Lines lm = read_lines(join(d, "latticeMesh"));
        std::memset(o, 0, sizeof(*o));
        const std::string cn = parameter_line(pc, "caseName");
        std::snprintf(o->case_name, sizeof(o->case_name), "%s", cn.c_str());
        o->Re = std::stod(parameter_line(pc, "Re"));
        o->u_inf = std::stod(parameter_line(pc, "u_inf"));
        auto integer = [&](const Lines &L, const char *n) {
            const std::string v = parameter_line(L, n);
            if (!is_number(v) || v.find('.') != std::string::npos)
                throw std::runtime_error(std::string("Parameter ") + n + " is not an integer");
            return (int64_t)std::stoll(v);
        };
        o->n_time_steps = integer(pc, "nTimeSteps");
        o->save_interval = integer(pc, "saveInterval");
        o->info_interval = integer(pc, "infoInterval");
        o->nx = (int)integer(lm, "nx");
        o->ny = (int)integer(lm, "ny");
        o->nz = (int)integer(lm, "nz");
        o->Lx = std::stod(parameter_line(lm, "Lx"));
        o->Ly = std::stod(parameter_line(lm, "Ly"));
        o->Lz = std::stod(parameter_line(lm, "Lz"));
        if (o->save_interval <= 0 || o->info_interval <= 0) throw std::runtime_error("intervals must be positive");
        o->latest_time = latest_time(d, cn);
    });
}

int lbm_case_initial_moments(const char *dir, int nx, int ny, int nz, float *const m[10]) {
    return guard([&] {
        if (!m || nx < 1 || ny < 1 || nz < 1) throw std::runtime_error("bad argument");
        const Lines ic = read_lines(join(dir ? dir : "", "initialConditions"));
        // region order of boundaryFields: North, South, East, West, Front, Back, internalField
        static const char *regions[7] = {"North", "South", "East", "West", "Front", "Back", "internalField"};
        for (int k = 0; k < 10; k++) {
            double r[7];
            for (int j = 0; j < 7; j++) r[j] = region_value(ic, kMomentNames[k], regions[j]);
            // hostArray.cuh:330-350 (z==0 is "front" there, z==nz-1 "back")
            for (int z = 0; z < nz; z++)
                for (int y = 0; y < ny; y++)
                    for (int x = 0; x < nx; x++) {
                        const int w = x == 0, e = x == nx - 1, s = y == 0, n = y == ny - 1, fr = z == 0,
                                  bk = z == nz - 1;
                        const int cnt = w + e + s + n + fr + bk;
                        const double sum = n * r[0] + s * r[1] + e * r[2] + w * r[3] + fr * r[4] + bk * r[5];
                        m[k][(size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * z)] =
                            (float)(cnt > 0 ? sum / cnt : r[6]);
                    }
        }
    });
}

int lbm_case_function_objects(const char *dir, unsigned *mask) {
    return guard([&] {
        if (!mask) throw std::runtime_error("null argument");
        *mask = 0;
        const std::string path = join(dir ? dir : "", "functionObjects");
        if (!fs::exists(path)) return;
        Lines b = extract_block(read_lines(path), "functionObjectList");
        // eraseBraces (strings.cuh:144-171): first line "{" and last "};" exactly
        if (b.size() < 2 || b.front() != "{" || b.back() != "};")
            throw std::runtime_error("functionObjectList must open with '{' and close with '};' on their own lines");
        std::vector<std::string> items;
        for (size_t i = 1; i + 1 < b.size(); i++) items.push_back(trim(b[i], true));
        auto has = [&](const char *n) { return std::find(items.begin(), items.end(), n) != items.end(); };
        if (has("k")) *mask |= LBM_FO_K;
        if (has("kMean")) *mask |= LBM_FO_KMEAN;
        if (has("S")) *mask |= LBM_FO_S;
        if (has("SMean")) *mask |= LBM_FO_SMEAN;
        if (has("momentsMean")) *mask |= 16u;
    });
}

int lbm_lbmbin_write(const char *path, int nx, int ny, int nz, int nvars, const char *const *names,
                     const float *const *fields, int64_t time_step, int time_average) {
    return guard([&] {
        if (!path || !names || !fields || nvars < 1) throw std::runtime_error("bad argument");
        if (nx % 8 || ny % 8 || nz % 8) throw std::runtime_error("LBMBin files need nx, ny, nz multiples of 8");
        const size_t n = (size_t)nx * ny * nz;
        std::ofstream out(path, std::ios::binary);
        if (!out) throw std::runtime_error(std::string("Cannot open file: ") + path);
        out << "systemInformation\n{\n\tbinaryType\tlittleEndian;\n\n\tscalarType\t32 bit;\n};\n\n";
        out << "fieldInformation\n{\n\ttimeStep\t" << time_step << ";\n\n\ttimeType\t"
            << (time_average ? "timeAverage" : "instantaneous") << ";\n\n\tfieldNames[" << nvars << "]\n\t{\n";
        for (int v = 0; v < nvars; v++) out << "\t\t" << names[v] << ";\n";
        out << "\t};\n};\n\n";
        out << "fieldData\n{\n\tfieldType\tnonUniform;\n\n\tfield[" << n * nvars << "][" << nvars << "][" << nz << "]["
            << ny << "][" << nx << "]\n\t{\n";
        std::vector<float> buf(n);
        for (int v = 0; v < nvars; v++) {
            for (int z = 0; z < nz; z++)
                for (int y = 0; y < ny; y++)
                    for (int x = 0; x < nx; x++)
                        buf[tiled8(x, y, z, nx, ny)] = fields[v][(size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * z)];
            out.write(reinterpret_cast<const char *>(buf.data()), (std::streamsize)(n * sizeof(float)));
        }
        out << "\n\t};\n};\n";
        if (!out) throw std::runtime_error(std::string("write failed: ") + path);
    });
}

int lbm_lbmbin_header(const char *path, int *nx, int *ny, int *nz, int *nvars, int *scalar_bytes,
                      int64_t *time_step, char *names, int names_cap) {
    return guard([&] {
        const Header h = parse_header(path ? path : "");
        if (nx) *nx = h.nx;
        if (ny) *ny = h.ny;
        if (nz) *nz = h.nz;
        if (nvars) *nvars = h.nvars;
        if (scalar_bytes) *scalar_bytes = h.scalar;
        if (time_step) *time_step = h.time_step;
        if (names && names_cap > 0) {
            std::string all;
            for (const auto &s : h.names) {
                all += s;
                all.push_back('\0');
            }
            if ((int)all.size() > names_cap) throw std::runtime_error("names buffer too small");
            std::memcpy(names, all.data(), all.size());
        }
    });
}

int lbm_lbmbin_read_field(const char *path, const char *name, float *out) {
    return guard([&] {
        if (!path || !name || !out) throw std::runtime_error("null argument");
        const Header h = parse_header(path);
        if (!h.little) throw std::runtime_error("big-endian files are not supported");
        const auto it = std::find(h.names.begin(), h.names.end(), std::string(name));
        if (it == h.names.end()) throw std::runtime_error(std::string("field ") + name + " not in " + path);
        if (h.nx % 8 || h.ny % 8 || h.nz % 8) throw std::runtime_error("tiled payload needs multiples of 8");
        const size_t idx = (size_t)(it - h.names.begin());
        const size_t n = (size_t)h.nx * h.ny * h.nz;
        std::ifstream in(path, std::ios::binary);
        in.seekg((std::streamoff)(h.data_pos + idx * n * h.scalar));
        std::vector<char> raw(n * h.scalar);
        in.read(raw.data(), (std::streamsize)raw.size());
        if ((size_t)in.gcount() != raw.size()) throw std::runtime_error("truncated payload");
        for (int z = 0; z < h.nz; z++)
            for (int y = 0; y < h.ny; y++)
                for (int x = 0; x < h.nx; x++) {
                    const size_t t = tiled8(x, y, z, h.nx, h.ny);
                    float v;
                    if (h.scalar == 4) {
                        std::memcpy(&v, raw.data() + 4 * t, 4);
                    } else {
                        double d;
                        std::memcpy(&d, raw.data() + 8 * t, 8);
                        v = (float)d;
                    }
                    out[(size_t)x + (size_t)h.nx * ((size_t)y + (size_t)h.ny * z)] = v;
                }
    });
}

}  // extern "C"
