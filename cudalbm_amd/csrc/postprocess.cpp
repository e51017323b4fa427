// postprocess.cpp -- host-only C ABI of include/lbm_post.h: .LBMBin time
// discovery, VTU/VTS/Tecplot writers, finite-difference derivatives, curl,
// divergence and cut planes, with the reference's conventions (cited per
// function in the header).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <limits>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "lbm_amd.h"
#include "lbm_post.h"

namespace lbm {
void case_set_error(const std::string &msg);
}

namespace fs = std::filesystem;

namespace {

template <typename F>
int guarded(F &&f) {
    try {
        return f();
    } catch (const std::invalid_argument &e) {
        lbm::case_set_error(e.what());
        return LBM_ERR_ARG;
    } catch (const std::exception &e) {
        lbm::case_set_error(e.what());
        return LBM_ERR_ARG;
    }
}

inline int64_t gidx(int64_t x, int64_t y, int64_t z, int64_t nx, int64_t ny) { return x + nx * (y + ny * z); }

// Node coordinates, interleaved x,y,z per node, FP32 (computed in FP64).
std::vector<float> coordinates(int nx, int ny, int nz, double Lx, double Ly, double Lz) {
    std::vector<float> c((size_t)nx * ny * nz * 3);
    auto pos = [](double L, int i, int n) { return n > 1 ? L * (double)i / (double)(n - 1) : 0.0; };
    size_t k = 0;
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                c[k++] = (float)pos(Lx, x, nx);
                c[k++] = (float)pos(Ly, y, ny);
                c[k++] = (float)pos(Lz, z, nz);
            }
    return c;
}

// Hexahedra over the nodes, corner order (i,j,k), (i+1,j,k), (i+1,j+1,k), (i,j+1,k), then k+1.
template <typename I>
std::vector<I> connectivity(int nx, int ny, int nz, I base) {
    std::vector<I> c;
    c.reserve((size_t)std::max(nx - 1, 0) * std::max(ny - 1, 0) * std::max(nz - 1, 0) * 8);
    const int64_t sy = nx, sz = (int64_t)nx * ny;
    for (int k = 0; k + 1 < nz; k++)
        for (int j = 0; j + 1 < ny; j++)
            for (int i = 0; i + 1 < nx; i++) {
                const int64_t b = gidx(i, j, k, nx, ny);
                const int64_t v[8] = {b, b + 1, b + sy + 1, b + sy, b + sz, b + sz + 1, b + sz + sy + 1, b + sz + sy};
                for (int64_t x : v) c.push_back((I)(x + base));
            }
    return c;
}

template <typename T>
void block(std::ofstream &f, const std::vector<T> &v) {
    const uint64_t n = v.size() * sizeof(T);
    f.write(reinterpret_cast<const char *>(&n), sizeof n);
    f.write(reinterpret_cast<const char *>(v.data()), (std::streamsize)n);
}

template <typename T>
const char *vtk_type();
template <>
const char *vtk_type<float>() { return "Float32"; }
template <>
const char *vtk_type<uint32_t>() { return "UInt32"; }
template <>
const char *vtk_type<uint64_t>() { return "UInt64"; }
template <>
const char *vtk_type<uint8_t>() { return "UInt8"; }

std::vector<std::vector<float>> copy_fields(int nvars, const float *const *fields, int64_t n) {
    std::vector<std::vector<float>> v((size_t)nvars);
    for (int i = 0; i < nvars; i++) {
        if (!fields[i]) throw std::invalid_argument("null field");
        v[i].assign(fields[i], fields[i] + n);
    }
    return v;
}

void point_data_header(std::ostringstream &x, int nvars, const char *const *names, int64_t n, uint64_t &off) {
    x << "      <PointData Scalars=\"" << (nvars > 0 ? names[0] : "") << "\">\n";
    for (int i = 0; i < nvars; i++) {
        x << "        <DataArray type=\"Float32\" Name=\"" << names[i] << "\" format=\"appended\" offset=\"" << off
          << "\"/>\n";
        off += sizeof(uint64_t) + (uint64_t)n * sizeof(float);
    }
    x << "      </PointData>\n";
}

void write_vts(const std::string &path, int nx, int ny, int nz, const std::vector<float> &pts, int nvars,
               const char *const *names, const std::vector<std::vector<float>> &f) {
    const int64_t n = (int64_t)nx * ny * nz;
    std::ostringstream x;
    uint64_t off = 0;
    x << "<?xml version=\"1.0\"?>\n";
    x << "<VTKFile type=\"StructuredGrid\" version=\"1.0\" byte_order=\"LittleEndian\" header_type=\"UInt64\">\n";
    x << "  <StructuredGrid WholeExtent=\"0 " << nx - 1 << " 0 " << ny - 1 << " 0 " << nz - 1 << "\">\n";
    x << "    <Piece Extent=\"0 " << nx - 1 << " 0 " << ny - 1 << " 0 " << nz - 1 << "\">\n";
    point_data_header(x, nvars, names, n, off);
    x << "      <Points>\n";
    x << "        <DataArray type=\"Float32\" Name=\"Coordinates\" NumberOfComponents=\"3\" format=\"appended\" offset=\""
      << off << "\"/>\n";
    x << "      </Points>\n";
    x << "    </Piece>\n";
    x << "  </StructuredGrid>\n";
    x << "  <AppendedData encoding=\"raw\">_";
    std::ofstream o(path, std::ios::binary);
    if (!o) throw std::runtime_error("Error opening file: " + path);
    o << x.str();
    for (const auto &v : f) block(o, v);
    block(o, pts);
    o << "</AppendedData>\n";
    o << "</VTKFile>\n";
}

template <typename I>
void write_vtu(const std::string &path, int nx, int ny, int nz, const std::vector<float> &pts, int nvars,
               const char *const *names, const std::vector<std::vector<float>> &f) {
    const int64_t n = (int64_t)nx * ny * nz;
    const std::vector<I> conn = connectivity<I>(nx, ny, nz, 0);
    const int64_t ne = (int64_t)conn.size() / 8;
    std::vector<I> offs((size_t)ne);
    for (int64_t e = 0; e < ne; e++) offs[(size_t)e] = (I)((e + 1) * 8);
    const std::vector<uint8_t> types((size_t)ne, 12);  // VTK hexahedron
    std::ostringstream x;
    uint64_t off = 0;
    x << "<?xml version=\"1.0\"?>\n";
    x << "<VTKFile type=\"UnstructuredGrid\" version=\"1.0\" byte_order=\"LittleEndian\" header_type=\"UInt64\">\n";
    x << "  <UnstructuredGrid>\n";
    x << "    <Piece NumberOfPoints=\"" << n << "\" NumberOfCells=\"" << ne << "\">\n";
    point_data_header(x, nvars, names, n, off);
    x << "      <Points>\n";
    x << "        <DataArray type=\"Float32\" Name=\"Coordinates\" NumberOfComponents=\"3\" format=\"appended\" offset=\""
      << off << "\"/>\n";
    x << "      </Points>\n";
    off += sizeof(uint64_t) + pts.size() * sizeof(float);
    x << "      <Cells>\n";
    x << "        <DataArray type=\"" << vtk_type<I>() << "\" Name=\"connectivity\" format=\"appended\" offset=\"" << off
      << "\"/>\n";
    off += sizeof(uint64_t) + conn.size() * sizeof(I);
    x << "        <DataArray type=\"" << vtk_type<I>() << "\" Name=\"offsets\" format=\"appended\" offset=\"" << off
      << "\"/>\n";
    off += sizeof(uint64_t) + offs.size() * sizeof(I);
    x << "        <DataArray type=\"UInt8\" Name=\"types\" format=\"appended\" offset=\"" << off << "\"/>\n";
    x << "      </Cells>\n";
    x << "    </Piece>\n";
    x << "  </UnstructuredGrid>\n";
    x << "  <AppendedData encoding=\"raw\">_";
    std::ofstream o(path, std::ios::binary);
    if (!o) throw std::runtime_error("Error opening file: " + path);
    o << x.str();
    for (const auto &v : f) block(o, v);
    block(o, pts);
    block(o, conn);
    block(o, offs);
    block(o, types);
    o << "</AppendedData>\n";
    o << "</VTKFile>\n";
}

void write_tecplot(const std::string &path, int nx, int ny, int nz, const std::vector<float> &pts, int nvars,
                   const char *const *names, const std::vector<std::vector<float>> &f) {
    const int64_t n = (int64_t)nx * ny * nz;
    std::ofstream o(path);
    if (!o) throw std::runtime_error("Error opening file: " + path);
    o.precision(std::numeric_limits<float>::max_digits10);
    o << "VARIABLES = \"X\" \"Y\" \"Z\" ";
    for (int i = 0; i < nvars; i++) o << "\"" << names[i] << "\" ";
    o << "\n";
    const std::vector<uint64_t> conn = connectivity<uint64_t>(nx, ny, nz, 1);
    const int64_t ne = (int64_t)conn.size() / 8;
    o << "ZONE T=\"Hexahedral Zone\", NODES=" << n << ", ELEMENTS=" << ne << ", DATAPACKING=BLOCK, ZONETYPE=FEBRICK\n";
    for (int c = 0; c < 3; c++)
        for (int64_t k = 0; k < n; k++) o << pts[(size_t)(3 * k + c)] << "\n";
    for (const auto &v : f)
        for (float val : v) o << val << "\n";
    for (int64_t e = 0; e < ne; e++)
        for (int k = 0; k < 8; k++) o << conn[(size_t)(e * 8 + k)] << (k < 7 ? " " : "\n");
}

// One line of n values extended by `pad` nodes each side with the odd
// reflection about the end nodes: g(-j) = -f(j), g(n-1+j) = -f(n-1-j).
double deriv_at(const std::vector<double> &g, int pad, int x, int order, bool ref) {
    const int c = pad + x;
    auto d = [&](int k) { return g[(size_t)(c + k)] - g[(size_t)(c - k)]; };
    const double s = ref ? 1.0 : -1.0;  // sign of the even-k terms
    switch (order) {
    case 2: return d(1) / 2.0;
    case 4: return 2.0 / 3.0 * d(1) + s * (1.0 / 12.0) * d(2);
    case 6: return 3.0 / 4.0 * d(1) + s * (3.0 / 20.0) * d(2) + 1.0 / 60.0 * d(3);
    default: return 4.0 / 5.0 * d(1) + s * (1.0 / 5.0) * d(2) + 4.0 / 105.0 * d(3) + s * (1.0 / 280.0) * d(4);
    }
}

void derivative(int scheme, int order, int axis, int nx, int ny, int nz, const float *f, double *out) {
    if (scheme != LBM_POST_CENTRAL && scheme != LBM_POST_REFERENCE) throw std::invalid_argument("bad scheme");
    if (order != 2 && order != 4 && order != 6 && order != 8)
        throw std::invalid_argument("Invalid finite difference scheme order: valid orders are 2, 4, 6 and 8");
    if (axis < 0 || axis > 2) throw std::invalid_argument("bad axis");
    const int n = axis == 0 ? nx : axis == 1 ? ny : nz;
    const int pad = order - 1;
    if (n < pad + 1) throw std::invalid_argument("mesh too small for the derivative scheme");
    std::vector<double> g((size_t)(n + 2 * pad));
    const int na = axis == 0 ? ny : nx, nb = axis == 2 ? ny : nz;  // the two other axes
    for (int b = 0; b < nb; b++)
        for (int a = 0; a < na; a++) {
            auto at = [&](int i) -> int64_t {
                if (axis == 0) return gidx(i, a, b, nx, ny);
                if (axis == 1) return gidx(a, i, b, nx, ny);
                return gidx(a, b, i, nx, ny);
            };
            for (int i = 0; i < n; i++) g[(size_t)(pad + i)] = f[at(i)];
            for (int i = 0; i < pad; i++) g[(size_t)i] = -(double)f[at(pad - i)];
            for (int i = 0; i < pad; i++) g[(size_t)(pad + n + i)] = -(double)f[at(n - 2 - i)];
            for (int i = 0; i < n; i++) out[at(i)] = deriv_at(g, pad, i, order, scheme == LBM_POST_REFERENCE);
        }
}

}  // namespace

extern "C" {

int lbm_post_time_indices(const char *dir, const char *prefix, int64_t *out, int cap) {
    return guarded([&]() -> int {
        if (!dir || !prefix) throw std::invalid_argument("null argument");
        const std::string pre = std::string(prefix) + "_", ext = ".LBMBin";
        std::vector<int64_t> t;
        for (const auto &e : fs::directory_iterator(dir)) {
            if (!e.is_regular_file()) continue;
            const std::string name = e.path().filename().string();
            if (name.size() <= pre.size() + ext.size() || name.compare(0, pre.size(), pre) != 0 ||
                name.compare(name.size() - ext.size(), ext.size(), ext) != 0)
                continue;
            const std::string mid = name.substr(pre.size(), name.size() - pre.size() - ext.size());
            if (mid.empty() || !std::all_of(mid.begin(), mid.end(), [](char c) { return c >= '0' && c <= '9'; }))
                continue;
            t.push_back(std::stoll(mid));
        }
        std::sort(t.begin(), t.end());
        for (int i = 0; i < (int)t.size() && i < cap; i++) out[i] = t[(size_t)i];
        return (int)t.size();
    });
}

int lbm_post_write(int format, const char *path, int nx, int ny, int nz, double Lx, double Ly, double Lz, int nvars,
                   const char *const *names, const float *const *fields) {
    return guarded([&]() -> int {
        if (!path || nvars < 0 || (nvars > 0 && (!names || !fields))) throw std::invalid_argument("null argument");
        if (nx < 1 || ny < 1 || nz < 1) throw std::invalid_argument("bad mesh size");
        const int64_t n = (int64_t)nx * ny * nz;
        const auto f = copy_fields(nvars, fields, n);
        const auto pts = coordinates(nx, ny, nz, Lx, Ly, Lz);
        const std::string p(path);
        if (format == LBM_POST_VTS)
            write_vts(p + ".vts", nx, ny, nz, pts, nvars, names, f);
        else if (format == LBM_POST_VTU) {
            if (n >= (int64_t)std::numeric_limits<uint32_t>::max())
                write_vtu<uint64_t>(p + ".vtu", nx, ny, nz, pts, nvars, names, f);
            else
                write_vtu<uint32_t>(p + ".vtu", nx, ny, nz, pts, nvars, names, f);
        } else if (format == LBM_POST_TECPLOT)
            write_tecplot(p + ".dat", nx, ny, nz, pts, nvars, names, f);
        else
            throw std::invalid_argument("unknown file type");
        return LBM_OK;
    });
}

int lbm_post_derivative(int scheme, int order, int axis, int nx, int ny, int nz, const float *f, double *out) {
    return guarded([&]() -> int {
        if (!f || !out) throw std::invalid_argument("null argument");
        derivative(scheme, order, axis, nx, ny, nz, f, out);
        return LBM_OK;
    });
}

int lbm_post_curl(int scheme, int order, int nx, int ny, int nz, const float *u, const float *v, const float *w,
                  float *cx, float *cy, float *cz) {
    return guarded([&]() -> int {
        if (!u || !v || !w || !cx || !cy || !cz) throw std::invalid_argument("null argument");
        const size_t n = (size_t)nx * ny * nz;
        std::vector<double> dwdy(n), dvdz(n), dudz(n), dwdx(n), dvdx(n), dudy(n);
        derivative(scheme, order, 1, nx, ny, nz, w, dwdy.data());
        derivative(scheme, order, 2, nx, ny, nz, v, dvdz.data());
        derivative(scheme, order, 2, nx, ny, nz, u, dudz.data());
        derivative(scheme, order, 0, nx, ny, nz, w, dwdx.data());
        derivative(scheme, order, 0, nx, ny, nz, v, dvdx.data());
        derivative(scheme, order, 1, nx, ny, nz, u, dudy.data());
        for (size_t i = 0; i < n; i++) {
            cx[i] = (float)(dwdy[i] - dvdz[i]);
            cy[i] = (float)(dudz[i] - dwdx[i]);
            cz[i] = (float)(dvdx[i] - dudy[i]);
        }
        return LBM_OK;
    });
}

int lbm_post_div(int scheme, int order, int nx, int ny, int nz, const float *u, const float *v, const float *w,
                 float *out) {
    return guarded([&]() -> int {
        if (!u || !v || !w || !out) throw std::invalid_argument("null argument");
        const size_t n = (size_t)nx * ny * nz;
        std::vector<double> a(n), b(n), c(n);
        derivative(scheme, order, 0, nx, ny, nz, u, a.data());
        derivative(scheme, order, 1, nx, ny, nz, v, b.data());
        derivative(scheme, order, 2, nx, ny, nz, w, c.data());
        for (size_t i = 0; i < n; i++) out[i] = (float)(a[i] + b[i] + c[i]);
        return LBM_OK;
    });
}

int lbm_post_cut_plane(int axis, double c, int nx, int ny, int nz, double L, const float *f, float *out) {
    return guarded([&]() -> int {
        if (!f || !out) throw std::invalid_argument("null argument");
        if (axis < 0 || axis > 2) throw std::invalid_argument("Invalid cardinal direction");
        const int n = axis == 0 ? nx : axis == 1 ? ny : nz;
        // FP32 arithmetic, as the reference's scalar_t
        const float i = (float)(n - 1) * ((float)c * (float)L);
        const int64_t i0 = (int64_t)std::floor(i), i1 = (int64_t)std::ceil(i);
        if (i0 < 0 || i1 >= n) throw std::invalid_argument("cut plane outside the mesh");
        const float wgt = (float)c - (float)i0;
        auto lerp = [&](float f0, float f1) { return ((1.0f - wgt) * f0) + (wgt * f1); };
        if (axis == 0) {
            for (int z = 0; z < nz; z++)
                for (int y = 0; y < ny; y++)
                    out[y + (int64_t)z * ny] = lerp(f[gidx(i0, y, z, nx, ny)], f[gidx(i1, y, z, nx, ny)]);
        } else if (axis == 1) {
            for (int z = 0; z < nz; z++)
                for (int x = 0; x < nx; x++)
                    out[x + (int64_t)z * nx] = lerp(f[gidx(x, i0, z, nx, ny)], f[gidx(x, i1, z, nx, ny)]);
        } else {
            for (int y = 0; y < ny; y++)
                for (int x = 0; x < nx; x++)
                    out[x + (int64_t)y * nx] = lerp(f[gidx(x, y, i0, nx, ny)], f[gidx(x, y, i1, nx, ny)]);
        }
        return LBM_OK;
    });
}

}  // extern "C"
