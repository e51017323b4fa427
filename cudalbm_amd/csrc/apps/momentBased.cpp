// momentBased.cpp -- the momentBasedD3Q19 / momentBasedD3Q27 solver drivers.
//
// Same case directory, CLI and outputs as the reference drivers
// (applications/momentBasedD3Q19/momentBasedD3Q19.cu:56-140): run inside a case
// directory holding programControl, latticeMesh, initialConditions and an
// optional functionObjects file, with `-GPU <id>[,<id>...]` (the first id is used).
// Per time step t = latestTime .. nTimeSteps-1 the loop prints "Time: t" every
// infoInterval, writes <caseName>_<t>.LBMBin (and the enabled function objects)
// every saveInterval, then advances one step; the function objects are computed
// inside the step.  A run restarts from the latest <caseName>_<t>.LBMBin.
#include <chrono>
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "lbm_amd.h"
#include "lbm_case.h"

namespace {

const char *kNames[10] = {"rho", "u", "v", "w", "m_xx", "m_xy", "m_xz", "m_yy", "m_yz", "m_zz"};
const char *kS[6] = {"S_xx", "S_xy", "S_xz", "S_yy", "S_yz", "S_zz"};
const char *kSMean[6] = {"S_xxMean", "S_xyMean", "S_xzMean", "S_yyMean", "S_yzMean", "S_zzMean"};
// functionObjects.cuh:66
const char *kMomentsMean[10] = {"rhoMean", "uMean", "vMean", "wMean", "m_xxMean",
                                "m_xyMean", "m_xzMean", "m_yyMean", "m_yzMean", "m_zzMean"};

[[noreturn]] void die(int code, const std::string &msg) {
    std::fprintf(stderr, "%s\n", msg.c_str());
    std::exit(code < 0 ? -code : code);
}

int device_from_cli(int argc, char **argv) {
    for (int i = 1; i < argc; i++)
        if (std::strcmp(argv[i], "-GPU") == 0) {
            if (i + 1 >= argc)
                die(1, "Input argument -GPU has not been supplied with a value; the correct syntax is -GPU 0,1 for example");
            const std::string v = argv[i + 1];
            const std::string first = v.substr(0, v.find(','));
            for (char c : first)
                if (c < '0' || c > '9') die(1, "Value supplied to argument -GPU is not numeric");
            return std::atoi(first.c_str());
        }
    die(1, "Input argument -GPU has not been supplied");
}

void check(lbm_handle *h, int rc, const char *what) {
    if (rc) die(rc, std::string(what) + ": " + lbm_last_error(h));
}

void save_moments(lbm_handle *h, const lbm_case_info &ci, int64_t t, std::vector<std::vector<float>> &buf) {
    float *p[10];
    for (int k = 0; k < 10; k++) p[k] = buf[k].data();
    check(h, lbm_get_moments(h, p, LBM_LAYOUT_GLOBAL), "lbm_get_moments");
    const float *cp[10];
    for (int k = 0; k < 10; k++) cp[k] = buf[k].data();
    const std::string f = std::string(ci.case_name) + "_" + std::to_string(t) + ".LBMBin";
    if (lbm_lbmbin_write(f.c_str(), ci.nx, ci.ny, ci.nz, 10, kNames, cp, t, 0))
        die(1, lbm_case_last_error());
}

void save_fo(lbm_handle *h, const lbm_case_info &ci, unsigned mask, int64_t t) {
    const size_t n = (size_t)ci.nx * ci.ny * ci.nz;
    auto write = [&](const char *base, int field, int ncomp, const char *const *names, int avg) {
        std::vector<double> d(n * ncomp);
        check(h, lbm_fo_get(h, field, d.data(), LBM_LAYOUT_GLOBAL), "lbm_fo_get");
        std::vector<std::vector<float>> f(ncomp, std::vector<float>(n));
        const float *p[10];
        for (int c = 0; c < ncomp; c++) {
            for (size_t i = 0; i < n; i++) f[c][i] = (float)d[c * n + i];
            p[c] = f[c].data();
        }
        const std::string path = std::string(base) + "_" + std::to_string(t) + ".LBMBin";
        if (lbm_lbmbin_write(path.c_str(), ci.nx, ci.ny, ci.nz, ncomp, names, p, t, avg)) die(1, lbm_case_last_error());
    };
    static const char *kk[1] = {"k"};
    static const char *km[1] = {"kMean"};
    if (mask & LBM_FO_K) write("k", LBM_FIELD_K, 1, kk, 0);
    if (mask & LBM_FO_KMEAN) write("kMean", LBM_FIELD_KMEAN, 1, km, 1);
    if (mask & LBM_FO_S) write("S", LBM_FIELD_S, 6, kS, 0);
    if (mask & LBM_FO_SMEAN) write("SMean", LBM_FIELD_SMEAN, 6, kSMean, 1);
    if (mask & LBM_FO_MOMENTSMEAN) write("momentsMean", LBM_FIELD_MOMENTSMEAN, 10, kMomentsMean, 1);
}

}  // namespace

int run_moment_based(int Q, int argc, char **argv) {
    const int device = device_from_cli(argc, argv);
    lbm_case_info ci;
    if (lbm_case_read(".", &ci)) die(1, lbm_case_last_error());
    std::printf("programControl:\n{\n    programName: %s;\n    deviceList: [%d];\n    caseName: %s;\n"
                "    Re = %g;\n    u_inf = %g;\n    nTimeSteps = %" PRId64 ";\n    saveInterval = %" PRId64
                ";\n    infoInterval = %" PRId64 ";\n    latestTime = %" PRId64 ";\n    scalarType: 32 bit;\n};\n\n",
                argv[0], device, ci.case_name, ci.Re, ci.u_inf, ci.n_time_steps, ci.save_interval, ci.info_interval,
                ci.latest_time);
    std::printf("latticeMesh:\n{\n    nx = %d;\n    ny = %d;\n    nz = %d;\n    Lx = %g;\n    Ly = %g;\n    Lz = %g;\n};\n\n",
                ci.nx, ci.ny, ci.nz, ci.Lx, ci.Ly, ci.Lz);
    if (ci.nx % 8 || ci.ny % 8 || ci.nz % 8)
        die(LBM_ERR_SIZE, "Mesh dimensions should be multiples of 8 (the .LBMBin tile)");

    lbm_config cfg{};
    cfg.velocity_set = Q;
    cfg.storage = LBM_STORE_FP32;
    cfg.nx = ci.nx;
    cfg.ny = ci.ny;
    cfg.nz = ci.nz;
    cfg.Re = ci.Re;
    cfg.u_inf = ci.u_inf;
    cfg.device = device;
    if (const char *s = std::getenv("LBM_STORAGE"))
        if (std::strcmp(s, "fp16") == 0) cfg.storage = LBM_STORE_FP16;
    lbm_handle *h = nullptr;
    if (int rc = lbm_create(&cfg, &h)) die(rc, std::string("lbm_create: ") + lbm_create_error());
    int tx, ty, tz;
    double omega, tau;
    lbm_get_tiling(h, &tx, &ty, &tz);
    lbm_get_omega(h, &omega, &tau);
    std::printf("D3Q%d, workgroup footprint %dx%d, %d planes per workgroup; omega = %.9f, tau = %.9f\n\n", Q, tx, ty, tz,
                omega, tau);

    const size_t n = (size_t)ci.nx * ci.ny * ci.nz;
    std::vector<std::vector<float>> buf(10, std::vector<float>(n));
    float *p[10];
    for (int k = 0; k < 10; k++) p[k] = buf[k].data();
    const std::string restart = std::string(ci.case_name) + "_" + std::to_string(ci.latest_time) + ".LBMBin";
    FILE *probe = std::fopen(restart.c_str(), "rb");
    if (probe) {
        std::fclose(probe);
        std::printf("Restarting from %s\n\n", restart.c_str());
        for (int k = 0; k < 10; k++)
            if (lbm_lbmbin_read_field(restart.c_str(), kNames[k], p[k])) die(1, lbm_case_last_error());
    } else if (lbm_case_initial_moments(".", ci.nx, ci.ny, ci.nz, p)) {
        die(1, lbm_case_last_error());
    }
    const float *cp[10];
    for (int k = 0; k < 10; k++) cp[k] = p[k];
    check(h, lbm_set_moments(h, cp, LBM_LAYOUT_GLOBAL), "lbm_set_moments");
    check(h, lbm_set_time(h, ci.latest_time), "lbm_set_time");

    unsigned fo = 0;
    if (lbm_case_function_objects(".", &fo)) die(1, lbm_case_last_error());
    if (fo) check(h, lbm_fo_enable(h, fo), "lbm_fo_enable");

    const auto t0 = std::chrono::steady_clock::now();
    int64_t t = ci.latest_time;
    while (t < ci.n_time_steps) {
        if (t % ci.info_interval == 0) std::printf("Time: %" PRId64 "\n", t);
        if (t % ci.save_interval == 0) {
            save_moments(h, ci, t, buf);
            if (fo) save_fo(h, ci, fo, t);
        }
        // advance to the next print/save event in one batch (graph-replayed steps)
        int64_t next = ci.n_time_steps;
        const int64_t ni = (t / ci.info_interval + 1) * ci.info_interval;
        const int64_t ns = (t / ci.save_interval + 1) * ci.save_interval;
        next = std::min(next, std::min(ni, ns));
        check(h, lbm_step(h, (int)(next - t)), "lbm_step");
        t = next;
    }
    check(h, lbm_synchronize(h), "lbm_synchronize");
    const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    const int64_t steps = ci.n_time_steps - ci.latest_time;
    std::printf("\nElapsed time: %.3f s\n", sec);
    // runTimeIO.cuh:144-165: nx ny nz (nt - latestTime - 1) / (1e6 * whole seconds)
    const int64_t whole = (int64_t)sec;
    if (whole > 0)
        std::printf("MLUPS (reference formula): %.1f\n", (double)n * (double)(steps - 1) / (1e6 * (double)whole));
    std::printf("MLUPS: %.1f\n", (double)n * (double)steps / (1e6 * sec));
    lbm_destroy(h);
    return 0;
}
