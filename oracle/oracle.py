"""ctypes binding of the FP64 CPU oracle (oracle/lbm_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the solver package.  Arrays are numpy
float64 in global x-fastest order (index = x + nx*(y + ny*z)); moments are the
reference's stored values (rho-1, 3u, 3v, 3w, 4.5(Pxx/rho-cs2), 9Pxy/rho, ...)
as in globalFunctions.cuh:482-503 of the reference.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "liblbm_oracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build():
    """Compile the oracle with its Makefile (gcc + OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.lbmo_velocity_set.argtypes = [ctypes.c_int, _ip, _ip, _ip, _dp]
        L.lbmo_qf.argtypes = [ctypes.c_int]
        L.lbmo_face_slots.argtypes = [ctypes.c_int, _ip]
        L.lbmo_reconstruct.argtypes = [ctypes.c_int, _dp, _dp]
        L.lbmo_moments_interior.argtypes = [ctypes.c_int, _dp, _dp]
        L.lbmo_known.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.lbmo_moments_boundary.argtypes = [ctypes.c_int, _dp, ctypes.c_int, ctypes.c_double, _dp, _dp]
        L.lbmo_scale.argtypes = [_dp]
        L.lbmo_scale.restype = None
        L.lbmo_collide.argtypes = [_dp, ctypes.c_double]
        L.lbmo_collide.restype = None
        L.lbmo_node_type.argtypes = [ctypes.c_int] * 6
        L.lbmo_omega.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int]
        L.lbmo_omega.restype = ctypes.c_double
        pp = ctypes.POINTER(_dp)
        L.lbmo_step.argtypes = [ctypes.c_int] * 4 + [ctypes.c_double, ctypes.c_double, pp, pp, ctypes.c_int]
        L.lbmo_run.argtypes = [ctypes.c_int] * 4 + [ctypes.c_double, ctypes.c_double, pp, ctypes.c_int, ctypes.c_int]
        L.lbmo_run_means.argtypes = [ctypes.c_int] * 4 + [ctypes.c_double, ctypes.c_double, pp, ctypes.c_int,
                                                         ctypes.c_int, ctypes.c_int64, ctypes.c_double, _dp, _dp, _dp]
        L.lbmo_initial_field.argtypes = [ctypes.c_int] * 3 + [_dp, _dp]
        L.lbmo_initial_field.restype = None
        L.lbmo_k.argtypes = [pp, ctypes.c_int64, _dp]
        L.lbmo_k.restype = None
        L.lbmo_S.argtypes = [pp, ctypes.c_int64, ctypes.c_double, _dp]
        L.lbmo_S.restype = None
        L.lbmo_time_average.argtypes = [_dp, _dp, ctypes.c_int64, ctypes.c_int64]
        L.lbmo_time_average.restype = None
        L.lbmo_tile_source_map.argtypes = [ctypes.c_int] * 4 + [ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_uint8)]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _ptrs(arrs):
    P = (_dp * len(arrs))()
    for i, a in enumerate(arrs):
        assert a.dtype == np.float64 and a.flags.c_contiguous
        P[i] = _d(a)
    return P


def velocity_set(Q):
    cx = (ctypes.c_int * Q)()
    cy = (ctypes.c_int * Q)()
    cz = (ctypes.c_int * Q)()
    w = np.zeros(Q)
    assert lib().lbmo_velocity_set(Q, cx, cy, cz, _d(w)) == 0
    c = np.array([list(cx), list(cy), list(cz)], dtype=np.int64).T
    return c, w


def face_slots(Q):
    qf = lib().lbmo_qf(Q)
    out = (ctypes.c_int * (6 * qf))()
    assert lib().lbmo_face_slots(Q, out) == 0
    return np.array(list(out)).reshape(6, qf)


def reconstruct(Q, m):
    m = np.ascontiguousarray(m, dtype=np.float64)
    f = np.zeros(Q)
    assert lib().lbmo_reconstruct(Q, _d(m), _d(f)) == 0
    return f


def moments_interior(Q, f):
    f = np.ascontiguousarray(f, dtype=np.float64)
    m = np.zeros(10)
    assert lib().lbmo_moments_interior(Q, _d(f), _d(m)) == 0
    return m


def known(Q, q, bits):
    return lib().lbmo_known(Q, q, bits)


def moments_boundary(Q, f, bits, omega, u_lid):
    f = np.ascontiguousarray(f, dtype=np.float64)
    U = np.zeros(18)
    U[6] = u_lid
    m = np.zeros(10)
    rc = lib().lbmo_moments_boundary(Q, _d(f), bits, omega, _d(U), _d(m))
    assert rc == 0, rc
    return m


def scale(m):
    m = np.array(m, dtype=np.float64)
    lib().lbmo_scale(_d(m))
    return m


def collide(m, omega):
    m = np.array(m, dtype=np.float64)
    lib().lbmo_collide(_d(m), omega)
    return m


def node_type(x, y, z, nx, ny, nz):
    return lib().lbmo_node_type(x, y, z, nx, ny, nz)


def omega(u_inf, Re, nx):
    return lib().lbmo_omega(u_inf, Re, nx)


def rest_state(nx, ny, nz):
    """Ten zero moment arrays: rho = 1, u = 0, Pi = cs2 I everywhere."""
    n = nx * ny * nz
    return [np.zeros(n) for _ in range(10)]


def run(Q, nx, ny, nz, omega_, u_lid, mom, nsteps, nthreads=0):
    """Advance `mom` (list of 10 float64 arrays, x-fastest) by nsteps in place."""
    rc = lib().lbmo_run(Q, nx, ny, nz, omega_, u_lid, _ptrs(mom), nsteps, nthreads)
    assert rc == 0, rc
    return mom


def run_means(Q, nx, ny, nz, omega_, u_lid, mom, nsteps, tau, means=None, t0=0, nthreads=0,
              fields=("kMean", "SMean", "momentsMean")):
    """Advance `mom` by nsteps in place, updating the FP64 running means after
    every step.  means: dict with any of "kMean" (n), "SMean" (6, n) and
    "momentsMean" (10, n) float64 arrays; the `fields` not in it are created as
    zeros, the others are not computed.  Returns the dict."""
    n = nx * ny * nz
    means = {} if means is None else means
    ptr = {}
    for key, shape in (("kMean", (n,)), ("SMean", (6, n)), ("momentsMean", (10, n))):
        if key in fields:
            means.setdefault(key, np.zeros(shape))
        if key in means:
            assert means[key].dtype == np.float64 and means[key].flags.c_contiguous and means[key].shape == shape
            ptr[key] = _d(means[key])
        else:
            ptr[key] = _dp()
    rc = lib().lbmo_run_means(Q, nx, ny, nz, omega_, u_lid, _ptrs(mom), nsteps, nthreads, t0, tau,
                              ptr["kMean"], ptr["SMean"], ptr["momentsMean"])
    assert rc == 0, rc
    return means


def step(Q, nx, ny, nz, omega_, u_lid, mom_in, nthreads=0):
    out = [np.zeros_like(a) for a in mom_in]
    rc = lib().lbmo_step(Q, nx, ny, nz, omega_, u_lid, _ptrs(mom_in), _ptrs(out), nthreads)
    assert rc == 0, rc
    return out


def initial_field(nx, ny, nz, region7):
    r = np.ascontiguousarray(region7, dtype=np.float64)
    out = np.zeros(nx * ny * nz)
    lib().lbmo_initial_field(nx, ny, nz, _d(r), _d(out))
    return out


def k_field(mom):
    n = mom[1].size
    k = np.zeros(n)
    lib().lbmo_k(_ptrs(mom), n, _d(k))
    return k


def S_field(mom, tau):
    n = mom[1].size
    S = np.zeros(6 * n)
    lib().lbmo_S(_ptrs(mom), n, tau, _d(S))
    return S.reshape(6, n)


def time_average(mean, x, t):
    mean = np.ascontiguousarray(mean, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    lib().lbmo_time_average(_d(mean), _d(x), mean.size, t)
    return mean


def tile_source_map(Q, nx, ny, nz):
    n = nx * ny * nz
    src = np.zeros(n * Q, dtype=np.int64)
    kn = np.zeros(n * Q, dtype=np.uint8)
    rc = lib().lbmo_tile_source_map(Q, nx, ny, nz, src.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                    kn.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
    assert rc == 0, rc
    return src.reshape(n, Q), kn.reshape(n, Q).astype(bool)


def lid_cavity_rest(nx, ny, nz):
    """Rest start: every region value 0 (the reference's initialConditions with
    North u set to 0)."""
    return rest_state(nx, ny, nz)


def lid_cavity_reference_start(nx, ny, nz, u_north):
    """The shipped case's start: North u = u_north, North m_xx = equilibrium
    (4.5 u^2), all else 0 (lidDrivenCavity/D3Q19/initialConditions:59,173),
    averaged over boundary count on edges and corners."""
    mom = rest_state(nx, ny, nz)
    z7 = np.zeros(7)
    r_u = z7.copy()
    r_u[0] = 3.0 * u_north
    r_mxx = z7.copy()
    r_mxx[0] = 4.5 * u_north * u_north
    mom[1] = initial_field(nx, ny, nz, r_u)
    mom[4] = initial_field(nx, ny, nz, r_mxx)
    return mom
