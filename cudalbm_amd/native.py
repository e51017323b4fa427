"""ctypes binding of liblbm_amd.so (the C ABI in include/lbm_amd.h).

The solver runs only through the HIP library: if it is missing or cannot be
loaded this module raises -- there is no CPU fallback in the product path.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblbm_amd.so")

LBM_STORE_FP32 = 0
LBM_STORE_FP16 = 1
LBM_LAYOUT_GLOBAL = 0
LBM_LAYOUT_TILED8 = 1
LBM_FO_K, LBM_FO_KMEAN, LBM_FO_S, LBM_FO_SMEAN = 1, 2, 4, 8
LBM_FIELD_K, LBM_FIELD_KMEAN, LBM_FIELD_S, LBM_FIELD_SMEAN = 0, 1, 2, 3

# Every symbol include/lbm_amd.h declares (checked by tests/test_native_abi.py).
EXPORTED = [
    "lbm_create", "lbm_destroy", "lbm_last_error", "lbm_create_error", "lbm_get_tiling", "lbm_get_omega",
    "lbm_set_moments", "lbm_get_moments", "lbm_set_moments_f64", "lbm_get_moments_f64", "lbm_step",
    "lbm_synchronize", "lbm_step_timed", "lbm_get_time", "lbm_set_time", "lbm_fo_enable", "lbm_fo_get",
    "lbm_reduce", "lbm_probe_streaming",
]


class LBMError(RuntimeError):
    pass


class lbm_config(ctypes.Structure):
    _fields_ = [
        ("velocity_set", ctypes.c_int),
        ("storage", ctypes.c_int),
        ("nx", ctypes.c_int),
        ("ny", ctypes.c_int),
        ("nz", ctypes.c_int),
        ("Re", ctypes.c_double),
        ("u_inf", ctypes.c_double),
        ("device", ctypes.c_int),
        ("stream", ctypes.c_void_p),
        ("tile_x", ctypes.c_int),
        ("tile_y", ctypes.c_int),
        ("tile_z", ctypes.c_int),
    ]


_lib = None


def lib():
    """Load the native library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LBMError(f"{LIB_PATH} is missing: build it with `python -m cudalbm_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    H = ctypes.c_void_p
    fp = ctypes.POINTER(ctypes.c_float)
    dp = ctypes.POINTER(ctypes.c_double)
    L.lbm_create.argtypes = [ctypes.POINTER(lbm_config), ctypes.POINTER(H)]
    L.lbm_destroy.argtypes = [H]
    L.lbm_last_error.argtypes = [H]
    L.lbm_last_error.restype = ctypes.c_char_p
    L.lbm_create_error.argtypes = []
    L.lbm_create_error.restype = ctypes.c_char_p
    L.lbm_get_tiling.argtypes = [H] + [ctypes.POINTER(ctypes.c_int)] * 3
    L.lbm_get_omega.argtypes = [H, dp, dp]
    L.lbm_set_moments.argtypes = [H, ctypes.POINTER(fp), ctypes.c_int]
    L.lbm_get_moments.argtypes = [H, ctypes.POINTER(fp), ctypes.c_int]
    L.lbm_set_moments_f64.argtypes = [H, ctypes.POINTER(dp), ctypes.c_int]
    L.lbm_get_moments_f64.argtypes = [H, ctypes.POINTER(dp), ctypes.c_int]
    L.lbm_step.argtypes = [H, ctypes.c_int]
    L.lbm_synchronize.argtypes = [H]
    L.lbm_step_timed.argtypes = [H, ctypes.c_int, dp]
    L.lbm_get_time.argtypes = [H, ctypes.POINTER(ctypes.c_int64)]
    L.lbm_set_time.argtypes = [H, ctypes.c_int64]
    L.lbm_fo_enable.argtypes = [H, ctypes.c_uint]
    L.lbm_fo_get.argtypes = [H, ctypes.c_int, dp, ctypes.c_int]
    L.lbm_reduce.argtypes = [H, dp, dp]
    L.lbm_probe_streaming.argtypes = [H, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint8)]
    _lib = L
    return L


def _ptr_array(arrs, ctype):
    P = (ctypes.POINTER(ctype) * len(arrs))()
    for i, a in enumerate(arrs):
        P[i] = a.ctypes.data_as(ctypes.POINTER(ctype))
    return P


class Solver:
    """One cavity on one GPU: a handle of the native library.

    Mirrors the reference driver (applications/momentBasedD3Q19/momentBasedD3Q19.cu:56-140):
    configuration = latticeMesh + programControl values, state = the 10 stored moments.
    """

    def __init__(self, nx, ny, nz, Re, u_inf, velocity_set=19, storage="fp32", device=0,
                 tile=(0, 0, 0)):
        L = lib()
        cfg = lbm_config()
        cfg.velocity_set = velocity_set
        cfg.storage = {"fp32": LBM_STORE_FP32, "fp16": LBM_STORE_FP16}[storage]
        cfg.nx, cfg.ny, cfg.nz = nx, ny, nz
        cfg.Re, cfg.u_inf = float(Re), float(u_inf)
        cfg.device = device
        cfg.stream = None
        cfg.tile_x, cfg.tile_y, cfg.tile_z = tile
        h = ctypes.c_void_p()
        rc = L.lbm_create(ctypes.byref(cfg), ctypes.byref(h))
        if rc != 0:
            raise LBMError(f"lbm_create failed ({rc}): {L.lbm_create_error().decode()}")
        self._h = h
        self.nx, self.ny, self.nz = nx, ny, nz
        self.n = nx * ny * nz
        self.Q = velocity_set
        self.storage = storage

    def _check(self, rc):
        if rc != 0:
            raise LBMError(f"({rc}) {lib().lbm_last_error(self._h).decode()}")

    def close(self):
        if getattr(self, "_h", None):
            lib().lbm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def tiling(self):
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._check(lib().lbm_get_tiling(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    @property
    def omega(self):
        o, t = ctypes.c_double(), ctypes.c_double()
        self._check(lib().lbm_get_omega(self._h, ctypes.byref(o), ctypes.byref(t)))
        return o.value

    @property
    def tau(self):
        o, t = ctypes.c_double(), ctypes.c_double()
        self._check(lib().lbm_get_omega(self._h, ctypes.byref(o), ctypes.byref(t)))
        return t.value

    def set_moments(self, mom, layout=LBM_LAYOUT_GLOBAL):
        arrs = [np.ascontiguousarray(m, dtype=np.float32).reshape(-1) for m in mom]
        assert len(arrs) == 10 and all(a.size == self.n for a in arrs)
        self._check(lib().lbm_set_moments(self._h, _ptr_array(arrs, ctypes.c_float), layout))

    def get_moments(self, layout=LBM_LAYOUT_GLOBAL):
        arrs = [np.empty(self.n, dtype=np.float32) for _ in range(10)]
        self._check(lib().lbm_get_moments(self._h, _ptr_array(arrs, ctypes.c_float), layout))
        return arrs

    def step(self, n=1):
        self._check(lib().lbm_step(self._h, int(n)))

    def synchronize(self):
        self._check(lib().lbm_synchronize(self._h))

    def step_timed(self, n):
        ms = ctypes.c_double()
        self._check(lib().lbm_step_timed(self._h, int(n), ctypes.byref(ms)))
        return ms.value

    @property
    def time(self):
        t = ctypes.c_int64()
        self._check(lib().lbm_get_time(self._h, ctypes.byref(t)))
        return t.value

    @time.setter
    def time(self, t):
        self._check(lib().lbm_set_time(self._h, int(t)))

    def fo_enable(self, mask):
        self._check(lib().lbm_fo_enable(self._h, int(mask)))

    def fo_get(self, field, layout=LBM_LAYOUT_GLOBAL):
        ncomp = 6 if field in (LBM_FIELD_S, LBM_FIELD_SMEAN) else 1
        out = np.empty(ncomp * self.n, dtype=np.float64)
        self._check(lib().lbm_fo_get(self._h, field, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), layout))
        return out.reshape(ncomp, self.n) if ncomp > 1 else out

    def reduce(self):
        ke, r = ctypes.c_double(), ctypes.c_double()
        self._check(lib().lbm_reduce(self._h, ctypes.byref(ke), ctypes.byref(r)))
        return ke.value, r.value

    def probe_streaming(self):
        src = np.empty(self.n * self.Q, dtype=np.uint32)
        typ = np.empty(self.n, dtype=np.uint8)
        self._check(lib().lbm_probe_streaming(self._h, src.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                              typ.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))))
        return src.reshape(self.n, self.Q), typ


# ------------------------------------------------------------------ case I/O
class lbm_case_info(ctypes.Structure):
    _fields_ = [
        ("case_name", ctypes.c_char * 256),
        ("Re", ctypes.c_double),
        ("u_inf", ctypes.c_double),
        ("n_time_steps", ctypes.c_int64),
        ("save_interval", ctypes.c_int64),
        ("info_interval", ctypes.c_int64),
        ("nx", ctypes.c_int),
        ("ny", ctypes.c_int),
        ("nz", ctypes.c_int),
        ("Lx", ctypes.c_double),
        ("Ly", ctypes.c_double),
        ("Lz", ctypes.c_double),
        ("latest_time", ctypes.c_int64),
    ]


EXPORTED += ["lbm_case_last_error", "lbm_case_read", "lbm_case_parameter", "lbm_case_initial_moments",
             "lbm_case_function_objects", "lbm_lbmbin_write", "lbm_lbmbin_header", "lbm_lbmbin_read_field"]


def _case_lib():
    L = lib()
    if not getattr(L, "_case_ready", False):
        fp = ctypes.POINTER(ctypes.c_float)
        L.lbm_case_last_error.restype = ctypes.c_char_p
        L.lbm_case_read.argtypes = [ctypes.c_char_p, ctypes.POINTER(lbm_case_info)]
        L.lbm_case_parameter.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
        L.lbm_case_initial_moments.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                               ctypes.POINTER(fp)]
        L.lbm_case_function_objects.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_uint)]
        L.lbm_lbmbin_write.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(fp), ctypes.c_int64, ctypes.c_int]
        L.lbm_lbmbin_header.argtypes = [ctypes.c_char_p] + [ctypes.POINTER(ctypes.c_int)] * 5 + [
            ctypes.POINTER(ctypes.c_int64), ctypes.c_char_p, ctypes.c_int]
        L.lbm_lbmbin_read_field.argtypes = [ctypes.c_char_p, ctypes.c_char_p, fp]
        L._case_ready = True
    return L


def _case_check(rc):
    if rc != 0:
        raise LBMError(_case_lib().lbm_case_last_error().decode())


def read_case(directory="."):
    """programControl + latticeMesh values and the latest saved time of a case."""
    info = lbm_case_info()
    _case_check(_case_lib().lbm_case_read(str(directory).encode(), ctypes.byref(info)))
    d = {name: getattr(info, name) for name, _ in lbm_case_info._fields_}
    d["case_name"] = info.case_name.decode()
    return d


def case_parameter(path, name):
    buf = ctypes.create_string_buffer(512)
    _case_check(_case_lib().lbm_case_parameter(str(path).encode(), name.encode(), buf, 512))
    return buf.value.decode()


def initial_moments(directory, nx, ny, nz):
    arrs = [np.empty(nx * ny * nz, dtype=np.float32) for _ in range(10)]
    _case_check(_case_lib().lbm_case_initial_moments(str(directory).encode(), nx, ny, nz,
                                                     _ptr_array(arrs, ctypes.c_float)))
    return arrs


def function_objects(directory):
    m = ctypes.c_uint()
    _case_check(_case_lib().lbm_case_function_objects(str(directory).encode(), ctypes.byref(m)))
    return m.value


def write_lbmbin(path, nx, ny, nz, names, fields, time_step, time_average=False):
    arrs = [np.ascontiguousarray(f, dtype=np.float32).reshape(-1) for f in fields]
    cn = (ctypes.c_char_p * len(names))(*[n.encode() for n in names])
    _case_check(_case_lib().lbm_lbmbin_write(str(path).encode(), nx, ny, nz, len(names), cn,
                                             _ptr_array(arrs, ctypes.c_float), int(time_step), int(time_average)))


def lbmbin_header(path):
    v = [ctypes.c_int() for _ in range(5)]
    t = ctypes.c_int64()
    names = ctypes.create_string_buffer(4096)
    _case_check(_case_lib().lbm_lbmbin_header(str(path).encode(), *[ctypes.byref(x) for x in v], ctypes.byref(t),
                                              names, 4096))
    nv = v[3].value
    return {"nx": v[0].value, "ny": v[1].value, "nz": v[2].value, "nvars": nv, "scalar_bytes": v[4].value,
            "time_step": t.value, "names": names.raw.split(b"\0")[:nv] and [s.decode() for s in names.raw.split(b"\0")[:nv]]}


def read_lbmbin_field(path, name):
    h = lbmbin_header(path)
    out = np.empty(h["nx"] * h["ny"] * h["nz"], dtype=np.float32)
    _case_check(_case_lib().lbm_lbmbin_read_field(str(path).encode(), name.encode(),
                                                  out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
    return out
