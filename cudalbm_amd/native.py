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
LBM_FO_K, LBM_FO_KMEAN, LBM_FO_S, LBM_FO_SMEAN, LBM_FO_MOMENTSMEAN = 1, 2, 4, 8, 16
LBM_FIELD_K, LBM_FIELD_KMEAN, LBM_FIELD_S, LBM_FIELD_SMEAN, LBM_FIELD_MOMENTSMEAN = 0, 1, 2, 3, 4

# Every symbol include/lbm_amd.h declares (checked by tests/test_native_abi.py).
EXPORTED = [
    "lbm_create", "lbm_destroy", "lbm_last_error", "lbm_create_error", "lbm_get_tiling", "lbm_get_omega",
    "lbm_set_moments", "lbm_get_moments", "lbm_set_moments_f64", "lbm_get_moments_f64", "lbm_step",
    "lbm_synchronize", "lbm_step_timed", "lbm_get_time", "lbm_set_time", "lbm_fo_enable", "lbm_fo_get",
    "lbm_reduce", "lbm_reduce_noalloc", "lbm_probe_streaming",
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
    L.lbm_reduce_noalloc.argtypes = [H, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_void_p]
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
        ncomp = 6 if field in (LBM_FIELD_S, LBM_FIELD_SMEAN) else (10 if field == LBM_FIELD_MOMENTSMEAN else 1)
        out = np.empty(ncomp * self.n, dtype=np.float64)
        self._check(lib().lbm_fo_get(self._h, field, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), layout))
        return out.reshape(ncomp, self.n) if ncomp > 1 else out

    def reduce(self):
        ke, r = ctypes.c_double(), ctypes.c_double()
        self._check(lib().lbm_reduce(self._h, ctypes.byref(ke), ctypes.byref(r)))
        return ke.value, r.value

    def reduce_noalloc(self, uprev_ptr, uprev_count, first, out_ptr):
        """lbm_reduce_noalloc: device pointers (ints) to the caller's previous-
        velocity buffer (uprev_count floats) and to 2 doubles of output;
        asynchronous on the solver's stream."""
        self._check(lib().lbm_reduce_noalloc(self._h, ctypes.c_void_p(uprev_ptr), int(uprev_count), int(first),
                                             ctypes.c_void_p(out_ptr)))

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


# ------------------------------------------------------------------ batches
class lbm_batch_case(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("Re", ctypes.c_double), ("u_inf", ctypes.c_double)]


class lbm_batch_config(ctypes.Structure):
    _fields_ = [("velocity_set", ctypes.c_int), ("device", ctypes.c_int), ("stream", ctypes.c_void_p),
                ("tile_x", ctypes.c_int), ("tile_y", ctypes.c_int), ("tile_z", ctypes.c_int),
                ("persistent", ctypes.c_int)]


EXPORTED += ["lbm_autotune", "lbm_batch_create", "lbm_batch_create_error", "lbm_batch_destroy",
             "lbm_batch_last_error", "lbm_batch_get_tiling", "lbm_batch_tuning_report", "lbm_batch_work_items",
             "lbm_batch_set_active", "lbm_batch_step", "lbm_batch_step_timed", "lbm_batch_synchronize",
             "lbm_batch_get_steps", "lbm_batch_set_moments", "lbm_batch_get_moments", "lbm_batch_reduce",
             "lbm_batch_run"]


def _batch_lib():
    L = lib()
    if not getattr(L, "_batch_ready", False):
        B = ctypes.c_void_p
        ip = ctypes.POINTER(ctypes.c_int)
        dp = ctypes.POINTER(ctypes.c_double)
        i64p = ctypes.POINTER(ctypes.c_int64)
        fpp = ctypes.POINTER(ctypes.POINTER(ctypes.c_float))
        L.lbm_autotune.argtypes = [ctypes.POINTER(lbm_config), ctypes.c_int, ip, ip, ip, dp]
        L.lbm_batch_create.argtypes = [ctypes.POINTER(lbm_batch_config), ctypes.POINTER(lbm_batch_case),
                                       ctypes.c_int, ctypes.POINTER(B)]
        L.lbm_batch_create_error.restype = ctypes.c_char_p
        L.lbm_batch_destroy.argtypes = [B]
        L.lbm_batch_last_error.argtypes = [B]
        L.lbm_batch_last_error.restype = ctypes.c_char_p
        L.lbm_batch_get_tiling.argtypes = [B, ip, ip, ip, ip]
        L.lbm_batch_tuning_report.argtypes = [B, ctypes.c_char_p, ctypes.c_int]
        L.lbm_batch_work_items.argtypes = [B, ip]
        L.lbm_batch_set_active.argtypes = [B, ctypes.POINTER(ctypes.c_uint8)]
        L.lbm_batch_step.argtypes = [B, ctypes.c_int]
        L.lbm_batch_step_timed.argtypes = [B, ctypes.c_int, dp]
        L.lbm_batch_synchronize.argtypes = [B]
        L.lbm_batch_get_steps.argtypes = [B, i64p]
        L.lbm_batch_set_moments.argtypes = [B, ctypes.c_int, fpp, ctypes.c_int]
        L.lbm_batch_get_moments.argtypes = [B, ctypes.c_int, fpp, ctypes.c_int]
        L.lbm_batch_reduce.argtypes = [B, dp, dp]
        L.lbm_batch_run.argtypes = [B, ctypes.c_double, ctypes.c_int, ctypes.c_int64, i64p, dp, dp]
        L._batch_ready = True
    return L


def autotune(nx, ny, nz, Re, u_inf, velocity_set=19, storage="fp32", steps=20, device=0):
    """Fastest (tile_x, tile_y, tile_z) for one case and its ms per step."""
    L = _batch_lib()
    cfg = lbm_config()
    cfg.velocity_set = velocity_set
    cfg.storage = {"fp32": LBM_STORE_FP32, "fp16": LBM_STORE_FP16}[storage]
    cfg.nx, cfg.ny, cfg.nz = nx, ny, nz
    cfg.Re, cfg.u_inf = float(Re), float(u_inf)
    cfg.device = device
    t = [ctypes.c_int() for _ in range(3)]
    ms = ctypes.c_double()
    rc = L.lbm_autotune(ctypes.byref(cfg), int(steps), *[ctypes.byref(x) for x in t], ctypes.byref(ms))
    if rc != 0:
        raise LBMError(f"lbm_autotune failed ({rc}): {L.lbm_batch_create_error().decode()}")
    return (t[0].value, t[1].value, t[2].value), ms.value


class Batch:
    """Independent cavities advancing together, one launch per step (lbm_batch_* of include/lbm_amd.h).

    cases: sequence of (nx, ny, nz, Re, u_inf).  tile = (tile_x, tile_y, tile_z), 0 = autotune;
    persistent = 0/1, -1 = autotune.
    """

    def __init__(self, cases, velocity_set=19, device=0, tile=(0, 0, 0), persistent=-1):
        L = _batch_lib()
        arr = (lbm_batch_case * len(cases))()
        for i, (nx, ny, nz, Re, u) in enumerate(cases):
            arr[i].nx, arr[i].ny, arr[i].nz = nx, ny, nz
            arr[i].Re, arr[i].u_inf = float(Re), float(u)
        cfg = lbm_batch_config()
        cfg.velocity_set = velocity_set
        cfg.device = device
        cfg.stream = None
        cfg.tile_x, cfg.tile_y, cfg.tile_z = tile
        cfg.persistent = persistent
        b = ctypes.c_void_p()
        rc = L.lbm_batch_create(ctypes.byref(cfg), arr, len(cases), ctypes.byref(b))
        if rc != 0:
            raise LBMError(f"lbm_batch_create failed ({rc}): {L.lbm_batch_create_error().decode()}")
        self._b = b
        self.cases = [tuple(c) for c in cases]
        self.ncases = len(cases)
        self.Q = velocity_set

    def _check(self, rc):
        if rc != 0:
            raise LBMError(f"({rc}) {_batch_lib().lbm_batch_last_error(self._b).decode()}")

    def close(self):
        if getattr(self, "_b", None):
            _batch_lib().lbm_batch_destroy(self._b)
            self._b = None

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
        v = [ctypes.c_int() for _ in range(4)]
        self._check(_batch_lib().lbm_batch_get_tiling(self._b, *[ctypes.byref(x) for x in v]))
        return tuple(x.value for x in v)

    @property
    def tuning_report(self):
        buf = ctypes.create_string_buffer(1 << 16)
        self._check(_batch_lib().lbm_batch_tuning_report(self._b, buf, len(buf)))
        return buf.value.decode()

    @property
    def work_items(self):
        n = ctypes.c_int()
        self._check(_batch_lib().lbm_batch_work_items(self._b, ctypes.byref(n)))
        return n.value

    def set_active(self, active):
        a = np.ascontiguousarray(np.asarray(active, dtype=np.uint8).reshape(-1))
        assert a.size == self.ncases
        self._check(_batch_lib().lbm_batch_set_active(self._b, a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))))

    def step(self, n=1):
        self._check(_batch_lib().lbm_batch_step(self._b, int(n)))

    def step_timed(self, n):
        ms = ctypes.c_double()
        self._check(_batch_lib().lbm_batch_step_timed(self._b, int(n), ctypes.byref(ms)))
        return ms.value

    def synchronize(self):
        self._check(_batch_lib().lbm_batch_synchronize(self._b))

    @property
    def steps(self):
        out = np.empty(self.ncases, dtype=np.int64)
        self._check(_batch_lib().lbm_batch_get_steps(self._b, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
        return out

    def _n(self, cs):
        # an index out of range is left to the library, which reports it
        if not 0 <= cs < self.ncases:
            return 1
        nx, ny, nz = self.cases[cs][:3]
        return nx * ny * nz

    def set_moments(self, cs, mom, layout=LBM_LAYOUT_GLOBAL):
        arrs = [np.ascontiguousarray(m, dtype=np.float32).reshape(-1) for m in mom]
        assert len(arrs) == 10 and all(a.size == self._n(cs) for a in arrs)
        self._check(_batch_lib().lbm_batch_set_moments(self._b, int(cs), _ptr_array(arrs, ctypes.c_float), layout))

    def get_moments(self, cs, layout=LBM_LAYOUT_GLOBAL):
        arrs = [np.empty(self._n(cs), dtype=np.float32) for _ in range(10)]
        self._check(_batch_lib().lbm_batch_get_moments(self._b, int(cs), _ptr_array(arrs, ctypes.c_float), layout))
        return arrs

    def reduce(self):
        ke = np.empty(self.ncases)
        res = np.empty(self.ncases)
        dp = ctypes.POINTER(ctypes.c_double)
        self._check(_batch_lib().lbm_batch_reduce(self._b, ke.ctypes.data_as(dp), res.ctypes.data_as(dp)))
        return ke, res

    def run(self, tol, check, max_steps):
        """Steps every active case until its residual < tol or max_steps: (steps, residual, wall ms)."""
        steps = np.empty(self.ncases, dtype=np.int64)
        res = np.empty(self.ncases)
        ms = ctypes.c_double()
        dp = ctypes.POINTER(ctypes.c_double)
        self._check(_batch_lib().lbm_batch_run(self._b, float(tol), int(check), int(max_steps),
                                               steps.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                               res.ctypes.data_as(dp), ctypes.byref(ms)))
        return steps, res, ms.value


# ------------------------------------------------------------------ surrogate
EXPORTED += ["lbm_sur_create", "lbm_sur_destroy", "lbm_sur_last_error", "lbm_sur_create_error",
             "lbm_sur_num_params", "lbm_sur_get_params", "lbm_sur_set_params", "lbm_sur_encode",
             "lbm_sur_gradient", "lbm_sur_load_data", "lbm_sur_train", "lbm_sur_predict", "lbm_sur_warm_start"]


def _sur_lib():
    L = lib()
    if not getattr(L, "_sur_ready", False):
        P = ctypes.c_void_p
        fp = ctypes.POINTER(ctypes.c_float)
        dp = ctypes.POINTER(ctypes.c_double)
        L.lbm_sur_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(P)]
        L.lbm_sur_destroy.argtypes = [P]
        L.lbm_sur_last_error.argtypes = [P]
        L.lbm_sur_last_error.restype = ctypes.c_char_p
        L.lbm_sur_create_error.restype = ctypes.c_char_p
        L.lbm_sur_num_params.argtypes = [P]
        L.lbm_sur_num_params.restype = ctypes.c_int64
        L.lbm_sur_get_params.argtypes = [P, fp]
        L.lbm_sur_set_params.argtypes = [P, fp]
        L.lbm_sur_encode.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, fp]
        L.lbm_sur_gradient.argtypes = [P, ctypes.c_int, fp, fp, dp, fp]
        L.lbm_sur_load_data.argtypes = [P, ctypes.c_int, fp, fp]
        L.lbm_sur_train.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_uint64, dp]
        L.lbm_sur_predict.argtypes = [P, ctypes.c_int, fp, fp]
        L.lbm_sur_warm_start.argtypes = [P, ctypes.c_void_p, ctypes.c_double, ctypes.c_double]
        L._sur_ready = True
    return L


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def encode_inputs(S, Re, N):
    """Input channels of one case on an S^3 grid: [S^3][3] (lid mask, c_Re, c_N)."""
    out = np.empty((S ** 3, 3), dtype=np.float32)
    rc = _sur_lib().lbm_sur_encode(int(S), float(Re), int(N), _fptr(out))
    if rc != 0:
        raise LBMError(f"lbm_sur_encode failed ({rc})")
    return out


class Surrogate:
    """The 3D U-Net warm-start network of include/lbm_surrogate.h (GPU only)."""

    def __init__(self, S=32, seed=0, device=0):
        L = _sur_lib()
        p = ctypes.c_void_p()
        rc = L.lbm_sur_create(int(device), int(S), int(seed), ctypes.byref(p))
        if rc != 0:
            raise LBMError(f"lbm_sur_create failed ({rc}): {L.lbm_sur_create_error().decode()}")
        self._p = p
        self.S = S

    def _check(self, rc):
        if rc != 0:
            raise LBMError(f"({rc}) {_sur_lib().lbm_sur_last_error(self._p).decode()}")

    def close(self):
        if getattr(self, "_p", None):
            _sur_lib().lbm_sur_destroy(self._p)
            self._p = None

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
    def num_params(self):
        return int(_sur_lib().lbm_sur_num_params(self._p))

    def get_params(self):
        out = np.empty(self.num_params, dtype=np.float32)
        self._check(_sur_lib().lbm_sur_get_params(self._p, _fptr(out)))
        return out

    def set_params(self, params):
        a = _f32(params).reshape(-1)
        assert a.size == self.num_params
        self._check(_sur_lib().lbm_sur_set_params(self._p, _fptr(a)))

    def gradient(self, inputs, targets):
        """(loss, grad) of the MSE for inputs [n][S^3][3], targets [n][S^3][4]."""
        x, t = _f32(inputs), _f32(targets)
        n = x.shape[0]
        g = np.empty(self.num_params, dtype=np.float32)
        loss = ctypes.c_double()
        self._check(_sur_lib().lbm_sur_gradient(self._p, n, _fptr(x), _fptr(t), ctypes.byref(loss), _fptr(g)))
        return loss.value, g

    def load_data(self, inputs, targets):
        x, t = _f32(inputs), _f32(targets)
        self._check(_sur_lib().lbm_sur_load_data(self._p, x.shape[0], _fptr(x), _fptr(t)))

    def train(self, epochs, batch=8, lr=1e-3, seed=0):
        losses = np.empty(max(1, epochs))
        self._check(_sur_lib().lbm_sur_train(self._p, int(epochs), int(batch), float(lr), int(seed),
                                             losses.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return losses[:epochs]

    def predict(self, inputs):
        x = _f32(inputs)
        n = x.shape[0]
        out = np.empty((n, self.S ** 3, 4), dtype=np.float32)
        self._check(_sur_lib().lbm_sur_predict(self._p, n, _fptr(x), _fptr(out)))
        return out

    def warm_start(self, solver, Re, u_inf):
        self._check(_sur_lib().lbm_sur_warm_start(self._p, solver._h, float(Re), float(u_inf)))


class lbm_solve_info(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_int64), ("residual", ctypes.c_double), ("ms", ctypes.c_double),
                ("ms_warm", ctypes.c_double), ("first_residual", ctypes.c_double), ("warm", ctypes.c_int),
                ("fell_back", ctypes.c_int)]


EXPORTED += ["lbm_solve_to_residual"]


def solve_to_residual(solver, Re, u_inf, tol=1e-6, check=100, max_steps=200000, surrogate=None):
    """Steps `solver` until its residual < tol (optionally warm-started by `surrogate`); returns a dict."""
    L = _sur_lib()
    if not getattr(L, "_solve_ready", False):
        L.lbm_solve_to_residual.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                                            ctypes.c_double, ctypes.c_int, ctypes.c_int64,
                                            ctypes.POINTER(lbm_solve_info)]
        L._solve_ready = True
    info = lbm_solve_info()
    rc = L.lbm_solve_to_residual(solver._h, surrogate._p if surrogate is not None else None, float(Re),
                                 float(u_inf), float(tol), int(check), int(max_steps), ctypes.byref(info))
    if rc != 0:
        where = surrogate if surrogate is not None else solver
        msg = (L.lbm_sur_last_error(where._p) if surrogate is not None else L.lbm_last_error(solver._h)).decode()
        raise LBMError(f"lbm_solve_to_residual failed ({rc}): {msg}")
    return {name: getattr(info, name) for name, _ in lbm_solve_info._fields_}


# ------------------------------------------------------------------ post-processing (host only)
LBM_POST_VTU, LBM_POST_VTS, LBM_POST_TECPLOT = 0, 1, 2
LBM_POST_CENTRAL, LBM_POST_REFERENCE = 0, 1
EXPORTED += ["lbm_post_time_indices", "lbm_post_write", "lbm_post_derivative", "lbm_post_curl", "lbm_post_div",
             "lbm_post_cut_plane"]


def _post_lib():
    L = _case_lib()
    if not getattr(L, "_post_ready", False):
        fp = ctypes.POINTER(ctypes.c_float)
        dp = ctypes.POINTER(ctypes.c_double)
        i = ctypes.c_int
        L.lbm_post_time_indices.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int64), i]
        L.lbm_post_write.argtypes = [i, ctypes.c_char_p, i, i, i, ctypes.c_double, ctypes.c_double, ctypes.c_double, i,
                                     ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(fp)]
        L.lbm_post_derivative.argtypes = [i, i, i, i, i, i, fp, dp]
        L.lbm_post_curl.argtypes = [i, i, i, i, i, fp, fp, fp, fp, fp, fp]
        L.lbm_post_div.argtypes = [i, i, i, i, i, fp, fp, fp, fp]
        L.lbm_post_cut_plane.argtypes = [i, ctypes.c_double, i, i, i, ctypes.c_double, fp, fp]
        L._post_ready = True
    return L


def post_time_indices(directory, prefix):
    L = _post_lib()
    n = L.lbm_post_time_indices(str(directory).encode(), prefix.encode(), None, 0)
    _case_check(0 if n >= 0 else n)
    out = np.empty(n, dtype=np.int64)
    if n:
        L.lbm_post_time_indices(str(directory).encode(), prefix.encode(),
                                out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), n)
    return list(out)


def post_write(fmt, path, shape, L_xyz, names, fields):
    nx, ny, nz = shape
    arrs = [_f32(f).reshape(-1) for f in fields]
    cn = (ctypes.c_char_p * len(names))(*[n.encode() for n in names])
    _case_check(_post_lib().lbm_post_write(fmt, str(path).encode(), nx, ny, nz, *map(float, L_xyz), len(names), cn,
                                           _ptr_array(arrs, ctypes.c_float)))


def post_derivative(f, shape, axis, order=8, scheme=LBM_POST_CENTRAL):
    nx, ny, nz = shape
    a = _f32(f).reshape(-1)
    out = np.empty(a.size)
    _case_check(_post_lib().lbm_post_derivative(scheme, order, axis, nx, ny, nz, _fptr(a),
                                                out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return out


def post_curl(u, v, w, shape, order=8, scheme=LBM_POST_CENTRAL):
    nx, ny, nz = shape
    a, b, c = (_f32(x).reshape(-1) for x in (u, v, w))
    o = [np.empty(a.size, dtype=np.float32) for _ in range(3)]
    _case_check(_post_lib().lbm_post_curl(scheme, order, nx, ny, nz, _fptr(a), _fptr(b), _fptr(c), *map(_fptr, o)))
    return o


def post_div(u, v, w, shape, order=8, scheme=LBM_POST_CENTRAL):
    nx, ny, nz = shape
    a, b, c = (_f32(x).reshape(-1) for x in (u, v, w))
    o = np.empty(a.size, dtype=np.float32)
    _case_check(_post_lib().lbm_post_div(scheme, order, nx, ny, nz, _fptr(a), _fptr(b), _fptr(c), _fptr(o)))
    return o


def post_cut_plane(f, shape, axis, coord, L=1.0):
    nx, ny, nz = shape
    a = _f32(f).reshape(-1)
    m = [nx, ny, nz]
    m[axis] = 1
    out = np.empty(m[0] * m[1] * m[2], dtype=np.float32)
    _case_check(_post_lib().lbm_post_cut_plane(axis, float(coord), nx, ny, nz, float(L), _fptr(a), _fptr(out)))
    return out
