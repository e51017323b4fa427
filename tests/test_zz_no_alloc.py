"""lbm_reduce_noalloc (the entry point that allocates nothing) against
lbm_reduce on the same states: two solvers in a row, the second smaller, each
reduced twice (first call, then against the stored previous field).  The
partial sums run through the same blocks in the same order, so kinetic energy
and residual must be equal bit for bit.  The caller's device buffers are
allocated here through the HIP runtime the library itself links (a second HIP
runtime in the process, such as torch's, would not see the GPU once the
library holds it); the library allocates nothing."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime liblbm_amd.so is linked to."""

    def __init__(self):
        self.L = ctypes.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.L.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]

    def malloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), nbytes) == 0
        return p.value

    def free(self, p):
        assert self.L.hipFree(ctypes.c_void_p(p)) == 0

    def memset(self, p, byte, nbytes):
        assert self.L.hipMemset(ctypes.c_void_p(p), byte, nbytes) == 0

    def to_host(self, p, count, dtype):
        out = np.empty(count, dtype=dtype)
        assert self.L.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(p), out.nbytes, 2) == 0
        return out


def _check_pair(hip, n, Q, storage):
    from cudalbm_amd import Solver

    print(f"no-alloc reduce: D3Q{Q} {storage} {n}^3", flush=True)
    with Solver(n, n, n, 400.0, 0.05, velocity_set=Q, storage=storage) as s:
        count = 3 * n ** 3
        uprev = hip.malloc(4 * count)
        out = hip.malloc(16)
        try:
            hip.memset(uprev, 0xFF, 4 * count)  # NaN: the first call must not read it
            s.step(20)
            s.reduce_noalloc(uprev, count, 1, out)
            s.synchronize()
            ke_a, r_a = hip.to_host(out, 2, np.float64).tolist()
            ke_b, r_b = s.reduce()
            assert r_a == 1.0 and r_b == 1.0
            assert ke_a == ke_b, (ke_a, ke_b)
            s.step(7)
            s.reduce_noalloc(uprev, count, 0, out)
            s.synchronize()
            ke_a, r_a = hip.to_host(out, 2, np.float64).tolist()
            ke_b, r_b = s.reduce()
            print(f"  ke {ke_a!r} / {ke_b!r}, r {r_a!r} / {r_b!r}", flush=True)
            assert ke_a == ke_b, (ke_a, ke_b)
            assert r_a == r_b, (r_a, r_b)
            assert 0.0 < r_a < 1.0 and np.isfinite(ke_a) and ke_a > 0
            # a buffer short of 3 * nodes is refused before anything is launched
            with pytest.raises(Exception, match="previous-velocity buffer"):
                s.reduce_noalloc(uprev, count - 1, 0, out)
        finally:
            s.synchronize()
            hip.free(uprev)
            hip.free(out)


@pytest.mark.parametrize("Q,storage", [(19, "fp32"), (27, "fp16")])
def test_reduce_noalloc_matches_reduce(native_lib, Q, storage):
    hip = _Hip()
    _check_pair(hip, 48, Q, storage)
    _check_pair(hip, 32, Q, storage)  # second call in a row, smaller mesh: the shared scratch is reused
