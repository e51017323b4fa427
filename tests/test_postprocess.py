"""Host-only post-processing (include/lbm_post.h, cudalbm_amd/bin/fieldConvert,
fieldCalculate): schemes pinned by convergence order and closed forms, writers
by parsing their output back, the drivers end to end on a written case."""
import os
import re
import shutil
import struct
import subprocess

import numpy as np
import pytest

from cudalbm_amd import native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "cudalbm_amd", "bin")
GOLDEN = os.path.join(ROOT, "tests", "golden", "cases", "ComputeSanitizerTest")  # 64^3


def line_field(f1d, n, axis, other=4):
    shape = [other, other, other]
    shape[2 - axis] = n  # arrays are z, y, x
    sl = [None, None, None]
    sl[2 - axis] = slice(None)
    F = np.broadcast_to(f1d[tuple(sl)], shape).copy()
    dims = [other, other, other]
    dims[axis] = n
    return F, tuple(dims)


@pytest.mark.parametrize("order,ratio", [(2, 3.5), (4, 12.0), (6, 40.0), (8, 150.0)])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_central_scheme_converges_at_its_order(order, ratio, axis):
    # sin(2 pi m x/(n-1)) is odd about both end nodes, so the odd extension is exact
    errs = []
    for n in (33, 65):
        x = np.arange(n)
        k = 2 * np.pi * 4 / (n - 1)
        F, dims = line_field(np.sin(k * x), n, axis)
        d = N.post_derivative(F, dims, axis, order).reshape(F.shape)
        exact = np.broadcast_to((k * np.cos(k * x))[tuple(slice(None) if i == 2 - axis else None for i in range(3))],
                                F.shape)
        errs.append(np.abs(d - exact).max())
    assert errs[0] / errs[1] > ratio, errs


@pytest.mark.parametrize("order,slope", [(2, 1.0), (4, 5.0 / 3.0), (6, 2.2), (8, 4 / 5 * 2 + 1 / 5 * 4 + 4 / 105 * 6 + 1 / 280 * 8)])
def test_reference_scheme_sums(order, slope):
    """The reference's stencils add every term: d/dx of f = x at interior nodes is sum_k 2k c_k."""
    n = 24
    F, dims = line_field(np.arange(n, dtype=float), n, 0)
    d = N.post_derivative(F, dims, 0, order, scheme=N.LBM_POST_REFERENCE).reshape(F.shape)[0, 0]
    c = N.post_derivative(F, dims, 0, order).reshape(F.shape)[0, 0]
    assert np.allclose(d[order:n - order], slope, rtol=1e-12)
    assert np.allclose(c[order:n - order], 1.0, rtol=1e-12)


def test_odd_extension_at_the_end_nodes():
    # f = x^2: g(-1) = -f(1) = -1, g(n) = -f(n-2); order 2 at x=0: (f(1) - g(-1))/2 = 1
    n = 10
    F, dims = line_field(np.arange(n, dtype=float) ** 2, n, 0)
    d = N.post_derivative(F, dims, 0, 2).reshape(F.shape)[0, 0]
    assert d[0] == 1.0
    assert d[-1] == ((-(n - 2) ** 2) - (n - 2) ** 2) / 2.0
    assert np.allclose(d[1:-1], 2 * np.arange(1, n - 1))


def test_curl_and_div_of_linear_fields():
    n = 20
    z, y, x = np.meshgrid(*(np.arange(n, dtype=float),) * 3, indexing="ij")
    u, v, w = -y + 0.5 * z, x, 2 * x - y  # curl = (-1 - 0, 0.5 - 2, 1 + 1), div = 0
    cx, cy, cz = (a.reshape(n, n, n) for a in N.post_curl(u, v, w, (n, n, n)))
    dv = N.post_div(u + 3 * x, v - y, w + z, (n, n, n)).reshape(n, n, n)
    inner = (slice(8, n - 8),) * 3
    assert np.allclose(cx[inner], -1.0) and np.allclose(cy[inner], -1.5) and np.allclose(cz[inner], 2.0)
    assert np.allclose(dv[inner], 3.0 - 1.0 + 1.0)


def parse_vtk_appended(path):
    raw = open(path, "rb").read()
    head, data = raw.split(b'<AppendedData encoding="raw">_', 1)
    arrays = {}
    for m in re.finditer(rb'<DataArray type="(\w+)" Name="([^"]+)"[^>]*offset="(\d+)"', head):
        t, name, off = m.group(1).decode(), m.group(2).decode(), int(m.group(3))
        nbytes = struct.unpack_from("<Q", data, off)[0]
        dt = {"Float32": "<f4", "UInt32": "<u4", "UInt64": "<u8", "UInt8": "u1"}[t]
        arrays[name] = np.frombuffer(data, dtype=dt, count=nbytes // np.dtype(dt).itemsize, offset=off + 8)
    return head.decode(), arrays


def test_vts_and_vtu_round_trip(tmp_path):
    nx, ny, nz = 3, 4, 5
    rng = np.random.default_rng(0)
    a, b = rng.normal(size=(2, nx * ny * nz)).astype(np.float32)
    for fmt, ext in ((N.LBM_POST_VTS, ".vts"), (N.LBM_POST_VTU, ".vtu")):
        N.post_write(fmt, tmp_path / "f", (nx, ny, nz), (1.0, 2.0, 4.0), ["a", "b"], [a, b])
        head, arr = parse_vtk_appended(tmp_path / ("f" + ext))
        assert np.array_equal(arr["a"], a) and np.array_equal(arr["b"], b)
        pts = arr["Coordinates"].reshape(-1, 3)
        assert np.allclose(pts[-1], (1.0, 2.0, 4.0)) and np.allclose(pts[1], (0.5, 0.0, 0.0))
        assert np.allclose(pts[nx], (0.0, 2.0 / 3.0, 0.0)) and np.allclose(pts[nx * ny], (0.0, 0.0, 1.0))
        if ext == ".vts":
            assert 'WholeExtent="0 2 0 3 0 4"' in head
        else:
            ne = (nx - 1) * (ny - 1) * (nz - 1)
            assert f'NumberOfCells="{ne}"' in head
            conn = arr["connectivity"].reshape(-1, 8)
            assert list(conn[0]) == [0, 1, nx + 1, nx, nx * ny, nx * ny + 1, nx * ny + nx + 1, nx * ny + nx]
            assert list(arr["offsets"][:3]) == [8, 16, 24] and np.all(arr["types"] == 12)


def test_tecplot_layout(tmp_path):
    nx, ny, nz = 2, 2, 3
    f = np.arange(nx * ny * nz, dtype=np.float32) * 0.5
    N.post_write(N.LBM_POST_TECPLOT, tmp_path / "t", (nx, ny, nz), (1.0, 1.0, 1.0), ["f"], [f])
    lines = open(tmp_path / "t.dat").read().splitlines()
    assert lines[0].startswith('VARIABLES = "X" "Y" "Z" "f"')
    assert "NODES=12, ELEMENTS=2, DATAPACKING=BLOCK, ZONETYPE=FEBRICK" in lines[1]
    vals = [float(v) for v in lines[2:2 + 4 * 12]]
    assert vals[12 * 3:] == list(f)  # after the X, Y, Z blocks
    assert lines[2 + 48].split() == ["1", "2", "4", "3", "5", "6", "8", "7"]  # one-based first hexahedron


def test_cut_plane_follows_the_reference_rule():
    nx, ny, nz = 5, 3, 2
    f = np.arange(nx * ny * nz, dtype=np.float32)
    F = f.reshape(nz, ny, nx)
    # c = 0.5: i = 4 * 0.5 = 2 exactly -> plane x = 2
    assert np.array_equal(N.post_cut_plane(f, (nx, ny, nz), 0, 0.5), F[:, :, 2].reshape(-1))
    # c = 0.3: i = 1.2, i0 = 1, i1 = 2, weight = c - i0 = -0.7 (the reference's rule)
    got = N.post_cut_plane(f, (nx, ny, nz), 0, 0.3)
    assert np.allclose(got, (1.7 * F[:, :, 1] - 0.7 * F[:, :, 2]).reshape(-1), rtol=1e-6)
    # z planes come out x fastest
    assert np.array_equal(N.post_cut_plane(f, (nx, ny, nz), 2, 1.0), F[1].reshape(-1))


def test_time_indices(tmp_path):
    for name in ("k_0.LBMBin", "k_100.LBMBin", "k_20.LBMBin", "k_x.LBMBin", "kMean_5.LBMBin", "k_7.txt"):
        (tmp_path / name).write_bytes(b"")
    assert N.post_time_indices(tmp_path, "k") == [0, 20, 100]


@pytest.fixture
def written_case(tmp_path):
    for f in ("programControl", "latticeMesh"):
        shutil.copyfile(os.path.join(GOLDEN, f), tmp_path / f)
    info = N.read_case(tmp_path)
    nx, ny, nz = info["nx"], info["ny"], info["nz"]
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    names = ["rho", "u", "v", "w", "m_xx", "m_xy", "m_xz", "m_yy", "m_yz", "m_zz"]
    fields = [np.full(nx * ny * nz, 0.01 * k, np.float32) for k in range(10)]
    fields[1] = (-(y - ny / 2.0)).astype(np.float32).reshape(-1) * 1e-3
    fields[2] = (x - nx / 2.0).astype(np.float32).reshape(-1) * 1e-3
    for t in (0, 100):
        N.write_lbmbin(tmp_path / f"{info['case_name']}_{t}.LBMBin", nx, ny, nz, names, fields, t)
    return tmp_path, info, fields


def run(tool, cwd, *args):
    return subprocess.run([os.path.join(BIN, tool), *args], cwd=cwd, capture_output=True, text=True, check=True).stdout


@pytest.mark.skipif(not os.path.exists(os.path.join(BIN, "fieldConvert")), reason="drivers not built")
def test_field_tools_end_to_end(native_lib, written_case):
    d, info, fields = written_case
    run("fieldConvert", d, "-fileType", "vts", "-latestTime")
    _, arr = parse_vtk_appended(d / "postProcess" / f"{info['case_name']}_100.vts")
    assert np.array_equal(arr["u"], fields[1])
    run("fieldConvert", d, "-fileType", "vtu", "-cutPlane", "z=0.5")
    assert (d / "postProcess" / f"{info['case_name']}CutPlane_z=0.5_0.vtu").exists()
    out = run("fieldCalculate", d, "-calculationType", "containsNaN")
    assert out.count("No NaN values detected in any field.") == 2
    out = run("fieldCalculate", d, "-calculationType", "spatialMean", "-latestTime")
    # the reference's FP32 running sum in index order, printed with 6 significant digits
    for k, name in ((4, "m_xx"), (1, "u")):
        mean = np.add.accumulate(fields[k], dtype=np.float32)[-1] / np.float32(fields[k].size)
        assert f"mean({name}): {float(mean):.6g};" in out
    run("fieldCalculate", d, "-calculationType", "vorticity", "-fileType", "vts", "-scheme", "central")
    _, arr = parse_vtk_appended(d / "postProcess" / "vorticity_0.vts")
    nx, ny, nz = info["nx"], info["ny"], info["nz"]
    oz = arr["omega_z"].reshape(nz, ny, nx)
    assert np.allclose(oz[8:-8, 8:-8, 8:-8], 2e-3, rtol=1e-4)  # dv/dx - du/dy
