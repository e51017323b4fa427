"""Case files and .LBMBin format (host only, no GPU).

Fixtures: tests/golden/cases/ holds the reference's shipped lid-driven cavity
case (lidDrivenCavity/D3Q19) and the case of its only test
(unitTests/ComputeSanitizerTest), copied verbatim as input data.
"""
import os
import shutil
import struct

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = os.path.join(HERE, "golden", "cases")


def test_read_shipped_case(native_lib):
    from cudalbm_amd import native as N

    c = N.read_case(os.path.join(CASES, "lidDrivenCavity_D3Q19"))
    assert c["case_name"] == "lidDrivenCavity"
    assert (c["Re"], c["u_inf"]) == (10000.0, 0.0256)
    assert (c["n_time_steps"], c["save_interval"], c["info_interval"]) == (200001, 20000, 2000)
    assert (c["nx"], c["ny"], c["nz"]) == (256, 256, 256)
    assert (c["Lx"], c["Ly"], c["Lz"]) == (1.0, 1.0, 1.0)
    assert c["latest_time"] == 0
    t = N.read_case(os.path.join(CASES, "ComputeSanitizerTest"))
    assert (t["nx"], t["n_time_steps"], t["info_interval"]) == (64, 201, 20)


def test_parameter_rule_first_substring_match(tmp_path, native_lib):
    """strings.cuh:561-595: the first line containing the name wins, even a
    longer name; the value is the 2nd space-separated token minus its last char."""
    from cudalbm_amd import native as N

    f = tmp_path / "p"
    f.write_text("nxx  12;\nnx    64;\n  Re 250.5;\n")
    assert N.case_parameter(f, "nx") == "12"
    assert N.case_parameter(f, "Re") == "250.5"
    with pytest.raises(N.LBMError):
        N.case_parameter(f, "ny")


def test_initial_conditions_shipped_case(native_lib):
    """hostArray.cuh:304-359: North u = 0.0256 (stored x3), North m_xx =
    equilibrium 4.5 u^2, averaged over the regions of edge and corner nodes."""
    from cudalbm_amd import native as N

    n = 16
    m = N.initial_moments(os.path.join(CASES, "lidDrivenCavity_D3Q19"), n, n, n)
    u = m[1].reshape(n, n, n)  # [z][y][x]
    mxx = m[4].reshape(n, n, n)
    U = np.float32(3 * 0.0256)
    assert u[5, n - 1, 5] == U
    assert u[5, n - 1, 0] == np.float32(3 * 0.0256 / 2)
    assert u[0, n - 1, 0] == np.float32(3 * 0.0256 / 3)
    assert u[5, 5, 5] == 0 and u[5, 0, 5] == 0
    assert mxx[5, n - 1, 5] == np.float32(4.5 * 0.0256 ** 2)
    for k in (0, 2, 3, 5, 6, 7, 8, 9):
        assert not m[k].any()


def test_initial_conditions_rules(tmp_path, native_lib):
    from cudalbm_amd import native as N

    src = os.path.join(CASES, "lidDrivenCavity_D3Q19", "initialConditions")
    text = open(src).read()
    # m_xy equilibrium uses the 4.5 factor (boundaryValue.cuh:205-210)
    text2 = text.replace("field m_xy\n{\n    internalField\n    {\n        value   0;",
                         "field m_xy\n{\n    internalField\n    {\n        value   equilibrium;")
    assert text2 != text
    text2 = text2.replace("field v\n{\n    internalField\n    {\n        value   0;",
                          "field v\n{\n    internalField\n    {\n        value   0.5;")
    text2 = text2.replace("field u\n{\n    internalField\n    {\n        value   0;",
                          "field u\n{\n    internalField\n    {\n        value   0.25;")
    (tmp_path / "initialConditions").write_text(text2)
    n = 8
    m = N.initial_moments(tmp_path, n, n, n)
    mid = lambda a: a.reshape(n, n, n)[4, 4, 4]
    assert mid(m[1]) == np.float32(0.75) and mid(m[2]) == np.float32(1.5)
    assert mid(m[5]) == np.float32(4.5 * 0.25 * 0.5)
    # a value that is neither a number nor "equilibrium" is an error
    (tmp_path / "initialConditions").write_text(text.replace("value   0.0256;", "value   1e-3;"))
    with pytest.raises(N.LBMError):
        N.initial_moments(tmp_path, n, n, n)


def test_function_objects_switches(tmp_path, native_lib):
    from cudalbm_amd import native as N

    assert N.function_objects(os.path.join(CASES, "lidDrivenCavity_D3Q19")) == 0  # all commented out
    assert N.function_objects(tmp_path) == 0  # no file: nothing enabled
    (tmp_path / "functionObjects").write_text(
        "functionObjectList\n{\n    k;\n    // S;\n    SMean;\n    kMean;\n};\n")
    assert N.function_objects(tmp_path) == (N.LBM_FO_K | N.LBM_FO_KMEAN | N.LBM_FO_SMEAN)


def expected_header(t, names, nx, ny, nz, time_type="instantaneous"):
    """Appendix-D layout of fileOutput.cuh:123-182, typed from the format description."""
    h = "systemInformation\n{\n\tbinaryType\tlittleEndian;\n\n\tscalarType\t32 bit;\n};\n\n"
    h += f"fieldInformation\n{{\n\ttimeStep\t{t};\n\n\ttimeType\t{time_type};\n\n\tfieldNames[{len(names)}]\n\t{{\n"
    h += "".join(f"\t\t{n};\n" for n in names)
    h += "\t};\n};\n\n"
    h += f"fieldData\n{{\n\tfieldType\tnonUniform;\n\n\tfield[{len(names) * nx * ny * nz}][{len(names)}][{nz}][{ny}][{nx}]\n\t{{\n"
    return h.encode()


def test_lbmbin_bytes_and_round_trip(tmp_path, native_lib):
    from cudalbm_amd import native as N

    nx, ny, nz = 16, 8, 24
    rng = np.random.default_rng(0)
    names = ["rho", "u", "v"]
    fields = [rng.normal(size=nx * ny * nz).astype(np.float32) for _ in names]
    p = tmp_path / "cavity_120.LBMBin"
    N.write_lbmbin(p, nx, ny, nz, names, fields, 120)
    raw = p.read_bytes()
    hdr = expected_header(120, names, nx, ny, nz)
    assert raw.startswith(hdr)
    payload = raw[len(hdr):]
    assert payload.endswith(b"\n\t};\n};\n")
    data = np.frombuffer(payload[: 4 * len(names) * nx * ny * nz], dtype="<f4")
    # payload order: var-major, reference 8^3 tiles (memory.cuh:157-195)
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    x, y, z = x.ravel(), y.ravel(), z.ravel()
    t8 = (x % 8) + 8 * ((y % 8) + 8 * ((z % 8) + 8 * ((x // 8) + (nx // 8) * ((y // 8) + (ny // 8) * (z // 8)))))
    assert len(np.unique(t8)) == nx * ny * nz  # bijection
    n = nx * ny * nz
    for i, f in enumerate(fields):
        assert np.array_equal(data[i * n + t8], f)
    h = N.lbmbin_header(p)
    assert (h["nx"], h["ny"], h["nz"], h["nvars"], h["scalar_bytes"], h["time_step"]) == (nx, ny, nz, 3, 4, 120)
    assert h["names"] == names
    for i, nm in enumerate(names):
        assert np.array_equal(N.read_lbmbin_field(p, nm), fields[i])


def test_lbmbin_reader_errors(tmp_path, native_lib):
    from cudalbm_amd import native as N

    nx = ny = nz = 8
    p = tmp_path / "a_0.LBMBin"
    N.write_lbmbin(p, nx, ny, nz, ["rho"], [np.zeros(512, np.float32)], 0)
    raw = p.read_bytes()
    (tmp_path / "trunc.LBMBin").write_bytes(raw[:-600])
    with pytest.raises(N.LBMError):
        N.read_lbmbin_field(tmp_path / "trunc.LBMBin", "rho")
    with pytest.raises(N.LBMError):
        N.read_lbmbin_field(p, "nope")
    (tmp_path / "nohdr.LBMBin").write_bytes(raw.replace(b"systemInformation", b"somethingElse"))
    with pytest.raises(N.LBMError):
        N.lbmbin_header(tmp_path / "nohdr.LBMBin")


def test_latest_time(tmp_path, native_lib):
    from cudalbm_amd import native as N

    shutil.copytree(os.path.join(CASES, "ComputeSanitizerTest"), tmp_path / "c", copy_function=shutil.copyfile)
    d = tmp_path / "c"
    for t in (0, 40, 200, 35):
        N.write_lbmbin(d / f"lidDrivenCavity_{t}.LBMBin", 8, 8, 8, ["rho"], [np.zeros(512, np.float32)], t)
    (d / "lidDrivenCavity_x9.LBMBin").write_bytes(b"")
    (d / "other_900.LBMBin").write_bytes(b"")
    assert N.read_case(d)["latest_time"] == 200
