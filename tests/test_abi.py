"""C-ABI checks that need no GPU: every entry point declared in
include/*.h is exported by libchanflow.so, host-only helpers, the
reference's file formats and index maps, and argument validation before any
device use."""
import os
import re
import subprocess

import numpy as np
import pytest

import chanflow as cf

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def declared_functions():
    """Every function declared in the public headers: a return type at the
    start of a declaration, then the name and its argument list."""
    names = set()
    for h in sorted(os.listdir(INCLUDE)):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(INCLUDE, h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"^\s*#.*$", "", text, flags=re.M)
        names |= set(re.findall(r"(?:^|;|})\s*(?:const\s+)?(?:void|int|double|cf_status|char|decomp_plan)\s*\**\s*"
                                r"([A-Za-z_][A-Za-z0-9_]*)\s*\(", text, flags=re.M))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 60
    for must in ("cf_run_create", "divergence4", "solve_helmholtz1d_w4", "pressure_poisson2", "create_decomp_plan",
                 "cf_pencil_gather", "cf_server_request_state"):
        assert must in names, must
    out = subprocess.run(["nm", "-D", "--defined-only", cf.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT ([A-Za-z_][A-Za-z0-9_]*)", out))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    lib = cf.lib()
    for n in names:
        assert hasattr(lib, n)


def test_version_string():
    assert "gfx950" in cf.version()


def test_cosine_grid_matches_reference_formula():
    """chanfast.c:184-189."""
    cfg = cf.Config(nz=65, lz=2.0, stretch="cosine")
    dz = cf.grid_dz(cfg)
    iz = np.arange(65)
    ref = 2.0 * 0.5 * (np.cos(np.pi * iz / 65) - np.cos(np.pi * (iz + 1) / 65))
    assert np.allclose(dz, ref, rtol=1e-15, atol=0)
    assert abs(dz.sum() - 2.0) <= 1e-14


@pytest.mark.parametrize("nz", [65, 129, 193])
def test_tanh_grid(nz):
    cfg = cf.Config(nz=nz, lz=2.0, stretch="tanh", stretch_beta=2.0)
    dz = cf.grid_dz(cfg)
    assert (dz > 0).all()
    assert abs(dz.sum() - 2.0) <= 1e-14
    assert np.allclose(dz, dz[::-1], rtol=1e-13, atol=0)          # symmetric about the centreline
    j = np.arange(nz + 1)
    zf = 1.0 + np.tanh(2.0 * (2.0 * j / nz - 1.0)) / np.tanh(2.0)
    assert np.allclose(dz, np.diff(zf), rtol=1e-12, atol=1e-15)


def test_initial_fields_deterministic_and_shaped():
    cfg = cf.Config(nx=16, ny=12, nz=17, stretch="tanh")
    a = cf.initial_fields(cfg, seed=1234)
    b = cf.initial_fields(cfg, seed=1234)
    c = cf.initial_fields(cfg, seed=1235)
    for n in "uvwc":
        assert a[n].shape == cfg.shape
        assert np.array_equal(a[n], b[n])
        assert not np.array_equal(a[n], c[n])
    assert np.all(a["w"][:, :, -1] == 0.0)                         # top-wall face
    assert np.abs(a["v"]).max() <= 0.1 and np.abs(a["w"]).max() <= 0.1
    dz = cf.grid_dz(cfg)
    zc = np.cumsum(dz) - 0.5 * dz
    lam = 1.5 * (1 - (zc - 1.0) ** 2)
    assert np.abs(a["u"] - lam[None, None, :]).max() <= 0.1 + 1e-15
    assert abs((a["u"] - lam[None, None, :]).mean()) < 0.01        # zero-mean noise


def test_io_roundtrip_bit_exact_case14(tmp_path):
    """case14 (cases.c:4139-4290): io_write -> io_read returns the same
    values; the file is raw native doubles in C order [nx][ny][nz]."""
    rng = np.random.default_rng(0)
    a = rng.standard_normal((8, 6, 5))
    path = str(tmp_path / "u_it0.dat")
    cf.io_write(path, a)
    assert os.path.getsize(path) == a.size * 8
    assert np.array_equal(np.fromfile(path, dtype=np.float64).reshape(a.shape), a)
    assert np.array_equal(cf.io_read(path, a.shape), a)


def test_append_ascii_format(tmp_path):
    """io.c:5-20: "%+e" values separated by one space, one line per call."""
    path = str(tmp_path / "cfl.dat")
    vals = np.array([1.0, 1e-3, -0.25])
    lib = cf.lib()
    assert lib.cf_io_append_ascii(path.encode(), cf._dptr(vals), 3) == 0
    assert lib.cf_io_append_ascii(path.encode(), cf._dptr(vals), 3) == 0
    lines = open(path).read().splitlines()
    assert lines == ["+1.000000e+00 +1.000000e-03 -2.500000e-01"] * 2


@pytest.mark.parametrize("bad,code", [
    (dict(nx=1), cf.CF_ERR_ARG),
    (dict(order=3), cf.CF_ERR_ARG),
    (dict(nu=-1.0), cf.CF_ERR_ARG),
    (dict(betg_z=1.0), cf.CF_ERR_ARG),
    (dict(nx=14), cf.CF_ERR_UNSUPPORTED),        # 7 is not a built FFT radix
])
def test_create_rejects_bad_config_before_device_use(bad, code):
    cfg = cf.Config(nx=16, ny=12, nz=17)
    for k, v in bad.items():
        setattr(cfg, k, v)
    with pytest.raises(cf.ChanflowError) as e:
        cf.Run(cfg)
    assert e.value.status == code


def test_footprint_scales_with_grid():
    small = cf.footprint(cf.Config(nx=64, ny=64, nz=65))
    big = cf.footprint(cf.Config(nx=256, ny=256, nz=193))
    assert big > 40 * (small - (16 << 20)) and big < 400e9


@pytest.mark.parametrize("ng", [1, 3])
def test_grid_z_closed_form(ng):
    """zp / zw of get_z_at_p / get_z_at_w (get.c:458-488): on the cosine grid
    the faces are lz/2 (1 - cos(pi k / nz)) (chanfast.c:184-189), measured
    from the bottom wall; cell centres sit midway between faces; the
    mirrored ghost levels are odd about each wall."""
    nz = 17
    cfg = cf.Config(nx=8, ny=8, nz=nz, stretch="cosine", lz=2.0)
    zp, zw = cf.grid_z(cfg, ng)
    faces = 1.0 - np.cos(np.pi * np.arange(nz + 1) / nz)
    assert np.allclose(zw[ng - 1:ng + nz], faces, rtol=0, atol=1e-14)
    assert np.allclose(zp[ng:ng + nz], 0.5 * (faces[:-1] + faces[1:]), rtol=0, atol=1e-14)
    for m in range(ng):
        assert abs(zp[ng - 1 - m] + zp[ng + m]) <= 1e-14
        assert abs((zp[ng + nz + m] - 2.0) + (zp[ng + nz - 1 - m] - 2.0)) <= 1e-14
    for m in range(1, ng):
        assert abs(zw[ng - 1 - m] + zw[ng - 1 + m]) <= 1e-14
    with pytest.raises(cf.ChanflowError):
        cf.grid_z(cfg, 0)


def test_write_ascii_format(tmp_path):
    """io_write_ascii (io.c:22-34): one "%+e" value per line."""
    p = str(tmp_path / "prof.dat")
    vals = np.array([1.5, -2.25e-7, 0.0])
    assert cf.lib().cf_io_write_ascii(p.encode(), vals.ctypes.data_as(cf.C.POINTER(cf.C.c_double)), 3) == 0
    assert open(p).read() == "+1.500000e+00\n-2.250000e-07\n+0.000000e+00\n"
