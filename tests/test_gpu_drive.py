"""The reference-format driver (chanflow_drive) on the GPU: reads
u/v/w_it<it>.dat, steps to nt, writes the field dumps every istat steps and
the cfl.dat / pstat.dat records (chanfast.c:392-393, :526-529, :658-877);
its output equals the same run through the library."""
import os
import subprocess

import numpy as np
import pytest

import chanflow as cf

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVE = os.path.join(ROOT, "ubiquitous-spoon_amd", "chanflow_drive")


def test_driver_matches_library(tmp_path):
    cfg = cf.Config(nx=32, ny=24, nz=33, stretch="tanh", nu=1.0 / 2800.0, dt=1e-3)
    f = cf.initial_fields(cfg, seed=11)
    with cf.Run(cfg) as r:
        for n in "uvw":
            r.upload(n, f[n])
        r.project(1.0)
        init = {n: r.download(n) for n in "uvw"}
    d = str(tmp_path)
    for n in "uvw":
        cf.io_write(os.path.join(d, f"{n}_it100.dat"), init[n])
    subprocess.run([DRIVE, "--dir", d, "--nx", "32", "--ny", "24", "--nz", "33", "--it", "100", "--nt", "130",
                    "--istat", "10", "--stretch", "tanh", "--nu", str(1.0 / 2800.0), "--dt", "1e-3"], check=True)
    with cf.Run(cfg) as r:
        for n in "uvw":
            r.upload(n, init[n])
        r.step(30)
        r.sync()
        want = {n: r.download(n) for n in "uvw"}
        lg = r.log()
    for it in (110, 120, 130):
        for n in "uvw":
            assert os.path.exists(os.path.join(d, f"{n}_it{it}.dat"))
    for n in "uvw":
        got = cf.io_read(os.path.join(d, f"{n}_it130.dat"), cfg.shape)
        assert np.array_equal(got, want[n]), n
    cfl = np.loadtxt(os.path.join(d, "cfl.dat"))
    pst = np.loadtxt(os.path.join(d, "pstat.dat"))
    assert cfl.shape == (30, 3) and pst.shape == (30, 6)
    assert np.array_equal(cfl[:, 0], np.arange(100, 130))
    assert np.array_equal(pst[:, 0], np.arange(101, 131))
    # "%+e" keeps 7 significant digits
    assert np.allclose(cfl[:, 2], lg[:, 2], rtol=1e-6, atol=0)
    assert np.allclose(pst[:, 4], lg[:, 6], rtol=1e-6, atol=0)


def test_driver_reference_main_run_writes_statistics(tmp_path):
    """The reference's production parameter set (main.c:45-61: chanfast4
    with the scalar, cflmode 1 at want_cfl 1, constant bulk u in the frame
    moving with it, constant bulk c, dpdx = -1, nu = 1/180) through the
    driver on a small grid, with the istat block of chanfast.c:658-856:
    every statistics file is written with the reference's n2 = nz + 6
    entries, the averages and correlations equal the library's statistics
    of the final state, and zp.dat / zw.dat are the padded-grid coordinates
    of get.c:458-488."""
    nx, ny, nz = 24, 16, 25
    kw = dict(nx=nx, ny=ny, nz=nz, order=4, scalar=1, stretch="cosine", lx=4 * np.pi, ly=4.0 / 3.0 * np.pi, lz=2.0,
              dt=2e-3, cflmode=1, want_cfl=1.0, nu=1.0 / 180.0, prandtl=1.0, dpdx=-1.0, want_uvolavg=1.0,
              uvolavgmode=1, want_cvolavg=0.0, cvolavgmode=1, uframe=1.0)
    cfg = cf.Config(**kw)
    f = cf.initial_fields(cfg, seed=21)
    with cf.Run(cfg) as r:
        for n in "uvwc":
            r.upload(n, f[n])
        r.project(1.0)
        init = {n: r.download(n) for n in "uvwc"}
    d = str(tmp_path)
    for n in "uvwc":
        cf.io_write(os.path.join(d, f"{n}_it0.dat"), init[n])
    args = [DRIVE, "--dir", d, "--nx", str(nx), "--ny", str(ny), "--nz", str(nz), "--it", "0", "--nt", "20",
            "--istat", "20", "--order", "4", "--scalar", "--stretch", "cosine", "--lx", repr(4 * np.pi),
            "--ly", repr(4.0 / 3.0 * np.pi), "--dt", "2e-3", "--cflmode", "1", "--want-cfl", "1.0",
            "--nu", repr(1.0 / 180.0), "--prandtl", "1.0", "--dpdx", "-1.0", "--ubulk", "1.0",
            "--uvolavgmode", "1", "--cbulk", "0.0", "--cvolavgmode", "1", "--uframe", "1.0"]
    subprocess.run(args, check=True)
    with cf.Run(cfg) as r:
        for n in "uvwc":
            r.upload(n, init[n])
        r.step(20)
        r.sync()
        lg = r.log()
        for n in "uvwc":
            got = cf.io_read(os.path.join(d, f"{n}_it20.dat"), cfg.shape)
            assert np.array_equal(got, r.download(n)), n
        ng, n2 = 3, nz + 6
        for name, kind, a, b in (("u_avg", "avg", "u", "u"), ("p_avg", "avg", "p", "p"), ("c_avg", "avg", "c", "c"),
                                 ("wu_corr", "corr", "w", "u"), ("wc_corr", "corr", "w", "c"),
                                 ("pp_corr", "corr", "p", "p")):
            prof = np.loadtxt(os.path.join(d, f"{name}_it20.dat"))
            assert prof.shape == (n2,), name
            want = r.xy_stats(kind, a, b)
            assert np.allclose(prof[ng:ng + nz], want, rtol=1e-6, atol=1e-12 * np.abs(want).max()), name
            assert np.all(prof[:ng] == 0) and np.all(prof[ng + nz:] == 0), name
        for m in (3, 4):
            prof = np.loadtxt(os.path.join(d, f"u_mom{m}_it20.dat"))
            want = r.xy_stats("moment", "u", en=float(m))
            assert np.allclose(prof[ng:ng + nz], want, rtol=1e-6, atol=1e-12 * np.abs(want).max())
        for pair in ("uu", "vv", "ww", "cc", "pp", "wu", "uc", "wc"):
            sx = np.loadtxt(os.path.join(d, f"{pair}_specx_it20.dat"))
            sy = np.loadtxt(os.path.join(d, f"{pair}_specy_it20.dat"))
            assert sx.shape == ((nx // 2 + 1) * n2 * 2,) and sy.shape == ((ny // 2 + 1) * n2 * 2,), pair
    # cfl.dat: dt rescaled every step, cfl logged as want_cfl (chanfast.c:385-393)
    cfl = np.loadtxt(os.path.join(d, "cfl.dat"))
    assert cfl.shape == (20, 3) and np.all(cfl[:, 2] == 1.0)
    assert np.allclose(cfl[:, 1], lg[:, 1], rtol=1e-6, atol=0) and len(np.unique(lg[:, 1])) > 10
    # zp / zw: get_z_at_p / get_z_at_w on the mirrored padded dz
    dz = cf.grid_dz(cfg)
    dzp = np.concatenate([dz[:ng][::-1], dz, dz[-ng:][::-1]])
    zw = np.cumsum(dzp)
    zw -= zw[ng - 1]
    zp = np.concatenate([[0.5 * dzp[0]], 0.5 * dzp[0] + np.cumsum(0.5 * (dzp[:-1] + dzp[1:]))])
    zp -= 0.5 * (zp[ng - 1] + zp[ng])
    assert np.allclose(np.loadtxt(os.path.join(d, "zp.dat")), zp, rtol=1e-6, atol=1e-12)
    assert np.allclose(np.loadtxt(os.path.join(d, "zw.dat")), zw, rtol=1e-6, atol=1e-12)
