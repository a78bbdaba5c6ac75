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
