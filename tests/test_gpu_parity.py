"""GPU parity of the chanflow HIP path against the CPU oracle (FP64, order 2).

Every test calls the product through the C ABI (libchanflow.so via the
chanflow ctypes mirror) and checks it against the oracle on identical
seeded inputs: bit-exact for data movement, <= 1e-12 relative L2 for
FP64 fields, divergence at round-off.
"""
import numpy as np
import pytest

import chanflow as cf
import oracle_lib as ol

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def check_log(glog, olog):
    """Per-step cfl/pstat records.  mdpdx = (Ub - U0) / U1 divides a
    cancellation (Ub - U0 ~ U1 mdpdx ~ 1e-7) by U1 ~ 5e-4, so a 1e-15
    round-off in the bulk velocity U0 moves it by ~1e-12 absolute: it is
    checked to 1e-10 absolute, the other columns to 1e-10 relative."""
    assert glog.shape == olog.shape
    assert np.array_equal(glog[:, 0], olog[:, 0])
    for col in (1, 2, 4, 5, 6, 7):
        assert np.allclose(glog[:, col], olog[:, col], rtol=1e-10, atol=1e-14), cf.LOG_COLUMNS[col]
    assert np.abs(glog[:, 3] - olog[:, 3]).max() <= 1e-10, "mdpdx"


def projected_case(nx, ny, nz, seed=1234, **kw):
    cfg = cf.Config(nx=nx, ny=ny, nz=nz, stretch=kw.pop("stretch", "tanh"), **kw)
    dz = cf.grid_dz(cfg)
    f = cf.initial_fields(cfg, seed=seed)
    orc = ol.OracleRun(ol.oracle_cfg(cfg, dz))
    for n in "uvw":
        orc.put(n, f[n])
    orc.project_ic(1.0)
    init = {n: orc.get(n) for n in "uvw"}
    return cfg, dz, orc, init, f


@pytest.mark.parametrize("shape", [(8, 6, 5), (36, 20, 41), (64, 64, 65), (96, 40, 7)])
def test_transpose_roundtrip_bit_exact(shape):
    cfg = cf.Config(nx=shape[0], ny=shape[1], nz=shape[2], stretch="uniform")
    rng = np.random.default_rng(7)
    a = rng.standard_normal(cfg.shape)
    with cf.Run(cfg) as run:
        dev, back = run.transpose_roundtrip(a)
    assert np.array_equal(dev, np.transpose(a, (2, 1, 0)))
    assert np.array_equal(back, a)


@pytest.mark.parametrize("shape", [(16, 16, 9), (24, 30, 5), (64, 96, 7), (256, 192, 3), (20, 10, 4)])
def test_fft3d_matches_numpy(shape):
    cfg = cf.Config(nx=shape[0], ny=shape[1], nz=shape[2], stretch="uniform")
    rng = np.random.default_rng(3)
    a = rng.standard_normal(cfg.shape)
    ref = np.fft.fft(np.fft.rfft(a, axis=0) / cfg.nx, axis=1) / cfg.ny   # fft.c:161-202 scaling
    with cf.Run(cfg) as run:
        spec = run.fft3d_fwd(a)
        back = run.fft3d_bwd(spec)
    assert spec.shape == ref.shape
    assert np.abs(spec - ref).max() <= 1e-14 * np.abs(ref).max() * np.log2(cfg.nx * cfg.ny)
    assert np.abs(back - a).max() <= 1e-13 * np.abs(a).max()


@pytest.mark.parametrize("dt", [1.0, 0.37e-3])
def test_poisson_matches_oracle(dt):
    cfg = cf.Config(nx=24, ny=20, nz=33, stretch="tanh")
    dz = cf.grid_dz(cfg)
    rng = np.random.default_rng(11)
    nkx = cfg.nx // 2 + 1
    spec = rng.standard_normal((nkx, cfg.ny, cfg.nz)) + 1j * rng.standard_normal((nkx, cfg.ny, cfg.nz))
    pg = ol.PaddedGrid(cfg.nx, cfg.ny, cfg.nz, cfg.lx, cfg.ly, dz)
    ng, n2 = pg.ng, pg.g.n2
    cout = np.zeros((nkx, cfg.ny, n2, 2))
    cout[:, :, ng:ng + cfg.nz, 0] = spec.real
    cout[:, :, ng:ng + cfg.nz, 1] = spec.imag
    ol.lib().orc_poisson(ol.dptr(cout), pg.ref, 2, dt)
    want = cout[:, :, ng:ng + cfg.nz, 0] + 1j * cout[:, :, ng:ng + cfg.nz, 1]
    with cf.Run(cfg) as run:
        got = run.poisson(spec, dt)
    assert rel(got, want) <= 1e-12


def test_projection_matches_oracle():
    cfg, dz, orc, init, f = projected_case(32, 24, 33)
    with cf.Run(cfg) as run:
        for n in "uvw":
            run.upload(n, f[n])
        run.project(1.0)
        for n in "uvw":
            assert rel(run.download(n), init[n]) <= 1e-12, n
        assert run.max_divergence() <= 1e-12


@pytest.mark.parametrize("shape,steps", [((16, 12, 17), 3), ((32, 24, 33), 10), ((32, 24, 129), 5),
                                         ((256, 16, 17), 3), ((64, 256, 9), 3)])
def test_step_parity_small(shape, steps):
    """Includes nz = 129, whose z solve runs as six spike chunks of 24 rows
    with a short last chunk; nx = 256 (the divergence + x FFT taking its
    first radix-4 pass in registers), and ny = 256 with nx = 64 (the 16 x 16
    two-pass y FFT, the two-pass inverse x FFT at 64)."""
    cfg, dz, orc, init, f = projected_case(*shape)
    olog = orc.steps(steps)
    with cf.Run(cfg) as run:
        for n in "uvw":
            run.upload(n, init[n])
        run.step(steps)
        run.sync()
        for n in "uvw":
            assert rel(run.download(n), orc.get(n)) <= 1e-12, n
        check_log(run.log(), olog)
        assert run.max_divergence() <= 1e-12


def test_step_parity_64x65x64_100_steps():
    """Config 1: 64 x 65 x 64 (x, wall-normal, span), nu = 1/2800, dt = 1e-3,
    constant mass flux Ub = 1, 100 RK3 steps from projected laminar + noise."""
    cfg, dz, orc, init, f = projected_case(64, 64, 65, nu=1.0 / 2800.0, dt=1e-3)
    olog = orc.steps(100)
    with cf.Run(cfg) as run:
        for n in "uvw":
            run.upload(n, init[n])
        run.step(100)
        run.sync()
        errs = {n: rel(run.download(n), orc.get(n)) for n in "uvw"}
        print("rel L2 after 100 steps:", errs)
        for n in "uvw":
            assert errs[n] <= 1e-12, (n, errs[n])
        check_log(run.log(), olog)
        div = run.max_divergence()
        umax = np.abs(orc.get("u")).max()
        assert div <= 1e-12 * umax / (cfg.lx / cfg.nx)


def test_restart_is_bit_identical():
    """RK3 is self-starting (zeta_0 = 0): stopping, downloading and
    restarting from the io-layout fields continues bit-identically."""
    cfg, dz, orc, init, f = projected_case(32, 24, 33)
    with cf.Run(cfg) as a:
        for n in "uvw":
            a.upload(n, init[n])
        a.step(6)
        full = {n: a.download(n) for n in "uvw"}
    with cf.Run(cfg) as b:
        for n in "uvw":
            b.upload(n, init[n])
        b.step(3)
        mid = {n: b.download(n) for n in "uvw"}
    with cf.Run(cfg) as c:
        for n in "uvw":
            c.upload(n, mid[n])
        c.step(3)
        for n in "uvw":
            assert np.array_equal(c.download(n), full[n]), n


def test_repeatable_bitwise():
    cfg, dz, orc, init, f = projected_case(48, 32, 33)
    outs = []
    for _ in range(2):
        with cf.Run(cfg) as r:
            for n in "uvw":
                r.upload(n, init[n])
            r.step(4)
            outs.append([r.download(n) for n in "uvw"])
    for a, b in zip(*outs):
        assert np.array_equal(a, b)
