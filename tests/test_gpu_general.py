"""GPU parity of the general path: 4th-order stencils (fourth.c,
heptadiagonal implicit z solves) and the passive scalar (QUICK, scalar-mean
constraint, buoyancy), against the CPU oracle through the C ABI.

FP64 fields to <= 1e-12 relative L2, per-step logs (cfl, pstat, cstat) to
1e-10, divergence at round-off.
"""
import numpy as np
import pytest

import chanflow as cf
import oracle_lib as ol

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def check_log(glog, olog, scalar):
    assert glog.shape == olog.shape
    assert np.array_equal(glog[:, 0], olog[:, 0])
    cols = [1, 2, 4, 5, 6, 7] + ([9, 10, 11, 12] if scalar else [])
    for col in cols:
        assert np.allclose(glog[:, col], olog[:, col], rtol=1e-10, atol=1e-13), cf.LOG_COLUMNS[col]
    # forcings divide a cancellation by the unit response (see test_gpu_parity.check_log)
    assert np.abs(glog[:, 3] - olog[:, 3]).max() <= 1e-10, "mdpdx"
    if scalar:
        assert np.abs(glog[:, 8] - olog[:, 8]).max() <= 1e-10, "dcdt"


def general_case(nx, ny, nz, order, scalar, seed=1234, **kw):
    cfg = cf.Config(nx=nx, ny=ny, nz=nz, order=order, scalar=scalar, stretch=kw.pop("stretch", "tanh"),
                    prandtl=kw.pop("prandtl", 0.71), **kw)
    dz = cf.grid_dz(cfg)
    f = cf.initial_fields(cfg, seed=seed)
    names = "uvwc" if scalar else "uvw"
    orc = ol.OracleRun(ol.oracle_cfg(cfg, dz))
    for n in names:
        orc.put(n, f[n])
    orc.project_ic(1.0)
    init = {n: orc.get(n) for n in names}
    return cfg, orc, init, f, names


def run_and_compare(cfg, orc, init, names, steps, tol=1e-12):
    olog = orc.steps(steps)
    with cf.Run(cfg) as run:
        for n in names:
            run.upload(n, init[n])
        run.step(steps)
        run.sync()
        errs = {n: rel(run.download(n), orc.get(n)) for n in names}
        print(f"{cfg.shape} order {cfg.order} scalar {cfg.scalar}: rel L2 after {steps} steps {errs}")
        for n in names:
            assert errs[n] <= tol, (n, errs[n])
        check_log(run.log(), olog, cfg.scalar)
        umax = np.abs(orc.get("u")).max()
        assert run.max_divergence() <= 1e-12 * umax / (cfg.lx / cfg.nx)


@pytest.mark.parametrize("dt", [1.0, 0.37e-3])
def test_poisson_order4_matches_oracle(dt):
    cfg = cf.Config(nx=24, ny=20, nz=33, order=4, stretch="tanh")
    dz = cf.grid_dz(cfg)
    rng = np.random.default_rng(11)
    nkx = cfg.nx // 2 + 1
    spec = rng.standard_normal((nkx, cfg.ny, cfg.nz)) + 1j * rng.standard_normal((nkx, cfg.ny, cfg.nz))
    pg = ol.PaddedGrid(cfg.nx, cfg.ny, cfg.nz, cfg.lx, cfg.ly, dz)
    ng, n2 = pg.ng, pg.g.n2
    cout = np.zeros((nkx, cfg.ny, n2, 2))
    cout[:, :, ng:ng + cfg.nz, 0] = spec.real
    cout[:, :, ng:ng + cfg.nz, 1] = spec.imag
    ol.lib().orc_poisson(ol.dptr(cout), pg.ref, 4, dt)
    want = cout[:, :, ng:ng + cfg.nz, 0] + 1j * cout[:, :, ng:ng + cfg.nz, 1]
    with cf.Run(cfg) as run:
        got = run.poisson(spec, dt)
    assert rel(got, want) <= 1e-12


def test_projection_order4_matches_oracle():
    cfg, orc, init, f, names = general_case(32, 24, 33, 4, 0)
    with cf.Run(cfg) as run:
        for n in "uvw":
            run.upload(n, f[n])
        run.project(1.0)
        for n in "uvw":
            assert rel(run.download(n), init[n]) <= 1e-12, n
        assert run.max_divergence() <= 1e-12


@pytest.mark.parametrize("order,scalar,shape,steps", [
    (4, 0, (16, 12, 17), 3),
    (2, 1, (16, 12, 17), 3),
    (4, 1, (16, 12, 17), 3),
    (4, 1, (32, 24, 33), 10),
    (2, 1, (20, 10, 9), 6),
])
def test_general_step_parity_small(order, scalar, shape, steps):
    cfg, orc, init, f, names = general_case(*shape, order, scalar, nu=1.0 / 3300.0, dt=5e-4)
    run_and_compare(cfg, orc, init, names, steps)


def test_order4_scalar_100_steps():
    """The 4th-order + scalar scheme of the 192 x 129 x 192 workload (Pr 0.71,
    nu 1/3300, dt 5e-4, walls theta = 0 / 1 as c = -1/2 / +1/2) on a
    48 x 33 x 32 grid the oracle steps in seconds: 100 RK3 steps."""
    cfg, orc, init, f, names = general_case(48, 32, 33, 4, 1, nu=1.0 / 3300.0, dt=5e-4)
    run_and_compare(cfg, orc, init, names, 100)


def test_scalar_mean_constraint_and_buoyancy():
    """cvolavgmode = 1 holds the bulk scalar (chanfast.c:572-617); Boussinesq
    forcing on u, v, w (scalar.c:6-76)."""
    for order in (2, 4):
        cfg, orc, init, f, names = general_case(16, 12, 17, order, 1, nu=1.0 / 2000.0, dt=5e-4,
                                                cvolavgmode=1, want_cvolavg=0.1,
                                                betg_x=0.3, betg_y=-0.2, betg_z=0.5)
        run_and_compare(cfg, orc, init, names, 5)


def test_general_restart_is_bit_identical():
    cfg, orc, init, f, names = general_case(32, 24, 33, 4, 1, nu=1.0 / 3300.0, dt=5e-4)
    with cf.Run(cfg) as a:
        for n in names:
            a.upload(n, init[n])
        a.step(6)
        full = {n: a.download(n) for n in names}
    with cf.Run(cfg) as b:
        for n in names:
            b.upload(n, init[n])
        b.step(3)
        mid = {n: b.download(n) for n in names}
    with cf.Run(cfg) as c:
        for n in names:
            c.upload(n, mid[n])
        c.step(3)
        for n in names:
            assert np.array_equal(c.download(n), full[n]), n


def test_order4_scalar_192x129x192_two_steps():
    """Full-size 4th-order + scalar workload (192 x 129 x 192): two RK3
    steps against the oracle, then 40 more on the GPU alone with the
    divergence at round-off, the bulk velocity held and the scalar inside
    its wall values."""
    cfg, orc, init, f, names = general_case(192, 192, 129, 4, 1, nu=1.0 / 3300.0, dt=5e-4)
    olog = orc.steps(2)
    with cf.Run(cfg) as run:
        for n in names:
            run.upload(n, init[n])
        run.step(2)
        run.sync()
        for n in names:
            e = rel(run.download(n), orc.get(n))
            print(n, e)
            assert e <= 1e-12, (n, e)
        check_log(run.log(), olog, 1)
        orc.close()
        run.step(40)
        run.sync()
        lg = run.log()
        assert np.allclose(lg[:, 6], 1.0, rtol=0, atol=1e-12)
        assert run.max_divergence() <= 1e-11
        c = run.download("c")
        assert c.min() > -0.5 - 0.1 and c.max() < 0.5 + 0.1
