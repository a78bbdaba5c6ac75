"""GPU tests of uneven batches: independent runs of different grids and
viscosities stepped together in persistent launches (cf_batch_*).

Each run in a batch keeps the semantics of the same run on its own: FP64
runs match the oracle to <= 1e-12 relative L2 and the per-step logs to
1e-10; a run gives the same fields alone and in a mix (<= 1e-13 relative,
reported when bit-identical); FP32 runs stay within the FP32 form's 5e-5.
"""
import numpy as np
import pytest

import chanflow as cf
import oracle_lib as ol

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def projected(cfg, seed):
    dz = cf.grid_dz(cfg)
    f = cf.initial_fields(cfg, seed=seed)
    names = "uvwc" if cfg.scalar else "uvw"
    orc = ol.OracleRun(ol.oracle_cfg(cfg, dz))
    for n in names:
        orc.put(n, f[n])
    orc.project_ic(1.0)
    return orc, {n: orc.get(n) for n in names}, names


def check_batch_vs_oracle(cfgs, steps, tol=1e-12, seeds=None):
    cases = [projected(c, 100 + q if seeds is None else seeds[q]) for q, c in enumerate(cfgs)]
    with cf.Batch(cfgs) as b:
        for q, (orc, init, names) in enumerate(cases):
            for n in names:
                b.upload(q, n, init[n])
        b.step(steps)
        b.sync()
        print("autotuner:", b.tuning())
        for q, (orc, init, names) in enumerate(cases):
            olog = orc.steps(steps)
            errs = {n: rel(b.download(q, n), orc.get(n)) for n in names}
            print(cfgs[q].shape, "order", cfgs[q].order, "prec", cfgs[q].precision, errs)
            for n in names:
                assert errs[n] <= tol, (q, n, errs[n])
            glog = b.log(q)
            assert np.array_equal(glog[:, 0], olog[:, 0])
            rt = 1e-10 if tol < 1e-9 else 1e-4
            for col in (1, 2, 6):
                assert np.allclose(glog[:, col], olog[:, col], rtol=rt, atol=1e-13), (q, cf.LOG_COLUMNS[col])
            umax = np.abs(orc.get("u")).max()
            assert b.max_divergence(q) <= (1e-12 if tol < 1e-9 else 1e-5) * umax / (cfgs[q].lx / cfgs[q].nx)


def test_uneven_batch_order2_matches_oracle():
    cfgs = [cf.Config(nx=16, ny=12, nz=17, stretch="tanh", nu=1.0 / 2000.0, dt=5e-4),
            cf.Config(nx=20, ny=10, nz=9, stretch="tanh", nu=1.0 / 4000.0, dt=5e-4),
            cf.Config(nx=32, ny=24, nz=33, stretch="tanh", nu=1.0 / 6000.0, dt=5e-4)]
    check_batch_vs_oracle(cfgs, 10)


def test_uneven_batch_order4_scalar_matches_oracle():
    cfgs = [cf.Config(nx=16, ny=12, nz=17, order=4, scalar=1, prandtl=0.71, stretch="tanh", nu=1.0 / 3300.0, dt=5e-4),
            cf.Config(nx=24, ny=20, nz=33, order=4, scalar=1, prandtl=0.71, stretch="tanh", nu=1.0 / 2500.0, dt=5e-4)]
    check_batch_vs_oracle(cfgs, 8)


def test_uneven_batch_fp32_against_fp64_oracle():
    cfgs = [cf.Config(nx=32, ny=24, nz=33, stretch="tanh", nu=1.0 / 2800.0, dt=1e-3, precision=cf.CF_FP32),
            cf.Config(nx=48, ny=32, nz=33, stretch="tanh", nu=1.0 / 2000.0, dt=1e-3, precision=cf.CF_FP32)]
    check_batch_vs_oracle(cfgs, 30, tol=5e-5)


@pytest.mark.parametrize("prec", [cf.CF_FP64, cf.CF_FP32])
def test_same_alone_and_in_the_mix(prec):
    mine = cf.Config(nx=32, ny=24, nz=33, stretch="tanh", nu=1.0 / 3000.0, dt=5e-4, precision=prec)
    others = [cf.Config(nx=64, ny=48, nz=65, stretch="tanh", nu=1.0 / 5000.0, dt=5e-4, precision=prec),
              cf.Config(nx=16, ny=12, nz=17, stretch="tanh", nu=1.0 / 2000.0, dt=5e-4, precision=prec)]
    f = cf.initial_fields(mine, seed=9)
    res = []
    for cfgs, idx in (([mine], 0), ([others[0], mine, others[1]], 1)):
        with cf.Batch(cfgs) as b:
            for q, c in enumerate(cfgs):
                g = f if q == idx else cf.initial_fields(c, seed=50 + q)
                for n in "uvw":
                    b.upload(q, n, g[n])
            b.project(1.0)
            b.step(12)
            b.sync()
            res.append(({n: b.download(idx, n) for n in "uvw"}, b.log(idx)))
    (a, la), (m, lm) = res
    for n in "uvw":
        print(n, "alone vs in the mix: rel L2", rel(m[n], a[n]), "bit-identical", np.array_equal(a[n], m[n]))
    for n in "uvw":
        assert np.array_equal(a[n], m[n]), n
    assert np.array_equal(la, lm)


def test_fp32_batch_matches_fp32_single_runs():
    """FP32 order-2 runs of a batch take the batched fast path; each matches
    the same run stepped alone on the single-run FP32 fast path to 1e-5
    relative L2 after 15 steps.  Not bit for bit in general: the two may use
    different FFT instantiations (a batch compiles fewer lengths in; nx = 16
    here takes the runtime plan), which round differently in FP32 (about
    1e-6 after 15 steps; the FP64 instantiations agree to 1e-12)."""
    cfgs = [cf.Config(nx=32, ny=24, nz=33, stretch="tanh", nu=1.0 / 2800.0, dt=1e-3, precision=cf.CF_FP32),
            cf.Config(nx=48, ny=32, nz=49, stretch="tanh", nu=1.0 / 2000.0, dt=1e-3, precision=cf.CF_FP32),
            cf.Config(nx=16, ny=12, nz=17, stretch="tanh", nu=1.0 / 4000.0, dt=1e-3, precision=cf.CF_FP32)]
    fields = [cf.initial_fields(c, seed=70 + q) for q, c in enumerate(cfgs)]
    with cf.Batch(cfgs) as b:
        assert any("fast path" in t for t in b.tuning())
        for q in range(len(cfgs)):
            for n in "uvw":
                b.upload(q, n, fields[q][n])
        b.project(1.0)
        b.step(15)
        b.sync()
        got = [{n: b.download(q, n) for n in "uvw"} for q in range(len(cfgs))]
    for q, c in enumerate(cfgs):
        with cf.Run(c) as r:
            for n in "uvw":
                r.upload(n, fields[q][n])
            r.project(1.0)
            r.step(15)
            r.sync()
            for n in "uvw":
                e = rel(got[q][n], r.download(n))
                print("run", q, c.shape, n, e)
                assert e <= 1e-5, (q, n, e)


def test_batch_48_runs_of_three_sizes():
    """The uneven-batch workload: 16 runs each of 64x65x64, 96x97x96 and
    128x129x128 (x, wall-normal, span), nu seeded uniform in [1/6000,
    1/2000], dt = 5e-4, order 2 FP64, 100 steps together.  Every run keeps
    its bulk velocity and divergence at round-off; sampled runs match the
    same run stepped alone on the order-2 fast path to <= 1e-12."""
    rng = np.random.default_rng(2024)
    cfgs = []
    for n in (64, 96, 128):
        for _ in range(16):
            cfgs.append(cf.Config(nx=n, ny=n, nz=n + 1, stretch="tanh", nu=float(rng.uniform(1 / 6000, 1 / 2000)),
                                  dt=5e-4))
    fields = [cf.initial_fields(c, seed=300 + q) for q, c in enumerate(cfgs)]
    with cf.Batch(cfgs) as b:
        for q in range(len(cfgs)):
            for n in "uvw":
                b.upload(q, n, fields[q][n])
        b.project(1.0)
        init = {q: {n: b.download(q, n) for n in "uvw"} for q in (0, 17, 47)}
        b.step(100)
        b.sync()
        for q in range(len(cfgs)):
            lg = b.log(q)
            assert lg.shape[0] == 100
            assert np.allclose(lg[:, 6], 1.0, rtol=0, atol=1e-12), q
            assert b.max_divergence(q) <= 1e-11, q
        out = {q: {n: b.download(q, n) for n in "uvw"} for q in init}
    for q in init:
        with cf.Run(cfgs[q]) as r:
            for n in "uvw":
                r.upload(n, init[q][n])
            r.step(100)
            r.sync()
            for n in "uvw":
                e = rel(out[q][n], r.download(n))
                print("run", q, cfgs[q].shape, n, e)
                assert e <= 1e-12, (q, n, e)


def test_order2_scalar_batch_solves_each_run_as_alone():
    """An order-2 scalar batch holding a run whose z systems the spike
    covers (nz = 33) and one it does not (nz = 300 > 256 rows: Thomas):
    each run keeps its own solve, so it matches the oracle and the same run
    stepped alone bit for bit."""
    cfgs = [cf.Config(nx=8, ny=8, nz=33, scalar=1, prandtl=0.71, stretch="tanh", nu=1.0 / 2000.0, dt=5e-4),
            cf.Config(nx=8, ny=8, nz=300, scalar=1, prandtl=0.71, stretch="tanh", nu=1.0 / 2500.0, dt=5e-4)]
    check_batch_vs_oracle(cfgs, 6)
    cases = [projected(c, 100 + q) for q, c in enumerate(cfgs)]
    alone = []
    for c, (orc, init, names) in zip(cfgs, cases):
        with cf.Run(c) as r:
            for n in names:
                r.upload(n, init[n])
            r.step(6)
            r.sync()
            alone.append({n: r.download(n) for n in names})
    with cf.Batch(cfgs) as b:
        for q, (orc, init, names) in enumerate(cases):
            for n in names:
                b.upload(q, n, init[n])
        b.step(6)
        b.sync()
        for q, (orc, init, names) in enumerate(cases):
            for n in names:
                assert np.array_equal(b.download(q, n), alone[q][n]), (q, n)
