"""GPU tests of the FP32 precision form: stencils, FFTs and implicit solves
in FP32 storage and arithmetic, with FP64 kept for the Poisson modes with
|kx|, |ky| <= 4, the plane sums and mass-flux / scalar-mean constraints and
the divergence check.  Errors are stated against the FP64 oracle (small
grids) or the FP64 GPU path (the 256 x 193 x 256 workload).

Tolerances (FP32 storage rounds every substep, so errors grow about
linearly with the step count; measured on MI355X: u 4e-6..2e-5, v, w
1e-5..3e-5 relative to their smaller norms, p 4e-3, wall shear 3e-4 after
200 steps of the 256 x 193 x 256 workload): per-field relative L2 <= 5e-5,
pressure <= 1e-2, wall shear <= 1e-3 relative, divergence at FP32 round-off
(<= 1e-5 max|u| / dx).
"""
import numpy as np
import pytest

import chanflow as cf
import oracle_lib as ol

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def fp_case(nx, ny, nz, order, scalar, steps, **kw):
    cfg64 = cf.Config(nx=nx, ny=ny, nz=nz, order=order, scalar=scalar, stretch="tanh", prandtl=0.71, **kw)
    dz = cf.grid_dz(cfg64)
    f = cf.initial_fields(cfg64, seed=4321)
    names = "uvwc" if scalar else "uvw"
    orc = ol.OracleRun(ol.oracle_cfg(cfg64, dz))
    for n in names:
        orc.put(n, f[n])
    orc.project_ic(1.0)
    init = {n: orc.get(n) for n in names}
    olog = orc.steps(steps)
    want = {n: orc.get(n) for n in names}
    cfg32 = cf.Config(**{**cfg64.__dict__, "precision": cf.CF_FP32})
    with cf.Run(cfg32) as run:
        for n in names:
            run.upload(n, init[n])
        run.step(steps)
        run.sync()
        got = {n: run.download(n) for n in names}
        glog = run.log()
        div = run.max_divergence()
    return cfg64, got, want, glog, olog, div


@pytest.mark.parametrize("order,scalar,shape,steps", [
    (2, 0, (64, 64, 65), 100),
    (4, 0, (32, 24, 33), 50),
    (2, 1, (32, 24, 33), 50),
    (4, 1, (48, 32, 33), 50),
])
def test_fp32_against_fp64_oracle(order, scalar, shape, steps):
    cfg, got, want, glog, olog, div = fp_case(*shape, order, scalar, steps, nu=1.0 / 2800.0, dt=1e-3)
    errs = {n: rel(got[n], want[n]) for n in got}
    print(f"FP32 {shape} order {order} scalar {scalar}, {steps} steps: rel L2 {errs}")
    for n, e in errs.items():
        assert e <= 5e-5, (n, e)
    for col in (4, 5):
        assert np.allclose(glog[:, col], olog[:, col], rtol=1e-3, atol=1e-8), cf.LOG_COLUMNS[col]
    assert np.allclose(glog[:, 6], olog[:, 6], rtol=1e-5, atol=1e-8), "ubulk"
    assert np.allclose(glog[:, 7], olog[:, 7], rtol=5e-5, atol=1e-8), "ucl"
    umax = np.abs(want["u"]).max()
    assert div <= 1e-5 * umax / (cfg.lx / cfg.nx)


def test_fp32_uploads_round_to_nearest():
    """v and w are stored rounded to nearest FP32; u is stored as its
    deviation from its FP64 plane-mean profile, rounded to nearest, so a u
    with a large mean keeps far more digits than FP32 rounding of u itself."""
    cfg = cf.Config(nx=16, ny=12, nz=17, stretch="tanh", precision=cf.CF_FP32)
    rng = np.random.default_rng(2)
    a = rng.standard_normal(cfg.shape)
    u = 1.5 + np.linspace(0.0, 0.3, cfg.nz)[None, None, :] + 0.01 * rng.standard_normal(cfg.shape)
    with cf.Run(cfg) as run:
        run.upload("v", a)
        assert np.array_equal(run.download("v"), a.astype(np.float32).astype(np.float64))
        run.upload("u", u)
        back = run.download("u")
    dev = u - u.mean(axis=(0, 1))[None, None, :]
    # at most half an FP32 ulp of the deviation (plus the profile's own FP64 rounding)
    assert np.all(np.abs(back - u) <= np.abs(dev) * 2.0 ** -24 + 1e-15)
    plain = np.abs(u.astype(np.float32).astype(np.float64) - u).max()
    assert np.abs(back - u).max() < plain / 20


def test_fp32_form_vs_fp64_256x193x256_200_steps():
    """The precision-form workload: 256 x 193 x 256, order 2, nu = 1/6900,
    dt = 2.5e-4, 200 steps, FP32 against the FP64 path on the same GPU
    (itself held to the oracle at 1e-12): per-field relative L2 of u, v,
    w, p and the wall-shear error of both walls."""
    base = dict(nx=256, ny=256, nz=193, order=2, stretch="tanh", stretch_beta=2.0, nu=1.0 / 6900.0, dt=2.5e-4)
    cfg64 = cf.Config(**base)
    f = cf.initial_fields(cfg64, seed=1234)
    res = {}
    for prec in (cf.CF_FP64, cf.CF_FP32):
        cfg = cf.Config(**base, precision=prec)
        with cf.Run(cfg) as run:
            for n in "uvw":
                run.upload(n, f[n])
            run.project(1.0)
            run.step(200)
            run.sync()
            res[prec] = ({n: run.download(n) for n in "uvwp"}, run.log())
    (a, la), (b, lb) = res[cf.CF_FP64], res[cf.CF_FP32]
    errs = {n: rel(b[n], a[n]) for n in "uvwp"}
    shear = {w: float(np.max(np.abs(lb[:, c] - la[:, c]) / np.abs(la[:, c]))) for w, c in (("bottom", 4), ("top", 5))}
    print("FP32 vs FP64, 256x193x256, 200 steps: rel L2", errs, "wall-shear rel err", shear)
    for n in "uvw":
        assert errs[n] <= 5e-5, (n, errs[n])
    assert errs["p"] <= 1e-2, errs["p"]
    for w, e in shear.items():
        assert e <= 1e-3, (w, e)
