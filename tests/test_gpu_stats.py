"""GPU tests of the plane statistics (cf_run_xy_stats, the istat block of
chanfast.c:658-861): x-y means, correlations, central moments and
one-dimensional cross-spectra of the run's own fields, against the
oracle's restatement of get.c on the downloaded fields (<= 1e-12
relative), and the reference's invariants case15 (first moment 0, second
moment = autocorrelation, cases.c:4293-4424) and case16 (the spectra sum to
the correlation, cases.c:4427-4732) on the device results."""
import numpy as np
import pytest

import chanflow as cf
import oracle_lib as ol

pytestmark = pytest.mark.gpu
dp = ol.dptr


def padded(pg, io):
    a = pg.zeros()
    a[3:-3, 3:-3, 3:-3] = io
    return a


def run_case(cfg, steps, names):
    f = cf.initial_fields(cfg, seed=77)
    r = cf.Run(cfg)
    for n in names:
        r.upload(n, f[n])
    r.project(1.0)
    r.step(steps)
    r.sync()
    return r, {n: r.download(n) for n in names + "p"}


def close(a, b, tol=1e-12, floor=1e-300):
    """|a - b| within tol relative to b, or to floor where b is a
    round-off zero (the plane mean of w is 0 by continuity: both sides
    then hold sums of ~1e-19 whose order of summation differs)"""
    return np.abs(a - b).max() <= tol * max(np.abs(b).max(), floor)


CASES = [
    (cf.Config(nx=32, ny=24, nz=33, stretch="tanh", nu=1 / 2800, dt=1e-3), "uvw"),
    (cf.Config(nx=24, ny=20, nz=17, order=4, scalar=1, prandtl=0.71, stretch="tanh", nu=1 / 3300, dt=5e-4), "uvwc"),
    (cf.Config(nx=30, ny=25, nz=21, stretch="tanh", nu=1 / 2000, dt=1e-3, precision=cf.CF_FP32), "uvw"),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_plane_statistics_match_oracle(case):
    cfg, names = CASES[case]
    r, f = run_case(cfg, 4, names)
    pg = ol.PaddedGrid(cfg.nx, cfg.ny, cfg.nz, cfg.lx, cfg.ly, cf.grid_dz(cfg))
    n2 = pg.g.n2
    P = {n: padded(pg, f[n]) for n in f}
    with r:
        for n in f:
            avg = np.zeros(n2)
            ol.lib().orc_xy_avg(dp(P[n]), dp(avg), pg.ref)
            assert close(r.xy_stats("avg", n), avg[3:-3], floor=np.abs(f[n]).max()), n
        pairs = [("u", "u"), ("w", "u"), ("p", "p")] + ([("u", "c"), ("w", "c")] if "c" in names else [])
        for a, b in pairs:
            corr = np.zeros(n2)
            ol.lib().orc_xy_corr(dp(P[a]), dp(P[b]), dp(corr), pg.ref)
            assert close(r.xy_stats("corr", a, b), corr[3:-3]), (a, b)
        for en in (1.0, 2.0, 3.0, 4.0):
            mom = np.zeros(n2)
            ol.lib().orc_xy_moment(dp(P["u"]), dp(mom), en, pg.ref)
            got = r.xy_stats("moment", "u", en=en)
            if en == 1.0:    # case15: the first central moment vanishes
                assert np.abs(got).max() < 1e-13 and np.abs(mom[3:-3]).max() < 1e-13
            else:
                # relative to the moment's own scale (the largest |u - mean|^en):
                # odd moments are small differences of large terms
                dev = np.abs(f["u"] - f["u"].mean(axis=(0, 1))[None, None, :]).max()
                assert close(got, mom[3:-3], 1e-11, floor=dev ** en), en
        assert close(r.xy_stats("moment", "u", en=2.0), r.xy_stats("corr", "u", "u"), 1e-13)    # case15


@pytest.mark.parametrize("case", range(len(CASES)))
def test_spectra_match_oracle_and_parseval(case):
    cfg, names = CASES[case]
    r, f = run_case(cfg, 3, names)
    nx, ny, nz = cfg.nx, cfg.ny, cfg.nz
    pg = ol.PaddedGrid(nx, ny, nz, cfg.lx, cfg.ly, cf.grid_dz(cfg))

    def fft3d(io):
        s = np.fft.fft(np.fft.rfft(io, axis=0) / nx, axis=1) / ny
        return np.ascontiguousarray(np.stack([s.real, s.imag], -1))
    with r:
        for a, b in (("u", "u"), ("w", "u"), ("p", "p")):
            sx, sy = r.xy_stats("spec", a, b)
            ox = np.zeros((nx // 2 + 1, nz, 2))
            oy = np.zeros((ny // 2 + 1, nz, 2))
            ol.lib().orc_xy_spec(dp(fft3d(f[a])), dp(fft3d(f[b])), dp(ox), dp(oy), nx, ny, nz)
            scale = max(np.abs(ox).max(), np.abs(oy).max())
            assert np.abs(sx - (ox[..., 0] + 1j * ox[..., 1])).max() <= 1e-12 * scale, (a, b)
            assert np.abs(sy - (oy[..., 0] + 1j * oy[..., 1])).max() <= 1e-12 * scale, (a, b)
            corr = r.xy_stats("corr", a, b)

            def total(spec, n):
                w = np.full(spec.shape[0], 2.0)
                w[0] = 1.0
                if n % 2 == 0:
                    w[-1] = 1.0
                return (w[:, None] * spec.real).sum(0)
            cs = max(np.abs(corr).max(), 1e-300)
            assert np.abs(total(sx, nx) - corr).max() <= 1e-12 * cs       # case16
            assert np.abs(total(sy, ny) - corr).max() <= 1e-12 * cs
