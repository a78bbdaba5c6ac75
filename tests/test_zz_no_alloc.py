"""cf_run_xy_stats_noalloc: the plane statistics with their work space in a
fixed device array (no allocation per call) give the same bits as
cf_run_xy_stats, on two calls in a row, the second on a smaller run; and a
run with more levels than the fixed work space is refused."""
import ctypes as C

import numpy as np
import pytest

import chanflow as cf

pytestmark = pytest.mark.gpu


def noalloc(run, kind, a, b=None, en=0.0):
    L = cf.lib()
    fn = L.cf_run_xy_stats_noalloc
    fn.restype = C.c_int
    fn.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.POINTER(C.c_double)]
    out = np.full(run.cfg.nz, np.nan)
    st = fn(run._h, cf.Run.STATS[kind], cf.FIELDS[a], cf.FIELDS[b if b is not None else a], float(en),
            out.ctypes.data_as(C.POINTER(C.c_double)))
    return st, out


def stepped_run(nx, ny, nz, order, scalar, prec, steps, seed):
    cfg = cf.Config(nx=nx, ny=ny, nz=nz, order=order, scalar=scalar, precision=prec, stretch="tanh")
    f = cf.initial_fields(cfg, seed=seed)
    run = cf.Run(cfg)
    for n in ("uvwc" if scalar else "uvw"):
        run.upload(n, f[n])
    run.project(1.0)
    run.step(steps)
    run.sync()
    return run


CASES = [("avg", "u", None, 0.0), ("avg", "p", None, 0.0), ("corr", "w", "u", 0.0), ("corr", "v", "v", 0.0),
         ("moment", "u", None, 3.0), ("moment", "w", None, 4.0)]


@pytest.mark.parametrize("shapes", [((48, 32, 49, 2, 0, cf.CF_FP64), (16, 12, 17, 2, 0, cf.CF_FP64)),
                                    ((32, 24, 33, 4, 1, cf.CF_FP64), (24, 16, 25, 2, 0, cf.CF_FP32))])
def test_noalloc_stats_match_allocating_entry(shapes):
    for q, (nx, ny, nz, order, scalar, prec) in enumerate(shapes):
        with stepped_run(nx, ny, nz, order, scalar, prec, 3, 40 + q) as run:
            cases = CASES + ([("corr", "w", "c", 0.0)] if scalar else [])
            for kind, a, b, en in cases:
                st, got = noalloc(run, kind, a, b, en)
                assert st == 0, cf.lib().cf_last_error().decode()
                want = run.xy_stats(kind, a, b, en=en)
                assert got.shape == (nz,)
                assert np.array_equal(got, want), (q, kind, a, b)


def test_noalloc_refuses_beyond_the_work_space():
    cfg = cf.Config(nx=8, ny=8, nz=2100, stretch="uniform")
    with cf.Run(cfg) as run:
        st, _ = noalloc(run, "avg", "u")
        assert st == cf.CF_ERR_ARG
        assert "exceeds the fixed work space" in cf.lib().cf_last_error().decode()
        st, _ = noalloc(run, "spec", "u", "u")
        assert st != 0
