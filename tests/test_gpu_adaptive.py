"""GPU parity of the adaptive time step (cflmode 1, chanfast.c:383-393):
every step starts with get_cfl of the current state at the current dt,
then dt *= want_cfl / cfl, and the step runs with every dt-dependent table
(z-solve factors, unit-forcing profiles, boundary lifts, Poisson tables)
rebuilt for the new dt.

FP64 fields to <= 1e-12 relative L2 against the CPU oracle after up to
100 steps with dt changing every step; the cfl.dat column is want_cfl
exactly, as the reference logs it.
"""
import numpy as np
import pytest

import chanflow as cf
import oracle_lib as ol
from test_gpu_general import check_log, general_case, rel

pytestmark = pytest.mark.gpu


def adaptive_compare(cfg, orc, init, names, steps, pieces=(None,)):
    olog = orc.steps(steps)
    with cf.Run(cfg) as run:
        for n in names:
            run.upload(n, init[n])
        done = 0
        for p in pieces:
            k = steps - done if p is None else p
            run.step(k)
            done += k
        assert done == steps
        run.sync()
        errs = {n: rel(run.download(n), orc.get(n)) for n in names}
        glog = run.log()
        print(f"{cfg.shape} order {cfg.order} scalar {cfg.scalar} cflmode 1: rel L2 after {steps} steps {errs}; "
              f"dt {glog[0, 1]:.6e} .. {glog[-1, 1]:.6e}")
        for n in names:
            assert errs[n] <= 1e-12, (n, errs[n])
        check_log(glog, olog, cfg.scalar)
        # dt really changes from step to step, and the logged cfl is want_cfl
        assert np.all(glog[:, 2] == cfg.want_cfl)
        assert len(np.unique(glog[:, 1])) > steps // 2
        assert abs(glog[-1, 1] - ol.lib().orc_run_dt(orc._h)) <= 1e-12 * glog[-1, 1]
        umax = np.abs(orc.get("u")).max()
        assert run.max_divergence() <= 1e-12 * umax / (cfg.lx / cfg.nx)
        return glog


def test_adaptive_dt_order2_fast_path_100_steps():
    cfg, orc, init, _, names = general_case(32, 24, 33, 2, 0, dt=1e-3, cflmode=1, want_cfl=0.4,
                                            nu=1.0 / 2800.0)
    adaptive_compare(cfg, orc, init, names, 100)


def test_adaptive_dt_split_calls_match_one_call():
    """Steps issued in several cf_run_step calls (the pending projection of
    one call meets the next call's new dt) give the same bits as one call."""
    cfg, orc, init, _, names = general_case(24, 16, 25, 2, 0, dt=1e-3, cflmode=1, want_cfl=0.4)
    outs = []
    for pieces in ((None,), (1, 4, 7, None)):
        with cf.Run(cfg) as run:
            for n in names:
                run.upload(n, init[n])
            done = 0
            for p in pieces:
                k = 20 - done if p is None else p
                run.step(k)
                done += k
            run.sync()
            outs.append({n: run.download(n) for n in names})
    for n in names:
        assert np.array_equal(outs[0][n], outs[1][n]), n


def test_adaptive_dt_reference_main_parameters():
    """The reference's own production run (main.c:45-61): chanfast4 with the
    scalar, cflmode 1, constant bulk velocity in a frame moving with it
    (uframe = want_uvolavg), constant bulk scalar, dpdx = -1, nu = 1/180,
    want_cfl = 1; on a small grid, 60 steps."""
    cfg, orc, init, _, names = general_case(24, 20, 25, 4, 1, dt=2e-3, cflmode=1, want_cfl=1.0, nu=1.0 / 180.0,
                                            prandtl=1.0, dpdx=-1.0, want_uvolavg=1.0, uvolavgmode=1,
                                            want_cvolavg=0.0, cvolavgmode=1, uframe=1.0, stretch="cosine")
    adaptive_compare(cfg, orc, init, names, 60, pieces=(7, None))


def test_adaptive_dt_order4_scalar_100_steps():
    cfg, orc, init, _, names = general_case(24, 16, 33, 4, 1, dt=5e-4, cflmode=1, want_cfl=0.5,
                                            nu=1.0 / 3300.0)
    adaptive_compare(cfg, orc, init, names, 100)


def test_batch_rejects_adaptive_dt():
    cfg = cf.Config(nx=16, ny=12, nz=17, cflmode=1, want_cfl=0.5)
    with pytest.raises(cf.ChanflowError):
        cf.Batch([cfg, cfg])
