"""GPU tests of the served stream (cf_server_*): continuous batching of
independent runs under an HBM budget.

A run's result is bit-identical alone and in a mix that it joins while
other runs are mid-flight (per-run reductions, fixed solve shape), and
matches the oracle (FP64, <= 1e-12 relative L2); admission never exceeds
the budget; a mixed stream of orders and precisions completes with its
latencies reported.
"""
import numpy as np
import pytest

import chanflow as cf
from chanflow import serve
import oracle_lib as ol

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def cfg_of(shape, order=2, prec=cf.CF_FP64, **kw):
    base = dict(stretch="tanh", nu=1.0 / 2800.0, dt=5e-4)
    base.update(kw)
    return cf.Config(nx=shape[0], ny=shape[1], nz=shape[2], order=order, precision=prec, **base)


def drain(srv):
    while True:
        a, q, d = srv.tick()
        if a == 0 and q == 0:
            return d


def test_same_alone_and_in_the_mix():
    mine = cfg_of((32, 24, 33))
    f = cf.initial_fields(mine, seed=5)
    with serve.Server(64 << 30) as srv:
        a = srv.submit(mine, f, 30, retain=True)
        drain(srv)
        alone = {n: srv.result(a, n) for n in "uvw"}
        alone_log = srv.log(a)
    with serve.Server(64 << 30, quantum=7) as srv:
        b = srv.submit(cfg_of((48, 32, 33), nu=1.0 / 4000.0), cf.initial_fields(cfg_of((48, 32, 33)), seed=6), 60)
        c = srv.submit(cfg_of((24, 20, 33), order=4, prec=cf.CF_FP32), cf.initial_fields(cfg_of((24, 20, 33)), seed=7), 40)
        srv.tick()
        srv.tick()
        a2 = srv.submit(mine, f, 30, retain=True)        # joins while b and c are mid-flight
        d = srv.submit(cfg_of((16, 12, 17), nu=1.0 / 2000.0), cf.initial_fields(cfg_of((16, 12, 17)), seed=8), 25)
        drain(srv)
        mixed = {n: srv.result(a2, n) for n in "uvw"}
        mixed_log = srv.log(a2)
        for rid in (b, c, d):
            assert srv.info(rid)[3] > 0
    for n in "uvw":
        assert np.array_equal(alone[n], mixed[n]), n
    assert np.array_equal(alone_log, mixed_log)
    dz = cf.grid_dz(mine)
    orc = ol.OracleRun(ol.oracle_cfg(mine, dz))
    for n in "uvw":
        orc.put(n, f[n])
    orc.project_ic(1.0)
    orc.steps(30)
    for n in "uvw":
        assert rel(alone[n], orc.get(n)) <= 1e-12, n


def test_admission_respects_the_budget():
    cfgs = [cfg_of((48, 32, 33)) for _ in range(5)]
    per = cf.footprint(cfgs[0])
    with serve.Server(2 * per + per // 2) as srv:          # room for two at a time
        for q, c in enumerate(cfgs):
            srv.submit(c, cf.initial_fields(c, seed=q), 15 + 5 * q)
        while True:
            a, qd, d = srv.tick()
            _, _, used = srv.stats()
            assert used <= 2 * per + per // 2
            if a == 0 and qd == 0:
                break
        assert d == 5
        # tick() reports the count after its retirements, so read concurrency
        # from the residency intervals [t_admit, t_done] instead
        ev = []
        for q in range(5):
            _, ta, td, n = srv.info(q)
            assert n == 15 + 5 * q
            ev += [(ta, 1), (td, -1)]
        peak = cur = 0
        for _, e in sorted(ev, key=lambda x: (x[0], x[1])):
            cur += e
            peak = max(peak, cur)
        assert peak == 2


def test_mixed_stream_completes():
    reqs = serve.poisson_stream(12, rate=50.0, seed=3, grids=((32, 24, 33), (48, 32, 33), (64, 48, 65)),
                                steps=(20, 60))
    res = serve.serve(reqs, budget=64 << 30, quantum=10, retain=True)
    srv = res["server"]
    s = serve.summary(res)
    print("served stream:", s)
    assert s["requests"] == 12 and s["p99_s"] >= s["p50_s"] > 0
    for q, r in enumerate(reqs):
        lg = srv.log(res["ids"][q])
        assert lg.shape[0] == r.nsteps
        assert np.allclose(lg[:, 6], 1.0, rtol=0, atol=1e-5 if r.cfg.precision == cf.CF_FP32 else 1e-12)
        u = srv.result(res["ids"][q], "u")
        assert np.isfinite(u).all()
    srv.close()


def test_served_workload_grids_same_alone_and_in_the_mix():
    """The served stream's own grids (x, wall-normal, span = 64x65x64,
    128x97x128, 192x129x192, 256x193x256), orders 2 and 4, FP32 and FP64:
    two large runs routed to their own graphs (solo), small and medium runs
    in group batches, an FP32 run and an FP64 run joining while the others
    are mid-flight.  Each result is bit-identical to the same request served
    alone (every reduction is per run and in a fixed order; each group's
    solve shape and each run's FFT lengths depend on the run alone), keeps
    its bulk velocity and ends divergence free."""
    def req(n, nzw, order, prec, seed, nu):
        c = cf.Config(nx=n, ny=n, nz=nzw, order=order, precision=prec, stretch="tanh", lx=2 * np.pi,
                      ly=np.pi, lz=2.0, nu=nu, dt=2.5e-4)
        return c, cf.initial_fields(c, seed=seed)

    first = [req(256, 193, 2, cf.CF_FP64, 1, 1.0 / 6900.0), req(192, 129, 4, cf.CF_FP32, 2, 1.0 / 3300.0),
             req(64, 65, 4, cf.CF_FP64, 3, 1.0 / 2800.0), req(128, 97, 2, cf.CF_FP64, 4, 1.0 / 4000.0)]
    late = [req(128, 97, 2, cf.CF_FP32, 5, 1.0 / 3000.0), req(64, 65, 2, cf.CF_FP64, 6, 1.0 / 2000.0)]
    steps = [12, 8, 30, 20, 24, 36]
    budget = 200 << 30
    alone = []
    for q, (c, f) in enumerate(first + late):
        with serve.Server(budget, quantum=5) as srv:
            rid = srv.submit(c, f, steps[q], retain=True)
            drain(srv)
            alone.append(({n: srv.result(rid, n) for n in "uvw"}, srv.log(rid)))
    with serve.Server(budget, quantum=5) as srv:
        ids = [srv.submit(c, f, steps[q], retain=True) for q, (c, f) in enumerate(first)]
        srv.tick()
        srv.tick()
        for q, (c, f) in enumerate(late):          # join while the first four are mid-flight
            ids.append(srv.submit(c, f, steps[len(first) + q], retain=True))
        drain(srv)
        for q, rid in enumerate(ids):
            c = (first + late)[q][0]
            lg = srv.log(rid)
            assert lg.shape[0] == steps[q]
            for n in "uvw":
                got = srv.result(rid, n)
                assert np.array_equal(got, alone[q][0][n]), (q, c.shape, c.order, c.precision, n)
            assert np.array_equal(lg, alone[q][1]), q
            tol = 1e-5 if c.precision == cf.CF_FP32 else 1e-12
            assert np.allclose(lg[:, 6], 1.0, rtol=0, atol=tol), q


def test_solo_runs_shortest_remaining_work_first():
    """With one solo run advancing at a time, the quantum goes to the solo
    run with the least remaining cells x steps: a short request submitted
    after a long one finishes first, and both results are the same as when
    each is served alone."""
    long_c, short_c = cfg_of((32, 24, 33)), cfg_of((24, 16, 25))
    fl, fs = cf.initial_fields(long_c, seed=11), cf.initial_fields(short_c, seed=12)
    alone = []
    for c, f, n in ((long_c, fl, 40), (short_c, fs, 10)):
        with serve.Server(8 << 30, quantum=5, solo_points=0) as srv:
            rid = srv.submit(c, f, n, retain=True)
            drain(srv)
            alone.append(srv.result(rid, "u"))
    with serve.Server(8 << 30, quantum=5, solo_points=0, solo_limit=1) as srv:
        a = srv.submit(long_c, fl, 40, retain=True)
        b = srv.submit(short_c, fs, 10, retain=True)
        drain(srv)
        ta_done, tb_done = srv.info(a)[2], srv.info(b)[2]
        assert tb_done < ta_done
        assert np.array_equal(srv.result(a, "u"), alone[0])
        assert np.array_equal(srv.result(b, "u"), alone[1])
