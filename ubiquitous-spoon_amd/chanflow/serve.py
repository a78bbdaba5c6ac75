"""Served streams of channel runs: the cf_server_* continuous-batching
scheduler, a seeded Poisson request stream and a real-time driver that
reports p50 / p99 completion latency and aggregate throughput; the plain
per-run loop (one run at a time, start to finish) is the baseline.
"""
import ctypes as C
import time
from dataclasses import dataclass, replace

import numpy as np

from . import (CF_FP32, CF_FP64, Config, FIELDS, Run, _CConfig, _check, _dptr, initial_fields, lib, LOG_N)

GRIDS = ((64, 64, 65), (128, 128, 97), (192, 192, 129), (256, 256, 193))   # (nx, ny span, nz wall-normal)


def _bind():
    L = lib()
    if getattr(L, "_srv_bound", False):
        return L
    P = C.c_void_p
    dp = C.POINTER(C.c_double)
    sig = {
        "cf_server_create": (C.c_int, [C.c_size_t, C.c_int, C.c_int, C.POINTER(P)]),
        "cf_server_destroy": (C.c_int, [P]),
        "cf_server_set_solo_points": (C.c_int, [P, C.c_size_t]),
        "cf_server_set_solo_limit": (C.c_int, [P, C.c_int]),
        "cf_server_submit": (C.c_int, [P, C.POINTER(_CConfig), dp, dp, dp, dp, C.c_int, C.c_int, C.POINTER(C.c_int)]),
        "cf_server_tick": (C.c_int, [P, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "cf_server_result": (C.c_int, [P, C.c_int, C.c_int, dp]),
        "cf_server_log": (C.c_int, [P, C.c_int, dp, C.c_int, C.POINTER(C.c_int)]),
        "cf_server_release": (C.c_int, [P, C.c_int]),
        "cf_server_request_info": (C.c_int, [P, C.c_int, dp, dp, dp, C.POINTER(C.c_int)]),
        "cf_server_stats": (C.c_int, [P, dp, dp, C.POINTER(C.c_size_t)]),
        "cf_server_request_state": (C.c_int, [P, C.c_int, C.POINTER(C.c_int), C.c_char_p, C.c_size_t]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    L._srv_bound = True
    return L


class Server:
    """A cf_server handle: HBM-budgeted admission, resident state,
    continuous batching per (order, scalar, precision) group; runs of at
    least solo_points cells (None: the library default) step through their
    own graphs beside the batches."""

    def __init__(self, hbm_budget_bytes=200 << 30, device=0, quantum=20, solo_points=None, solo_limit=None):
        L = _bind()
        h = C.c_void_p()
        _check(L.cf_server_create(int(hbm_budget_bytes), device, quantum, C.byref(h)))
        self._h = h
        self.cfgs = {}
        if solo_points is not None:
            _check(L.cf_server_set_solo_points(h, int(solo_points)))
        if solo_limit is not None:
            _check(L.cf_server_set_solo_limit(h, int(solo_limit)))

    def close(self):
        if getattr(self, "_h", None):
            _bind().cf_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit(self, cfg, fields, nsteps, retain=False):
        c, _keep = cfg.to_c()
        arrs = [np.ascontiguousarray(fields[n], dtype=np.float64) if n in fields else None for n in "uvwc"]
        rid = C.c_int(-1)
        _check(_bind().cf_server_submit(self._h, C.byref(c), _dptr(arrs[0]), _dptr(arrs[1]), _dptr(arrs[2]),
                                        _dptr(arrs[3]), int(nsteps), 1 if retain else 0, C.byref(rid)))
        self.cfgs[rid.value] = cfg
        return rid.value

    def tick(self):
        a, q, d = C.c_int(0), C.c_int(0), C.c_int(0)
        _check(_bind().cf_server_tick(self._h, C.byref(a), C.byref(q), C.byref(d)))
        return a.value, q.value, d.value

    def result(self, rid, name):
        out = np.zeros(self.cfgs[rid].shape)
        _check(_bind().cf_server_result(self._h, rid, FIELDS[name], _dptr(out)))
        return out

    def log(self, rid):
        n = C.c_int(0)
        _check(_bind().cf_server_log(self._h, rid, None, 1 << 20, C.byref(n)))
        out = np.zeros((max(n.value, 1), LOG_N))
        _check(_bind().cf_server_log(self._h, rid, _dptr(out), n.value, C.byref(n)))
        return out[: n.value]

    def release(self, rid):
        _check(_bind().cf_server_release(self._h, rid))

    def info(self, rid):
        ts, ta, td = C.c_double(), C.c_double(), C.c_double()
        n = C.c_int(0)
        _check(_bind().cf_server_request_info(self._h, rid, C.byref(ts), C.byref(ta), C.byref(td), C.byref(n)))
        return ts.value, ta.value, td.value, n.value

    STATES = ("queued", "active", "done", "released", "failed")

    def state(self, rid):
        """(state name, failure message) of a request."""
        st = C.c_int(0)
        msg = C.create_string_buffer(512)
        _check(_bind().cf_server_request_state(self._h, rid, C.byref(st), msg, 512))
        return self.STATES[st.value], msg.value.decode()

    def stats(self):
        sec, cs = C.c_double(), C.c_double()
        used = C.c_size_t()
        _check(_bind().cf_server_stats(self._h, C.byref(sec), C.byref(cs), C.byref(used)))
        return sec.value, cs.value, used.value


@dataclass
class Request:
    t_arrive: float
    cfg: Config
    nsteps: int
    seed: int


def poisson_stream(n, rate=20.0, seed=7, grids=GRIDS, steps=(50, 500), p_order4=0.5, p_fp32=0.5, base=None):
    """n requests with exponential inter-arrival times (mean 1/rate s),
    grid uniform over `grids`, order 4 with probability p_order4, FP32 with
    p_fp32, steps uniform in [steps[0], steps[1]]; channel parameters of
    the served workload (tanh beta = 2, box 2 pi x 2 x pi)."""
    rng = np.random.default_rng(seed)
    base = base or dict(stretch="tanh", stretch_beta=2.0, lx=2 * np.pi, ly=np.pi, lz=2.0, dt=2.5e-4,
                        nu=1.0 / 2800.0)
    t = 0.0
    out = []
    for q in range(n):
        t += float(rng.exponential(1.0 / rate))
        nx, ny, nz = grids[int(rng.integers(len(grids)))]
        order = 4 if rng.random() < p_order4 else 2
        prec = CF_FP32 if rng.random() < p_fp32 else CF_FP64
        nsteps = int(rng.integers(steps[0], steps[1] + 1))
        cfg = Config(nx=nx, ny=ny, nz=nz, order=order, precision=prec, **base)
        out.append(Request(t, cfg, nsteps, 1000 + q))
    return out


def serve(requests, budget=200 << 30, quantum=20, retain=False, time_scale=1.0, solo_points=None, solo_limit=None):
    """Drive a Server in real time: each request is submitted once its
    arrival time (scaled by time_scale) has passed; tick until all are
    done.  Returns per-request latencies (completion - arrival, s), total
    wall time, cell-steps and the server for retained results."""
    srv = Server(budget, quantum=quantum, solo_points=solo_points, solo_limit=solo_limit)
    pending = sorted(range(len(requests)), key=lambda q: requests[q].t_arrive)
    ids = {}
    t0 = time.perf_counter()
    offset = srv.stats()[0]                  # server clock at the driver's time zero
    done = 0
    while done < len(requests):
        now = time.perf_counter() - t0
        while pending and requests[pending[0]].t_arrive * time_scale <= now:
            q = pending.pop(0)
            r = requests[q]
            f = initial_fields(r.cfg, seed=r.seed)
            ids[q] = srv.submit(r.cfg, f, r.nsteps, retain=retain)
        active, queued, done = srv.tick()
        if active == 0 and queued == 0 and pending and done < len(requests):
            time.sleep(max(0.0, requests[pending[0]].t_arrive * time_scale - (time.perf_counter() - t0)))
    wall = time.perf_counter() - t0
    lat = []
    for q in range(len(requests)):
        ts, ta, td, n = srv.info(ids[q])
        lat.append(td - offset - requests[q].t_arrive * time_scale)
    _, cell_steps, _ = srv.stats()
    return {"latency_s": lat, "wall_s": wall, "cell_steps": cell_steps, "ids": ids, "server": srv}


def plain_loop(requests):
    """The per-run baseline: requests in arrival order, one at a time,
    each created, projected and stepped to completion on its own."""
    t0 = time.perf_counter()
    lat = []
    cell_steps = 0.0
    for r in sorted(requests, key=lambda r: r.t_arrive):
        now = time.perf_counter() - t0
        if now < r.t_arrive:
            time.sleep(r.t_arrive - now)
        f = initial_fields(r.cfg, seed=r.seed)
        with Run(r.cfg) as run:
            for n in "uvw":
                run.upload(n, f[n])
            run.project(1.0)
            run.step(r.nsteps)
            run.sync()
        lat.append(time.perf_counter() - t0 - r.t_arrive)
        cell_steps += r.cfg.nx * r.cfg.ny * r.cfg.nz * r.nsteps
    return {"latency_s": lat, "wall_s": time.perf_counter() - t0, "cell_steps": cell_steps}


def summary(res):
    lat = np.array(res["latency_s"])
    return {"p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "value": res["cell_steps"] / 1e6 / res["wall_s"],
            "unit": "Mpt*RK3-steps/s", "wall_s": res["wall_s"], "requests": len(lat)}
