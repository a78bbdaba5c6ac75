#!/usr/bin/env python3
"""Benchmark: RK3 steps of the channel DNS at 256 x 193 x 256 on one MI355X.

Workload (x, wall-normal, span = 256, 193, 256 cells; nx, nz, ny in the
solver's axis names): FP64, 2nd order, Lx = 2 pi, Ly(span) = pi, Lz = 2,
tanh wall stretching beta = 2, nu = 1/6900, dt = 2.5e-4, constant mass
flux Ub = 1, seeded laminar + 0.1 noise projected to zero divergence.

Prints one JSON line: Mpt.RK3-steps/s (whole job, inputs resident in HBM),
the dominant kernel's HBM roofline fraction from HIP-event kernel times,
and the single-thread CPU baseline (the oracle built -O3 -march=native on
the host; 1 warm-up step, median of 3) on a bounded sample.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-steps 3]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ubiquitous-spoon_amd"))

import numpy as np  # noqa: E402

import chanflow as cf  # noqa: E402

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, 8 TB/s
def step_bytes(N, M):
    """Algorithmic bytes of one order-2 FP64 RK3 step: the sum of alg_bytes
    over its launches (3 RHS, 3 spike solves, 3 divergence + x-FFT, 6
    y-FFTs, 3 Poisson, 3 inverse x-FFT; the plane sums and scalars are
    negligible)."""
    return 3 * alg_bytes("rhs", N, M) + 3 * alg_bytes("spike", N, M) + 3 * alg_bytes("div_fftx", N, M) \
        + 3 * alg_bytes("ffty_fwd", N, M) + 3 * alg_bytes("ffty_inv", N, M) + 3 * alg_bytes("poisson", N, M) \
        + 3 * alg_bytes("fftx_inv", N, M)


def alg_bytes(kernel, N, M):
    """Algorithmic bytes of one launch (each array read or written once)."""
    table = {
        # read u* v* w* p g_u g_v g_w, write f_u f_v f_w r_u r_v r_w: 104 N in the second
        # substep; the first reads no g (zeta = 0) and the last stores no f (never read),
        # 80 N each, so a launch averages 88 N
        "rhs": 88 * N,
        "spike": 48 * N,            # read r (3 fields), write solution
        "thomas": 48 * N,           # read r (3 fields), write solution
        "plane_sum": 8 * N,         # read u*
        "div_fftx": 24 * N + 16 * M,
        "ffty_fwd": 32 * M,
        "ffty_inv": 32 * M,
        "poisson": 32 * M,
        "fftx_inv": 16 * M + 8 * N,
        "project": 56 * N,          # read u* v* w* p, write u v w
        "cfl": 24 * N,
    }
    return table.get(kernel)


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_step_fp64_o2_256x193x256.json")


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed FETCH_SIZE /
    WRITE_SIZE counter runs of this workload, scaled by the same counters'
    reading for a plain copy of known size; None if not measured."""
    try:
        with open(TRAFFIC_FILE) as fh:
            t = json.load(fh)
        return t["kernels"][kernel]["hbm_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def workload_config():
    return cf.Config(nx=256, ny=256, nz=193, order=2, stretch="tanh", stretch_beta=2.0,
                     lx=2 * np.pi, ly=np.pi, lz=2.0, nu=1.0 / 6900.0, dt=2.5e-4,
                     want_uvolavg=1.0, uvolavgmode=1)


def time_form(name, cfg, steps, warmup, names="uvw", bytes_per_pt=None):
    """Throughput of one more form of the step (same metric), from a seeded
    projected field; per-kernel HIP-event times of a few extra steps."""
    f = cf.initial_fields(cfg, seed=1234)
    N = cfg.nx * cfg.ny * cfg.nz
    with cf.Run(cfg) as run:
        for n in names:
            run.upload(n, f[n])
        run.project(1.0)
        run.step(warmup)
        run.sync()
        t0 = time.perf_counter()
        run.step(steps)
        run.sync()
        secs = time.perf_counter() - t0
        run.set_profiling(True)
        run.step(2)
        run.sync()
        kt = {}
        for k in ("rhs", "cfl_finish", "solve", "spike", "plane_sum", "mass", "div_fftx", "ffty_fwd", "poisson",
                  "ffty_inv", "fftx_inv", "project"):
            ms, n = run.kernel_time(k)
            if n:
                kt[k] = round(ms * n / 2, 4)
        div = run.max_divergence()
    out = {"workload": name, "value": N / 1e6 * steps / secs, "unit": "Mpt*RK3-steps/s",
           "ms_per_step": 1e3 * secs / steps, "steps": steps, "max_div_end": div, "kernel_ms_per_step": kt}
    if bytes_per_pt:
        out["step_hbm_frac"] = bytes_per_pt * N * steps / secs / (HBM_PEAK_GBS * 1e9)
    print(f"[bench] form {name}: {out['ms_per_step']:.3f} ms/step", file=sys.stderr, flush=True)
    return out


def other_forms(steps, only=None):
    forms = {}
    if only == "order4":
        return {"order4_scalar_192x129x192": time_form("FP64 order 4 + scalar (Pr 0.71), 192x129x192", ORDER4_CFG(),
                                                       steps, 5, names="uvwc", bytes_per_pt=816.0)}
    c3 = cf.Config(nx=256, ny=256, nz=193, order=2, stretch="tanh", stretch_beta=2.0, lx=2 * np.pi, ly=np.pi,
                   lz=2.0, nu=1.0 / 6900.0, dt=2.5e-4, precision=cf.CF_FP32)
    forms["fp32_256x193x256"] = time_form("FP32 form (FP64 islands), 256x193x256 order 2", c3, steps, 5,
                                          bytes_per_pt=360.0)
    if only == "fp32":
        return forms
    forms["order4_scalar_192x129x192"] = time_form("FP64 order 4 + scalar (Pr 0.71), 192x129x192", ORDER4_CFG(),
                                                   steps, 5, names="uvwc", bytes_per_pt=816.0)
    return forms


def ORDER4_CFG():
    """The 4th-order + scalar workload: 192 x 129 x 192, Pr 0.71, nu 1/3300, dt 5e-4."""
    return cf.Config(nx=192, ny=192, nz=129, order=4, scalar=1, prandtl=0.71, stretch="tanh", stretch_beta=2.0,
                     lx=2 * np.pi, ly=np.pi, lz=2.0, nu=1.0 / 3300.0, dt=5e-4)


def run_batch(cfgs, fields, steps, total=None):
    """Create (and autotune) a batch, project, 2 warm-up steps, then time
    `steps`, then step on untimed to `total` steps in all (if given);
    returns seconds, creation time, autotuner lines and the fields."""
    t_create = time.perf_counter()
    with cf.Batch(cfgs) as b:
        t_create = time.perf_counter() - t_create
        for q in range(len(cfgs)):
            for n in "uvw":
                b.upload(q, n, fields[q][n])
        b.project(1.0)
        b.step(2)
        b.sync()
        t0 = time.perf_counter()
        b.step(steps)
        b.sync()
        secs = time.perf_counter() - t0
        if total is not None and total > steps + 2:
            b.step(total - steps - 2)
            b.sync()
        tuning = b.tuning()
        out = [{n: b.download(q, n) for n in "uvw"} for q in range(len(cfgs))]
    return secs, t_create, tuning, out


def batch48(steps):
    """The uneven batch: 16 runs each of 64x65x64, 96x97x96, 128x129x128
    (x, wall-normal, span), nu seeded uniform in [1/6000, 1/2000], dt 5e-4,
    order 2, stepped together in FP64 and in FP32 storage; against the same
    48 runs stepped one at a time in FP64 on the single-run path.  The FP32
    batch's error is measured against the FP64 batch after the same steps
    (the FP64 batch matches the oracle to 1e-12 in the tests)."""
    rng = np.random.default_rng(2024)
    cfgs = [cf.Config(nx=n, ny=n, nz=n + 1, stretch="tanh", nu=float(rng.uniform(1 / 6000, 1 / 2000)), dt=5e-4)
            for n in (64, 96, 128) for _ in range(16)]
    cells = sum(c.nx * c.ny * c.nz for c in cfgs)
    fields = [cf.initial_fields(c, seed=300 + q) for q, c in enumerate(cfgs)]
    total = 100                              # the error is stated after the workload's 100 steps
    secs, t_create, tuning, ref = run_batch(cfgs, fields, steps, total)
    c32 = []
    for c in cfgs:
        d = cf.Config(nx=c.nx, ny=c.ny, nz=c.nz, stretch="tanh", nu=c.nu, dt=c.dt, precision=cf.CF_FP32)
        c32.append(d)
    secs32, t_create32, tuning32, got = run_batch(c32, fields, steps, total)
    err = {n: max(float(np.linalg.norm(got[q][n] - ref[q][n]) / np.linalg.norm(ref[q][n])) for q in range(len(cfgs)))
           for n in "uvw"}
    t1 = 0.0
    for q, c in enumerate(cfgs):
        with cf.Run(c) as r:
            for n in "uvw":
                r.upload(n, fields[q][n])
            r.project(1.0)
            r.step(2)
            r.sync()
            t0 = time.perf_counter()
            r.step(steps)
            r.sync()
            t1 += time.perf_counter() - t0
    out = {"workload": "48 runs: 16 each of 64x65x64, 96x97x96, 128x129x128, FP64 order 2, one batch",
           "value": cells / 1e6 * steps / secs, "unit": "Mpt*RK3-steps/s", "steps": steps,
           "ms_per_step": 1e3 * secs / steps, "create_and_autotune_s": t_create,
           "one_at_a_time": {"value": cells / 1e6 * steps / t1, "ms_per_step": 1e3 * t1 / steps,
                             "precision": "FP64"},
           "speedup_vs_one_at_a_time": t1 / secs, "autotuner": tuning,
           "fp32_batch": {"value": cells / 1e6 * steps / secs32, "ms_per_step": 1e3 * secs32 / steps,
                          "create_and_autotune_s": t_create32,
                          "speedup_vs_fp64_one_at_a_time": t1 / secs32,
                          "max_rel_l2_vs_fp64_batch": err, "after_steps": total}}
    print(f"[bench] batch48: {out['ms_per_step']:.3f} ms/step, FP32 {1e3 * secs32 / steps:.3f} ms/step "
          f"(err {max(err.values()):.2e}), one at a time {1e3 * t1 / steps:.3f} ms/step", file=sys.stderr, flush=True)
    return out


def served_stream(nreq):
    """The served stream: Poisson arrivals at 20 req/s (seeded) of runs with
    grid uniform over 64x65x64, 128x97x128, 192x129x192, 256x193x256, order
    2 or 4 and FP32 or FP64 with p = 1/2, 50..500 steps, 200 GB HBM budget;
    against the plain per-run loop on the same requests."""
    from chanflow import serve
    reqs = serve.poisson_stream(nreq, rate=20.0, seed=7)
    res = serve.serve(reqs, budget=200 << 30, quantum=20)
    res["server"].close()
    s = serve.summary(res)
    base = serve.summary(serve.plain_loop(reqs))
    out = {"workload": f"{nreq} Poisson arrivals at 20 req/s, mixed grids / order 2|4 / FP32|FP64, 50-500 steps",
           **s, "plain_per_run_loop": base, "throughput_vs_plain": s["value"] / base["value"]}
    print(f"[bench] served: p50 {s['p50_s']:.2f} s p99 {s['p99_s']:.2f} s {s['value']:.0f} Mpt*steps/s; "
          f"plain loop {base['value']:.0f}", file=sys.stderr, flush=True)
    return out


def native_oracle():
    """Build the oracle on this host with the baseline's flags (-O3
    -march=native, as the reference's Makefile builds for its own host);
    returns (library path, flags) or (None, why not)."""
    import subprocess
    import tempfile
    flags = ["-O3", "-march=native", "-std=gnu99", "-fPIC", "-shared"]
    out = os.path.join(tempfile.mkdtemp(prefix="cf_cpu_baseline_"), "liboracle_native.so")
    src = [os.path.join(ROOT, "oracle", f) for f in ("orc_core.c", "orc_step.c")]
    try:
        subprocess.run(["gcc", *flags, "-o", out, *src, "-lm"], check=True, capture_output=True, timeout=300)
    except (OSError, subprocess.SubprocessError) as e:
        return None, f"native build failed ({e.__class__.__name__}); prebuilt -O3 -ffp-contract=off oracle"
    return out, "gcc " + " ".join(flags)


def host_cpu():
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "nproc": len(os.sched_getaffinity(0)), "os_cpu_count": os.cpu_count()}


def cpu_baseline(cfg, init, nsteps):
    """The CPU oracle, single-threaded, on a bounded sample of the same
    workload: one warm-up step, then the median of `nsteps` timed steps."""
    lib_path, flags = native_oracle()
    if lib_path:
        os.environ["ORACLE_LIB_PATH"] = lib_path
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    orc = ol.OracleRun(ol.oracle_cfg(cfg, cf.grid_dz(cfg)))
    for n in "uvw":
        orc.put(n, init[n])
    orc.steps(1)
    times = []
    for _ in range(nsteps):
        t0 = time.perf_counter()
        orc.steps(1)
        times.append(time.perf_counter() - t0)
    orc.close()
    med = float(np.median(times))
    npts = cfg.nx * cfg.ny * cfg.nz
    return {"value": npts / 1e6 / med, "unit": "Mpt*RK3-steps/s", "cores": 1, "threads": 1,
            "kind": "port", "compile": flags, "host": host_cpu(),
            "step_seconds": times,
            "sample": f"1 warm-up + median of {nsteps} RK3 steps of the same 256x193x256 FP64 workload on the "
                      f"single-threaded C oracle ({med:.2f} s per step)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--profile-steps", type=int, default=10)
    ap.add_argument("--cpu-steps", type=int, default=3, help="timed CPU steps (median; after 1 warm-up)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--form-steps", type=int, default=50, help="timed steps of each other form (0: skip)")
    ap.add_argument("--served", type=int, default=100, help="requests of the served stream (0: skip)")
    ap.add_argument("--copy-calibrate", type=int, default=0,
                    help="only run one device-to-device copy of this many bytes (HBM counter calibration)")
    ap.add_argument("--form", choices=["fp32", "order4"], default=None,
                    help="time only this other form and print its record (for profiling it alone)")
    args = ap.parse_args()
    if args.form:
        print(json.dumps(other_forms(max(args.form_steps, 1), only=args.form)))
        return
    if args.copy_calibrate:
        import torch
        n = args.copy_calibrate // 8
        a = torch.full((n,), 1.0, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 8 * n, "copy_bytes_written": 8 * n}))
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for {args.gpus}); "
              "multi-GPU sharding is a design note, not built", file=sys.stderr)
        sys.exit(2)

    cfg = workload_config()
    N = cfg.nx * cfg.ny * cfg.nz
    M = (cfg.nx // 2 + 1) * cfg.ny * cfg.nz
    f = cf.initial_fields(cfg, seed=1234, amp=0.1)
    run = cf.Run(cfg)
    for n in "uvw":
        run.upload(n, f[n])
    run.project(1.0)
    init = {n: run.download(n) for n in "uvw"}
    print(f"[bench] setup done, max div {run.max_divergence():.2e}", file=sys.stderr, flush=True)

    run.step(args.warmup)
    run.sync()
    t0 = time.perf_counter()
    run.step(args.steps)
    run.sync()
    t1 = time.perf_counter()
    secs = t1 - t0
    ms_per_step = 1e3 * secs / args.steps
    value = N / 1e6 * args.steps / secs
    print(f"[bench] {args.steps} steps in {secs:.3f} s ({ms_per_step:.3f} ms/step)", file=sys.stderr, flush=True)

    # per-kernel HIP-event times on the run's stream, same workload, right after
    run.set_profiling(True)
    run.step(args.profile_steps)
    run.sync()
    kt = {}
    for name in ("cfl", "rhs", "spike", "thomas", "plane_sum", "mass", "div_fftx", "ffty_fwd", "poisson",
                 "ffty_inv", "fftx_inv", "project"):
        ms, n = run.kernel_time(name)
        if n:
            kt[name] = (ms, n)
    run.set_profiling(False)
    per_step = {k: v[0] * v[1] / args.profile_steps for k, v in kt.items()}
    dom = max((k for k in per_step if alg_bytes(k, N, M)), key=lambda k: per_step[k])
    dom_ms = kt[dom][0]
    dom_bytes = alg_bytes(dom, N, M)
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    print("[bench] kernel ms/launch: " + ", ".join(f"{k}={v[0]:.4f}" for k, v in kt.items()),
          file=sys.stderr, flush=True)
    div = run.max_divergence()
    run.close()

    forms = other_forms(args.form_steps) if args.form_steps > 0 else {}
    if args.form_steps > 0:
        forms["batch48"] = batch48(20)
    if args.served > 0:
        forms["served_stream"] = served_stream(args.served)
    cpu = None if args.no_cpu else cpu_baseline(cfg, init, args.cpu_steps)

    out = {
        "metric": "Mpt*RK3-steps/s (256x193x256 channel, FP64 order 2)",
        "value": value,
        "unit": "Mpt*RK3-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded laminar + 0.1 noise, projected)",
        "config": {"workload": "channel DNS RK3 step, 256x193x256 (x, wall-normal, span), FP64, order 2, "
                               "tanh beta=2, nu=1/6900, dt=2.5e-4, Ub=1",
                   "nx": cfg.nx, "ny_span": cfg.ny, "nz_wall": cfg.nz, "order": 2,
                   "step_hbm_frac": step_bytes(N, M) * args.steps / secs / (HBM_PEAK_GBS * 1e9),
                   "max_div_end": div},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(dom),
                     "alg_bytes_per_launch": dom_bytes, "avg_launch_ms": dom_ms},
        "kernel_ms_per_step": per_step,
        "forms": forms,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
