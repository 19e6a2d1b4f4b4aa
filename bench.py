#!/usr/bin/env python3
"""Hermite-4 block-step force throughput on one MI355X (Ginteractions/s), end to end through
the GRAPE-6 C-ABI.

Workload (one "step" = one sweep of 18 block steps): Plummer N = 131072, equal masses, Henon
units, eps2 = 1e-6; j-particles carry acc/jerk and times t_j on 20 power-of-two levels; each
block step re-sends the previous active set (the corrector's update), sets the block time and
evaluates acc, jerk, pot, nearest neighbour and neighbour lists (mean ~32, cap 256) on an
active set of 2^k particles, k = 0..17, in calls of <= 16384 i-particles.

The timed loop is the compiled GRAPE-6 caller sapporo2_amd/bench/g6_blocksteps (no Python in
it), linked against libsapporo_amd.so as an external integrator would be:

value       = interactions per step / median over the timed steps of the end-to-end API time
              (g6_set_ti_ + g6calc_firsthalf_ -> g6calc_lasthalf2_ + g6_read_neighbour_list_
              for every call of the step; PCIe, host packing and result/list read-back
              included), the reference's convention (test_performance_blockStep.cpp:160-183);
              the engine runs as a user runs it, without per-call HIP events.  The run
              times twice the steps, alternating between that and the engine with its HIP
              events on its stream around every call and kernel.
device_value= the same interactions over the summed device windows of the calls (first kernel
              after the call's input copy -> last kernel of the call), secondary; median of
              the instrumented steps (instrumented_value = their API rate: the events cost
              the host ~16 us per call).
roofline    = dominant kernel (force, P = 64 lane shape): 60 flop per interaction (reference
              convention) over its summed launch times, HIP events on its stream, in the
              instrumented steps of the same run.
cpu_baseline= the FP64 direct summation of the oracle (OpenMP, compiled for this host with
              -march=native) on a sample of 16384 i-particles against all j, on this host's
              cores; the same sample is the accuracy reference for the reported errors.
configs     = the other configurations side by side (mixed, FP64, CPU sample, errors,
              nearest-neighbour / list mismatches).
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import sapporo2_amd as sa  # noqa: E402
from sapporo2_amd import ics  # noqa: E402

METRIC = "Ginteractions/s Hermite-4 N=128K, mixed vs FP64 GPU at rel. force err ≤1e-6"
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "force_h4_mixed_traffic.md")
DRIVER = os.path.join(ROOT, "sapporo2_amd", "bench", "g6_blocksteps")
FLOP_PER_INTERACTION = 60.0          # test_gravity_block.cpp:271
# Algorithmic bytes of the HBM-bound kernels (Hermite-4).  Predictor, per j-particle: reads
# t, m, x, v, a, j (14 doubles) and the id; writes the working-precision record (48 B mixed,
# 64 B FP64) and, mixed only, the FP64 predicted position and velocity kept for read-back.
# Scatter, per staged update followed by a prediction in the same call: reads the staged row
# (16 doubles and the address/id pair), writes the 16 raw fields and the id.
PREDICT_BYTES = {"mixed": 116 + 48 + 48, "fp64": 116 + 64}
SCATTER_BYTES = 136 + 128 + 4
PEAK_TFLOPS = {"mixed": 157.3, "fp64": 78.6}   # MI355X FP32 / FP64 vector
PRECS = (("mixed", sa.DOUBLESINGLE), ("fp64", sa.DOUBLE))
EPS2 = 1e-6
N_DEFAULT = 131072
SEED = 2024
CPU_SAMPLE_S = 10.0        # CPU baseline: repeat the oracle sample until this much time has elapsed
CPU_SAMPLE_MAX_REPS = 40


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def measured_traffic():
    """HBM bytes per launch of the mixed P = 64 force kernel over one step, as measured by
    rocprofv3 FETCH_SIZE / WRITE_SIZE (calibrated on a known-size copy) and recorded in the
    committed profile summary; None when that summary is absent."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(line.split(":")[1])
    except OSError:
        pass
    return None


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def host_predict(t, tj, x, v, a, j):
    """Caller-side (integrator) prediction of the i-particles to the block time."""
    dt = (t - tj)[:, None]
    px = x + dt * (v + dt * 0.5 * (a + dt * (1.0 / 3.0) * j))
    pv = v + dt * (a + dt * 0.5 * j)
    return px, pv


def init_derivatives(m, x, v, ids, eps2):
    """acc/jerk of every particle from one all-i evaluation (FP64 kernel): the j-particles'
    Taylor coefficients for the predictor."""
    n = len(m)
    with sa.Engine(sa.FOURTH, sa.DOUBLE) as eng:
        eng.set_j_particles(np.arange(n), ids, m, x, v)
        eng.set_time(0.0)
        r = eng.calc(n, ids, x, v, eps2, None, ngb=False)
    return r["acc"], r["jrk"]


class Workload:
    """The block-step workload of the metric (Plummer N, 20 time levels, active 2^0..2^17)."""

    def __init__(self, n, seed):
        self.n = n
        self.eps2 = EPS2
        self.m, self.x, self.v = ics.plummer(n, seed=seed)
        self.ids = np.arange(n, dtype=np.int32)
        self.h2 = ics.neighbour_h2(self.x, 32.0)
        rng = np.random.Generator(np.random.PCG64(seed + 1))
        lev = ics.block_levels(n, 20, seed=seed + 2)
        self.tj = -rng.uniform(0.0, 1.0, n) * 2.0 ** -(3.0 + lev)
        self.counts = [2 ** k for k in range(18) if 2 ** k <= n]
        self.blocks = ics.active_subsets(n, self.counts, seed=seed + 3)
        self.acc, self.jrk = init_derivatives(self.m, self.x, self.v, self.ids, EPS2)
        self.px, self.pv = host_predict(0.0, self.tj, self.x, self.v, self.acc, self.jrk)
        self.resend = True
        self.ngb = True

    @property
    def interactions_per_step(self):
        return sum(len(b) for b in self.blocks) * self.n


def write_workload(path, n, ids, m, x, v, a, j, tj, h2, px, pv, blocks, eps2, resend, ngb, tblock=0.0):
    """The driver's input file (layout in g6_blocksteps.cpp read_workload)."""
    with open(path, "wb") as f:
        f.write(b"SAPG6WL1")
        f.write(np.array([n, len(blocks), int(resend), int(ngb)], np.int32).tobytes())
        f.write(np.array([eps2, tblock], np.float64).tobytes())
        f.write(np.ascontiguousarray(ids, np.int32).tobytes())
        if n % 2:
            f.write(b"\0\0\0\0")
        for arr in (m, x, v, a, j, tj, h2, px, pv):
            f.write(np.ascontiguousarray(arr, np.float64).tobytes())
        f.write(np.array([len(b) for b in blocks], np.int32).tobytes())
        for b in blocks:
            f.write(np.ascontiguousarray(b, np.int32).tobytes())


def read_results(path):
    with open(path, "rb") as f:
        buf = f.read()
    assert buf[:8] == b"SAPG6RS1", "bad results file"
    ns, nl, n, ngb = np.frombuffer(buf, np.int32, 4, 8)
    o = 24
    out = {}

    def take(name, dtype, count, shape=None):
        nonlocal o
        a = np.frombuffer(buf, dtype, count, o).copy()
        o += a.nbytes
        out[name] = a if shape is None else a.reshape(shape)
    take("acc", np.float64, 3 * ns, (ns, 3))
    take("jrk", np.float64, 3 * ns, (ns, 3))
    take("pot", np.float64, ns)
    take("nnb", np.int32, ns)
    take("cnt", np.int32, ns)
    take("list", np.int32, nl)
    take("ppos", np.float64, 3 * n, (n, 3))
    take("pvel", np.float64, 3 * n, (n, 3))
    return out


def run_driver(wl, res, prec, steps, warmup, sample_block, timing=1):
    """Runs the compiled g6 caller in a child process with the chosen working precision."""
    if not os.path.exists(DRIVER):
        raise RuntimeError(f"{DRIVER} is missing: build it with __graft_entry__.build()")
    env = dict(os.environ, SAPPORO2_AMD_PRECISION=prec)
    cmd = [DRIVER, wl, str(steps), str(warmup), res, str(sample_block), str(timing)]
    p = subprocess.run(cmd, env=env, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, cwd=ROOT)
    if p.returncode != 0:
        raise RuntimeError(f"{' '.join(cmd)} failed ({p.returncode}): {p.stderr[-2000:]}")
    return json.loads(p.stdout.strip().splitlines()[-1])


def measure(wl, res, prec, steps, warmup, sample_block):
    """One run of the compiled caller with 2 x steps timed steps alternating between the
    engine as a user runs it (no HIP-event instrumentation: the API rate) and the engine with
    its per-call events on (device windows, per-kernel times), so that both see the same
    process and GPU state."""
    return summarize(run_driver(wl, res, prec, steps, warmup, sample_block, timing=2), prec)


def summarize(js, prec):
    """value (API, median of the uninstrumented steps); device window, roofline of the P = 64
    force kernel and the per-block device times from the instrumented steps."""
    ips = js["interactions_per_step"]
    api = float(np.median(js["step_api_s"]))
    api_timed = float(np.median(js["step_api_instrumented_s"]))
    dev_ms = float(np.median(js["step_device_ms"]))
    f = js["force64"]
    achieved = FLOP_PER_INTERACTION * f["interactions"] / (f["ms"] * 1e-3) / 1e12 if f["ms"] > 0 else 0.0
    pr, sc = js.get("predict", {}), js.get("scatter", {})
    pred = scat = None
    if pr.get("launches"):
        t = pr["ms"] * 1e-3 / pr["launches"]
        pred = dict(launches=pr["launches"], mean_us=round(t * 1e6, 2), bytes_per_j=PREDICT_BYTES[prec],
                    nj=pr["nj"], achieved_GBps=round(PREDICT_BYTES[prec] * pr["nj"] / t / 1e9, 1))
    if sc.get("launches"):
        t = sc["ms"] * 1e-3 / sc["launches"]
        u = sc["updates"] / sc["launches"]
        scat = dict(launches=sc["launches"], mean_us=round(t * 1e6, 2), updates_per_launch=round(u, 1),
                    bytes_per_update=SCATTER_BYTES, achieved_GBps=round(SCATTER_BYTES * u / t / 1e9, 1))
    return dict(value=ips / api / 1e9, ms_per_step=api * 1e3, predictor=pred, scatter=scat,
                instrumented_value=ips / api_timed / 1e9,
                device_value=ips / (dev_ms * 1e-3) / 1e9 if dev_ms > 0 else None, device_ms_per_step=dev_ms,
                force_mean_ms=f["ms"] / max(1, f["launches"]), force_launches=f["launches"],
                achieved_tflops=achieved, peak=PEAK_TFLOPS[prec],
                host_ms_per_step={k: round(1e3 * float(np.median(js[k])), 3) for k in
                                  ("step_setti_s", "step_firsthalf_s", "step_lasthalf_s", "step_read_ngb_s")},
                outside_api_ms_per_step={"j_resend": round(1e3 * float(np.median(js["step_resend_s"])), 3),
                                         "list_copy_out": round(1e3 * float(np.median(js["step_list_copy_s"])), 3)},
                blocks=js["blocks"])


def run_engine_loop(w, name, prec, steps, warmup):
    """The same block-step sweep through the engine handle API inside this process (no
    child process; the form profiled with rocprofv3): per-call device windows and force
    kernel times from the engine's HIP events; prints one JSON line."""
    eng = sa.Engine(sa.FOURTH, prec)
    eng.set_timing(True)
    n = w.n
    eng.set_j_particles(np.arange(n), w.ids, w.m, w.x, w.v, w.acc, w.jrk, tj=w.tj)
    api, dev, f_ms, f_int = [], [], 0.0, 0.0
    for step in range(warmup + steps):
        timed = step >= warmup
        a_s, d_ms = 0.0, 0.0
        for b, sel in enumerate(w.blocks):
            prev = w.blocks[b - 1]
            eng.set_j_particles(prev, w.ids[prev], w.m[prev], w.px[prev], w.pv[prev], w.acc[prev], w.jrk[prev],
                                tj=0.0)
            eng.set_time(0.0)
            for c0 in range(0, len(sel), sa.NPIPES):
                ch = sel[c0:c0 + sa.NPIPES]
                t0 = time.perf_counter()
                eng.start_grav_calc(n, w.ids[ch], w.px[ch], w.pv[ch], EPS2, w.h2[ch])
                eng.get_grav_results(n, len(ch), ngb=True)
                eng.read_ngb_list()
                a_s += time.perf_counter() - t0
                d_ms += eng.last_device_ms()
                if timed and eng.last_config()[0] == 64:
                    f_ms += eng.last_force_ms()
                    f_int += len(ch) * n
        if timed:
            api.append(a_s)
            dev.append(d_ms)
    eng.close()
    out = {"mode": "engine-loop", "kernel": name, "steps": steps, "warmup": warmup,
           "ms_per_step_api": round(1e3 * float(np.median(api)), 4),
           "ms_per_step_device": round(float(np.median(dev)), 4),
           "device_value": round(w.interactions_per_step / (float(np.median(dev)) * 1e-3) / 1e9, 3),
           "force64_ginter": round(f_int / (f_ms * 1e-3) / 1e9, 3) if f_ms > 0 else None}
    print(json.dumps(out), flush=True)
    return out


def err_stats(got, ref, keys):
    import oracle
    out = {}
    for k in keys:
        e = oracle.rel_err(got[k], ref[k])
        out[k] = dict(median=float(np.median(e)), p90=float(np.percentile(e, 90)), max=float(e.max()))
    return out


def check_nn_lists(res, ids_j, nn_addr, nn_ds2, cnt, lst, with_lists):
    """Nearest-neighbour ids (g6calc_lasthalf2_'s inn) and the lists read through
    g6_get_neighbour_list_ against the oracle: mismatch counts."""
    import oracle
    nnb, _, _ = oracle.api_nn(nn_addr, nn_ds2, ids_j)
    ns = len(nnb)
    nn_bad = int(np.sum(res["nnb"][:ns] != nnb))
    lb = 0
    if with_lists:
        offs = np.concatenate([[0], np.cumsum(np.minimum(res["cnt"], 256))])
        for i in range(ns):
            c = res["cnt"][i]
            if c != cnt[i] or (c < 256 and not np.array_equal(res["list"][offs[i]:offs[i] + c], lst[i, :c])):
                lb += 1
    return dict(nn_mismatch=nn_bad, ngb_list_mismatch=lb if with_lists else None,
                ngb_mean=float(np.mean(cnt)) if with_lists else None)


def timed_cpu(fn, budget_s, max_reps=CPU_SAMPLE_MAX_REPS):
    reps, t, out = 0, 0.0, None
    while (t < budget_s and reps < max_reps) or reps == 0:
        t0 = time.perf_counter()
        r = fn()
        t += time.perf_counter() - t0
        reps += 1
        out = out if out is not None else r
    return out, reps, t


def accuracy_and_cpu(sample_idx, xi, vi, ids_i, h2i, pj, vj, m, ids, eps2, res_by_prec, nthreads, budget_s,
                     with_lists):
    """FP64 oracle on the sample (timed on the host cores) and both kernels' errors and
    NN/list mismatches against it."""
    import oracle
    ref, reps, cpu_s = timed_cpu(
        lambda: oracle.force(1, xi, vi, ids_i, pj, vj, m, ids, eps2, h2=h2i if with_lists else None,
                             nthreads=nthreads), budget_s)
    mix = oracle.neighbours_mixed(xi, ids_i, pj, ids, h2i if with_lists else np.zeros(len(xi)),
                                  nthreads=nthreads, lists=with_lists)
    acc, checks = {}, {}
    ns = len(xi)
    for name, res in res_by_prec.items():
        got = {k: res[k][:ns] for k in ("acc", "jrk", "pot")}
        acc[name] = err_stats(got, ref, ("acc", "jrk", "pot"))
        src = mix if name == "mixed" else ref
        checks[name] = check_nn_lists(res, ids, src["nn_addr"], src["nn_ds2"], src["ngb_cnt"], src["ngb_list"],
                                      with_lists)
    cpu = dict(value=reps * ns * len(pj) / cpu_s / 1e9, unit="Ginteractions/s", cores=nthreads, seconds=cpu_s,
               passes=reps, sample_i=ns)
    return acc, checks, cpu


def g6_config(name, tmp, n, ids, m, x, v, a, j, tj, h2, px, pv, blocks, eps2, resend, ngb, steps, warmup,
              sample_block, sample_n, nthreads, cpu_budget):
    """One configuration through the g6 driver in both precisions, with its accuracy, NN/list
    checks and CPU sample rate."""
    wl = os.path.join(tmp, f"{name}.wl")
    write_workload(wl, n, ids, m, x, v, a, j, tj, h2, px, pv, blocks, eps2, resend, ngb)
    runs, res = {}, {}
    for prec, _ in PRECS:
        rp = os.path.join(tmp, f"{name}.{prec}.res")
        runs[prec] = measure(wl, rp, prec, steps, warmup, sample_block)
        res[prec] = read_results(rp)
        log(f"{name} {prec}: {runs[prec]['value']:.1f} Ginter/s API, {runs[prec]['device_value'] or 0:.1f} device")
    assert np.array_equal(res["mixed"]["ppos"], res["fp64"]["ppos"]), "both kernels must see the same predicted j"
    sel = np.asarray(blocks[sample_block])[:sample_n]
    acc, checks, cpu = accuracy_and_cpu(sel, px[sel], pv[sel], ids[sel], h2[sel], res["mixed"]["ppos"],
                                        res["mixed"]["pvel"], m, ids, eps2, res, nthreads, cpu_budget, ngb)
    return runs, acc, checks, cpu


def config_line(runs, acc, checks, cpu):
    mx, fd = runs["mixed"], runs["fp64"]
    return {"mixed": round(mx["value"], 2), "fp64": round(fd["value"], 2),
            "mixed_over_fp64": round(mx["value"] / fd["value"], 3),
            "mixed_device": round(mx["device_value"] or 0, 2), "fp64_device": round(fd["device_value"] or 0, 2),
            "cpu_fp64": round(cpu["value"], 3), "cpu_cores": cpu["cores"],
            "cpu_sample": f"{cpu['passes']} x {cpu['sample_i']} i", "errors": acc, "checks": checks,
            "blocks": {p: [{k: b[k] for k in ("ni", "calls", "api_ms", "device_ms")} for b in runs[p]["blocks"]]
                       for p in runs}}


def extra_h6_king(nthreads, cpu_budget):
    """Hermite-6, King W0 = 6, N = 262144, eps2 = 1e-6: all-i sweep (16 calls of 16384) through
    the engine handle API, end-to-end wall time and device windows; errors of acc, jerk, snap
    and pot on a 1024-particle sample against the FP64 oracle."""
    import oracle
    n = 262144
    m, x, v = ics.king(n, 6.0, seed=81)
    ids = np.arange(n, dtype=np.int32)
    z = np.zeros((n, 3))
    sample = np.sort(np.random.Generator(np.random.PCG64(82)).choice(n, 1024, replace=False))
    out, res = {}, {}
    for name, prec in PRECS:
        with sa.Engine(sa.SIXTH, prec) as eng:
            eng.set_timing(True)
            eng.set_j_particles(np.arange(n), ids, m, x, v, z, z, z, z)
            eng.set_time(0.0)
            walls, devs = [], []
            for rep in range(3):
                t0 = time.perf_counter()
                dev = 0.0
                for s0 in range(0, n, sa.NPIPES):
                    sl = slice(s0, s0 + sa.NPIPES)
                    eng.start_grav_calc(n, ids[sl], x[sl], v[sl], 1e-6, None, ai=z[sl])
                    r = eng.get_grav_results(n, len(ids[sl]), ngb=True)
                    dev += eng.last_device_ms()
                    if rep == 0:
                        for k in ("acc", "jrk", "snp", "pot", "nnb"):
                            res.setdefault(name, {}).setdefault(k, []).append(r[k])
                walls.append(time.perf_counter() - t0)
                devs.append(dev)
            out[name] = dict(value=n * n / np.median(walls[1:]) / 1e9, device=n * n / (np.median(devs[1:]) * 1e-3) / 1e9)
        res[name] = {k: np.concatenate(v)[sample] for k, v in res[name].items()}
    ref, reps, cpu_s = timed_cpu(lambda: oracle.force(2, x[sample], v[sample], ids[sample], x, v, m, ids, 1e-6,
                                                      ai=z[sample], aj=z, nthreads=nthreads), cpu_budget)
    errs = {name: err_stats(res[name], ref, ("acc", "jrk", "snp", "pot")) for name in res}
    mix = oracle.neighbours_mixed(x[sample], ids[sample], x, ids, np.zeros(len(sample)), nthreads=nthreads,
                                  lists=False)
    checks = {}
    for name, src in (("mixed", mix), ("fp64", ref)):
        nnb, _, _ = oracle.api_nn(src["nn_addr"], src["nn_ds2"], ids)
        checks[name] = dict(nn_mismatch=int(np.sum(res[name]["nnb"] != nnb)))
    return {"mixed": round(out["mixed"]["value"], 2), "fp64": round(out["fp64"]["value"], 2),
            "mixed_over_fp64": round(out["mixed"]["value"] / out["fp64"]["value"], 3),
            "mixed_device": round(out["mixed"]["device"], 2), "fp64_device": round(out["fp64"]["device"], 2),
            "flop_per_interaction": 100, "cpu_fp64": round(reps * len(sample) * n / cpu_s / 1e9, 3),
            "cpu_sample": f"{reps} x {len(sample)} i", "errors": errs, "checks": checks}


def extra_ensemble(nthreads, cpu_budget):
    """512 Plummer clusters, N log-uniform in [256, 16384] (seed 4), eps2 = 1e-4: eight calls
    (active fractions 2^0..2^-7 of every cluster), one ensemble launch each, through the
    engine handle API, forces and nearest neighbours (no neighbour lists, as the
    configuration has none); errors and NN checks on the first two i of every cluster in the
    all-active call against the FP64 oracle, and the lists of those i-particles from one
    further, untimed all-active call with neighbour radii."""
    import oracle
    sizes = ics.ensemble_sizes(512, 256, 16384, seed=4)
    cl = [ics.plummer(int(k), seed=1000 + c) for c, k in enumerate(sizes)]
    h2c = [ics.neighbour_h2(c[1], 32.0) for c in cl]
    rng = np.random.Generator(np.random.PCG64(5))
    fracs = [2.0 ** -k for k in range(8)]
    picks = [[np.sort(rng.choice(int(k), max(1, int(round(k * f))), replace=False)) for k in sizes] for f in fracs]
    calls = []
    for pk in picks:
        cidx = np.concatenate([np.full(len(p), c, np.int32) for c, p in enumerate(pk)])
        X = np.concatenate([cl[c][1][p] for c, p in enumerate(pk)])
        V = np.concatenate([cl[c][2][p] for c, p in enumerate(pk)])
        H = np.concatenate([h2c[c][p] for c, p in enumerate(pk)])
        I = np.concatenate([p for p in pk]).astype(np.int32)
        calls.append((cidx, I, X, V, H, sum(len(p) * int(sizes[c]) for c, p in enumerate(pk))))
    inter = sum(c[5] for c in calls)
    eps = np.full(len(sizes), 1e-4)
    out, kept = {}, {}
    for name, prec in PRECS:
        with sa.Engine(sa.FOURTH, prec) as eng:
            eng.set_timing(True)
            offs = eng.set_clusters(sizes)
            for c, (mc, xc, vc) in enumerate(cl):
                eng.set_j_particles(offs[c] + np.arange(len(mc)), np.arange(len(mc)), mc, xc, vc)
            eng.set_time(0.0)
            walls, devs = [], []
            buf = None   # result arrays kept from call to call, as a compiled caller keeps them
            first = np.concatenate([[0], np.cumsum(sizes)[:-1]])
            smp = np.concatenate([first + q for q in (0, 1)])
            for rep in range(3):
                t0 = time.perf_counter()
                dev = 0.0
                for k, (cidx, I, X, V, H, _) in enumerate(calls):
                    eng.start_grav_calc_clusters(cidx, I, X, V, eps)
                    # a compiled caller's arrays: dsmin untouched where no neighbour is found
                    r = eng.get_grav_results(1, len(I), ngb=True, out=buf, keep_dsmin=buf is not None)
                    if buf is None:
                        buf = r      # the first call is the largest (all clusters fully active)
                    dev += eng.last_device_ms()
                    if rep == 0 and k == 0:
                        kept[name] = dict(smp=smp, res={q: r[q][smp] for q in ("acc", "jrk", "pot", "nnb")})
                walls.append(time.perf_counter() - t0)
                devs.append(dev)
            cidx, I, X, V, H, _ = calls[0]
            eng.start_grav_calc_clusters(cidx, I, X, V, eps, H)
            eng.get_grav_results(1, len(I), ngb=True)
            eng.read_ngb_list()
            kept[name]["lists"] = [eng.get_ngb_list(int(i)) for i in smp]
            out[name] = dict(value=inter / np.median(walls[1:]) / 1e9, device=inter / (np.median(devs[1:]) * 1e-3) / 1e9)
    # oracle per cluster on its two sampled i-particles
    errs_in = {name: {q: [] for q in ("acc", "jrk", "pot")} for name in out}
    ref_all = {q: [] for q in ("acc", "jrk", "pot")}
    bad = {name: dict(nn_mismatch=0, ngb_list_mismatch=0) for name in out}
    t_cpu, n_int = 0.0, 0
    for c, (mc, xc, vc) in enumerate(cl):
        ii = np.array([0, 1])
        ids_c = np.arange(len(mc), dtype=np.int32)
        t0 = time.perf_counter()
        ref = oracle.force(1, xc[ii], vc[ii], ids_c[ii], xc, vc, mc, ids_c, 1e-4, h2=h2c[c][ii], nthreads=nthreads)
        t_cpu += time.perf_counter() - t0
        n_int += 2 * len(mc)
        mix = oracle.neighbours_mixed(xc[ii], ids_c[ii], xc, ids_c, h2c[c][ii], nthreads=nthreads)
        for q in ref_all:
            ref_all[q].append(ref[q])
        for name in out:
            smp_pos = [c, len(sizes) + c]   # the sample is [first i of each cluster | second i of each]
            src = mix if name == "mixed" else ref
            nnb, _, _ = oracle.api_nn(src["nn_addr"], src["nn_ds2"], ids_c)
            for k, sp in enumerate(smp_pos):
                for q in ("acc", "jrk", "pot"):
                    errs_in[name][q].append(kept[name]["res"][q][sp])
                bad[name]["nn_mismatch"] += int(kept[name]["res"]["nnb"][sp] != nnb[k])
                lst, cnt, _ = kept[name]["lists"][sp]
                if cnt != src["ngb_cnt"][k] or (cnt < 256 and not np.array_equal(lst, src["ngb_list"][k, :cnt])):
                    bad[name]["ngb_list_mismatch"] += 1
    ref_cat = {q: np.concatenate(ref_all[q]) for q in ref_all}
    # reorder the reference into [first i of each cluster | second i of each] sample order
    ref_cat = {q: np.concatenate([v[0::2], v[1::2]]) for q, v in ref_cat.items()}
    errs = {}
    for name in out:
        got = {q: np.array(errs_in[name][q]).reshape(len(sizes), 2, *np.shape(errs_in[name][q][0]))
               for q in errs_in[name]}
        got = {q: np.concatenate([v[:, 0], v[:, 1]]) for q, v in got.items()}
        errs[name] = err_stats(got, ref_cat, ("acc", "jrk", "pot"))
    return {"mixed": round(out["mixed"]["value"], 2), "fp64": round(out["fp64"]["value"], 2),
            "mixed_over_fp64": round(out["mixed"]["value"] / out["fp64"]["value"], 3),
            "mixed_device": round(out["mixed"]["device"], 2), "fp64_device": round(out["fp64"]["device"], 2),
            "particles": int(sizes.sum()), "launches_per_sweep": len(fracs), "interactions_per_sweep": int(inter),
            "cpu_fp64": round(n_int / t_cpu / 1e9, 3), "cpu_sample": f"2 i per cluster x {len(sizes)} clusters",
            "errors": errs, "checks": bad}


def copy_calibration(mb):
    """Three device-to-device copies of `mb` MiB through the HIP runtime: the known-size
    reference that scales rocprofv3's FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = int(mb) << 20
    a = ctypes.c_void_p(); b = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(n)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(n)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3) == 0   # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a); hip.hipFree(b)
    print(json.dumps({"copy_calibration_bytes": n, "copies": 3}), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=N_DEFAULT)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline / accuracy leg")
    ap.add_argument("--no-extras", action="store_true", help="skip the other configurations")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MB",
                    help="only run the known-size copy used to calibrate traffic counters")
    ap.add_argument("--only", choices=["mixed", "fp64"], default=None,
                    help="run one kernel only (profiling)")
    ap.add_argument("--engine-loop", action="store_true",
                    help="run the block-step sweep through the engine handle API in this process only")
    ap.add_argument("--write-workload", metavar="PATH", default=None,
                    help="only write the block-step workload file for the g6 driver (to profile it directly)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this engine drives one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU j-split is a design note only)", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.write_workload:
        w = Workload(args.n, SEED)
        write_workload(args.write_workload, w.n, w.ids, w.m, w.x, w.v, w.acc, w.jrk, w.tj, w.h2, w.px, w.pv,
                       w.blocks, EPS2, w.resend, w.ngb)
        print(json.dumps({"workload": args.write_workload, "blocks": len(w.blocks),
                          "sample_block": len(w.blocks) - 1}), flush=True)
        return
    steps, warmup = max(1, args.steps), max(0, args.warmup)
    if args.engine_loop:
        w = Workload(args.n, SEED)
        for name, prec in PRECS:
            if not args.only or args.only == name:
                run_engine_loop(w, name, prec, steps, warmup)
        return
    nthreads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    tmp = tempfile.mkdtemp(prefix="sapamd_bench_")
    oracle_build = None
    if not args.no_cpu:
        import oracle
        oracle_build = oracle.build_native(tmp) or "oracle/Makefile (portable x86-64-v3 build)"
    t_setup = time.perf_counter()
    w = Workload(args.n, SEED)
    wl = os.path.join(tmp, "c2.wl")
    write_workload(wl, w.n, w.ids, w.m, w.x, w.v, w.acc, w.jrk, w.tj, w.h2, w.px, w.pv, w.blocks, EPS2,
                   w.resend, w.ngb)
    log(f"setup {time.perf_counter() - t_setup:.1f}s; interactions/step {w.interactions_per_step:.3e}")
    sample_block = len(w.blocks) - 1
    res, out = {}, {}
    for name, _ in PRECS:
        if args.only and args.only != name:
            continue
        rp = os.path.join(tmp, f"c2.{name}.res")
        res[name] = measure(wl, rp, name, steps, warmup, sample_block)
        out[name] = read_results(rp)
        log(f"{name}: value {res[name]['value']:.1f} Ginter/s ({res[name]['ms_per_step']:.2f} ms/step API, "
            f"{res[name]['device_ms_per_step']:.2f} ms device), host {res[name]['host_ms_per_step']}")
    cpu = accuracy = checks = None
    if not args.no_cpu and "mixed" in out and "fp64" in out:
        assert np.array_equal(out["mixed"]["ppos"], out["fp64"]["ppos"]), "both kernels must see the same predicted j"
        sel = np.asarray(w.blocks[sample_block])[:sa.NPIPES]
        accuracy, checks, c = accuracy_and_cpu(sel, w.px[sel], w.pv[sel], w.ids[sel], w.h2[sel],
                                               out["mixed"]["ppos"], out["mixed"]["pvel"], w.m, w.ids, EPS2,
                                               out, nthreads, CPU_SAMPLE_S, True)
        cpu = {"value": c["value"], "unit": "Ginteractions/s", "cores": nthreads, "kind": "port",
               "cpu_model": cpu_model(), "build": oracle_build,
               "sample": f"{c['passes']} passes of {c['sample_i']} i-particles x {w.n} j-particles (the device's "
                         f"predicted j-state), FP64 acc/jerk/pot + NN + neighbour lists (list bookkeeping "
                         f"included), oracle/oracle.c orc_force, {c['seconds']:.2f} s"}
    main_name = "mixed" if "mixed" in res else "fp64"
    r = res[main_name]
    line = {
        "metric": METRIC,
        "value": round(r["value"], 3),
        "unit": "Ginteractions/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": round(r["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pair terms, f64 positions/accumulation" if main_name == "mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": f"Plummer N={w.n} Hermite-4 block steps, active 2^0..2^{len(w.counts) - 1}, "
                               f"eps2={EPS2}, neighbour lists (mean ~32, cap 256) + NN, through the g6 C-ABI",
                   "n": w.n, "eps2": EPS2, "active_counts": w.counts, "npipes": sa.NPIPES,
                   "kernel": main_name, "timed": "g6_set_ti_ + g6calc_firsthalf_ .. g6calc_lasthalf2_ + "
                                                 "g6_read_neighbour_list_, median of the timed steps run with the "
                                                 "engine's instrumentation off (device window and kernel times "
                                                 "from instrumented steps interleaved with them)",
                   "driver": "sapporo2_amd/bench/g6_blocksteps"},
        "roofline": {"bound": "valu", "achieved": round(r["achieved_tflops"], 3),
                     "peak": r["peak"], "unit": "TFLOP/s",
                     "frac": round(r["achieved_tflops"] / r["peak"], 4),
                     "traffic": None,
                     "traffic_note": "not measured for the current launch shapes (rocprofv3 counter collection "
                                     "ends in a host SIGSEGV in a profiler thread with this build); last "
                                     "measurement, at S = 256 j-splits per 16384-particle call: "
                                     f"{measured_traffic()} B per launch, "
                                     f"{os.path.relpath(TRAFFIC_PROFILE, ROOT)}",
                     "kernel": "k_force<H4,%s,P=64>" % main_name,
                     "mean_launch_ms": round(r["force_mean_ms"], 4),
                     "launches": r["force_launches"],
                     "flop_per_interaction": FLOP_PER_INTERACTION},
        "cpu_baseline": cpu,
        "device_value": round(r["device_value"], 3) if r["device_value"] else None,
        "instrumented_value": round(r["instrumented_value"], 3),
        "device_ms_per_step": round(r["device_ms_per_step"], 4),
        "host_ms_per_step": r["host_ms_per_step"],
        "outside_api_ms_per_step": r["outside_api_ms_per_step"],
        "blocks": r["blocks"],
        "predictor": r["predictor"],
        "scatter": r["scatter"],
    }
    if "fp64" in res and main_name == "mixed":
        d = res["fp64"]
        line["fp64_value"] = round(d["value"], 3)
        line["mixed_over_fp64"] = round(r["value"] / d["value"], 3)
        line["fp64_device_value"] = round(d["device_value"], 3) if d["device_value"] else None
        line["fp64_roofline"] = {"achieved": round(d["achieved_tflops"], 3), "peak": d["peak"],
                                 "frac": round(d["achieved_tflops"] / d["peak"], 4),
                                 "mean_launch_ms": round(d["force_mean_ms"], 4)}
    if accuracy:
        line["accuracy"] = accuracy
        line["checks"] = checks
    if not args.no_extras and not args.only:
        cfgs = {}
        budget = 2.0
        # C1: Plummer 8192, eps2 = 1e-4, one all-i evaluation through the g6 ABI (no lists)
        m, x, v = ics.plummer(8192, seed=11)
        ids = np.arange(8192, dtype=np.int32)
        z = np.zeros((8192, 3))
        cfgs["c1_plummer8192"] = config_line(*g6_config(
            "c1", tmp, 8192, ids, m, x, v, z, z, np.zeros(8192), np.zeros(8192), x, v, [ids], 1e-4, False,
            False, 5, 1, 0, 8192, nthreads, budget))
        log(f"c1: {cfgs['c1_plummer8192']}")
        cfgs["c3_h6_king262144"] = extra_h6_king(nthreads, budget)
        log(f"c3: {cfgs['c3_h6_king262144']}")
        cfgs["c4_ensemble512"] = extra_ensemble(nthreads, budget)
        log(f"c4: {cfgs['c4_ensemble512']}")
        # C5: Plummer 2^20 + 1% central mass at the origin, eps2 = 1e-8, t_j on 28 levels down to
        # 2^-30, active sets: the central mass and its 1023 nearest stars, 32768 random, all
        nstar = 1048576
        m, x, v = ics.plummer_bh(nstar, seed=91, mbh=0.01)
        n = nstar + 1
        ids = np.arange(n, dtype=np.int32)
        h2 = np.concatenate([ics.neighbour_h2(x[:nstar], 32.0, mtot=0.99), [1e-4]])
        a, j = init_derivatives(m, x, v, ids, 1e-8)
        rng = np.random.Generator(np.random.PCG64(92))
        lev = ics.block_levels(n, 28, seed=93)
        tj = -rng.uniform(0.0, 1.0, n) * 2.0 ** -(3.0 + lev)
        px, pv = host_predict(0.0, tj, x, v, a, j)
        near = np.sort(np.argsort(np.sum(x * x, axis=1))[:1024])
        blocks = [near, np.sort(rng.choice(n, 32768, replace=False)), np.arange(n, dtype=np.int32)]
        cfgs["c5_plummer2e20_central_mass"] = config_line(*g6_config(
            "c5", tmp, n, ids, m, x, v, a, j, tj, h2, px, pv, blocks, 1e-8, False, True, 2, 1, 0, 1024,
            nthreads, budget))
        log(f"c5: {cfgs['c5_plummer2e20_central_mass']}")
        line["configs"] = cfgs
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
