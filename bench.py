#!/usr/bin/env python3
"""Hermite-4 block-step force throughput on one MI355X (Ginteractions/s).

Workload (one "step" = one sweep of 18 block steps): Plummer N = 131072, equal masses,
Henon units, eps2 = 1e-6; j-particles carry acc/jerk and times t_j on 20 power-of-two
levels; each block step re-sends the previous active set (the corrector's update),
sets the block time and evaluates acc, jerk, pot, nearest neighbour and neighbour
lists (mean ~32, cap 256) on an active set of 2^k particles, k = 0..17, through the
engine C-ABI in calls of <= 16384 i-particles (the g6 firsthalf/lasthalf sequence).

value       = sum(ni * nj) / device time, inputs resident in HBM: each call's device
              window runs from the first kernel after its host->device copies to the end
              of the finalize kernel (scatter, predictor, force, finalize), HIP events on
              the engine's stream.
api_value   = the same interactions over the wall time of firsthalf..lasthalf2 plus the
              neighbour-list read (PCIe and host included).
roofline    = dominant kernel (force, P = 64 lane shape): 60 flop per interaction
              (reference convention) over its mean launch time, HIP events on its stream.
cpu_baseline= the FP64 direct summation of the oracle (OpenMP) on a sample of 16384
              i-particles against all j, timed on this host's cores; the same sample is
              the accuracy reference for the reported errors.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import sapporo2_amd as sa  # noqa: E402
from sapporo2_amd import ics  # noqa: E402

METRIC = "Ginteractions/s Hermite-4 N=128K, mixed vs FP64 GPU at rel. force err ≤1e-6"
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "force_h4_mixed_traffic.md")


def measured_traffic():
    """HBM bytes per launch of the mixed P = 64 force kernel over one step, as measured
    by rocprofv3 FETCH_SIZE / WRITE_SIZE (calibrated on a known-size copy) and recorded
    in the committed profile summary; None when that summary is absent."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(line.split(":")[1])
    except OSError:
        pass
    return None
FLOP_PER_INTERACTION = 60.0          # test_gravity_block.cpp:271
PEAK_TFLOPS = {sa.DOUBLESINGLE: 157.3, sa.DOUBLE: 78.6}   # MI355X FP32 / FP64 vector
EPS2 = 1e-6
N_DEFAULT = 131072
SEED = 2024
CPU_SAMPLE_S = 10.0        # CPU baseline: repeat the oracle sample until this much time has elapsed
CPU_SAMPLE_MAX_REPS = 40


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_predict(t, tj, x, v, a, j):
    """Caller-side (integrator) prediction of the i-particles to the block time."""
    dt = (t - tj)[:, None]
    px = x + dt * (v + dt * 0.5 * (a + dt * (1.0 / 3.0) * j))
    pv = v + dt * (a + dt * 0.5 * j)
    return px, pv


class Workload:
    def __init__(self, n, seed):
        self.n = n
        self.m, self.x, self.v = ics.plummer(n, seed=seed)
        self.ids = np.arange(n, dtype=np.int32)
        self.h2 = ics.neighbour_h2(self.x, 32.0)
        rng = np.random.Generator(np.random.PCG64(seed + 1))
        lev = ics.block_levels(n, 20, seed=seed + 2)
        self.tj = -rng.uniform(0.0, 1.0, n) * 2.0 ** -(3.0 + lev)
        self.counts = [2 ** k for k in range(18) if 2 ** k <= n]
        self.subsets = ics.active_subsets(n, self.counts, seed=seed + 3)
        self.acc = np.zeros((n, 3))
        self.jrk = np.zeros((n, 3))

    def init_derivatives(self):
        """Initial acc/jerk of every particle from an all-i evaluation (FP64 kernel)."""
        with sa.Engine(sa.FOURTH, sa.DOUBLE) as eng:
            eng.set_j_particles(np.arange(self.n), self.ids, self.m, self.x, self.v)
            eng.set_time(0.0)
            r = eng.calc(self.n, self.ids, self.x, self.v, EPS2, None, ngb=False)
        self.acc, self.jrk = r["acc"], r["jrk"]
        self.px, self.pv = host_predict(0.0, self.tj, self.x, self.v, self.acc, self.jrk)

    @property
    def interactions_per_step(self):
        return sum(self.counts) * self.n


def run_engine(w: Workload, prec, steps, warmup, keep_chunk=None):
    """Runs warmup + timed sweeps; returns timing records and the kept results."""
    eng = sa.Engine(sa.FOURTH, prec)
    eng.set_timing(True)
    n = w.n
    eng.set_j_particles(np.arange(n), w.ids, w.m, w.x, w.v, w.acc, w.jrk, tj=w.tj)
    recs = []
    kept = {}
    t_wall = 0.0
    for step in range(warmup + steps):
        timed = step >= warmup
        for b, sel in enumerate(w.subsets):
            prev = w.subsets[b - 1]
            eng.set_j_particles(prev, w.ids[prev], w.m[prev], w.px[prev], w.pv[prev],
                                w.acc[prev], w.jrk[prev], tj=0.0)
            eng.set_time(0.0)
            for s in range(0, len(sel), sa.NPIPES):
                ch = sel[s:s + sa.NPIPES]
                ni = len(ch)
                t0 = time.perf_counter()
                eng.start_grav_calc(n, w.ids[ch], w.px[ch], w.pv[ch], EPS2, w.h2[ch])
                res = eng.get_grav_results(n, ni, ngb=True)
                eng.read_ngb_list()
                t1 = time.perf_counter()
                if timed:
                    P, S, jc = eng.last_config()
                    recs.append(dict(ni=ni, P=P, S=S, dev_ms=eng.last_device_ms(),
                                     force_ms=eng.last_force_ms(), pred_ms=eng.last_predict_ms(),
                                     api_s=t1 - t0))
                    t_wall += t1 - t0
                    if step == warmup + steps - 1 and ni == sa.NPIPES and b == len(w.subsets) - 1:
                        kept.setdefault("full", {})[s] = res
                        if s == 0 and keep_chunk is not None:
                            got = [eng.get_ngb_list(i) for i in range(ni)]
                            kept["sample"] = dict(sel=ch, res=res, lists=[g[0] for g in got],
                                                  cnts=np.array([g[1] for g in got]))
    # predicted j state actually used on the device (for bit-exact checks)
    if keep_chunk is not None:
        pj = np.empty((n, 3))
        for a in range(n):
            pj[a] = eng.retrieve_predicted_j_particle(a)["pos"]
        kept["pj"] = pj
    eng.close()
    return recs, kept


def summarize(recs, w, prec, steps):
    inter = np.array([r["ni"] * w.n for r in recs], dtype=np.float64)
    dev = np.array([r["dev_ms"] for r in recs])
    api = np.array([r["api_s"] for r in recs])
    value = inter.sum() / (dev.sum() * 1e-3) / 1e9
    api_value = inter.sum() / api.sum() / 1e9
    dom = [r for r in recs if r["P"] == 64]
    f_int = sum(r["ni"] * w.n for r in dom)
    f_ms = sum(r["force_ms"] for r in dom)
    achieved = FLOP_PER_INTERACTION * f_int / (f_ms * 1e-3) / 1e12
    pred = [r["pred_ms"] for r in recs if r["pred_ms"] > 0]
    return dict(value=value, api_value=api_value, ms_per_step=dev.sum() / steps,
                api_ms_per_step=api.sum() * 1e3 / steps,
                force_mean_ms=f_ms / max(1, len(dom)), force_launches=len(dom),
                force_ginter=f_int / (f_ms * 1e-3) / 1e9,
                achieved_tflops=achieved, peak=PEAK_TFLOPS[prec],
                pred_mean_ms=float(np.mean(pred)) if pred else 0.0, pred_launches=len(pred))


def cpu_baseline_and_accuracy(w, kept_mix, kept_d, nthreads):
    """FP64 oracle on the kept sample (timed on the host cores) and the accuracy of
    both GPU kernels against it (the oracle's results are the reference)."""
    import oracle
    smp = kept_mix["sample"]
    sel = smp["sel"]
    pj = kept_mix["pj"]
    assert np.array_equal(pj, kept_d["pj"]), "both engines must see the same predicted j"
    xi, vi = w.px[sel], w.pv[sel]
    # timed over repeated passes of the sample until CPU_SAMPLE_S seconds have elapsed
    reps, cpu_s, ref = 0, 0.0, None
    while cpu_s < CPU_SAMPLE_S and reps < CPU_SAMPLE_MAX_REPS:
        t0 = time.perf_counter()
        r0 = oracle.force(1, xi, vi, w.ids[sel], pj, w.pv, w.m, w.ids, EPS2, h2=w.h2[sel],
                          nthreads=nthreads)
        cpu_s += time.perf_counter() - t0
        reps += 1
        ref = ref if ref is not None else r0
    # the FP64 oracle's j velocities: the device predicted them too; read back not needed for
    # forces at this tolerance (velocities enter jerk only).
    acc = {}
    for name, kept in (("mixed", kept_mix), ("fp64", kept_d)):
        r = kept["sample"]["res"]
        e = {k: oracle.rel_err(r[k], ref[k]) for k in ("acc", "jrk", "pot")}
        acc[name] = {k: dict(median=float(np.median(v)), p90=float(np.percentile(v, 90)),
                             max=float(v.max())) for k, v in e.items()}
    # nearest neighbour and lists, bit-exact under each kernel's distance contract
    mix = oracle.neighbours_mixed(xi, w.ids[sel], pj, w.ids, w.h2[sel], nthreads=nthreads)
    out = {}
    for name, kept, nn_addr, nn_ds2, cnt, lst in (
            ("mixed", kept_mix, mix["nn_addr"], mix["nn_ds2"], mix["ngb_cnt"], mix["ngb_list"]),
            ("fp64", kept_d, ref["nn_addr"], ref["nn_ds2"], ref["ngb_cnt"], ref["ngb_list"])):
        nnb, dsmin, found = oracle.api_nn(nn_addr, nn_ds2, w.ids)
        r = kept["sample"]["res"]
        nn_bad = int(np.sum(r["nnb"] != nnb)) + int(np.sum(r["dsmin"][found] != dsmin[found]))
        lb = 0
        for i in range(len(sel)):
            c = kept["sample"]["cnts"][i]
            if c != cnt[i] or (c < 256 and not np.array_equal(kept["sample"]["lists"][i], lst[i, :c])):
                lb += 1
        out[name] = dict(nn_mismatch=nn_bad, ngb_list_mismatch=lb, ngb_mean=float(cnt.mean()))
    return cpu_s, reps, acc, out


def _timed_calls(eng, fn, reps=2):
    """Runs fn() reps+1 times; returns the median summed device time (ms) of the last reps."""
    out = []
    for r in range(reps + 1):
        out.append(fn())
    return float(np.median(out[1:]))


def extra_workloads(log):
    """The other configurations, one measurement each (device-resident Ginteractions/s,
    mixed and FP64 kernels): Hermite-6 King W0=6 N=262144; a 512-cluster ensemble with
    one launch per block step; Plummer N=2^20 + 1% central mass on three active sets."""
    out = {}
    # Hermite-6, King W0 = 6, N = 262144, all-i (16 calls of 16384)
    n = 262144
    m, x, v = ics.king(n, 6.0, seed=81)
    ids = np.arange(n, dtype=np.int32)
    z = np.zeros((n, 3))
    r6 = {}
    for name, prec in (("mixed", sa.DOUBLESINGLE), ("fp64", sa.DOUBLE)):
        with sa.Engine(sa.SIXTH, prec) as eng:
            eng.set_timing(True)
            eng.set_j_particles(np.arange(n), ids, m, x, v, z, z, z, z)
            eng.set_time(0.0)

            def sweep():
                t = 0.0
                for s0 in range(0, n, sa.NPIPES):
                    sl = slice(s0, s0 + sa.NPIPES)
                    eng.start_grav_calc(n, ids[sl], x[sl], v[sl], 1e-6, None, ai=z[sl])
                    eng.get_grav_results(n, len(ids[sl]), ngb=True)
                    t += eng.last_device_ms()
                return t
            t = _timed_calls(eng, sweep, 1)
            r6[name] = n * n / (t * 1e-3) / 1e9
    out["h6_king_w6_n262144"] = {"mixed": round(r6["mixed"], 2), "fp64": round(r6["fp64"], 2),
                                 "mixed_over_fp64": round(r6["mixed"] / r6["fp64"], 3),
                                 "flop_per_interaction": 100}
    log(f"extra h6 king: {out['h6_king_w6_n262144']}")
    # ensemble: 512 Plummer clusters, N log-uniform in [256, 16384], eps2 = 1e-4
    sizes = ics.ensemble_sizes(512, 256, 16384, seed=4)
    cl = [ics.plummer(int(k), seed=1000 + c) for c, k in enumerate(sizes)]
    rng = np.random.Generator(np.random.PCG64(5))
    fracs = [2.0 ** -k for k in range(8)]
    picks = [[np.sort(rng.choice(int(k), max(1, int(round(k * f))), replace=False)) for k in sizes] for f in fracs]
    re = {}
    for name, prec in (("mixed", sa.DOUBLESINGLE), ("fp64", sa.DOUBLE)):
        with sa.Engine(sa.FOURTH, prec) as eng:
            eng.set_timing(True)
            offs = eng.set_clusters(sizes)
            for c, (mc, xc, vc) in enumerate(cl):
                eng.set_j_particles(offs[c] + np.arange(len(mc)), np.arange(len(mc)), mc, xc, vc)
            eng.set_time(0.0)
            eps = np.full(len(sizes), 1e-4)
            calls = []
            for pk in picks:
                cidx = np.concatenate([np.full(len(p), c, np.int32) for c, p in enumerate(pk)])
                X = np.concatenate([cl[c][1][p] for c, p in enumerate(pk)])
                V = np.concatenate([cl[c][2][p] for c, p in enumerate(pk)])
                I = np.concatenate([p for p in pk]).astype(np.int32)
                inter = sum(len(p) * int(sizes[c]) for c, p in enumerate(pk))
                calls.append((cidx, I, X, V, inter))

            def sweep():
                t = 0.0
                for cidx, I, X, V, inter in calls:
                    eng.start_grav_calc_clusters(cidx, I, X, V, eps)
                    eng.get_grav_results(1, len(I), ngb=True)
                    t += eng.last_device_ms()
                return t
            t = _timed_calls(eng, sweep, 1)
            re[name] = sum(c[4] for c in calls) / (t * 1e-3) / 1e9
            re["interactions"] = sum(c[4] for c in calls)
    out["ensemble_512_plummer"] = {"mixed": round(re["mixed"], 2), "fp64": round(re["fp64"], 2),
                                   "mixed_over_fp64": round(re["mixed"] / re["fp64"], 3),
                                   "particles": int(sizes.sum()), "launches_per_sweep": len(fracs),
                                   "interactions_per_sweep": int(re["interactions"])}
    log(f"extra ensemble: {out['ensemble_512_plummer']}")
    del cl
    # Plummer 2^20 + 1% central mass, eps2 = 1e-8, active sets 1024, 32768, all
    nstar = 1048576
    m, x, v = ics.plummer_bh(nstar, seed=91, mbh=0.01)
    n = nstar + 1
    ids = np.arange(n, dtype=np.int32)
    h2 = np.concatenate([ics.neighbour_h2(x[:nstar], 32.0, mtot=0.99), [1e-4]])
    rng = np.random.Generator(np.random.PCG64(92))
    sets = {"1024": np.sort(rng.choice(n, 1024, replace=False)),
            "32768": np.sort(rng.choice(n, 32768, replace=False)), "all": np.arange(n)}
    rb = {}
    for name, prec in (("mixed", sa.DOUBLESINGLE), ("fp64", sa.DOUBLE)):
        with sa.Engine(sa.FOURTH, prec) as eng:
            eng.set_timing(True)
            eng.set_j_particles(np.arange(n), ids, m, x, v)
            eng.set_time(0.0)
            for key, sel in sets.items():
                def sweep():
                    t = 0.0
                    for s0 in range(0, len(sel), sa.NPIPES):
                        ch = sel[s0:s0 + sa.NPIPES]
                        eng.start_grav_calc(n, ids[ch], x[ch], v[ch], 1e-8, h2[ch])
                        eng.get_grav_results(n, len(ch), ngb=True)
                        t += eng.last_device_ms()
                    return t
                t = _timed_calls(eng, sweep, 1 if key == "all" else 2)
                rb.setdefault(key, {})[name] = round(len(sel) * n / (t * 1e-3) / 1e9, 2)
    for key in rb:
        rb[key]["mixed_over_fp64"] = round(rb[key]["mixed"] / rb[key]["fp64"], 3)
    out["plummer_2e20_central_mass"] = rb
    log(f"extra plummer+BH: {rb}")
    return out


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
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=N_DEFAULT)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline / accuracy leg")
    ap.add_argument("--no-extras", action="store_true", help="skip the other configurations")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MB",
                    help="only run the known-size copy used to calibrate traffic counters")
    ap.add_argument("--only", choices=["mixed", "fp64"], default=None,
                    help="run one kernel only (profiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this engine drives one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU j-split is a design note only)", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    steps, warmup = max(1, args.steps), max(0, args.warmup)
    t_setup = time.perf_counter()
    w = Workload(args.n, SEED)
    w.init_derivatives()
    log(f"setup {time.perf_counter() - t_setup:.1f}s; interactions/step {w.interactions_per_step:.3e}")
    keep = None if args.no_cpu else 0
    res = {}
    kept = {}
    for name, prec in (("mixed", sa.DOUBLESINGLE), ("fp64", sa.DOUBLE)):
        if args.only and args.only != name:
            continue
        recs, k = run_engine(w, prec, steps, warmup, keep_chunk=keep)
        res[name] = summarize(recs, w, prec, steps)
        kept[name] = k
        log(f"{name}: {res[name]}")
    # full-size property: mixed vs FP64 GPU over all N i-particles (largest bucket)
    full_cmp = None
    if "mixed" in kept and "fp64" in kept and "full" in kept["mixed"]:
        import numpy.linalg  # noqa: F401
        da, dd = kept["mixed"]["full"], kept["fp64"]["full"]
        errs = []
        for s in sorted(da):
            a, b = da[s]["acc"], dd[s]["acc"]
            errs.append(np.sqrt(((a - b) ** 2).sum(1) / (b ** 2).sum(1)))
        e = np.concatenate(errs)
        full_cmp = dict(n=int(e.size), median=float(np.median(e)), p90=float(np.percentile(e, 90)))
    cpu = None
    accuracy = checks = None
    if not args.no_cpu and "mixed" in kept and "fp64" in kept:
        nthreads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        cpu_s, reps, accuracy, checks = cpu_baseline_and_accuracy(w, kept["mixed"], kept["fp64"], nthreads)
        ni_s = len(kept["mixed"]["sample"]["sel"])
        cpu = {"value": reps * ni_s * w.n / cpu_s / 1e9, "unit": "Ginteractions/s", "cores": nthreads,
               "kind": "port",
               "sample": f"{reps} passes of {ni_s} i-particles x {w.n} j-particles, FP64 acc/jerk/pot + NN + "
                         f"neighbour lists, oracle/oracle.c orc_force, {cpu_s:.2f} s"}
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
                               f"eps2={EPS2}, neighbour lists (mean ~32, cap 256) + NN",
                   "n": w.n, "eps2": EPS2, "active_counts": w.counts, "npipes": sa.NPIPES,
                   "kernel": main_name},
        "roofline": {"bound": "valu", "achieved": round(r["achieved_tflops"], 3),
                     "peak": r["peak"], "unit": "TFLOP/s",
                     "frac": round(r["achieved_tflops"] / r["peak"], 4),
                     "traffic": measured_traffic() if main_name == "mixed" else None,
                     "traffic_source": os.path.relpath(TRAFFIC_PROFILE, ROOT),
                     "kernel": "k_force<H4,%s,P=64>" % main_name,
                     "mean_launch_ms": round(r["force_mean_ms"], 4),
                     "launches": r["force_launches"],
                     "flop_per_interaction": FLOP_PER_INTERACTION},
        "cpu_baseline": cpu,
        "api_value": round(r["api_value"], 3),
        "api_ms_per_step": round(r["api_ms_per_step"], 3),
    }
    if "fp64" in res and main_name == "mixed":
        d = res["fp64"]
        line["fp64_value"] = round(d["value"], 3)
        line["mixed_over_fp64"] = round(r["value"] / d["value"], 3)
        line["fp64_roofline"] = {"achieved": round(d["achieved_tflops"], 3), "peak": d["peak"],
                                 "frac": round(d["achieved_tflops"] / d["peak"], 4),
                                 "mean_launch_ms": round(d["force_mean_ms"], 4)}
        line["fp64_api_value"] = round(d["api_value"], 3)
    # predictor algorithmic bytes per j: reads t, x, v, a, j, m (14 doubles) + id; writes the
    # working record (48 B mixed, 64 B FP64) + FP64 predicted position and velocity (mixed
    # only, 48 B, kept for read-back)
    pb = 116 + (48 + 48 if main_name == "mixed" else 64)
    line["predictor"] = {"mean_launch_ms": round(r["pred_mean_ms"], 4), "launches": r["pred_launches"],
                         "bytes_per_j": pb, "n": w.n,
                         "achieved_GBps": round(pb * w.n / (r["pred_mean_ms"] * 1e-3) / 1e9, 1)
                         if r["pred_mean_ms"] > 0 else None}
    if accuracy:
        line["accuracy"] = accuracy
        line["checks"] = checks
    if full_cmp:
        line["mixed_vs_fp64_acc_full_N"] = full_cmp
    if not args.no_extras and not args.only:
        line["extra_workloads"] = extra_workloads(log)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
