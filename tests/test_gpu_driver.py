"""The compiled GRAPE-6 caller that bench.py times (sapporo2_amd/bench/g6_blocksteps): a
small block-step workload (Plummer N = 8192, 20 time levels, active sets 2^0..2^13, the
previous active set re-sent before each block step, neighbour lists) in its interleaved mode,
half the timed steps with the engine's HIP events off and half with them on.  The JSON must
carry both kinds of step, and the sampled results (the all-active block of the last step,
read through g6calc_lasthalf2_ and g6_get_neighbour_list_) must match the oracle: forces
within tolerance, nearest neighbours and lists bit-exact."""
from __future__ import annotations

import os
import sys
import tempfile

import numpy as np
import pytest

import oracle
import sapporo2_amd as sa
from _util import check_forces

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402

pytestmark = pytest.mark.gpu


def test_driver_interleaved_steps_and_parity():
    print("g6_blocksteps: Plummer 8192, interleaved timing", flush=True)
    w = bench.Workload(8192, 31)
    tmp = tempfile.mkdtemp()
    wl = os.path.join(tmp, "w.wl")
    rp = os.path.join(tmp, "w.res")
    bench.write_workload(wl, w.n, w.ids, w.m, w.x, w.v, w.acc, w.jrk, w.tj, w.h2, w.px, w.pv, w.blocks,
                         bench.EPS2, True, True)
    steps = 2
    js = bench.run_driver(wl, rp, "mixed", steps, 1, len(w.blocks) - 1, timing=2)
    assert js["timing"] == 2 and js["steps"] == steps
    for k in ("step_api_s", "step_api_instrumented_s", "step_device_ms", "step_force64_ms", "step_firsthalf_s"):
        assert len(js[k]) == steps, k
    assert all(v > 0 for v in js["step_api_s"] + js["step_api_instrumented_s"] + js["step_device_ms"])
    assert js["force64"]["launches"] > 0 and js["force64"]["ms"] > 0
    assert all(b["device_ms"] > 0 and b["force_ms"] > 0 for b in js["blocks"])
    s = bench.summarize(js, "mixed")
    assert s["value"] > 0 and s["device_value"] > 0

    r = bench.read_results(rp)
    sel = np.asarray(w.blocks[-1])
    assert len(sel) == w.n
    # the predicted j-state the calls used, against the oracle predictor: the first block
    # step of every step re-sends the whole (all-active) previous block at the block time
    pj, vj, _ = oracle.predict(1, 0.0, np.zeros(w.n), w.px, w.pv, w.acc, w.jrk)
    assert np.array_equal(r["ppos"], pj) and np.array_equal(r["pvel"], vj), "predicted j-state"
    smp = np.sort(np.random.Generator(np.random.PCG64(5)).choice(len(sel), 512, replace=False))
    isel = sel[smp]
    ref = oracle.force(1, w.px[isel], w.pv[isel], w.ids[isel], pj, vj, w.m, w.ids, bench.EPS2, h2=w.h2[isel])
    check_forces({"acc": r["acc"][smp], "jrk": r["jrk"][smp], "pot": r["pot"][smp]}, ref, 1, 1e-6,
                 label="driver sample")
    nn = oracle.neighbours_mixed(w.px[isel], w.ids[isel], pj, w.ids, w.h2[isel])
    nnb, _, _ = oracle.api_nn(nn["nn_addr"], nn["nn_ds2"], w.ids)
    assert np.array_equal(r["nnb"][smp], nnb), "nearest neighbours"
    starts = np.concatenate([[0], np.cumsum(np.minimum(r["cnt"], 256))[:-1]])
    bad = 0
    for q, k in enumerate(smp):
        c = int(r["cnt"][k])
        got = r["list"][starts[k]:starts[k] + min(c, 256)]
        if c != nn["ngb_cnt"][q] or (c < 256 and not np.array_equal(got, nn["ngb_list"][q, :c])):
            bad += 1
    assert bad == 0, f"neighbour-list mismatches {bad}"


def test_driver_host_threads_do_not_change_results():
    """The host copies of each call (pack, unpack) on one thread or on the pool: the same
    bytes in every result and predicted-state field of the driver's output and in every
    neighbour list below the 256 cap.  (A list at the cap holds whichever 256 neighbours
    the appends stored first, in any run, as the reference's does.)"""
    print("g6_blocksteps: Plummer 8192, host copies on 1 thread vs the pool", flush=True)
    w = bench.Workload(8192, 32)
    tmp = tempfile.mkdtemp()
    wl = os.path.join(tmp, "w.wl")
    bench.write_workload(wl, w.n, w.ids, w.m, w.x, w.v, w.acc, w.jrk, w.tj, w.h2, w.px, w.pv, w.blocks,
                         bench.EPS2, True, True)
    res = {}
    old = os.environ.get("SAPPORO2_AMD_HOST_THREADS")
    try:
        for t in ("1", "8"):
            os.environ["SAPPORO2_AMD_HOST_THREADS"] = t
            rp = os.path.join(tmp, f"t{t}.res")
            bench.run_driver(wl, rp, "mixed", 1, 0, len(w.blocks) - 1, timing=0)
            res[t] = bench.read_results(rp)
    finally:
        if old is None:
            os.environ.pop("SAPPORO2_AMD_HOST_THREADS", None)
        else:
            os.environ["SAPPORO2_AMD_HOST_THREADS"] = old
    for k in res["1"]:
        if k != "list":
            assert np.array_equal(res["1"][k], res["8"][k]), f"field {k} differs with the host pool"
    cnt = res["1"]["cnt"]
    stored = np.minimum(cnt, 256)
    starts = np.concatenate([[0], np.cumsum(stored)[:-1]])
    below = np.flatnonzero(cnt < 256)
    assert len(below) > len(cnt) // 2
    for i in below:
        a, b = res["1"]["list"], res["8"]["list"]
        s = starts[i]
        assert np.array_equal(a[s:s + cnt[i]], b[s:s + cnt[i]]), f"list of i-particle {i} differs"
