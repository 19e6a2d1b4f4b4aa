"""Launch-shape autotuner for the force path.

For each system size N (a Plummer sphere), integration order, working precision,
neighbour-list mode and active-count bucket ni = 2^k (k = 0 .. 14), it times candidate
lane shapes P (i-particles per wave) and j-splits S, and keeps the shape with the
smallest device time of the whole call (scatter/reset, force, finalize: the engine's
device window).  The result is a text table the engine reads at open(); a call uses the
entry of its bucket at the nearest log2(nj):

    # order precision ngb log2(ni) log2(nj) P log2(j per split)
    1 1 1 14 17 64 10

usage (on the GPU): python -m sapporo2_amd.autotune [--n 8192,131072,262144,1048576] [--out PATH]
The default output is tune/<arch>.txt beside the library, the file the engine loads.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

import sapporo2_amd as sa
from sapporo2_amd import ics


MIN_WAVES = 1024   # four waves per CU of a 256-CU part


def candidates(ni, nj, min_waves=MIN_WAVES):
    """Lane shapes P (the active count's power of two and two narrower ones) and
    power-of-two j-splits S with at least max(64, 16 Q) j per split and between
    min_waves and 262144 waves; the wave floor is dropped when no shape meets it
    (few i-particles against few j)."""
    pnat = 1
    while pnat < min(ni, 64):
        pnat <<= 1
    out = []
    for P in sorted({pnat, max(1, pnat // 2), max(1, pnat // 4)}):
        Q = 64 // P
        G = (ni + P - 1) // P
        for k in range(0, 16):
            S = 1 << k
            jchunk = math.ceil(nj / S)
            if jchunk < max(64, 16 * Q):
                break
            waves = G * S
            if waves < min_waves or waves > 262144:
                continue
            out.append((P, S))
    if not out and min_waves > 1:
        return candidates(ni, nj, 1)
    return out


TIE = 0.01   # shapes within 1 % of the fastest count as equally fast


def partial_rows(S):
    """Partial-sum rows a call writes per i-particle (the 4 waves of a work-group
    pre-reduce 4 consecutive j-splits when S is a multiple of 4)."""
    return S // 4 if S % 4 == 0 else S


def select(cands):
    """The launch shape for one bucket from its measurements [(dev_ms, P, S, jchunk)]: among
    the shapes within TIE of the fastest, the one writing the fewest partial-sum rows per
    i-particle group (the finalize reads them back), then the fastest."""
    best = min(c[0] for c in cands)
    near = [c for c in cands if c[0] <= best * (1.0 + TIE)]
    return min(near, key=lambda c: (partial_rows(c[2]) / c[1], c[0]))


def time_call(eng, nj, ids, x, v, h2, sel, eps2, reps):
    ts = []
    for r in range(reps + 1):
        eng.start_grav_calc(nj, ids[sel], x[sel], v[sel], eps2, None if h2 is None else h2[sel],
                            ai=None if eng.order < sa.SIXTH else np.zeros((len(sel), 3)))
        eng.get_grav_results(nj, len(sel), ngb=True)
        if r:
            ts.append(eng.last_device_ms())
    return float(np.median(ts))


def tune(n=131072, orders=(sa.FOURTH, sa.SIXTH), precs=(sa.DOUBLESINGLE, sa.DOUBLE), reps=3,
         kmax=14, seed=1, log=print):
    """Table entries and every measurement for one system size n."""
    m, x, v = ics.plummer(n, seed=seed)
    ids = np.arange(n, dtype=np.int32)
    h2 = ics.neighbour_h2(x, 32.0)
    rng = np.random.Generator(np.random.PCG64(seed + 1))
    rows, table = [], []
    lj = int(math.ceil(math.log2(n)))
    kmax = min(kmax, int(math.floor(math.log2(n))))
    for order in orders:
        for prec in precs:
            eng = sa.Engine(order, prec)
            eng.load_tuning(None)
            eng.set_timing(True)
            eng.set_j_particles(np.arange(n), ids, m, x, v, np.zeros((n, 3)), np.zeros((n, 3)),
                                np.zeros((n, 3)) if order >= sa.SIXTH else None,
                                np.zeros((n, 3)) if order >= sa.SIXTH else None)
            eng.set_time(0.0)
            for ngb in (1, 0):
                for k in range(kmax + 1):
                    ni = 1 << k
                    sel = np.sort(rng.choice(n, ni, replace=False))
                    meas = []
                    for P, S in candidates(ni, n):
                        eng.set_force_override(P, S)
                        t = time_call(eng, n, ids, x, v, h2 if ngb else None, sel, 1e-6, reps)
                        _, S_used, jc = eng.last_config()
                        rows.append(dict(n=n, order=order, prec=prec, ngb=ngb, ni=ni, P=P, S=S_used,
                                         jchunk=jc, dev_ms=t))
                        meas.append((t, P, S_used, jc))
                    eng.set_force_override(0, 0)
                    t, P, _, jc = select(meas)
                    jlog = int(round(math.log2(jc)))
                    table.append((order, prec, ngb, k, lj, P, jlog))
                    log(f"n {n} order {order} prec {prec} ngb {ngb} ni {ni}: P {P} jchunk {jc} "
                        f"{t * 1e3:.1f} us ({ni * n / t / 1e6:.0f} Ginter/s)")
            eng.close()
    return table, rows


def table_from_rows(rows):
    """Rebuilds the table from saved measurements (the --json output) with select()."""
    groups = {}
    for r in rows:
        key = (r["order"], r["prec"], r["ngb"], int(round(math.log2(r["ni"]))), int(math.ceil(math.log2(r["n"]))))
        groups.setdefault(key, []).append((r["dev_ms"], r["P"], r["S"], r["jchunk"]))
    table = []
    for key in sorted(groups, key=lambda k: (k[4], k[0], k[1], -k[2], k[3])):
        t, P, _, jc = select(groups[key])
        table.append(key[:4] + (key[4], P, int(round(math.log2(jc)))))
    return table


def write_table(path, table, ns):
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as f:
        f.write("# launch shapes tuned by sapporo2_amd/autotune.py on Plummer systems of N = "
                + ", ".join(str(n) for n in ns) + "\n")
        f.write("# order precision ngb log2(ni) log2(nj) P log2(j per split)\n")
        for t in table:
            f.write(" ".join(str(int(v)) for v in t) + "\n")


def main():
    ap = argparse.ArgumentParser(description="force launch-shape autotuner")
    ap.add_argument("--n", default="8192,131072,262144,1048576",
                    help="comma-separated system sizes (one table block per log2 N)")
    ap.add_argument("--out", default=None)
    ap.add_argument("--json", default=None, help="also write every measurement here")
    ap.add_argument("--orders", default="1,2")
    ap.add_argument("--precs", default="1,2")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--from-json", default=None,
                    help="rebuild the table from saved measurements (no GPU) instead of measuring")
    args = ap.parse_args()
    t0 = time.time()
    if args.from_json:
        with open(args.from_json) as f:
            rows = json.load(f)
        ns = sorted({r["n"] for r in rows})
        out = args.out or os.path.join(sa.HERE, "tune", "gfx950.txt")
        write_table(out, table_from_rows(rows), ns)
        print(f"wrote {out} from {len(rows)} measurements", flush=True)
        return 0
    ns = [int(v) for v in args.n.split(",")]
    table, rows = [], []
    for n in ns:
        t, r = tune(n, tuple(int(o) for o in args.orders.split(",")),
                    tuple(int(p) for p in args.precs.split(",")), args.reps,
                    log=lambda s: print(s, flush=True))
        table += t
        rows += r
    out = args.out or os.path.join(sa.HERE, "tune", "gfx950.txt")
    write_table(out, table, ns)
    if args.json:
        with open(args.json, "w") as f:
            json.dump(rows, f)
    print(f"wrote {out} ({len(table)} entries) in {time.time() - t0:.1f} s", flush=True)


if __name__ == "__main__":
    sys.exit(main())
