"""sapporo2_amd -- direct-summation N-body gravity (Hermite-4/6) on one AMD MI355X.

The product is the pair of C-ABI shared libraries built from ``sapporo2_amd/csrc``:

* ``libsapporo_amd.so``     -- engine handle API (``include/sapporo_amd.h``) and the
                               GRAPE-6 ``g6_*`` ABI (``include/g6lib.h``)
* ``libsapporo6th_amd.so``  -- the 6th-order Hermite ABI (``include/sapporo6thlib.h``)

This module is a thin ctypes mirror of the reference's host class ``sapporo``
(``lib/include/sapporohostclass.h:121-176``) used by the tests and the benchmark.  It
never falls back to a CPU path: if the HIP library is missing it raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

__all__ = ["Engine", "lib", "lib6th", "FOURTH", "SIXTH", "DOUBLESINGLE", "DOUBLE",
           "NPIPES", "NGB_CAP", "LIB_PATH", "LIB6_PATH", "build"]

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsapporo_amd.so")
LIB6_PATH = os.path.join(HERE, "libsapporo6th_amd.so")

GRAPE5, FOURTH, SIXTH = 0, 1, 2          # reference defines.h:30
FLOAT, DOUBLESINGLE, DOUBLE = 0, 1, 2    # reference defines.h:31
MIXED = DOUBLESINGLE
NPIPES = 16384                           # reference defines.h:14
NGB_CAP = 256                            # reference defines.h:9

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p

_LIB = None
_LIB6 = None


def build(verbose=False):
    """Compile the HIP libraries in-tree (hipcc, gfx950)."""
    import subprocess
    cmd = ["make", "-s", "-C", os.path.join(HERE, "csrc"), "-j4"]
    subprocess.check_call(cmd, stdout=None if verbose else subprocess.DEVNULL)


# Symbols each library must export (checked by the ABI tests).
ENGINE_SYMBOLS = [
    "sapamd_create", "sapamd_destroy", "sapamd_open", "sapamd_close", "sapamd_get_n_pipes",
    "sapamd_set_time", "sapamd_set_no_time", "sapamd_set_j_particle", "sapamd_start_grav_calc",
    "sapamd_get_grav_results", "sapamd_set_j_particles", "sapamd_last_device_ms", "sapamd_read_ngb_list", "sapamd_get_ngb_list",
    "sapamd_force_prediction", "sapamd_retrieve_predicted_j_particle",
    "sapamd_retrieve_j_particle_state", "sapamd_set_force_override", "sapamd_set_split_major", "sapamd_set_ngb_pipeline", "sapamd_set_timing", "sapamd_set_graphs", "sapamd_graph_count", "sapamd_retrieve_predicted_j_particles",
    "sapamd_state_bytes", "sapamd_save_state", "sapamd_load_state",
    "sapamd_last_force_ms", "sapamd_last_predict_ms", "sapamd_last_scatter_ms", "sapamd_last_finalize_ms", "sapamd_last_config", "sapamd_load_tuning", "sapamd_tuning_entries", "sapamd_num_cus",
    "sapamd_synchronize", "sapamd_last_error", "sapamd_set_clusters", "sapamd_cluster_offset",
    "sapamd_start_grav_calc_clusters", "sapamd_g6_engine", "sapamd_grav_calc_noalloc",
    "sapamd_noalloc_scratch_bytes",
]


def lib():
    """The HIP engine library (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                               "(there is no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        L.sapamd_create.restype = _vp
        L.sapamd_destroy.argtypes = [_vp]
        L.sapamd_open.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.sapamd_close.argtypes = [_vp]
        L.sapamd_get_n_pipes.argtypes = [_vp]
        L.sapamd_set_time.argtypes = [_vp, ctypes.c_double]
        L.sapamd_set_no_time.argtypes = [_vp]
        L.sapamd_set_j_particle.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                            ctypes.c_double, ctypes.c_double, _dp, _dp, _dp, _dp,
                                            _dp, _dp, ctypes.c_double]
        L.sapamd_set_j_particles.argtypes = [_vp, ctypes.c_int, _ip, _ip, _dp, _dp, _dp, _dp, _dp,
                                             _dp, _dp, _dp, _dp, _dp]
        L.sapamd_last_device_ms.argtypes = [_vp]
        L.sapamd_last_device_ms.restype = ctypes.c_float
        L.sapamd_start_grav_calc.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _ip, _dp, _dp, _dp,
                                             ctypes.c_double, _dp]
        L.sapamd_get_grav_results.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp,
                                              _dp, _ip, _dp, ctypes.c_int]
        L.sapamd_set_clusters.argtypes = [_vp, ctypes.c_int, _ip]
        L.sapamd_cluster_offset.argtypes = [_vp, ctypes.c_int]
        L.sapamd_start_grav_calc_clusters.argtypes = [_vp, ctypes.c_int, _ip, _ip, _dp, _dp, _dp, _dp, _dp]
        L.sapamd_read_ngb_list.argtypes = [_vp]
        L.sapamd_get_ngb_list.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _ip, _ip]
        L.sapamd_force_prediction.argtypes = [_vp, ctypes.c_int]
        L.sapamd_retrieve_predicted_j_particle.argtypes = [_vp, ctypes.c_int, _dp, _dp, _dp, _dp,
                                                           _dp, _dp]
        L.sapamd_retrieve_j_particle_state.argtypes = [_vp, ctypes.c_int] + [_dp] * 10
        L.sapamd_set_force_override.argtypes = [_vp, ctypes.c_int, ctypes.c_int]
        L.sapamd_set_timing.argtypes = [_vp, ctypes.c_int]
        L.sapamd_set_graphs.argtypes = [_vp, ctypes.c_int]
        L.sapamd_retrieve_predicted_j_particles.argtypes = [_vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp]
        L.sapamd_state_bytes.argtypes = [_vp]
        L.sapamd_state_bytes.restype = ctypes.c_longlong
        L.sapamd_save_state.argtypes = [_vp, _vp, ctypes.c_longlong]
        L.sapamd_load_state.argtypes = [_vp, _vp, ctypes.c_longlong]
        L.sapamd_graph_count.argtypes = [_vp]
        L.sapamd_set_split_major.argtypes = [_vp, ctypes.c_int]
        L.sapamd_set_ngb_pipeline.argtypes = [_vp, ctypes.c_int]
        L.sapamd_last_force_ms.argtypes = [_vp]
        L.sapamd_last_force_ms.restype = ctypes.c_float
        L.sapamd_last_predict_ms.argtypes = [_vp]
        L.sapamd_last_predict_ms.restype = ctypes.c_float
        L.sapamd_last_scatter_ms.argtypes = [_vp, _ip]
        L.sapamd_last_scatter_ms.restype = ctypes.c_float
        L.sapamd_last_finalize_ms.argtypes = [_vp]
        L.sapamd_last_finalize_ms.restype = ctypes.c_float
        L.sapamd_last_config.argtypes = [_vp, _ip, _ip, _ip]
        L.sapamd_num_cus.argtypes = [_vp]
        L.sapamd_load_tuning.argtypes = [_vp, ctypes.c_char_p]
        L.sapamd_tuning_entries.argtypes = [_vp]
        L.sapamd_synchronize.argtypes = [_vp]
        L.sapamd_last_error.argtypes = [_vp]
        L.sapamd_last_error.restype = ctypes.c_char_p
        _LIB = L
    return _LIB


def lib6th():
    """The 6th-order ABI library (a separate .so, as in the reference)."""
    global _LIB6
    if _LIB6 is None:
        if not os.path.exists(LIB6_PATH):
            raise RuntimeError(f"{LIB6_PATH} is missing: build it first")
        _LIB6 = ctypes.CDLL(LIB6_PATH)
    return _LIB6


def _d(a, shape=None):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


def _pd(a):
    return None if a is None else a.ctypes.data_as(_dp)


def _pi(a):
    return None if a is None else a.ctypes.data_as(_ip)


class Engine:
    """One GPU engine (handle over ``sapamd_*``).  Methods mirror the reference's
    ``class sapporo``; arrays are numpy (n, 3) float64."""

    def __init__(self, order=FOURTH, precision=DOUBLESINGLE, device=0):
        L = lib()
        self._L = L
        self._h = L.sapamd_create()
        if not self._h:
            raise MemoryError("sapamd_create failed")
        self.order = order
        self.precision = precision
        self._check(L.sapamd_open(self._h, device, order, precision), "open")

    def _check(self, rc, what):
        if rc < 0:
            msg = self._L.sapamd_last_error(self._h)
            raise RuntimeError(f"sapporo2_amd {what} failed: {msg.decode() if msg else ''}")
        return rc

    def close(self):
        if self._h:
            self._L.sapamd_close(self._h)
            self._L.sapamd_destroy(self._h)
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

    @property
    def n_pipes(self):
        return self._L.sapamd_get_n_pipes(self._h)

    def set_time(self, t):
        return self._check(self._L.sapamd_set_time(self._h, float(t)), "set_time")

    def set_no_time(self):
        return self._check(self._L.sapamd_set_no_time(self._h), "set_no_time")

    def set_j_particle(self, address, id, tj, dtj, mass, acc, jrk, vel, pos, snp=None, crk=None,
                       eps=0.0):
        a = [_d(v, (3,)) for v in (acc, jrk, vel, pos, snp, crk)]
        return self._check(self._L.sapamd_set_j_particle(
            self._h, int(address), int(id), float(tj), float(dtj), float(mass),
            *[_pd(v) for v in a], float(eps)), "set_j_particle")

    def set_j_particles(self, addresses, ids, mass, pos, vel, acc=None, jrk=None, snp=None,
                        crk=None, tj=0.0, eps=None):
        """Array form (one C call): same per-particle semantics as set_j_particle."""
        n = len(addresses)
        addr = np.ascontiguousarray(addresses, np.int32)
        idv = np.ascontiguousarray(ids, np.int32)
        mass = np.ascontiguousarray(np.broadcast_to(np.asarray(mass, np.float64), (n,)))
        tjv = np.ascontiguousarray(np.broadcast_to(np.asarray(tj, np.float64), (n,)))
        arrs = [None if a is None else _d(a, (n, 3)) for a in (acc, jrk, vel, pos, snp, crk)]
        epsv = None if eps is None else np.ascontiguousarray(np.broadcast_to(np.asarray(eps, np.float64), (n,)))
        return self._check(self._L.sapamd_set_j_particles(
            self._h, n, _pi(addr), _pi(idv), _pd(tjv), None, _pd(mass), *[_pd(a) for a in arrs],
            _pd(epsv)), "set_j_particles")

    def start_grav_calc(self, nj, ids, xi, vi, eps2, h2=None, ai=None):
        ni = len(ids)
        self._keep = (np.ascontiguousarray(ids, np.int32), _d(xi, (ni, 3)), _d(vi, (ni, 3)),
                      None if ai is None else _d(ai, (ni, 3)), None if h2 is None else _d(h2, (ni,)))
        ids_, xi_, vi_, ai_, h2_ = self._keep
        return self._check(self._L.sapamd_start_grav_calc(
            self._h, int(nj), ni, _pi(ids_), _pd(xi_), _pd(vi_), _pd(ai_), float(eps2), _pd(h2_)),
            "start_grav_calc")

    def get_grav_results(self, nj, ni, ngb=True, out=None, keep_dsmin=False):
        """Returns dict acc, jrk, snp (SIXTH), pot, nnb, dsmin (NaN where not written).
        out: a dict of that form from an earlier call, with arrays of at least ni rows,
        whose first ni rows are filled and returned instead of new arrays.  keep_dsmin
        (with out): dsmin keeps the caller's values where no nearest neighbour was found,
        as the reference's getGravResults leaves them (sapporohostclass.cpp:705-802),
        instead of being refilled with NaN first."""
        if out is None:
            acc = np.zeros((ni, 3)); jrk = np.zeros((ni, 3)); pot = np.zeros(ni)
            snp = np.zeros((ni, 3)) if self.order >= SIXTH else None
            crk = np.zeros((ni, 3)) if self.order >= SIXTH else None
            nnb = np.zeros(ni, np.int32); dsmin = np.full(ni, np.nan)
        else:
            def rows(k, shape, dtype):
                a = out.get(k)
                if a is None or a.dtype != dtype or a.shape[1:] != shape or len(a) < ni or not a.flags.c_contiguous:
                    raise ValueError(f"get_grav_results: out[{k!r}] must be a contiguous {dtype} array of "
                                     f">= {ni} rows of shape {shape}")
                return a[:ni]
            acc = rows("acc", (3,), np.float64); jrk = rows("jrk", (3,), np.float64)
            pot = rows("pot", (), np.float64)
            snp = rows("snp", (3,), np.float64) if self.order >= SIXTH else None
            crk = rows("crk", (3,), np.float64) if self.order >= SIXTH else None
            nnb = rows("nnb", (), np.int32); dsmin = rows("dsmin", (), np.float64)
            if not keep_dsmin:
                dsmin.fill(np.nan)
            if not ngb:
                nnb.fill(0)
        self._check(self._L.sapamd_get_grav_results(
            self._h, int(nj), int(ni), _pd(acc), _pd(jrk), _pd(snp), _pd(crk), _pd(pot), _pi(nnb),
            _pd(dsmin), 1 if ngb else 0), "get_grav_results")
        return dict(acc=acc, jrk=jrk, snp=snp, crk=crk, pot=pot, nnb=nnb, dsmin=dsmin)

    def calc(self, nj, ids, xi, vi, eps2, h2=None, ai=None, ngb=True):
        """firsthalf + lasthalf in one call (any ni; chunked by n_pipes)."""
        ni = len(ids)
        out = None
        for s in range(0, max(ni, 1), NPIPES):
            e = min(ni, s + NPIPES)
            self.start_grav_calc(nj, ids[s:e], xi[s:e], vi[s:e], eps2,
                                 None if h2 is None else h2[s:e], None if ai is None else ai[s:e])
            r = self.get_grav_results(nj, e - s, ngb)
            if out is None:
                out = {k: (None if v is None else np.empty((ni,) + v.shape[1:], v.dtype))
                       for k, v in r.items()}
            for k, v in r.items():
                if v is not None:
                    out[k][s:e] = v
        return out

    # --- ensembles of independent clusters ----------------------------------------------
    def set_clusters(self, sizes):
        sizes = np.ascontiguousarray(sizes, np.int32)
        self._check(self._L.sapamd_set_clusters(self._h, len(sizes), _pi(sizes)), "set_clusters")
        return np.array([self._L.sapamd_cluster_offset(self._h, c) for c in range(len(sizes))])

    def start_grav_calc_clusters(self, cluster, ids, xi, vi, eps2, h2=None, ai=None):
        ni = len(ids)
        self._keep = (np.ascontiguousarray(cluster, np.int32), np.ascontiguousarray(ids, np.int32),
                      _d(xi, (ni, 3)), _d(vi, (ni, 3)), None if ai is None else _d(ai, (ni, 3)),
                      np.ascontiguousarray(eps2, np.float64), None if h2 is None else _d(h2, (ni,)))
        cl, ids_, xi_, vi_, ai_, e2, h2_ = self._keep
        return self._check(self._L.sapamd_start_grav_calc_clusters(
            self._h, ni, _pi(cl), _pi(ids_), _pd(xi_), _pd(vi_), _pd(ai_), _pd(e2), _pd(h2_)),
            "start_grav_calc_clusters")

    def read_ngb_list(self):
        return self._check(self._L.sapamd_read_ngb_list(self._h), "read_ngb_list")

    def get_ngb_list(self, ipipe, maxlen=NGB_CAP):
        buf = np.zeros(max(maxlen, 1), np.int32)
        n = ctypes.c_int(0)
        ov = self._check(self._L.sapamd_get_ngb_list(self._h, int(ipipe), int(maxlen),
                                                     ctypes.byref(n), _pi(buf)), "get_ngb_list")
        return buf[:min(n.value, maxlen, NGB_CAP)].copy(), n.value, ov

    def force_prediction(self, nj):
        return self._check(self._L.sapamd_force_prediction(self._h, int(nj)), "force_prediction")

    def retrieve_predicted_j_particle(self, addr):
        m = ctypes.c_double(); idd = ctypes.c_double(); e2 = ctypes.c_double()
        pos = np.zeros(3); vel = np.zeros(3); acc = np.zeros(3)
        self._check(self._L.sapamd_retrieve_predicted_j_particle(
            self._h, int(addr), ctypes.byref(m), ctypes.byref(idd), ctypes.byref(e2), _pd(pos),
            _pd(vel), _pd(acc)), "retrieve_predicted_j_particle")
        return dict(mass=m.value, id=int(idd.value), eps2=e2.value, pos=pos, vel=vel, acc=acc)

    def retrieve_predicted_j_particles(self, addresses):
        """Array form: mass, predicted pos, vel, acc of the listed addresses (device gather)."""
        a = np.ascontiguousarray(addresses, np.int32)
        n = len(a)
        m = np.zeros(n); pos = np.zeros((n, 3)); vel = np.zeros((n, 3)); acc = np.zeros((n, 3))
        self._check(self._L.sapamd_retrieve_predicted_j_particles(self._h, n, _pi(a), _pd(m), _pd(pos), _pd(vel),
                                                                  _pd(acc)), "retrieve_predicted_j_particles")
        return dict(mass=m, pos=pos, vel=vel, acc=acc)

    def save_state(self):
        """Checkpoint of the j-store as a uint8 array (see load_state)."""
        nb = self._check(self._L.sapamd_state_bytes(self._h), "state_bytes")
        buf = np.zeros(nb, np.uint8)
        self._check(self._L.sapamd_save_state(self._h, buf.ctypes.data, nb), "save_state")
        return buf

    def load_state(self, buf):
        """Replace the j-store by a checkpoint of an engine of the same order and precision."""
        b = np.ascontiguousarray(buf, np.uint8)
        self._check(self._L.sapamd_load_state(self._h, b.ctypes.data, len(b)), "load_state")

    def retrieve_j_particle_state(self, addr):
        m = ctypes.c_double(); idd = ctypes.c_double(); e2 = ctypes.c_double()
        arrs = [np.zeros(3) for _ in range(7)]
        self._check(self._L.sapamd_retrieve_j_particle_state(
            self._h, int(addr), ctypes.byref(m), ctypes.byref(idd), ctypes.byref(e2),
            *[_pd(a) for a in arrs]), "retrieve_j_particle_state")
        keys = ["pos", "vel", "acc", "jrk", "ppos", "pvel", "pacc"]
        d = dict(zip(keys, arrs))
        d.update(mass=m.value, id=int(idd.value), eps2=e2.value)
        return d

    # --- launch shape / instrumentation -------------------------------------------------
    def set_force_override(self, P=0, S=0):
        self._L.sapamd_set_force_override(self._h, int(P), int(S))

    def set_split_major(self, on=True):
        self._L.sapamd_set_split_major(self._h, 1 if on else 0)

    def set_ngb_pipeline(self, on=True):
        self._L.sapamd_set_ngb_pipeline(self._h, 1 if on else 0)

    def set_timing(self, on=True):
        self._L.sapamd_set_timing(self._h, 1 if on else 0)

    def set_graphs(self, on=True):
        """Capture each call's device work as a hipGraph per launch shape and replay it."""
        self._L.sapamd_set_graphs(self._h, 1 if on else 0)

    def graph_count(self):
        return int(self._L.sapamd_graph_count(self._h))

    def last_force_ms(self):
        return float(self._L.sapamd_last_force_ms(self._h))

    def last_device_ms(self):
        return float(self._L.sapamd_last_device_ms(self._h))

    def last_predict_ms(self):
        return float(self._L.sapamd_last_predict_ms(self._h))

    def last_finalize_ms(self):
        return float(self._L.sapamd_last_finalize_ms(self._h))

    def last_scatter_ms(self):
        """(scatter kernel ms, updates applied) of the last call (timing on)."""
        n = ctypes.c_int(0)
        ms = float(self._L.sapamd_last_scatter_ms(self._h, ctypes.byref(n)))
        return ms, n.value

    def last_config(self):
        P = ctypes.c_int(); S = ctypes.c_int(); jc = ctypes.c_int()
        self._L.sapamd_last_config(self._h, ctypes.byref(P), ctypes.byref(S), ctypes.byref(jc))
        return P.value, S.value, jc.value

    def load_tuning(self, path):
        """Load a launch-shape table (None or a missing file: heuristic shapes)."""
        return self._L.sapamd_load_tuning(self._h, None if path is None else str(path).encode())

    @property
    def tuning_entries(self):
        return self._L.sapamd_tuning_entries(self._h)

    @property
    def num_cus(self):
        return self._L.sapamd_num_cus(self._h)

    def synchronize(self):
        return self._check(self._L.sapamd_synchronize(self._h), "synchronize")
