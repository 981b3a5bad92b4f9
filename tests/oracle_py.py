"""ctypes binding of the CPU oracle (oracle/libref_cpu.so) for the tests.

The oracle is the checker: tests compare the HIP path against it; nothing in
the product imports this module.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from dataclasses import dataclass
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
LIB = ROOT / "oracle" / "libref_cpu.so"

i32, i64, f64 = C.c_int32, C.c_int64, C.c_double
P = C.POINTER


class P1Metrics(C.Structure):
    _fields_ = [("initial_edge_cut", i64), ("initial_vertex_balance", f64),
                ("initial_edge_balance", f64), ("total_vertices", i64), ("total_edges", i64),
                ("first_seed", i32), ("first_seed_degree", i32), ("n_hubs", i64),
                ("n_seeds_found", i32), ("n_seeds_fallback", i32), ("rounds", i64),
                ("fill_round", i32), ("reserved", i32)]


class P2Metrics(C.Structure):
    _fields_ = [("edge_cut", i64), ("vertex_balance", f64), ("edge_balance", f64),
                ("total_vertices", i64), ("total_edges", i64), ("rounds", i32),
                ("exit_reason", i32), ("lp_edges", i64), ("initial_cut", i64)]


TRACE_FIELDS = ("prev_count", "prev_deg", "cand_count", "scan_len", "bd_count", "b_count",
                "b_deg", "chg_count", "chg_deg", "cut")


class P2Trace(C.Structure):
    _fields_ = [(f, P(i64)) for f in TRACE_FIELDS]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], check=True)
        L = C.CDLL(str(LIB))
        VP = C.c_void_p
        sig = {
            "ref_probe_degree_map_order": [i32, P(i32)],
            "ref_first_seed": [i64, P(i64), P(i32), P(i32)],
            "ref_hubs": [i64, P(i64), P(i32), P(i64)],
            "ref_seeds": [i64, P(i64), P(i32), i32, P(i32), P(i32), P(i32)],
            "ref_ball2": [i64, P(i64), P(i32), i32, P(C.c_uint8)],
            "ref_expand": [i64, P(i64), P(i32), i32, i32, P(i32), P(i64), P(i32), P(i32)],
            "ref_phase1_metrics": [i64, P(i64), P(i32), i32, P(i32), P(i64), P(i64),
                                   P(P1Metrics)],
            "ref_phase1": [i64, P(i64), P(i32), i32, i32, P(i32), P(i32), P(i64), P(i64),
                           P(P1Metrics)],
            "ref_stats": [i64, P(i64), P(i32), i32, P(i32), P(i64), P(i64), P(i64)],
            "ref_select_boundary": [i64, P(i64), P(i32), P(i32), P(i32), i64, P(f64), i32, i32,
                                    P(i32), P(i64)],
            "ref_lp_round": [i64, P(i64), P(i32), P(i32), P(i32), i64, P(f64), i32, P(i32)],
            "ref_phase2": [i64, P(i64), P(i32), i32, i32, f64, P(i32), P(i64), P(i64),
                           P(P2Metrics), P(P2Trace)],
            "ref_phase2_ranks": [i64, P(i64), P(i32), i32, i32, i32, i32, f64, P(i32), P(i64),
                                 P(i64), P(P2Metrics), P(P2Trace), P(i64)],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.restype = C.c_int
            fn.argtypes = args
        L.ref_set_threads.restype = None
        L.ref_set_threads.argtypes = [i32]
        L.ref_get_threads.restype = i32
        L.ref_get_threads.argtypes = []
        L.ref_penalties.restype = None
        L.ref_penalties.argtypes = [i32, P(i64), P(i64), i32, P(f64), P(f64)]
        _ = VP
        _lib = L
    return _lib


def set_threads(t: int):
    """OpenMP threads of the oracle's sweeps (results do not change)."""
    lib().ref_set_threads(int(t))


def get_threads() -> int:
    return int(lib().ref_get_threads())


def _p(a, ct):
    return a.ctypes.data_as(P(ct))


def _ok(rc):
    if rc != 0:
        raise RuntimeError(f"oracle status {rc}")


def csr(off, col):
    return (np.ascontiguousarray(off, np.int64), np.ascontiguousarray(col, np.int32))


def degree_map_order(n):
    out = np.zeros(n, np.int32)
    _ok(lib().ref_probe_degree_map_order(n, _p(out, i32)))
    return out


def first_seed(off):
    off = np.ascontiguousarray(off, np.int64)
    s, d = i32(), i32()
    _ok(lib().ref_first_seed(len(off) - 1, _p(off, i64), C.byref(s), C.byref(d)))
    return s.value, d.value


def hubs(off):
    off = np.ascontiguousarray(off, np.int64)
    n = len(off) - 1
    out = np.zeros(max(n, 1), np.int32)
    nh = i64()
    _ok(lib().ref_hubs(n, _p(off, i64), _p(out, i32), C.byref(nh)))
    return out[:nh.value].copy()


def seeds(off, col, k):
    off, col = csr(off, col)
    out = np.zeros(k, np.int32)
    f, fb = i32(), i32()
    rc = lib().ref_seeds(len(off) - 1, _p(off, i64), _p(col, i32), k, _p(out, i32),
                         C.byref(f), C.byref(fb))
    _ok(rc)
    return out, f.value, fb.value


def ball2(off, col, s):
    off, col = csr(off, col)
    out = np.zeros(len(off) - 1, np.uint8)
    _ok(lib().ref_ball2(len(off) - 1, _p(off, i64), _p(col, i32), s, _p(out, C.c_uint8)))
    return out.astype(bool)


def expand(off, col, k, theta, seeds_arr):
    off, col = csr(off, col)
    seeds_arr = np.ascontiguousarray(seeds_arr, np.int32)
    labels = np.zeros(len(off) - 1, np.int32)
    rounds, fill = i64(), i32()
    _ok(lib().ref_expand(len(off) - 1, _p(off, i64), _p(col, i32), k, theta,
                         _p(seeds_arr, i32), C.byref(rounds), C.byref(fill), _p(labels, i32)))
    return labels, rounds.value, fill.value


@dataclass
class Phase1Out:
    labels: np.ndarray
    seeds: np.ndarray
    vc: np.ndarray
    ec: np.ndarray
    m: P1Metrics


def phase1(off, col, k, theta) -> Phase1Out:
    off, col = csr(off, col)
    n = len(off) - 1
    labels = np.zeros(n, np.int32)
    sd = np.zeros(k, np.int32)
    vc = np.zeros(k, np.int64)
    ec = np.zeros(k, np.int64)
    m = P1Metrics()
    _ok(lib().ref_phase1(n, _p(off, i64), _p(col, i32), k, theta, _p(labels, i32), _p(sd, i32),
                         _p(vc, i64), _p(ec, i64), C.byref(m)))
    return Phase1Out(labels, sd, vc, ec, m)


def phase1_metrics(off, col, k, labels):
    """ref_phase1_metrics: (vc, ec, P1Metrics) of a labelling."""
    off, col = csr(off, col)
    labels = np.ascontiguousarray(labels, np.int32)
    vc = np.zeros(k, np.int64)
    ec = np.zeros(k, np.int64)
    m = P1Metrics()
    _ok(lib().ref_phase1_metrics(len(off) - 1, _p(off, i64), _p(col, i32), k, _p(labels, i32),
                                 _p(vc, i64), _p(ec, i64), C.byref(m)))
    return vc, ec, m


def stats(off, col, k, labels):
    off, col = csr(off, col)
    labels = np.ascontiguousarray(labels, np.int32)
    vc = np.zeros(k, np.int64)
    ic = np.zeros(k, np.int64)
    cut = i64()
    rc = lib().ref_stats(len(off) - 1, _p(off, i64), _p(col, i32), k, _p(labels, i32),
                         _p(vc, i64), _p(ic, i64), C.byref(cut))
    return rc, vc, ic, cut.value


def penalties(k, vc, ic, t):
    vc = np.ascontiguousarray(vc, np.int64)
    ic = np.ascontiguousarray(ic, np.int64)
    pen = np.zeros(k)
    re = np.zeros(k)
    lib().ref_penalties(k, _p(vc, i64), _p(ic, i64), t, _p(pen, f64), _p(re, f64))
    return pen, re


def select_boundary(off, col, labels, prev, re, k, t):
    off, col = csr(off, col)
    labels = np.ascontiguousarray(labels, np.int32)
    prev = np.ascontiguousarray(prev, np.int32)
    re = np.ascontiguousarray(re, np.float64)
    out = np.zeros(len(off) - 1, np.int32)
    nout = i64()
    _ok(lib().ref_select_boundary(len(off) - 1, _p(off, i64), _p(col, i32), _p(labels, i32),
                                  _p(prev, i32), len(prev), _p(re, f64), k, t, _p(out, i32),
                                  C.byref(nout)))
    return out[:nout.value].copy()


def lp_round(off, col, labels, b, pen, k):
    off, col = csr(off, col)
    labels = np.ascontiguousarray(labels, np.int32)
    b = np.ascontiguousarray(b, np.int32)
    pen = np.ascontiguousarray(pen, np.float64)
    out = labels.copy()
    _ok(lib().ref_lp_round(len(off) - 1, _p(off, i64), _p(col, i32), _p(labels, i32),
                           _p(b, i32), len(b), _p(pen, f64), k, _p(out, i32)))
    return out


@dataclass
class Phase2Out:
    labels: np.ndarray
    vc: np.ndarray
    ic: np.ndarray
    m: P2Metrics
    trace: dict


def phase2(off, col, k, labels, max_rounds, epsilon) -> Phase2Out:
    off, col = csr(off, col)
    labels = np.array(labels, np.int32, copy=True)
    vc = np.zeros(k, np.int64)
    ic = np.zeros(k, np.int64)
    m = P2Metrics()
    arrays = {f: np.full(max(max_rounds, 1), -7, np.int64) for f in TRACE_FIELDS}
    tr = P2Trace(**{f: _p(a, i64) for f, a in arrays.items()})
    _ok(lib().ref_phase2(len(off) - 1, _p(off, i64), _p(col, i32), k, max_rounds, epsilon,
                         _p(labels, i32), _p(vc, i64), _p(ic, i64), C.byref(m), C.byref(tr)))
    trace = {f: a[:m.rounds].copy() for f, a in arrays.items()}
    return Phase2Out(labels, vc, ic, m, trace)


def phase2_ranks(off, col, k, labels, nranks, scope, max_rounds, epsilon) -> Phase2Out:
    """Phase 2 as nranks ranks on their shares (scope 0: per-rank selection
    as the reference; 1: selection over all ranks).  The trace gains
    'ghost_updates'."""
    off, col = csr(off, col)
    labels = np.array(labels, np.int32, copy=True)
    vc = np.zeros(k, np.int64)
    ic = np.zeros(k, np.int64)
    m = P2Metrics()
    arrays = {f: np.full(max(max_rounds, 1), -7, np.int64) for f in TRACE_FIELDS}
    tr = P2Trace(**{f: _p(a, i64) for f, a in arrays.items()})
    gu = np.full(max(max_rounds, 1), -7, np.int64)
    _ok(lib().ref_phase2_ranks(len(off) - 1, _p(off, i64), _p(col, i32), k, nranks, scope,
                               max_rounds, epsilon, _p(labels, i32), _p(vc, i64), _p(ic, i64),
                               C.byref(m), C.byref(tr), _p(gu, i64)))
    trace = {f: a[:m.rounds].copy() for f, a in arrays.items()}
    trace["ghost_updates"] = gu[:m.rounds].copy()
    return Phase2Out(labels, vc, ic, m, trace)
