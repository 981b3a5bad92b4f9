"""Python view of the C ABI: contexts, device graphs, phase 1 and phase 2.

Every call goes through libdmolp_hip.so; errors raise DmolpError with the
library's message.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _abi

GEN_RMAT, GEN_CHUNG_LU, GEN_GRID = 1, 2, 3
EXIT_NAMES = {0: "max_rounds", 1: "empty_boundary", 2: "no_change", 3: "converged"}


class DmolpError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"dmolp status {status}: {message}")
        self.status = status


def _p(a: np.ndarray, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def _check(rc: int, ctx=None):
    if rc != 0:
        msg = _abi.lib().dmolp_last_error(ctx)
        raise DmolpError(rc, (msg or b"").decode(errors="replace"))


def rmat_spec(scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=1) -> _abi.GenSpec:
    return _abi.GenSpec(kind=GEN_RMAT, scale=scale, edge_factor=edge_factor, a=a, b=b, c=c,
                        seed=seed)


def chung_lu_spec(n, m, gamma=2.0, dmax=3.0e6, seed=3) -> _abi.GenSpec:
    return _abi.GenSpec(kind=GEN_CHUNG_LU, n=n, m=m, gamma=gamma, dmax=dmax, seed=seed)


def grid_spec(rows, cols) -> _abi.GenSpec:
    return _abi.GenSpec(kind=GEN_GRID, rows=rows, cols=cols)


def _take_csr(csr: _abi.Csr):
    n, nnz = csr.n, csr.nnz
    off = np.ctypeslib.as_array(csr.off, shape=(n + 1,)).copy()
    col = (np.ctypeslib.as_array(csr.col, shape=(nnz,)).copy() if nnz > 0
           else np.zeros(0, np.int32))
    _abi.lib().dmolp_csr_free(C.byref(csr))
    return off, col


def gen_host(spec: _abi.GenSpec):
    """Generate a graph on the host: returns (off int64[n+1], col int32[nnz])."""
    csr = _abi.Csr()
    _check(_abi.lib().dmolp_gen_host(C.byref(spec), C.byref(csr)))
    return _take_csr(csr)


def read_adjacency_text(path: str):
    csr = _abi.Csr()
    _check(_abi.lib().dmolp_read_adjacency_text(str(path).encode(), C.byref(csr)))
    return _take_csr(csr)


def write_adjacency_text(path: str, off: np.ndarray, col: np.ndarray):
    off = np.ascontiguousarray(off, np.int64)
    col = np.ascontiguousarray(col, np.int32)
    _check(_abi.lib().dmolp_write_adjacency_text(str(path).encode(), len(off) - 1,
                                                 _p(off, C.c_int64), _p(col, C.c_int32)))


def hubs_host(deg: np.ndarray) -> np.ndarray:
    """The library's hub list of a degree sequence (host, no device)."""
    deg = np.ascontiguousarray(deg, np.int32)
    out = np.zeros(max(len(deg), 1), np.int32)
    nh = C.c_int64()
    _check(_abi.lib().dmolp_hubs_host(_p(deg, C.c_int32), len(deg), _p(out, C.c_int32),
                                      C.byref(nh)))
    return out[:nh.value].copy()


def first_seed_host(deg: np.ndarray) -> tuple[int, int]:
    """The library's first seed of a degree sequence and its degree (host,
    no device)."""
    deg = np.ascontiguousarray(deg, np.int32)
    s, d = C.c_int32(), C.c_int32()
    _check(_abi.lib().dmolp_first_seed_host(_p(deg, C.c_int32), len(deg), C.byref(s), C.byref(d)))
    return s.value, d.value


def write_csr_binary(path: str, off: np.ndarray, col: np.ndarray):
    off = np.ascontiguousarray(off, np.int64)
    col = np.ascontiguousarray(col, np.int32)
    _check(_abi.lib().dmolp_write_csr_binary(str(path).encode(), len(off) - 1,
                                             _p(off, C.c_int64), _p(col, C.c_int32)))


def read_csr_binary(path: str):
    csr = _abi.Csr()
    _check(_abi.lib().dmolp_read_csr_binary(str(path).encode(), C.byref(csr)))
    return _take_csr(csr)


def load_adjacency_cached(path: str):
    """(off, col, from_cache): the text file through its <path>.csr cache."""
    csr = _abi.Csr()
    hit = C.c_int32()
    _check(_abi.lib().dmolp_load_adjacency_cached(str(path).encode(), C.byref(csr),
                                                  C.byref(hit)))
    off, col = _take_csr(csr)
    return off, col, bool(hit.value)


@dataclass
class Phase1Result:
    labels: np.ndarray
    seeds: np.ndarray
    vc: np.ndarray
    ec: np.ndarray
    metrics: _abi.Phase1Metrics


@dataclass
class Phase2Result:
    labels: np.ndarray | None
    vc: np.ndarray
    ic: np.ndarray
    metrics: _abi.Phase2Metrics
    trace: dict


class Context:
    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        _check(_abi.lib().dmolp_ctx_create(device, C.byref(self._h)))

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            _abi.lib().dmolp_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def tuning(self) -> _abi.Tuning:
        t = _abi.Tuning()
        _check(_abi.lib().dmolp_ctx_get_tuning(self._h, C.byref(t)), self._h)
        return t

    def set_tuning(self, heavy_degree=0, chunk_edges=0, grid_blocks=0, pull_per_mille=0,
                   batch_work_per_block=0, hist_recount_per_mille=0, hist_push_per_mille=0,
                   light_hist=0):
        t = _abi.Tuning(heavy_degree=heavy_degree, chunk_edges=chunk_edges, block_threads=0,
                        grid_blocks=grid_blocks, pull_per_mille=pull_per_mille,
                        batch_work_per_block=batch_work_per_block,
                        hist_recount_per_mille=hist_recount_per_mille,
                        hist_push_per_mille=hist_push_per_mille, light_hist=light_hist)
        _check(_abi.lib().dmolp_ctx_set_tuning(self._h, C.byref(t)), self._h)

    def set_profiling(self, on: bool):
        _check(_abi.lib().dmolp_ctx_set_profiling(self._h, 1 if on else 0), self._h)

    def round_times(self, group: int) -> np.ndarray:
        n = C.c_int()
        _check(_abi.lib().dmolp_ctx_round_times(self._h, group, None, 0, C.byref(n)), self._h)
        out = np.zeros(n.value, np.float64)
        if n.value:
            _check(_abi.lib().dmolp_ctx_round_times(self._h, group, _p(out, C.c_double), n.value,
                                                    C.byref(n)), self._h)
        return out

    def set_hub_prefix(self, prefix: int):
        """Hubs phase 1 orders before its seed search (0: all); results do
        not depend on it (dmolp_ctx_set_hub_prefix)."""
        _check(_abi.lib().dmolp_ctx_set_hub_prefix(self._h, int(prefix)), self._h)

    def set_hub_device(self, dev_min: int, host_max: int):
        """Phase 1 searches the hubs of graphs with at least dev_min vertices
        on the device, finishing ranges shorter than host_max on the host;
        results do not depend on either (dmolp_ctx_set_hub_device)."""
        _check(_abi.lib().dmolp_ctx_set_hub_device(self._h, int(dev_min), int(host_max)), self._h)

    def top_degree_device(self, deg: np.ndarray, m: int, host_max: int = 1 << 15) -> np.ndarray:
        """The first min(m, n) vertices by (degree descending, id ascending),
        selected on the device (dmolp_top_degree_device)."""
        deg = np.ascontiguousarray(deg, np.int32)
        out = np.zeros(max(min(m, len(deg)), 1), np.int32)
        _check(_abi.lib().dmolp_top_degree_device(self._h, _p(deg, C.c_int32), len(deg), int(m),
                                                  int(host_max), _p(out, C.c_int32)), self._h)
        return out[:min(m, len(deg))].copy()

    def hubs_device(self, deg: np.ndarray, limit: int = -1, host_max: int = 1 << 15):
        """(hubs, n_hubs, first seed, its degree) of a degree sequence by the
        device hub search (dmolp_hubs_device)."""
        deg = np.ascontiguousarray(deg, np.int32)
        out = np.zeros(max(len(deg), 1), np.int32)
        nh, s, d = C.c_int64(), C.c_int32(), C.c_int32()
        _check(_abi.lib().dmolp_hubs_device(self._h, _p(deg, C.c_int32), len(deg), int(limit),
                                            int(host_max), _p(out, C.c_int32), C.byref(nh),
                                            C.byref(s), C.byref(d)), self._h)
        m = nh.value if limit < 0 else min(nh.value, limit)
        return out[:m].copy(), nh.value, s.value, d.value

    def probe_atomics(self, rows: int, share: int, reps: int = 32) -> tuple[float, int]:
        """(ms, adds) of scattered 32-bit adds into `rows` 64-byte rows, groups
        of `share` lanes sharing a row (dmolp_probe_atomics)."""
        ms = C.c_double()
        adds = C.c_int64()
        _check(_abi.lib().dmolp_probe_atomics(self._h, rows, share, reps, C.byref(ms), C.byref(adds)),
               self._h)
        return ms.value, adds.value

    def upload(self, off: np.ndarray, col: np.ndarray) -> "Graph":
        off = np.ascontiguousarray(off)
        col = np.ascontiguousarray(col, np.int32)
        ob = 8 if off.dtype == np.int64 else 4
        if ob == 4:
            off = off.astype(np.int32, copy=False)
        h = C.c_void_p()
        _check(_abi.lib().dmolp_graph_upload(self._h, len(off) - 1, len(col),
                                             off.ctypes.data_as(C.c_void_p), ob,
                                             _p(col, C.c_int32), C.byref(h)), self._h)
        return Graph(self, h)

    def generate(self, spec: _abi.GenSpec) -> "Graph":
        h = C.c_void_p()
        _check(_abi.lib().dmolp_graph_generate(self._h, C.byref(spec), C.byref(h)), self._h)
        return Graph(self, h)

    def penalties(self, k, vc, ic, t):
        vc = np.ascontiguousarray(vc, np.int64)
        ic = np.ascontiguousarray(ic, np.int64)
        pen = np.zeros(k, np.float64)
        re = np.zeros(k, np.float64)
        _check(_abi.lib().dmolp_penalties(self._h, k, _p(vc, C.c_int64), _p(ic, C.c_int64), t,
                                          _p(pen, C.c_double), _p(re, C.c_double)), self._h)
        return pen, re


@dataclass
class BatchJob:
    graph: "Graph"
    k: int
    theta: int
    max_rounds: int
    epsilon: float


@dataclass
class BatchResult:
    status: int
    rounds: int
    exit_reason: int
    n_seeds_fallback: int
    phase1_cut: int
    edge_cut: int
    lp_edges: int
    vertex_balance: float
    edge_balance: float
    labels: np.ndarray | None


def batch_run(ctx: "Context", jobs, want_labels=False):
    """Partition every job's graph (phase 1 + phase 2) in one job set.
    Returns (results, device_ms)."""
    n = len(jobs)
    arr = (_abi.Job * max(n, 1))()
    outs = []
    for i, j in enumerate(jobs):
        lab = np.zeros(j.graph.n, np.int32) if want_labels else None
        outs.append(lab)
        arr[i] = _abi.Job(graph=j.graph.handle, k=j.k, theta=j.theta, max_rounds=j.max_rounds,
                          reserved=0, epsilon=j.epsilon,
                          labels_out=_p(lab, C.c_int32) if lab is not None else None)
    res = (_abi.Result * max(n, 1))()
    ms = C.c_double()
    _check(_abi.lib().dmolp_batch_run(ctx.handle, n, arr, res, C.byref(ms)), ctx.handle)
    out = []
    for i in range(n):
        r = res[i]
        out.append(BatchResult(r.status, r.rounds, r.exit_reason, r.n_seeds_fallback, r.phase1_cut,
                               r.edge_cut, r.lp_edges, r.vertex_balance, r.edge_balance, outs[i]))
    return out, ms.value


def _job_array(jobs):
    arr = (_abi.Job * max(len(jobs), 1))()
    for i, j in enumerate(jobs):
        arr[i] = _abi.Job(graph=j.graph.handle, k=j.k, theta=j.theta, max_rounds=j.max_rounds,
                          reserved=0, epsilon=j.epsilon, labels_out=None)
    return arr


def batch_autotune(ctx: "Context", jobs, max_trials=0) -> _abi.Tuning:
    """Pick launch parameters for this batch (kept in ctx); returns them."""
    t = _abi.Tuning()
    _check(_abi.lib().dmolp_batch_autotune(ctx.handle, len(jobs), _job_array(jobs), max_trials,
                                           C.byref(t)), ctx.handle)
    return t


class Server:
    """Served mode: continuous batching of partition jobs on ctx's device.
    step() must be called from the thread that owns ctx."""

    def __init__(self, ctx: "Context", budget_bytes: int, max_active: int):
        self.ctx = ctx
        self._h = C.c_void_p()
        _check(_abi.lib().dmolp_server_create(ctx.handle, int(budget_bytes), int(max_active),
                                              C.byref(self._h)), ctx.handle)
        self._n = {}

    def close(self):
        if self._h:
            _abi.lib().dmolp_server_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def clock(self) -> float:
        return _abi.lib().dmolp_server_clock(self._h)

    def submit(self, k, theta, max_rounds, epsilon, spec=None, graph=None, want_labels=False,
               arrival_s=-1.0) -> int:
        req = _abi.Request(graph=graph.handle if graph is not None else None,
                           spec=spec if spec is not None else _abi.GenSpec(), k=k, theta=theta,
                           max_rounds=max_rounds, want_labels=1 if want_labels else 0,
                           epsilon=epsilon, arrival_s=arrival_s)
        t = C.c_int64()
        _check(_abi.lib().dmolp_server_submit(self._h, C.byref(req), C.byref(t)))
        if want_labels:
            self._n[t.value] = graph.n if graph is not None else _spec_n(spec)
        return t.value

    def step(self):
        a, q = C.c_int32(), C.c_int32()
        _check(_abi.lib().dmolp_server_step(self._h, C.byref(a), C.byref(q)))
        return a.value, q.value

    def take(self, ticket):
        """(Served, labels or None) when the job has finished, else None."""
        out = _abi.Served()
        n = self._n.get(ticket)
        lab = np.zeros(n, np.int32) if n is not None else None
        rc = _abi.lib().dmolp_server_take(self._h, ticket, C.byref(out),
                                          _p(lab, C.c_int32) if lab is not None else None)
        if rc < 0:
            _check(rc)
        if rc == 0:
            return None
        self._n.pop(ticket, None)
        return out, lab


def _spec_n(spec) -> int:
    if spec.kind == GEN_RMAT:
        return 1 << spec.scale
    if spec.kind == GEN_GRID:
        return spec.rows * spec.cols
    return spec.n


class Graph:
    def __init__(self, ctx: Context, h):
        self.ctx = ctx
        self._h = h
        n, nnz = C.c_int64(), C.c_int64()
        _check(_abi.lib().dmolp_graph_info(h, C.byref(n), C.byref(nnz)))
        self.n, self.nnz = n.value, nnz.value

    @property
    def handle(self):
        return self._h

    def free(self):
        if self._h:
            _abi.lib().dmolp_graph_free(self.ctx.handle, self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            if self.ctx.handle:
                self.free()
        except Exception:
            pass

    def download(self):
        off = np.zeros(self.n + 1, np.int64)
        col = np.zeros(self.nnz, np.int32)
        _check(_abi.lib().dmolp_graph_download(self.ctx.handle, self._h, _p(off, C.c_int64),
                                               _p(col, C.c_int32)), self.ctx.handle)
        return off, col

    def set_labels(self, labels):
        labels = np.ascontiguousarray(labels, np.int32)
        assert labels.shape == (self.n,)
        _check(_abi.lib().dmolp_graph_set_labels(self.ctx.handle, self._h, _p(labels, C.c_int32)),
               self.ctx.handle)

    def checkpoint_labels(self, restore: bool = False):
        _check(_abi.lib().dmolp_graph_labels_checkpoint(self.ctx.handle, self._h,
                                                        1 if restore else 0), self.ctx.handle)

    def bins(self):
        """(heavy rows, CSR entries they hold) under the context's tuning."""
        h, e = C.c_int64(), C.c_int64()
        _check(_abi.lib().dmolp_graph_bins(self.ctx.handle, self._h, C.byref(h), C.byref(e)),
               self.ctx.handle)
        return h.value, e.value

    def get_labels(self) -> np.ndarray:
        out = np.zeros(self.n, np.int32)
        _check(_abi.lib().dmolp_graph_get_labels(self.ctx.handle, self._h, _p(out, C.c_int32)),
               self.ctx.handle)
        return out

    def phase1(self, k: int, theta: int, want_labels=True) -> Phase1Result:
        labels = np.zeros(self.n, np.int32) if want_labels else None
        seeds = np.zeros(k, np.int32)
        vc = np.zeros(k, np.int64)
        ec = np.zeros(k, np.int64)
        m = _abi.Phase1Metrics()
        _check(_abi.lib().dmolp_phase1(
            self.ctx.handle, self._h, k, theta,
            _p(labels, C.c_int32) if want_labels else None, _p(seeds, C.c_int32),
            _p(vc, C.c_int64), _p(ec, C.c_int64), C.byref(m)), self.ctx.handle)
        return Phase1Result(labels, seeds, vc, ec, m)

    def phase1_random(self, k: int, want_labels=True) -> Phase1Result:
        """Random initial mode: label = vertex id mod k (random.cpp:71-74)."""
        labels = np.zeros(self.n, np.int32) if want_labels else None
        vc = np.zeros(k, np.int64)
        ec = np.zeros(k, np.int64)
        m = _abi.Phase1Metrics()
        _check(_abi.lib().dmolp_phase1_random(
            self.ctx.handle, self._h, k, _p(labels, C.c_int32) if want_labels else None,
            _p(vc, C.c_int64), _p(ec, C.c_int64), C.byref(m)), self.ctx.handle)
        return Phase1Result(labels, np.zeros(0, np.int32), vc, ec, m)

    def phase2(self, k: int, max_rounds: int, epsilon: float, want_labels=True,
               want_trace=True) -> Phase2Result:
        vc = np.zeros(k, np.int64)
        ic = np.zeros(k, np.int64)
        m = _abi.Phase2Metrics()
        tr_arrays = {}
        tr = None
        if want_trace and max_rounds > 0:
            tr = _abi.Trace()
            for name, _ in _abi.Trace._fields_:
                a = np.full(max_rounds, -7, np.int64)
                tr_arrays[name] = a
                setattr(tr, name, _p(a, C.c_int64))
        _check(_abi.lib().dmolp_phase2(self.ctx.handle, self._h, k, max_rounds, epsilon,
                                       _p(vc, C.c_int64), _p(ic, C.c_int64), C.byref(m),
                                       C.byref(tr) if tr is not None else None),
               self.ctx.handle)
        trace = {name: a[:m.rounds].copy() for name, a in tr_arrays.items()}
        labels = self.get_labels() if want_labels else None
        return Phase2Result(labels, vc, ic, m, trace)

    def stats(self, k, labels):
        labels = np.ascontiguousarray(labels, np.int32)
        vc = np.zeros(k, np.int64)
        ic = np.zeros(k, np.int64)
        cut = C.c_int64()
        _check(_abi.lib().dmolp_stats(self.ctx.handle, self._h, k, _p(labels, C.c_int32),
                                      _p(vc, C.c_int64), _p(ic, C.c_int64), C.byref(cut)),
               self.ctx.handle)
        return vc, ic, cut.value

    def stats_noalloc(self, k, labels):
        """stats() through the entry point that allocates no device memory."""
        labels = np.ascontiguousarray(labels, np.int32)
        vc = np.zeros(k, np.int64)
        ic = np.zeros(k, np.int64)
        cut = C.c_int64()
        _check(_abi.lib().dmolp_stats_noalloc(self.ctx.handle, self._h, k, _p(labels, C.c_int32),
                                              _p(vc, C.c_int64), _p(ic, C.c_int64),
                                              C.byref(cut)), self.ctx.handle)
        return vc, ic, cut.value

    def select_boundary(self, labels, prev, re, k, t):
        labels = np.ascontiguousarray(labels, np.int32)
        prev = np.ascontiguousarray(prev, np.int32)
        re = np.ascontiguousarray(re, np.float64)
        out = np.zeros(self.n, np.int32)
        nout = C.c_int64()
        _check(_abi.lib().dmolp_select_boundary(
            self.ctx.handle, self._h, _p(labels, C.c_int32), _p(prev, C.c_int32), len(prev),
            _p(re, C.c_double), k, t, _p(out, C.c_int32), C.byref(nout)), self.ctx.handle)
        return out[:nout.value].copy()

    def lp_round(self, labels, b, pen, k):
        labels = np.ascontiguousarray(labels, np.int32)
        b = np.ascontiguousarray(b, np.int32)
        pen = np.ascontiguousarray(pen, np.float64)
        out = np.zeros(self.n, np.int32)
        _check(_abi.lib().dmolp_lp_round(self.ctx.handle, self._h, _p(labels, C.c_int32),
                                         _p(b, C.c_int32), len(b), _p(pen, C.c_double), k,
                                         _p(out, C.c_int32)), self.ctx.handle)
        return out


class Shares:
    """P ranks' shares of a graph with on-device ghost labels (the halo);
    built from the graph's resident labels."""

    def __init__(self, g: "Graph", nranks: int):
        self.g = g
        self.ctx = g.ctx
        self.P = nranks
        self._h = C.c_void_p()
        _check(_abi.lib().dmolp_shares_create(self.ctx.handle, g.handle, nranks,
                                              C.byref(self._h)), self.ctx.handle)

    def close(self):
        if self._h:
            _abi.lib().dmolp_shares_destroy(self.ctx.handle, self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            if self.ctx.handle:
                self.close()
        except Exception:
            pass

    def info(self, r):
        nl, ng, nnz = C.c_int64(), C.c_int64(), C.c_int64()
        _check(_abi.lib().dmolp_shares_info(self._h, r, C.byref(nl), C.byref(ng), C.byref(nnz)))
        return nl.value, ng.value, nnz.value

    def download(self, r):
        """(off, col in unified ids, ghost global ids in slot order, labels
        of the owned vertices then of the ghosts) of share r."""
        nl, ng, nnz = self.info(r)
        off = np.zeros(nl + 1, np.int64)
        col = np.zeros(max(nnz, 1), np.int32)
        gg = np.zeros(max(ng, 1), np.int32)
        lab = np.zeros(max(nl + ng, 1), np.int32)
        _check(_abi.lib().dmolp_shares_download(self.ctx.handle, self._h, r, _p(off, C.c_int64),
                                                _p(col, C.c_int32), _p(gg, C.c_int32),
                                                _p(lab, C.c_int32)), self.ctx.handle)
        return off, col[:nnz], gg[:ng], lab[:nl + ng]

    def refresh(self, r, gid, label) -> int:
        gid = np.ascontiguousarray(gid, np.int32)
        label = np.ascontiguousarray(label, np.int32)
        u = C.c_int64()
        _check(_abi.lib().dmolp_halo_refresh(self.ctx.handle, self._h, r, _p(gid, C.c_int32),
                                             _p(label, C.c_int32), len(gid), C.byref(u)),
               self.ctx.handle)
        return u.value

    def phase2(self, k, max_rounds, epsilon, scope) -> Phase2Result:
        labels = np.zeros(self.g.n, np.int32)
        vc = np.zeros(k, np.int64)
        ic = np.zeros(k, np.int64)
        m = _abi.Phase2Metrics()
        tr = _abi.Trace()
        arrays = {}
        for name, _ in _abi.Trace._fields_:
            a = np.full(max(max_rounds, 1), -7, np.int64)
            arrays[name] = a
            setattr(tr, name, _p(a, C.c_int64))
        gu = np.full(max(max_rounds, 1), -7, np.int64)
        _check(_abi.lib().dmolp_phase2_shares(self.ctx.handle, self._h, k, max_rounds, epsilon,
                                              scope, _p(labels, C.c_int32), _p(vc, C.c_int64),
                                              _p(ic, C.c_int64), C.byref(m), C.byref(tr),
                                              _p(gu, C.c_int64)), self.ctx.handle)
        trace = {name: a[:m.rounds].copy() for name, a in arrays.items()}
        trace["ghost_updates"] = gu[:m.rounds].copy()
        return Phase2Result(labels, vc, ic, m, trace)
