"""ctypes declarations of include/dmolp_hip.h.

Loading fails loudly when libdmolp_hip.so is missing: there is no CPU
fallback for the partitioning path.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libdmolp_hip.so"

i32, i64, u64, f64 = C.c_int32, C.c_int64, C.c_uint64, C.c_double
P = C.POINTER


class Tuning(C.Structure):
    _fields_ = [("heavy_degree", i32), ("chunk_edges", i32), ("block_threads", i32),
                ("grid_blocks", i32), ("pull_per_mille", i32),
                ("batch_work_per_block", i32), ("hist_recount_per_mille", i32),
                ("hist_push_per_mille", i32), ("light_hist", i32)]


class GenSpec(C.Structure):
    _fields_ = [("kind", i32), ("scale", i32), ("edge_factor", i32), ("rows", i32),
                ("cols", i32), ("reserved", i32), ("a", f64), ("b", f64), ("c", f64),
                ("gamma", f64), ("dmax", f64), ("n", i64), ("m", i64), ("seed", u64)]


class Csr(C.Structure):
    _fields_ = [("n", i64), ("nnz", i64), ("off", P(i64)), ("col", P(i32))]


class Phase1Metrics(C.Structure):
    _fields_ = [("initial_edge_cut", i64), ("initial_vertex_balance", f64),
                ("initial_edge_balance", f64), ("total_vertices", i64), ("total_edges", i64),
                ("first_seed", i32), ("first_seed_degree", i32), ("n_hubs", i64),
                ("n_seeds_found", i32), ("n_seeds_fallback", i32), ("rounds", i64),
                ("fill_round", i32), ("reserved", i32), ("device_ms", f64), ("host_ms", f64)]


class Phase2Metrics(C.Structure):
    _fields_ = [("edge_cut", i64), ("vertex_balance", f64), ("edge_balance", f64),
                ("total_vertices", i64), ("total_edges", i64), ("rounds", i32),
                ("exit_reason", i32), ("lp_edges", i64), ("initial_cut", i64),
                ("device_ms", f64)]


class Trace(C.Structure):
    _fields_ = [(name, P(i64)) for name in (
        "prev_count", "prev_deg", "cand_count", "bd_count", "b_count", "b_deg",
        "chg_count", "chg_deg", "cut", "b_light_count", "b_light_deg", "chg_light_count",
        "chg_light_deg")]


class Job(C.Structure):
    _fields_ = [("graph", C.c_void_p), ("k", i32), ("theta", i32), ("max_rounds", i32),
                ("reserved", i32), ("epsilon", f64), ("labels_out", P(i32))]


class Result(C.Structure):
    _fields_ = [("status", i32), ("rounds", i32), ("exit_reason", i32),
                ("n_seeds_fallback", i32), ("phase1_cut", i64), ("edge_cut", i64),
                ("lp_edges", i64), ("vertex_balance", f64), ("edge_balance", f64)]


class Request(C.Structure):
    _fields_ = [("graph", C.c_void_p), ("spec", GenSpec), ("k", i32), ("theta", i32),
                ("max_rounds", i32), ("want_labels", i32), ("epsilon", f64), ("arrival_s", f64)]


class Served(C.Structure):
    _fields_ = [("result", Result), ("n", i64), ("nnz", i64), ("arrival_s", f64),
                ("admit_s", f64), ("done_s", f64)]


# Every function include/dmolp_hip.h declares: name -> (restype, argtypes).
VP = C.c_void_p
SIGNATURES = {
    "dmolp_abi_version": (C.c_int, []),
    "dmolp_status_string": (C.c_char_p, [C.c_int]),
    "dmolp_last_error": (C.c_char_p, [VP]),
    "dmolp_ctx_create": (C.c_int, [C.c_int, P(VP)]),
    "dmolp_ctx_destroy": (C.c_int, [VP]),
    "dmolp_ctx_stream": (VP, [VP]),
    "dmolp_ctx_get_tuning": (C.c_int, [VP, P(Tuning)]),
    "dmolp_ctx_set_profiling": (C.c_int, [VP, C.c_int]),
    "dmolp_ctx_round_times": (C.c_int, [VP, C.c_int, P(f64), C.c_int, P(C.c_int)]),
    "dmolp_graph_labels_checkpoint": (C.c_int, [VP, VP, C.c_int]),
    "dmolp_ctx_set_tuning": (C.c_int, [VP, P(Tuning)]),
    "dmolp_gen_host": (C.c_int, [P(GenSpec), P(Csr)]),
    "dmolp_csr_free": (None, [P(Csr)]),
    "dmolp_hubs_host": (C.c_int, [P(i32), i64, P(i32), P(i64)]),
    "dmolp_ctx_set_hub_prefix": (C.c_int, [VP, i64]),
    "dmolp_first_seed_host": (C.c_int, [P(i32), i64, P(i32), P(i32)]),
    "dmolp_hubs_device": (C.c_int, [VP, P(i32), i64, i64, i64, P(i32), P(i64), P(i32), P(i32)]),
    "dmolp_ctx_set_hub_device": (C.c_int, [VP, i64, i64]),
    "dmolp_top_degree_device": (C.c_int, [VP, P(i32), i64, i64, i64, P(i32)]),
    "dmolp_probe_atomics": (C.c_int, [VP, i64, i32, i32, P(f64), P(i64)]),
    "dmolp_stats_noalloc": (C.c_int, [VP, VP, i32, P(i32), P(i64), P(i64), P(i64)]),
    "dmolp_shares_create": (C.c_int, [VP, VP, i32, P(VP)]),
    "dmolp_shares_destroy": (C.c_int, [VP, VP]),
    "dmolp_shares_info": (C.c_int, [VP, i32, P(i64), P(i64), P(i64)]),
    "dmolp_shares_download": (C.c_int, [VP, VP, i32, P(i64), P(i32), P(i32), P(i32)]),
    "dmolp_halo_refresh": (C.c_int, [VP, VP, i32, P(i32), P(i32), i64, P(i64)]),
    "dmolp_phase2_shares": (C.c_int, [VP, VP, i32, i32, f64, i32, P(i32), P(i64), P(i64),
                                      P(Phase2Metrics), P(Trace), P(i64)]),
    "dmolp_read_adjacency_text": (C.c_int, [C.c_char_p, P(Csr)]),
    "dmolp_write_adjacency_text": (C.c_int, [C.c_char_p, i64, P(i64), P(i32)]),
    "dmolp_write_csr_binary": (C.c_int, [C.c_char_p, i64, P(i64), P(i32)]),
    "dmolp_read_csr_binary": (C.c_int, [C.c_char_p, P(Csr)]),
    "dmolp_load_adjacency_cached": (C.c_int, [C.c_char_p, P(Csr), P(i32)]),
    "dmolp_graph_upload": (C.c_int, [VP, i64, i64, VP, C.c_int, P(i32), P(VP)]),
    "dmolp_graph_generate": (C.c_int, [VP, P(GenSpec), P(VP)]),
    "dmolp_graph_info": (C.c_int, [VP, P(i64), P(i64)]),
    "dmolp_graph_download": (C.c_int, [VP, VP, P(i64), P(i32)]),
    "dmolp_graph_free": (C.c_int, [VP, VP]),
    "dmolp_graph_set_labels": (C.c_int, [VP, VP, P(i32)]),
    "dmolp_graph_get_labels": (C.c_int, [VP, VP, P(i32)]),
    "dmolp_phase1": (C.c_int, [VP, VP, i32, i32, P(i32), P(i32), P(i64), P(i64),
                               P(Phase1Metrics)]),
    "dmolp_phase1_random": (C.c_int, [VP, VP, i32, P(i32), P(i64), P(i64), P(Phase1Metrics)]),
    "dmolp_phase2": (C.c_int, [VP, VP, i32, i32, f64, P(i64), P(i64), P(Phase2Metrics),
                               P(Trace)]),
    "dmolp_stats": (C.c_int, [VP, VP, i32, P(i32), P(i64), P(i64), P(i64)]),
    "dmolp_penalties": (C.c_int, [VP, i32, P(i64), P(i64), i32, P(f64), P(f64)]),
    "dmolp_select_boundary": (C.c_int, [VP, VP, P(i32), P(i32), i64, P(f64), i32, i32,
                                        P(i32), P(i64)]),
    "dmolp_lp_round": (C.c_int, [VP, VP, P(i32), P(i32), i64, P(f64), i32, P(i32)]),
    "dmolp_batch_run": (C.c_int, [VP, i32, P(Job), P(Result), P(f64)]),
    "dmolp_graph_bins": (C.c_int, [VP, VP, P(i64), P(i64)]),
    "dmolp_batch_autotune": (C.c_int, [VP, i32, P(Job), i32, P(Tuning)]),
    "dmolp_server_create": (C.c_int, [VP, i64, i32, P(VP)]),
    "dmolp_server_destroy": (C.c_int, [VP]),
    "dmolp_server_clock": (f64, [VP]),
    "dmolp_server_submit": (C.c_int, [VP, P(Request), P(i64)]),
    "dmolp_server_step": (C.c_int, [VP, P(i32), P(i32)]),
    "dmolp_server_take": (C.c_int, [VP, i64, P(Served), P(i32)]),
}

_lib = None


def lib() -> C.CDLL:
    """Load libdmolp_hip.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    path = os.environ.get("DMOLP_LIB", str(LIB_PATH))
    if not Path(path).exists():
        raise RuntimeError(
            f"{path} is missing: build it with `python -m dmolp_amd.build` "
            "(there is no CPU fallback for the partitioning path)")
    L = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L
