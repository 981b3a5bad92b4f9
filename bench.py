"""Phase-2 label-propagation throughput on one MI355X.

A step is one full phase-2 run (device-resident boundary LP rounds until an
exit condition) over the graph's phase-1 labels, which are restored on the
device before each step.  value = LP edges/sec = sum over rounds of
sum_{b in B_t} deg b, divided by the device time (hipEvents on the library's
stream) of the timed steps.

Default workload: the power-law follower graph -- Chung-Lu, 41,652,230
vertices, 1,468,365,182 sampled directed edges, degree exponent 2.0, expected
maximum degree 3e6, seed 3, symmetrised (~2.43e9 CSR entries); k=32,
theta=10000, 30 rounds, epsilon 0.05.  --config c2 runs R-MAT scale 22
(k=16, 20 rounds), --config c1 R-MAT scale 16 (k=8, 10 rounds).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0

from dmolp_amd.workloads import (C4_WORKLOAD, C5_WORKLOAD, CONFIGS, THETA,  # noqa: E402
                                 c4_specs, c5_jobs, spec_for)

# Algorithmic bytes of the LP decision per boundary vertex b (int64 offsets,
# int32 columns and labels): 2 offsets + own label read + label write, and a
# column plus a label per neighbour.
O_BYTES, C_BYTES, L_BYTES = 8, 4, 4
FIXTURES = ROOT / "tests" / "golden"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def sha(a) -> str:
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def host_info():
    """The host the CPU baseline ran on: logical CPUs, the CPUs this process
    may use, physical cores and the CPU model (lscpu)."""
    import subprocess
    info = {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0))}
    try:
        out = subprocess.run(["lscpu", "-p=CORE,SOCKET"], capture_output=True, text=True,
                             timeout=10).stdout
        info["physical_cores"] = len({ln for ln in out.splitlines() if ln and ln[0] != "#"})
        for ln in subprocess.run(["lscpu"], capture_output=True, text=True,
                                 timeout=10).stdout.splitlines():
            if ln.startswith("Model name:"):
                info["model"] = ln.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return info


def cpu_threads():
    """Threads for the multi-threaded CPU baseline: the physical cores this
    process may use, at most 16 (the host share of one GPU)."""
    info = host_info()
    return max(1, min(16, info["affinity"], info.get("physical_cores") or info["affinity"])), info


def oracle():
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_py as O
    return O


def timed_median(fn, reps):
    ts, out = [], None
    for _ in range(reps):
        t0 = time.perf_counter()
        out = fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), out


def cpu_baseline_full(off, col, k, R, eps, threads, reps=3):
    """The CPU oracle's phase 1 and phase 2 (R rounds) on the same graph,
    timed separately, median of `reps` runs, on `threads` threads and on one
    thread.  Returns the baseline record and the oracle's results."""
    O = oracle()
    rec = {}
    p1 = p2 = None
    for T in sorted({threads, 1}, reverse=True):
        O.set_threads(T)
        t1, p1 = timed_median(lambda: O.phase1(off, col, k, THETA), reps)
        t2, p2 = timed_median(lambda: O.phase2(off, col, k, p1.labels, R, eps), reps)
        rec[T] = {"phase1_s": t1, "phase2_s": t2, "lp_edges_per_s": p2.m.lp_edges / t2}
    top = rec[threads]
    return {
        "value": top["lp_edges_per_s"], "unit": "LP edges/s", "cores": threads, "kind": "port",
        "sample": f"the whole workload: oracle phase 1 and phase 2 ({p2.m.rounds} rounds, "
                  f"{p2.m.lp_edges} LP edges), median of {reps} runs each",
        "phase1_s": top["phase1_s"], "phase2_s": top["phase2_s"],
        "one_thread": rec[1],
    }, p1, p2


def cpu_baseline_round1(off, col, k, labels0, eps, threads):
    """The CPU oracle's first phase-2 round (initial statistics, round 1,
    final statistics) from the device's phase-1 labels, on `threads` threads
    and on one thread (one run each).  Returns the record and the oracle's
    one-round result."""
    O = oracle()
    rec, res = {}, None
    for T in sorted({threads, 1}, reverse=True):
        O.set_threads(T)
        t0 = time.perf_counter()
        res = O.phase2(off, col, k, labels0, 1, eps)
        el = time.perf_counter() - t0
        rec[T] = {"phase2_round1_s": el, "lp_edges_per_s": res.m.lp_edges / el}
    return {
        "value": rec[threads]["lp_edges_per_s"], "unit": "LP edges/s", "cores": threads,
        "kind": "port",
        "sample": f"oracle phase 2 limited to round 1 (initial statistics sweep, boundary "
                  f"selection over all vertices, {res.m.lp_edges} LP edges, final statistics) "
                  f"on the same graph and phase-1 labels, one run per thread count",
        "phase2_round1_s": rec[threads]["phase2_round1_s"],
        "one_thread": rec[1],
    }, res


def oracle_jobs_baseline(specs, rounds, eps, budget_s, unit):
    """The CPU oracle's phase 1 + phase 2 on jobs taken in order until about
    budget_s seconds of work, on all usable cores (generation excluded)."""
    from dmolp_amd import api
    T, hinfo = cpu_threads()
    O = oracle()
    O.set_threads(T)
    done, el = 0, 0.0
    for spec, k in specs:
        off, col = api.gen_host(spec)
        t0 = time.perf_counter()
        p1 = O.phase1(off, col, k, THETA)
        O.phase2(off, col, k, p1.labels, rounds, eps)
        el += time.perf_counter() - t0
        done += 1
        if el >= budget_s:
            break
    return {"value": done / el, "unit": unit, "cores": T, "kind": "port",
            "sample": f"oracle phase 1 + phase 2 on the first {done} jobs of the same list "
                      f"(generation excluded), {el:.1f} s on {T} threads", "host": hinfo}


def job_parity(fixture, results):
    """Every job's labels (SHA-256) and metrics against the oracle fixture."""
    if fixture is None or len(fixture["jobs"]) != len(results):
        return None
    bad = []
    for i, (w, r) in enumerate(zip(fixture["jobs"], results)):
        res, labels = r
        ok = (res.status == 0 and sha(labels) == w["labels_sha256"] and res.edge_cut == w["cut"]
              and res.phase1_cut == w["phase1_cut"] and res.rounds == w["rounds"]
              and res.exit_reason == w["exit_reason"] and res.lp_edges == w["lp_edges"]
              and res.vertex_balance == w["vertex_balance"]
              and res.edge_balance == w["edge_balance"])
        if not ok:
            bad.append(i)
    return {"jobs": len(results), "mismatched": bad, "ok": not bad}


def run_batched(a):
    """Batched mode: every graph's phase 1 + 10 LP rounds in one job set,
    against the same graphs partitioned one at a time (plain mode)."""
    from dmolp_amd import api
    theta, rounds, eps = THETA, 10, 0.03
    ctx = api.Context(0)
    specs = c4_specs()
    t0 = time.perf_counter()
    graphs = [ctx.generate(sp) for sp, _ in specs]
    n_tot = sum(g.n for g in graphs)
    nnz_tot = sum(g.nnz for g in graphs)
    log(f"[bench] generated {len(graphs)} graphs: n={n_tot} nnz={nnz_tot} in "
        f"{time.perf_counter() - t0:.2f} s")
    jobs = [api.BatchJob(g, k, theta, rounds, eps) for g, (_, k) in zip(graphs, specs)]
    # Plain: one graph at a time (phase 1 + phase 2 each), after one warm pass.
    plain_walls, plain_res = [], None
    for it in range(1 + a.warmup):
        t0 = time.perf_counter()
        res = []
        for j in jobs:
            j.graph.phase1(j.k, j.theta, want_labels=False)
            res.append(j.graph.phase2(j.k, j.max_rounds, j.epsilon, want_labels=True))
        plain_walls.append(time.perf_counter() - t0)
        plain_res = res
    plain_s = min(plain_walls[1:] if len(plain_walls) > 1 else plain_walls)
    log(f"[bench] plain: {len(jobs)} graphs in {plain_s:.3f} s")
    for _ in range(a.warmup):
        api.batch_run(ctx, jobs, want_labels=False)
    walls, dev = [], []
    bres = None
    for _ in range(a.steps):
        t0 = time.perf_counter()
        bres, ms = api.batch_run(ctx, jobs, want_labels=True)
        walls.append(time.perf_counter() - t0)
        dev.append(ms)
    batch_s = float(np.mean(walls))
    same_cut = all(b.edge_cut == p.metrics.edge_cut for b, p in zip(bres, plain_res))
    same_labels = all(np.array_equal(b.labels, p.labels) for b, p in zip(bres, plain_res))
    log(f"[bench] batched: {len(jobs)} graphs in {batch_s:.3f} s (device {np.mean(dev):.1f} ms), "
        f"speedup {plain_s / batch_s:.2f}x, identical cut {same_cut}, identical labels {same_labels}")
    fx_path = FIXTURES / "large_c4.json"
    fx = json.loads(fx_path.read_text()) if fx_path.exists() else None
    parity = job_parity(fx, [(b, b.labels) for b in bres])
    if parity is not None:
        parity["fixture"] = str(fx_path.relative_to(ROOT))
    parity = parity or {"ok": None}
    parity["batched_equals_plain"] = bool(same_cut and same_labels)
    log(f"[bench] parity {json.dumps(parity)}")
    cpu = None
    if not a.no_cpu_baseline:
        cpu = oracle_jobs_baseline(specs, rounds, eps, 15.0, "graphs/s")
        log(f"[bench] cpu baseline {cpu['value']:.2f} graphs/s ({cpu['sample']})")
    out = {
        "metric": "batched graphs/sec (phase 1 + 10 LP rounds per graph)",
        "value": len(jobs) / batch_s,
        "unit": "graphs/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * batch_s,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (project generator, seeded)",
        "config": {"workload": C4_WORKLOAD, "graphs": len(jobs), "n_total": n_tot,
                   "nnz_total": nnz_tot, "theta": theta, "max_rounds": rounds, "epsilon": eps},
        "plain_graphs_per_s": len(jobs) / plain_s,
        "speedup_vs_plain": plain_s / batch_s,
        "identical_cut": same_cut,
        "identical_labels": same_labels,
        "batch_device_ms": float(np.mean(dev)),
        "roofline": None,
        "cpu_baseline": cpu,
        "parity": parity,
    }
    for g in graphs:
        g.free()
    ctx.close()
    return out


def serve_leg(srv, jobs, lam, rng, theta, rounds, eps, graphs=None, keep=False):
    """One arrival rate: Poisson arrivals of the jobs (resident graphs, or
    generator specs when graphs is None); returns (per-request records,
    labels of each request when keep, span in seconds, arrival offsets)."""
    gaps = rng.exponential(1.0 / lam, len(jobs))
    arr_rel = np.cumsum(gaps) - gaps[0]
    arr = srv.clock() + 0.01 + arr_rel
    tickets, done, kept, nxt = [], {}, {}, 0
    while len(done) < len(jobs):
        now = srv.clock()
        while nxt < len(jobs) and arr[nxt] <= now:
            sp, k = jobs[nxt]
            if graphs is None:
                t = srv.submit(k, theta, rounds, eps, spec=sp, arrival_s=float(arr[nxt]),
                               want_labels=keep)
            else:
                t = srv.submit(k, theta, rounds, eps, graph=graphs[nxt],
                               arrival_s=float(arr[nxt]), want_labels=keep)
            tickets.append(t)
            nxt += 1
        active, queued = srv.step()
        for t in tickets:
            if t not in done:
                r = srv.take(t)
                if r is not None:
                    done[t] = r[0]
                    if keep:
                        kept[t] = r[1]
        if active == 0 and queued == 0 and nxt < len(jobs) and len(done) == nxt:
            time.sleep(max(0.0, min(0.05, arr[nxt] - srv.clock())))
    recs = [done[t] for t in tickets]
    labels = [kept[t] for t in tickets] if keep else None
    span = max(r.done_s for r in recs) - float(arr[0])
    return recs, labels, span, arr_rel


def plain_fifo(arr_rel, service):
    """A one-job-at-a-time server over measured service times: (jobs/s,
    latencies)."""
    fin, lat = 0.0, []
    for x, sv in zip(arr_rel, service):
        fin = max(fin, x) + sv
        lat.append(fin - x)
    return len(service) / fin, np.array(lat)


def run_served(a):
    """Served mode: Poisson arrivals at several rates; continuous batching
    against a one-job-at-a-time plain server (FIFO over measured plain
    service times of the same jobs).  Requests name CSR graphs already
    resident in HBM when timing starts (as in the plain and batched lines);
    a second leg at the highest rate has the server generate each request's
    graph on admission, against plain service times that include the same
    generation."""
    from dmolp_amd import api
    theta, rounds, eps = THETA, 20, 0.03
    rates = [10, 30, 100, 300, 1000]
    jobs = c5_jobs(a.jobs)
    ctx = api.Context(0)
    t0 = time.perf_counter()
    graphs = [ctx.generate(sp) for sp, _ in jobs]
    log(f"[bench] generated {len(graphs)} graphs: n={sum(g.n for g in graphs)} "
        f"nnz={sum(g.nnz for g in graphs)} in {time.perf_counter() - t0:.2f} s")

    def plain_times(generate):
        # Second of two passes (the first allocates each graph's work arrays).
        for _ in range(2):
            service, dev_ms, host_ms = [], [], []
            for i, (sp, k) in enumerate(jobs):
                t0 = time.perf_counter()
                g = ctx.generate(sp) if generate else graphs[i]
                p1 = g.phase1(k, theta, want_labels=False)
                p2 = g.phase2(k, rounds, eps, want_labels=False, want_trace=False)
                if generate:
                    g.free()
                service.append(time.perf_counter() - t0)
                dev_ms.append(p1.metrics.device_ms + p2.metrics.device_ms)
                host_ms.append(p1.metrics.host_ms)
        return service, dev_ms, host_ms

    service, dev_ms, host_ms = plain_times(False)
    order = np.argsort(service)[::-1][:5]
    log(f"[bench] plain service: total {sum(service):.2f} s over {len(jobs)} jobs, "
        f"max {max(service) * 1e3:.1f} ms; phase 1 + 2 device time {sum(dev_ms) / 1e3:.2f} s, "
        f"hub sorts {sum(host_ms) / 1e3:.2f} s; slowest jobs (ms service/device/hub): " +
        ", ".join(f"{service[i] * 1e3:.0f}/{dev_ms[i]:.0f}/{host_ms[i]:.0f}" for i in order))
    srv = api.Server(ctx, 200 << 30, a.max_active)
    per_rate = {}
    rng = np.random.default_rng(5)
    # Timed legs return metrics only, as the plain service times above do;
    # the labels are checked in untimed legs at the highest rate below.
    for lam in rates:
        recs, _, span, arr_rel = serve_leg(srv, jobs, lam, rng, theta, rounds, eps, graphs=graphs)
        lat = np.array([r.done_s - r.arrival_s for r in recs])
        bad = [i for i, r in enumerate(recs) if r.result.status != 0]
        pj, plat = plain_fifo(arr_rel, service)
        per_rate[str(lam)] = {
            "jobs_per_s": len(jobs) / span, "p50_ms": 1e3 * float(np.percentile(lat, 50)),
            "p99_ms": 1e3 * float(np.percentile(lat, 99)),
            "plain_jobs_per_s": pj, "plain_p50_ms": 1e3 * float(np.percentile(plat, 50)),
            "plain_p99_ms": 1e3 * float(np.percentile(plat, 99)), "failed": len(bad)}
        log(f"[bench] rate {lam}/s: served {per_rate[str(lam)]['jobs_per_s']:.1f} jobs/s "
            f"p50 {per_rate[str(lam)]['p50_ms']:.1f} ms p99 {per_rate[str(lam)]['p99_ms']:.1f} ms | "
            f"plain {pj:.1f} jobs/s "
            f"p50 {per_rate[str(lam)]['plain_p50_ms']:.1f} ms p99 {per_rate[str(lam)]['plain_p99_ms']:.1f} ms")
    # The same requests with generation on admission, at the highest rate.
    gservice, _, _ = plain_times(True)
    recs, _, gspan, garr = serve_leg(srv, jobs, rates[-1], rng, theta, rounds, eps, graphs=None)
    gpj, gplat = plain_fifo(garr, gservice)
    glat = np.array([r.done_s - r.arrival_s for r in recs])
    generated = {
        "rate_per_s": rates[-1], "jobs_per_s": len(jobs) / gspan,
        "p50_ms": 1e3 * float(np.percentile(glat, 50)), "p99_ms": 1e3 * float(np.percentile(glat, 99)),
        "plain_jobs_per_s": gpj, "plain_p50_ms": 1e3 * float(np.percentile(gplat, 50)),
        "plain_p99_ms": 1e3 * float(np.percentile(gplat, 99)),
        "plain_service_s": sum(gservice)}
    generated["speedup_vs_plain"] = generated["jobs_per_s"] / gpj
    log(f"[bench] with generation on admission, rate {rates[-1]}/s: served "
        f"{generated['jobs_per_s']:.1f} jobs/s p50 {generated['p50_ms']:.1f} ms p99 "
        f"{generated['p99_ms']:.1f} ms | plain {gpj:.1f} jobs/s")
    # Parity legs (untimed): every request's labels and metrics, resident
    # and generated, against the oracle's fixture.
    recs, labels, _, _ = serve_leg(srv, jobs, rates[-1], rng, theta, rounds, eps, graphs=graphs,
                                   keep=True)
    last_results = [(r.result, lab) for r, lab in zip(recs, labels)]
    recs, glabels, _, _ = serve_leg(srv, jobs, rates[-1], rng, theta, rounds, eps, graphs=None,
                                    keep=True)
    srv.close()
    fx_path = FIXTURES / "large_c5.json"
    fx = json.loads(fx_path.read_text()) if fx_path.exists() else None
    parity = job_parity(fx, last_results)
    gparity = job_parity(fx, [(r.result, lab) for r, lab in zip(recs, glabels)])
    if parity is not None:
        parity["fixture"] = str(fx_path.relative_to(ROOT))
        parity["generated_leg_ok"] = gparity["ok"] if gparity else None
        parity["ok"] = bool(parity["ok"] and parity["generated_leg_ok"])
    parity = parity or {"ok": None}
    log(f"[bench] parity {json.dumps(parity)}")
    for g in graphs:
        g.free()
    cpu = None
    if not a.no_cpu_baseline:
        cpu = oracle_jobs_baseline(jobs, rounds, eps, 15.0, "jobs/s")
        log(f"[bench] cpu baseline {cpu['value']:.2f} jobs/s ({cpu['sample']})")
    top = per_rate[str(rates[-1])]
    out = {
        "metric": "served jobs/sec at saturation (Poisson arrivals), with p50/p99 latency",
        "value": top["jobs_per_s"],
        "unit": "jobs/s",
        "n_gpus": 1,
        "steps": len(rates),
        "warmup": 0,
        "ms_per_step": 1e3 / top["jobs_per_s"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (project generator, seeded)",
        "config": {"workload": C5_WORKLOAD, "jobs_per_rate": len(jobs), "rates_per_s": rates,
                   "theta": theta, "max_rounds": rounds, "epsilon": eps,
                   "max_active": a.max_active, "budget_gib": 200,
                   "inputs": "CSR resident in HBM before timing (generated_on_admission: the "
                             "server generates each graph)"},
        "per_rate": per_rate,
        "speedup_vs_plain_at_saturation": top["jobs_per_s"] / top["plain_jobs_per_s"],
        "generated_on_admission": generated,
        "plain_device_s": sum(dev_ms) / 1e3,
        "plain_service_s": sum(service),
        "roofline": None,
        "cpu_baseline": cpu,
        "parity": parity,
    }
    ctx.close()
    return out


def kernel_models(g, k, tr, per_round):
    """Algorithmic bytes of the hot kernels, per launch (one launch per round,
    the initial sweep once per step), from the step's trace, next to their
    measured per-launch times.  Offsets o = 8, columns c = 4, labels l = 4;
    a kept light-row histogram is 16 bytes per 8 labels below min(k, 32).
    Also returns the step priced with the reference algorithm's bytes."""
    o, c, l = O_BYTES, C_BYTES, L_BYTES
    E = min(k, 32)
    row_b = 16 * ((E + 7) // 8)
    h_rows, h_entries = g.bins()
    n, nnz = g.n, g.nnz
    light_rows, light_entries = n - h_rows, nnz - h_entries
    dl_n = tr["chg_light_count"].astype(np.float64)
    dl_deg = tr["chg_light_deg"].astype(np.float64)
    bl_n = tr["b_light_count"].astype(np.float64)
    bd = tr["bd_count"].astype(np.float64)
    pulled = tr["cand_count"] < 0
    kern = {
        # Per changed light row: offsets, old label, new label, its differing
        # count; per entry: the column, the neighbour's old|new label pair
        # (2 bytes) and one word of the neighbour's kept row read and written.
        "k_delta_light": (per_round["delta_light"], (2 * o + 3 * l) * dl_n + (c + 2 + 8) * dl_deg),
        # The selection's take pass fused with the light LP: per vertex its
        # record entry, per record its label byte, the B_t bitmap written
        # (n/8 bytes); per light member of B_t its kept row; per changed row
        # its new label, label pair and Delta entry.
        "k_lp_light_take": (per_round["lp_light"],
                            4 * n + 1 * bd + n / 8 + row_b * bl_n + (l + 2 + 4) * dl_n),
        # Pull rounds: per vertex its differing count, offset, witness (the
        # neighbour that last made it a candidate) and record entry (rows
        # scanned for pending vertices not counted).  The group holds the
        # round start and the other candidate kernels, which do no work in
        # pull rounds.
        "k_cand_pull": (per_round["candidates"], np.where(pulled, (3 * l + o) * n, 0.0)),
        # Per light row: offset, label, differing count and kept row written;
        # per light entry: the column and a label byte.
        "k_init_light": (per_round["init_light"],
                         np.array([(o + 2 * l + row_b) * light_rows + (c + 1) * light_entries],
                                  np.float64)),
    }
    # The reference algorithm's bytes for the whole step: candidates pushed
    # from B_{t-1}, the boundary test reading at least one entry per
    # candidate, selection keys, recounted LP histograms, incremental
    # statistics and the initial sweep.
    s0 = (n + 1) * o + (c + l) * nnz + l * n
    cand = np.maximum(tr["cand_count"], tr["bd_count"]).astype(np.float64)
    rounds = (2 * o * tr["prev_count"] + c * tr["prev_deg"]
              + (2 * o + l + (c + l)) * cand
              + 8 * bd
              + (2 * o + 2 * l) * tr["b_count"] + (c + l) * tr["b_deg"]
              + 2 * o * tr["chg_count"] + (c + l) * tr["chg_deg"])
    return kern, float(s0 + rounds.astype(np.float64).sum())


def measured_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC
    summary (FETCH_SIZE + WRITE_SIZE, copy-calibrated), if one exists."""
    p = ROOT / "profiles" / "traffic.json"
    if not p.exists():
        return None
    try:
        d = json.loads(p.read_text())
        return d[workload][kernel]["hbm_bytes_per_launch"]
    except (KeyError, ValueError):
        return None


def measured_atomics(workload, kernel):
    """L2 atomic requests per launch of `kernel` (rocprofv3 --pmc TCC_ATOMIC_sum)
    from the committed table, if one exists."""
    p = ROOT / "profiles" / "traffic.json"
    if not p.exists():
        return None
    try:
        return json.loads(p.read_text())[workload][kernel]["tcc_atomic_per_launch"]
    except (KeyError, ValueError):
        return None


def atomic_ceiling(ctx, rows):
    """The device's rate of scattered 32-bit atomic adds, measured now: groups
    of 1 and of 2 lanes add into random 64-byte rows of `rows` rows (the kept
    light-row array's size).  requests = adds / lanes per row."""
    out = {}
    for share in (1, 2):
        ms, adds = ctx.probe_atomics(rows, share)
        out[f"lanes_per_row_{share}"] = {"adds_per_s": adds / (ms * 1e-3),
                                         "requests_per_s": adds / share / (ms * 1e-3),
                                         "ms": ms, "adds": adds}
    return out


def atomic_roofline(ctx, rows, workload, kernel, avg_ms):
    """The Delta pass is bound by scattered atomic requests, not bytes: its L2
    atomic requests per launch against the request rate the probe sustains."""
    probe = atomic_ceiling(ctx, rows)
    ceiling = max(v["requests_per_s"] for v in probe.values())
    out = {"bound": "scattered atomic requests", "ceiling_requests_per_s": ceiling, "probe": probe}
    req = measured_atomics(workload, kernel)
    if req is not None:
        achieved = req / (avg_ms * 1e-3)
        out.update({"kernel_requests_per_launch": req, "achieved_requests_per_s": achieved,
                    "frac": achieved / ceiling,
                    "requests_source": "rocprofv3 --pmc TCC_ATOMIC_sum of this kernel on this "
                                       "workload (profiles/traffic.json), not measured in this run"})
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default=os.environ.get("DMOLP_BENCH_CONFIG", "c3"),
                    choices=sorted(CONFIGS) + ["c4", "c5"])
    ap.add_argument("--jobs", type=int, default=96, help="served mode: jobs per arrival rate")
    ap.add_argument("--max-active", type=int, default=96, help="served mode: admitted jobs")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--probe-atomics", action="store_true",
                    help="only measure the scattered-atomic ceiling at the config's vertex count")
    a = ap.parse_args(argv)
    if a.gpus != 1:
        print(json.dumps({"error": "the library runs on one GPU; sharding the path over several "
                                   "GPUs is not built",
                          "n_gpus": a.gpus}))
        return 2
    if a.probe_atomics:
        from dmolp_amd import api
        with api.Context(0) as ctx:
            n = CONFIGS[a.config if a.config in CONFIGS else "c3"]["gen"]
            rows = n[1] if n[0] == "chunglu" else 1 << n[1]
            print(json.dumps({"probe": "scattered atomic adds", "rows": rows,
                              "result": atomic_ceiling(ctx, rows)}))
        return 0
    if a.config in ("c4", "c5"):
        out = run_batched(a) if a.config == "c4" else run_served(a)
        print(json.dumps(out))
        return 0 if out["parity"].get("ok") is not False else 1
    from dmolp_amd import api
    cfg = CONFIGS[a.config]
    k, R, eps = cfg["k"], cfg["rounds"], cfg["eps"]
    ctx = api.Context(0)
    t0 = time.perf_counter()
    g = ctx.generate(spec_for(cfg))
    log(f"[bench] generated {cfg['workload']}: n={g.n} nnz={g.nnz} in "
        f"{time.perf_counter() - t0:.2f} s")
    fx_path = FIXTURES / f"large_{a.config}.json"
    fx = json.loads(fx_path.read_text()) if fx_path.exists() else None
    parity = {"fixture": str(fx_path.relative_to(ROOT)) if fx else None}
    t0 = time.perf_counter()
    p1 = g.phase1(k, THETA, want_labels=True)
    p1_wall = time.perf_counter() - t0
    labels0 = p1.labels
    log(f"[bench] phase 1: {p1_wall:.2f} s wall, device {p1.metrics.device_ms:.1f} ms, "
        f"hub sort {p1.metrics.host_ms:.1f} ms, rounds {p1.metrics.rounds}, "
        f"cut {p1.metrics.initial_edge_cut}")
    if fx:
        w1 = fx["phase1"]
        parity["phase1"] = bool(sha(labels0) == w1["labels_sha256"]
                                and p1.metrics.initial_edge_cut == w1["cut"]
                                and p1.seeds.tolist() == w1["seeds"]
                                and p1.vc.tolist() == w1["vc"] and p1.ec.tolist() == w1["ec"])
    g.checkpoint_labels()

    for _ in range(a.warmup):
        g.checkpoint_labels(restore=True)
        g.phase2(k, R, eps, want_labels=False, want_trace=False)
    dev_ms, lp_edges, walls, res = 0.0, 0, [], None
    for _ in range(a.steps):
        g.checkpoint_labels(restore=True)
        t0 = time.perf_counter()
        res = g.phase2(k, R, eps, want_labels=False, want_trace=False)
        walls.append(time.perf_counter() - t0)
        dev_ms += res.metrics.device_ms
        lp_edges += res.metrics.lp_edges
    value = lp_edges / (dev_ms / 1e3)
    log(f"[bench] phase 2: {a.steps} steps, cut {res.metrics.initial_cut} -> "
        f"{res.metrics.edge_cut}, rounds {res.metrics.rounds}, "
        f"device {dev_ms / a.steps:.2f} ms/step, wall {1e3 * np.mean(walls):.2f} ms/step")

    # One profiled step: HIP events on the library's stream around each
    # kernel group of every round (kernels launched directly, not replayed).
    ctx.set_profiling(True)
    g.checkpoint_labels(restore=True)
    pr = g.phase2(k, R, eps, want_labels=False, want_trace=True)
    ctx.set_profiling(False)
    names = ["lp_light", "lp_heavy", "candidates", "select", "delta_apply", "round", "hist_update",
             "delta_light", "init_light"]
    per_round = {nm: ctx.round_times(i) for i, nm in enumerate(names)}
    groups = {nm: float(v.sum()) for nm, v in per_round.items()}
    tr = pr.trace
    light_n, light_deg = tr["b_light_count"], tr["b_light_deg"]
    heavy_n, heavy_deg = tr["b_count"] - light_n, tr["b_deg"] - light_deg
    log(f"[bench] per round: |B light| {light_n.tolist()} light entries {light_deg.tolist()} "
        f"|B heavy| {heavy_n.tolist()} changed {tr['chg_count'].tolist()} "
        f"changed entries {tr['chg_deg'].tolist()}")
    log("[bench] per round ms: " + json.dumps(
        {nm: [round(float(x), 3) for x in v] for nm, v in per_round.items()}))
    kern, p2eq = kernel_models(g, k, tr, per_round)
    dom = max(kern, key=lambda kk: float(kern[kk][0].sum()))
    ms_r, bytes_r = kern[dom]
    live = bytes_r > 0
    achieved = float(bytes_r[live].sum()) / (float(ms_r[live].sum()) / 1e3) / 1e9
    p2eq_gbs = p2eq / (float(pr.metrics.device_ms) / 1e3) / 1e9
    log(f"[bench] profiled step: groups ms "
        f"{json.dumps({kk: round(v, 3) for kk, v in groups.items()})}; dominant {dom}; "
        f"kernel ms {json.dumps({kk: round(float(v[0].sum()), 3) for kk, v in kern.items()})}")

    # Parity of the timed run's final state: its labels against the
    # oracle's fixture, and its incrementally kept counts against a
    # from-scratch sweep of those labels.
    final = g.get_labels()
    svc, sic, scut = g.stats(k, final)
    parity["counts_vs_sweep"] = bool(scut == res.metrics.edge_cut and
                                     svc.tolist() == pr.vc.tolist() and
                                     sic.tolist() == pr.ic.tolist())
    if fx:
        w2 = fx["phase2"]
        parity["phase2"] = bool(sha(final) == w2["labels_sha256"]
                                and res.metrics.edge_cut == w2["cut"]
                                and res.metrics.rounds == w2["rounds"]
                                and res.metrics.exit_reason == w2["exit_reason"]
                                and res.metrics.lp_edges == w2["lp_edges"]
                                and pr.vc.tolist() == w2["vc"] and pr.ic.tolist() == w2["ic"])
    del final

    cpu = None
    if not a.no_cpu_baseline:
        T, hinfo = cpu_threads()
        off, col = g.download()
        if a.config == "c3":
            # Bounded sample: round 1 alone, from the device's phase-1 labels,
            # which also checks the device's first round against the oracle's.
            cpu, o1 = cpu_baseline_round1(off, col, k, labels0, eps, T)
            g.set_labels(labels0)
            d1 = g.phase2(k, 1, eps, want_labels=True, want_trace=False)
            parity["round1_vs_oracle"] = bool(np.array_equal(d1.labels, o1.labels)
                                              and d1.metrics.edge_cut == o1.m.edge_cut
                                              and d1.vc.tolist() == o1.vc.tolist()
                                              and d1.ic.tolist() == o1.ic.tolist())
        else:
            cpu, o1, o2 = cpu_baseline_full(off, col, k, R, eps, T)
            parity["phase1_vs_oracle"] = bool(np.array_equal(o1.labels, labels0))
            g.set_labels(labels0)
            d2 = g.phase2(k, R, eps, want_labels=True, want_trace=False)
            parity["phase2_vs_oracle"] = bool(np.array_equal(o2.labels, d2.labels)
                                              and o2.m.edge_cut == d2.metrics.edge_cut)
        cpu["host"] = hinfo
        log(f"[bench] cpu baseline {cpu['value']:.3e} LP edges/s on {T} threads "
            f"({cpu['sample']})")
        del off, col
    parity["ok"] = all(v for kk, v in parity.items() if kk != "fixture")
    log(f"[bench] parity {json.dumps(parity)}")

    traffic = measured_traffic(cfg["workload"], dom)
    atomic = atomic_roofline(ctx, g.n, cfg["workload"], dom, float(ms_r[live].mean())) \
        if dom == "k_delta_light" else None
    out = {
        "metric": "LP edges/sec (phase-2 boundary label propagation)",
        "value": value,
        "unit": "LP edges/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * float(np.mean(walls)),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (project generator, seeded)",
        "config": {"workload": cfg["workload"], "n": g.n, "nnz": g.nnz, "k": k, "theta": THETA,
                   "max_rounds": R, "epsilon": eps, "rounds_run": int(res.metrics.rounds),
                   "exit_reason": api.EXIT_NAMES[res.metrics.exit_reason],
                   "initial_cut": int(res.metrics.initial_cut),
                   "final_cut": int(res.metrics.edge_cut),
                   "lp_edges_per_step": int(res.metrics.lp_edges),
                   "phase2_device_ms": dev_ms / a.steps,
                   "phase1_device_ms": p1.metrics.device_ms, "phase1_wall_s": p1_wall,
                   "phase1_rounds": int(p1.metrics.rounds)},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic,
                     "traffic_source": "rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE of this "
                                       "kernel on this workload, copy-calibrated "
                                       "(profiles/traffic.json), not measured in this run"
                                       if traffic is not None else None,
                     "algorithmic_bytes_per_launch": float(bytes_r[live].mean()),
                     "avg_launch_ms": float(ms_r[live].mean()),
                     "launches": int(live.sum()),
                     "atomic": atomic},
        # The whole phase-2 step priced with the reference algorithm's bytes
        # (recounted LP histograms, candidates pushed from B_{t-1}): what a
        # streaming implementation of that algorithm would have to move.
        "phase2_equivalent": {"bytes_per_step": p2eq, "achieved": p2eq_gbs, "unit": "GB/s",
                              "frac": p2eq_gbs / HBM_PEAK_GBS,
                              "device_ms": float(pr.metrics.device_ms),
                              "basis": "bytes the reference algorithm would stream per step "
                                       "(LP histograms recounted from the rows, candidates "
                                       "pushed from the previous boundary, boundary test, "
                                       "selection keys, statistics, initial sweep); not this "
                                       "build's traffic, which the roofline object and the "
                                       "rocprofv3 PMC tables report"},
        "cpu_baseline": cpu,
        "parity": parity,
        "kernel_groups_ms": groups,
    }
    print(json.dumps(out))
    g.free()
    ctx.close()
    return 0 if parity["ok"] else 1


if __name__ == "__main__":
    sys.exit(main())
