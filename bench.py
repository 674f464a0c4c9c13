"""Benchmark: d3-force ticks/s at 1,000,000 nodes in exact mode on one MI355X.

Workload (the exact 1M-node configuration): 1,000,000 nodes from d3's
phyllotaxis start, 1% of them pinned with fx, fy uniform in [-500, 500];
2,000,000 links over distinct node pairs drawn with a seeded d3 LCG;
forceManyBody (strength -30, theta 0.9, distanceMin 1, distanceMax inf),
forceLink (distance 30, strength 1/min(degree), 1 iteration), forceCenter(0, 0);
FP64, bit-identical to the reference core. A step is one tick. The timed
region runs K ticks on arrays already resident in HBM (the facade's copy-in
happens before it, the copy-out after it).

Output: one JSON line with the metric, the roofline of the dominant kernel
(measured live with HIP events on the library's stream) and a CPU baseline:
the oracle (a C++ restatement of the reference core) timed on this host's
cores over a bounded sample of the same workload.
"""

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "d3-force-wasm_amd"), os.path.join(ROOT, "tests")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "ticks/s at 1M nodes: exact BH (bit-exact) vs fast FMM (RMS err ≤1e-6)"
FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 (vector and matrix)
HBM_PEAK_GBS = 8000.0     # MI355X HBM3E


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_workload(core, n, m, pin_fraction, seed=1):
    import d3force_amd as d3
    from scenarios import pinned, random_graph
    t0 = time.time()
    src, tgt = random_graph(n, m, seed=seed)
    nodes = d3.Nodes(n)
    if pin_fraction:
        idx, fx, fy = pinned(n, pin_fraction)
        nodes.fix(fx=fx, fy=fy, where=idx)
    sim = (d3.forceSimulation(nodes, core=core)
           .force("charge", d3.forceManyBody())
           .force("link", d3.forceLink((src, tgt)))
           .force("center", d3.forceCenter()))
    log("workload %d nodes, %d links ready in %.1fs" % (n, m, time.time() - t0))
    return sim, nodes


def make_fast_workload(core, n, seed=1, order=8, opening=1.5):
    """The fast-mode workload: phyllotaxis start, Barabasi-Albert links (m0 = 3,
    3 per new node), forceManyBody (FMM of the given order and opening; 8 and
    1.5 by default), forceLink (adjacency sweep), forceCollide with radius
    uniform in [2, 6]."""
    import numpy as np
    import d3force_amd as d3
    from scenarios import ba_graph
    t0 = time.time()
    src, tgt = ba_graph(n, seed=seed)
    radii = 2 + 4 * np.random.default_rng(seed).random(n)
    nodes = d3.Nodes(n)
    sim = (d3.forceSimulation(nodes, core=core)
           .force("charge", d3.forceManyBody())
           .force("link", d3.forceLink((src, tgt)))
           .force("collide", d3.forceCollide(radii)))
    sim.precision("fast", {"order": order, "opening": opening})
    log("fast workload %d nodes, %d links ready in %.1fs" % (n, len(src), time.time() - t0))
    return sim, nodes, len(src)


def fast_rms(core, nodes, sample_size=10000, seed=3, order=8, opening=1.5):
    """Relative RMS error of the fast many-body force on the current layout
    against the library's FP64 direct sum, over a fixed sample of nodes
    (d3force_amd.accuracy; the form of the reference's
    test/fast/error-test.js)."""
    from d3force_amd.accuracy import fast_force_error
    return fast_force_error(nodes.x, nodes.y, sample_size=sample_size, seed=seed, order=order, opening=opening,
                            core=core)


def run_fast(core, n, steps, warmup, order=8, opening=1.5):
    sim, nodes, m = make_fast_workload(core, n, order=order, opening=opening)
    ctx = sim.begin()
    log("fast: warmup %d ticks" % warmup)
    sim.iterate(warmup, ctx)
    device_sync(core)
    core.stats(ctx.h)
    core.core_set_timing(1)
    log("fast: timing %d ticks" % steps)
    device_sync(core)
    t0 = time.perf_counter()
    sim.iterate(steps, ctx)
    device_sync(core)
    dt = time.perf_counter() - t0
    core.core_set_timing(0)
    st = core.stats(ctx.h)
    sim.end(ctx)
    log("fast: %.3f ms/tick, phases %s" % (1e3 * dt / steps, st["ms"]))
    err, k = fast_rms(core, nodes, order=order, opening=opening)
    log("fast: rms vs FP64 direct sum %.3e on %d nodes" % (err, k))
    p = order
    launches = max(st["launches"]["fmm"], 1)
    t_fmm = st["ms"]["fmm"] / 1e3 / launches
    flops = (8 * (p + 1) ** 2 * st["m2l"] + 11 * st["p2p_pairs"]) / launches
    return {
        "value": steps / dt, "unit": "ticks/s", "ms_per_step": 1e3 * dt / steps, "steps": steps, "warmup": warmup,
        "dtype": "f64 positions and centres, f32 expansions / M2L / P2P",
        "config": {"workload": "fast mode, %d nodes, Barabasi-Albert graph (m0 = 3, 3 links per new node, %d links), "
                               "phyllotaxis start, manyBody FMM order %d opening %g + link sweep + grid collide "
                               "radius U[2, 6]" % (n, m, order, opening), "nodes": n, "links": m, "mode": "fast",
                   "order": order, "opening": opening,
                   "ticks": "%d-%d" % (warmup + 1, warmup + steps)},
        "rms_force_error_vs_fp64_direct_sum": err, "rms_sample_nodes": k,
        "phases_ms_per_tick": {ph: v / steps for ph, v in st["ms"].items() if v > 0},
        "fmm_work_per_tick": {"m2l_pairs": st["m2l"] / launches, "p2p_point_pairs": st["p2p_pairs"] / launches,
                              "cells": st["fmm_cells"]},
        "fmm_roofline": {"kernel": "fast many-body (all FMM kernels)", "bound": "fp32",
                         "pipe": "FP32: M2L on the matrix cores, P2P and expansions on the vector ALUs "
                                 "(MI355X FP32 peak 157.3 TFLOP/s, vector and matrix)",
                         "achieved": flops / t_fmm / 1e12, "peak": 157.3, "unit": "TFLOP/s",
                         "frac": flops / t_fmm / 1e12 / 157.3, "launch_ms": t_fmm * 1e3,
                         "work_per_launch": {"flops": flops, "m2l_flops_per_pair": 8 * (p + 1) ** 2,
                                             "p2p_flops_per_pair": 11}},
    }


def run_fast_same(core, n, steps, warmup, order=10, opening=1.5):
    """Fast precision on the exact leg's workload (the same graph, pins and
    forces: manyBody + link + center), at the lowest FMM order whose RMS force
    error against the FP64 direct sum is within 1e-6 (the metric's fast form).
    The sums of forceCenter run on the second stream as in exact mode."""
    m = 2 * n
    sim, nodes = make_workload(core, n, m, 0.01)
    sim.precision("fast", {"order": order, "opening": opening})
    ctx = sim.begin()
    log("fast (exact workload): warmup %d ticks" % warmup)
    sim.iterate(warmup, ctx)
    device_sync(core)
    core.stats(ctx.h)
    core.core_set_timing(1)
    device_sync(core)
    t0 = time.perf_counter()
    sim.iterate(steps, ctx)
    device_sync(core)
    dt = time.perf_counter() - t0
    core.core_set_timing(0)
    st = core.stats(ctx.h)
    sim.end(ctx)
    log("fast (exact workload): %.3f ms/tick, phases %s" % (1e3 * dt / steps, st["ms"]))
    err, k = fast_rms(core, nodes, order=order, opening=opening)
    log("fast (exact workload): rms vs FP64 direct sum %.3e on %d nodes" % (err, k))
    return {
        "value": steps / dt, "unit": "ticks/s", "ms_per_step": 1e3 * dt / steps, "steps": steps, "warmup": warmup,
        "dtype": "f64 positions and centres, f32 expansions / M2L / P2P",
        "config": {"workload": "fast mode on the exact leg's workload: %d nodes, %d links (distinct pairs), 1%% "
                               "pinned, phyllotaxis start, manyBody FMM order %d opening %g + link sweep + center"
                               % (n, m, order, opening), "nodes": n, "links": m, "mode": "fast", "order": order,
                   "opening": opening, "ticks": "%d-%d" % (warmup + 1, warmup + steps)},
        "rms_force_error_vs_fp64_direct_sum": err, "rms_sample_nodes": k,
        "phases_ms_per_tick": {ph: v / steps for ph, v in st["ms"].items() if v > 0},
    }


def run_batch(core, graphs, steps, warmup, seed=1):
    """The batched workload: `graphs` stochastic-block-model graphs (n
    log-uniform in [50, 20000], 2-16 blocks, p_in 0.05 / p_out 0.002 scaled to
    mean degree 6), each from d3's phyllotaxis start, ticked together with
    manyBody (FMM order 8, opening 1.5) + link + center. A step is one tick of
    every graph."""
    import math
    import numpy as np
    from d3force_amd.batch import (Batch, BS_ALPHA, BS_CENTER_STRENGTH, BS_LINK_ITERATIONS, BS_VELOCITY_DECAY)
    from scenarios import sbm_batch
    t0 = time.time()
    gs = sbm_batch(graphs, seed=seed)
    b = Batch(gs, core=core)
    n, m = b.n, b.m
    log("batch: %d graphs, %d nodes, %d links ready in %.1fs" % (graphs, n, m, time.time() - t0))
    tuned = b.autotune(2)  # 29 ticks, untimed: launch shapes only, results do not depend on them
    log("batch: autotuned %s" % tuned)
    b.tick(warmup)
    device_sync(core)
    b.stats()
    core.batch_set_timing(1)
    t0 = time.perf_counter()
    b.tick(steps)
    device_sync(core)
    dt = time.perf_counter() - t0
    core.batch_set_timing(0)
    st = b.stats()
    log("batch: %.3f ms/tick, phases %s" % (1e3 * dt / steps, st["ms"]))
    # The same ticks as one captured HIP graph replayed per tick (an eager
    # tick and the capture first, then one replay untimed).
    from d3force_amd.batch import BS_GRAPH
    b.scalar(BS_GRAPH, 1)
    b.tick(2)
    device_sync(core)
    t0 = time.perf_counter()
    b.tick(steps)
    device_sync(core)
    dt_graph = time.perf_counter() - t0
    log("batch: graph replay %.3f ms/tick" % (1e3 * dt_graph / steps))
    b.close()
    # Per-graph many-body error on the start layout of a sample of graphs: one
    # manyBody-only tick from v = 0, against an FP64 direct sum (numpy) on up
    # to 300 nodes of each sampled graph.
    rng = np.random.default_rng(7)
    pick = sorted(rng.choice(graphs, min(graphs, 16), replace=False).tolist())
    probe = Batch([gs[g] for g in pick], core=core)
    probe.scalar(BS_LINK_ITERATIONS, 0).scalar(BS_CENTER_STRENGTH, 0).scalar(BS_VELOCITY_DECAY, 1)
    s0 = probe.state()
    probe.tick(1)
    alpha = probe.scalar(BS_ALPHA)
    s1 = probe.state()
    errs = []
    for k in range(len(pick)):
        a, c = probe.graph_state(k, s0), probe.graph_state(k, s1)
        sample = rng.choice(a.shape[1], min(a.shape[1], 300), replace=False)
        e2 = r2 = 0.0
        for i in sample:
            dx, dy = a[0] - a[0][i], a[1] - a[1][i]
            l = dx * dx + dy * dy
            keep = l > 0
            dx, dy, l = dx[keep], dy[keep], l[keep]
            l = np.where(l < 1, np.sqrt(l), l)
            w = -30.0 * alpha / l
            fx, fy = np.sum(dx * w), np.sum(dy * w)
            e2 += (c[2][i] - fx) ** 2 + (c[3][i] - fy) ** 2
            r2 += fx * fx + fy * fy
        errs.append(math.sqrt(e2 / r2))
    probe.close()
    sizes = [gs[g][0] for g in pick]
    return {
        "value": steps / dt, "unit": "ticks/s (every graph one tick)", "ms_per_step": 1e3 * dt / steps,
        "graph_ticks_per_s": graphs * steps / dt, "node_ticks_per_s": n * steps / dt, "steps": steps,
        "warmup": warmup, "dtype": "f64 positions and centres, f32 expansions / M2L / P2P",
        "config": {"workload": "batched fast mode, %d stochastic-block-model graphs (n log-uniform in [50, 20000], "
                               "2-16 blocks, p_in 0.05 / p_out 0.002 scaled to mean degree 6), %d nodes, %d links, "
                               "phyllotaxis start per graph, manyBody FMM order 8 opening 1.5 + link + center"
                               % (graphs, n, m), "graphs": graphs, "nodes": n, "links": m,
                   "graphs_below_2000_nodes": int(sum(1 for g in gs if g[0] < 2000))},
        "phases_ms_per_tick": {ph: v / steps for ph, v in st["ms"].items() if v > 0},
        "autotuned_launch_shapes": tuned,
        "graph_replay": {"ms_per_step": 1e3 * dt_graph / steps, "ticks_per_s": steps / dt_graph,
                         "note": "the same tick captured once as a HIP graph and replayed per tick (bit-identical "
                                 "to the eager ticks); value above is the eager form"},
        "fmm_work_per_tick": {"m2l_pairs": st["m2l"] / steps, "p2p_point_pairs": st["p2p_pairs"] / steps,
                              "cells": st["fmm_cells"]},
        "rms_force_error_vs_fp64_direct_sum": {"sampled_graphs": len(pick), "nodes": sizes,
                                               "per_graph": errs, "max": max(errs)},
    }


def make_collide_workload(core, n, seed=5):
    """Exact forceCollide alone: n uniform points at 0.0556 nodes per unit
    area (mean nearest-neighbour distance 2.1), radii U[0.5, 1.5], strength
    0.7, one iteration per tick."""
    import numpy as np
    import d3force_amd as d3
    rng = np.random.default_rng(seed)
    half = 0.5 * math.sqrt(n / 0.0556)
    xy = rng.uniform(-half, half, size=(n, 2))
    radii = rng.uniform(0.5, 1.5, size=n)
    nodes = d3.Nodes(n, x=xy[:, 0].copy(), y=xy[:, 1].copy())
    sim = d3.forceSimulation(nodes, core=core).force("collide", d3.forceCollide(radii).strength(0.7)).stop()
    return sim


def run_collide(core, n, steps, warmup, cpu_ticks=1):
    sim = make_collide_workload(core, n)
    ctx = sim.begin()
    log("collide: warmup %d ticks" % warmup)
    sim.iterate(warmup, ctx)
    device_sync(core)
    core.stats(ctx.h)
    log("collide: timing %d ticks" % steps)
    t0 = time.perf_counter()
    sim.iterate(steps, ctx)
    device_sync(core)
    dt = time.perf_counter() - t0
    st = core.stats(ctx.h)
    sim.end(ctx)
    out = {"value": steps / dt, "unit": "ticks/s", "ms_per_step": 1e3 * dt / steps, "steps": steps,
           "warmup": warmup, "dtype": "f64",
           "config": {"workload": "exact forceCollide alone, %d nodes uniform at 0.0556 per unit area, radius "
                                  "U[0.5, 1.5], strength 0.7, 1 iteration" % n, "nodes": n, "mode": "exact"},
           "serial_iterations": st["collide_serial"], "stop_reasons": st["collide_status"]}
    if cpu_ticks:
        from oracle import oracle
        ocore = oracle.load()
        osim = make_collide_workload(ocore, n)
        octx = osim.begin()
        osim.iterate(warmup, octx)
        log("collide: oracle %d ticks" % cpu_ticks)
        t0 = time.perf_counter()
        osim.iterate(cpu_ticks, octx)
        cdt = time.perf_counter() - t0
        osim.end(octx)
        out["cpu_baseline"] = {"value": cpu_ticks / cdt, "unit": "ticks/s", "cores": 1, "kind": "port",
                               "sample": "%d ticks after the same %d warm-up ticks" % (cpu_ticks, warmup)}
    return out


def run_warm(core, nodes=100000):
    """The learned warm start with the committed network on held-out graphs of
    `nodes` nodes (two stochastic block models, one Barabasi-Albert graph):
    wall-clock time for the warm start (rows, features, network, scaling and
    jitter on the GPU from the links, then the refinement from alpha 0.5,
    layout back on the host) to come within 1% of the energy of 300 plain
    fast-mode ticks, against those 300 ticks on the same clock, and the
    energy gap of both at equal tick counts."""
    from d3force_amd import warmstart as ws
    from d3force_amd.graphs import ba_graph, sbm_batch
    net = ws.Network(core=core)
    net.load(ws.WEIGHTS)
    graphs = [("sbm-a", sbm_batch(1, seed=778, n_min=nodes, n_max=nodes)[0]),
              ("sbm-b", sbm_batch(1, seed=779, n_min=nodes, n_max=nodes)[0]),
              ("ba", (nodes,) + ba_graph(nodes, seed=777))]
    out = []
    for name, (n, src, tgt) in graphs:
        r = ws.compare(net, n, src, tgt, core=core, seed=99, short=(25, 50, 75, 100))
        hit = next((run for run in r["runs"] if abs(run["warm"]["relative_to_plain"]) <= 0.01), None)
        row = {"graph": name, "nodes": n, "links": len(src), "plain_ticks": 300,
               "plain_seconds": r["plain_seconds"], "predict_seconds": r["predict_seconds"],
               "predicted_layout_energy_gap": r["predicted_energy_gap"],
               "within_1pct": None if hit is None else {
                   "ticks": hit["ticks"], "seconds": hit["warm"]["seconds"],
                   "speedup": r["plain_seconds"] / hit["warm"]["seconds"]},
               "equal_ticks": [{"ticks": run["ticks"], "warm_gap": run["warm"]["gap"],
                                "plain_gap": run["plain_same_ticks"]["gap"],
                                "warm_stress": run["warm"]["stress"], "plain_stress": run["plain_same_ticks"]["stress"]}
                               for run in r["runs"]],
               "plain_300_stress": r["plain_stress"]}
        log("warm start %s: %s" % (name, json.dumps(row["within_1pct"])))
        out.append(row)
    net.close()
    return {"graphs": out, "network": "4 message-passing layers, width 64, BF16 on matrix cores",
            "clock": "wall-clock seconds on both sides, from layout (plain) or links (warm) on the host to the "
                     "refined layout back on the host; features and prediction on the GPU",
            "energy": "d3 potential: links 1/2 k (r - 30)^2 + 30 sum over node pairs log(1 / max(r, 1)), "
                      "pair sum from a fixed 1M-pair sample; gap = (E - E_plain300) / (E_start - E_plain300)"}


def device_sync(core):
    """Waits for the device (the library's own hipDeviceSynchronize: the
    library links ROCm's HIP runtime directly, so torch is not loaded)."""
    from d3force_amd.core import check
    check(core.core_synchronize())


def read_traffic(kernel):
    """HBM bytes per tick of `kernel` (all its dispatches in one tick) from the
    committed rocprofv3 PMC summary; the traversal is one dispatch per tick."""
    path = os.path.join(ROOT, "profiles", "exact_1m_tick_hbm_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = [p.strip() for p in line.strip().strip("|").split("|")]
        if len(parts) >= 2 and parts[0].split("<")[0] == kernel:
            try:
                return float(parts[-1])
            except ValueError:
                return None
    return None


def host_cpu():
    """The host's CPU as the box reports it: nproc, the CPUs this process may
    run on, and the model name."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = None
    return {"nproc": os.cpu_count(), "affinity": affinity, "cpu_model": model}


def cpu_baseline(n, m, pin_fraction, warmup, ticks, rounds, threads, single_ticks):
    """The oracle (a C++ restatement of the reference core, the reference's
    pool chunking for tree build, accumulate, traversal and integrate; link
    and center serial as in the reference) on `threads` host threads, by the
    reference's method (bench/threads.js:24-35): `warmup` ticks, then the
    fastest of `rounds` rounds of `ticks` ticks. Every round restarts from the
    state after the warm-up (positions, velocities, alpha, random source), so
    each times ticks warmup+1 .. warmup+ticks, the range the GPU times. Then
    a bounded single-thread sample of `single_ticks` ticks from that state."""
    from oracle import oracle
    core = oracle.load()
    core.set_threads(threads)
    sim, nodes = make_workload(core, n, m, pin_fraction)
    log("cpu baseline: %d warm-up ticks on %d threads" % (warmup, threads))
    sim.tick(warmup)
    state, alpha, rnd = nodes.state(), sim.alpha(), list(sim.random_source().state)

    def restore():
        nodes.x[:], nodes.y[:], nodes.vx[:], nodes.vy[:] = state
        sim.alpha(alpha)
        sim.random_source().state[:] = rnd

    def timed(k):
        restore()
        ctx = sim.begin()
        t0 = time.perf_counter()
        sim.iterate(k, ctx)
        dt = time.perf_counter() - t0
        sim.end(ctx)
        return dt

    best = min(timed(ticks) for _ in range(rounds))
    log("cpu baseline: pool %.3f ticks/s" % (ticks / best))
    core.set_threads(1)
    one = timed(single_ticks)
    log("cpu baseline: single thread %.3f ticks/s" % (single_ticks / one))
    span = "ticks %d-%d" % (warmup + 1, warmup + ticks)
    return {"value": ticks / best, "unit": "ticks/s", "cores": threads, "kind": "port",
            "sample": "%s of the same %d-node workload (the GPU's timed ticks): %d warm-up ticks, then the fastest "
                      "of %d rounds of %d ticks, each round restarted from the state after the warm-up "
                      "(bench/threads.js:24-35's method); C++ oracle (restatement of the reference core) with the "
                      "reference's pool chunking for tree build, accumulate, traversal and integrate, link and "
                      "center serial as in the reference, on %d threads (the box's CPU share for one GPU)"
                      % (span, n, warmup, rounds, ticks, threads),
            "host": host_cpu(),
            "single_thread": {"value": single_ticks / one, "unit": "ticks/s", "cores": 1,
                              "sample": "ticks %d-%d from the same post-warm-up state, one round"
                                        % (warmup + 1, warmup + single_ticks)}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--nodes", type=int, default=1000000)
    ap.add_argument("--cpu-rounds", type=int, default=3, help="rounds of the CPU pool baseline (fastest kept)")
    ap.add_argument("--cpu-single-ticks", type=int, default=3, help="ticks of the single-thread CPU sample")
    ap.add_argument("--cpu-threads", type=int, default=16, help="CPU pool threads (the box's share for one GPU)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--fast-nodes", type=int, default=1000000, help="nodes of the fast-mode leg (0: skip it)")
    ap.add_argument("--fast-only", action="store_true", help="run only the fast-mode leg (prints its object)")
    ap.add_argument("--fast-order", type=int, default=8, help="FMM order of the fast-only leg")
    ap.add_argument("--fast-opening", type=float, default=1.5, help="FMM opening of the fast-only leg")
    ap.add_argument("--fast-same-only", action="store_true",
                    help="run only fast precision on the exact leg's workload (prints its object)")
    ap.add_argument("--batch-graphs", type=int, default=4096, help="graphs of the batched leg (0: skip it)")
    ap.add_argument("--batch-only", action="store_true", help="run only the batched leg (prints its object)")
    ap.add_argument("--warm-nodes", type=int, default=100000, help="nodes of the warm-start graphs (0: skip)")
    ap.add_argument("--warm-only", action="store_true", help="run only the warm-start leg (prints its object)")
    ap.add_argument("--collide-nodes", type=int, default=1000000, help="nodes of the exact collide leg (0: skip)")
    ap.add_argument("--collide-only", action="store_true", help="run only the exact collide leg (prints its object)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs the tick on a single MI355X "
              "(sharding the tick over several GPUs is not implemented)", file=sys.stderr)
        sys.exit(2)

    import d3force_amd as d3
    if args.fast_only:
        print(json.dumps(run_fast(d3.gpu_core(), args.fast_nodes, args.steps, args.warmup, order=args.fast_order,
                                  opening=args.fast_opening)))
        return
    if args.fast_same_only:
        print(json.dumps(run_fast_same(d3.gpu_core(), args.nodes, args.steps, args.warmup)))
        return
    if args.batch_only:
        print(json.dumps(run_batch(d3.gpu_core(), args.batch_graphs, args.steps, args.warmup)))
        return
    if args.warm_only:
        print(json.dumps(run_warm(d3.gpu_core(), args.warm_nodes)))
        return
    if args.collide_only:
        print(json.dumps(run_collide(d3.gpu_core(), args.collide_nodes, args.steps, args.warmup,
                                     0 if args.no_cpu else 1)))
        return
    n = args.nodes
    m = 2 * n
    pin = 0.01
    core = d3.gpu_core()
    sim, nodes = make_workload(core, n, m, pin)
    ctx = sim.begin()
    log("warmup %d ticks" % args.warmup)
    sim.iterate(args.warmup, ctx)
    device_sync(core)
    # The headline: the facade's tick loop as it runs by default, each tick a
    # replay of the captured tick graph (sim_ticks).
    captured0 = ctx.tick_counters()
    log("timing %d ticks" % args.steps)
    t0 = time.perf_counter()
    sim.iterate(args.steps, ctx)
    device_sync(core)
    dt = time.perf_counter() - t0
    captured1 = ctx.tick_counters()
    ticks_per_s = args.steps / dt
    log("%.3f ms/tick (captured ticks %s -> %s)" % (1e3 * dt / args.steps, captured0, captured1))
    # Per-kernel times: the next `steps` ticks force by force, with HIP events
    # around each phase's launches on the stream they run on.
    core.stats(ctx.h)  # reset counters and timers
    core.core_set_timing(1)
    device_sync(core)
    t1 = time.perf_counter()
    sim.iterate(args.steps, ctx)
    device_sync(core)
    dt_timed = time.perf_counter() - t1
    core.core_set_timing(0)
    st = core.stats(ctx.h)
    sim.end(ctx)
    log("%.3f ms/tick force by force with phase events, phases %s" % (1e3 * dt_timed / args.steps, st["ms"]))
    # The facade's tick(k) with the boundary copies (copy-in of 7 arrays and
    # copy-out of 4 per call, over PCIe): k = 1 and k = steps, after the
    # timed region.
    with_copies = {}
    for k, calls in ((1, 10), (args.steps, 1)):
        t0 = time.perf_counter()
        for _ in range(calls):
            sim.tick(k)
        with_copies["tick(%d)" % k] = {"ms_per_tick": 1e3 * (time.perf_counter() - t0) / (k * calls),
                                       "calls": calls}
    log("with copies: %s" % with_copies)
    cells = int(core.tree_info(ctx.h, 0))  # cells of the last tree (host copy, after the timing)

    # Per-kernel rooflines from the live event timing; the dominant kernel is
    # the one with the most device time per tick. Algorithmic work per tick
    # (C = cells of the tree, read from the library after the timed region):
    #   traversal  8 V + 8 A + 6 P FP64 flops (V visits, A accepted cells,
    #              P leaf pairs, counted by the kernel itself)
    #   center_sum 16 B/node read; 2n dependent FP64 adds in d3's order
    #   tree       28 B/node + 88 B/cell;  accumulate 152 B/cell (88-byte cells read,
    #              64-byte walk records written) + 24 B/node
    #   link       128 B/link;  integrate 81 B/node
    phase_ms = st["ms"]
    C = cells
    hbm_bytes = {"tree": 28 * n + 88 * C, "accumulate": 152 * C + 24 * n, "link": 128 * m,
                 "integrate": 81 * n, "center_sum": 16 * n}
    names = {"traversal": "k_traverse<RECORD> (Barnes-Hut traversal)", "center_sum": "k_cs_chain (center sums, side stream)",
             "tree": "quadtree build (13 kernels)", "accumulate": "k_acc_subtrees+k_acc_top (two launches)",
             "link": "k_link_level+k_link_run (all levels)", "integrate": "k_integrate"}
    kernels = {}
    floors = {}
    for ph, ms in phase_ms.items():
        if ph == "center" or ms <= 0:
            continue
        launches = max(st["launches"][ph], 1)
        t = ms / 1e3 / launches
        if ph == "traversal":
            flops = (8 * st["visits"] + 8 * st["accepts"] + 6 * st["pairs"]) / launches
            a = flops / t / 1e12
            k = {"kernel": names[ph], "bound": "valu",
                 "pipe": "FP64 vector ALU (no matrix cores; MI355X FP64 peak 78.6 TFLOP/s counts an FMA as 2, "
                         "so with contraction off the ceiling is 39.3 T add/mul per s)",
                 "achieved": a, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": a / FP64_PEAK_TFLOPS,
                 "work_per_launch": {"flops": flops, "visits": st["visits"] / launches,
                                     "accepts": st["accepts"] / launches, "pairs": st["pairs"] / launches}}
            floors[ph] = flops / (FP64_PEAK_TFLOPS * 1e12) * launches
        else:
            a = hbm_bytes[ph] / t / 1e9
            k = {"kernel": names[ph], "bound": "hbm", "achieved": a, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                 "frac": a / HBM_PEAK_GBS, "bytes_per_launch": hbm_bytes[ph]}
            if ph == "center_sum":
                k["bound"] = "latency"
                k["note"] = ("second stream, overlapped with the tick: d3's serial sums by exact binade segments "
                             "(chunk transforms composed in parallel, one serial pass over the segment records); "
                             "%.2f ns per node from the side stream's start to its end" % (t * 1e9 / n))
            else:
                floors[ph] = hbm_bytes[ph] / (HBM_PEAK_GBS * 1e9) * launches
        k["launch_ms"] = t * 1e3
        parts = [read_traffic(kn.split("<")[0]) for kn in names[ph].split(" ")[0].split("+")]
        k["traffic"] = None if None in parts else sum(parts)
        kernels[ph] = k
    # The dominant kernel is taken on the tick's critical path: the center sums
    # run on a second stream, overlapped with the tree build and traversal.
    dominant = max((ph for ph in kernels if ph != "center_sum"),
                   key=lambda ph: kernels[ph]["launch_ms"] * max(st["launches"][ph], 1))
    roof = dict(kernels[dominant])
    floor_s = sum(floors.values()) / args.steps if floors else 0.0
    tick_roofline = {"floors_ms_per_tick": {ph: 1e3 * v / args.steps for ph, v in floors.items()},
                     "sum_of_floors_ms": 1e3 * floor_s, "tick_ms": 1e3 * dt / args.steps,
                     "frac": floor_s / (dt / args.steps),
                     "note": "per phase floor = algorithmic bytes / 8 TB/s (HBM phases) or flops / 78.6 TFLOP/s "
                             "(traversal); the center sums overlap on a second stream and are left out"}

    fast = run_fast(core, args.fast_nodes, args.steps, args.warmup) if args.fast_nodes else None
    # Order 8 measures ~1.7e-6 against the direct sum on this workload; order
    # 10 is the lowest order within the metric's 1e-6.
    fast_1e6 = None
    if args.fast_nodes:
        fast_1e6 = run_fast(core, args.fast_nodes, args.steps, args.warmup, order=10)
        fast_1e6["note"] = ("the same workload at FMM order 10, the lowest order whose RMS force error against "
                            "the FP64 direct sum is within 1e-6")
    fast_same = run_fast_same(core, n, args.steps, args.warmup) if args.fast_nodes else None
    batch = run_batch(core, args.batch_graphs, args.steps, args.warmup) if args.batch_graphs else None
    warm = run_warm(core, args.warm_nodes) if args.warm_nodes else None
    collide = (run_collide(core, args.collide_nodes, min(args.steps, 10), args.warmup, 0 if args.no_cpu else 1)
               if args.collide_nodes else None)

    cpu = None
    if not args.no_cpu:
        threads = max(1, min(args.cpu_threads, os.cpu_count() or 1))
        cpu = cpu_baseline(n, m, pin, args.warmup, args.steps, args.cpu_rounds, threads, args.cpu_single_ticks)

    out = {
        "metric": METRIC,
        "value": ticks_per_s,
        "unit": "ticks/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * dt / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "exact mode, %d nodes, %d links (distinct pairs), 1%% pinned, phyllotaxis start, "
                               "manyBody + link + center with d3 defaults" % (n, m),
                   "nodes": n, "links": m, "pinned": int(round(n * pin)), "mode": "exact",
                   "ticks": "%d-%d" % (args.warmup + 1, args.warmup + args.steps)},
        "captured_ticks": {"replayed_in_timed_region": captured1[1] - captured0[1],
                           "rolled_back": captured1[2] - captured0[2],
                           "note": "the timed ticks are replays of one captured HIP graph per tick (sim_ticks); "
                                   "the per-kernel figures below come from the next %d ticks run force by force "
                                   "with HIP events around each phase (%.3f ms per tick)"
                                   % (args.steps, 1e3 * dt_timed / args.steps)},
        "phases_ms_per_tick": {k: v / args.steps for k, v in phase_ms.items()},
        "roofline": roof,
        "tick_roofline": tick_roofline,
        "with_copies": with_copies,
        "cells": cells,
        "rooflines": kernels,
        "cpu_baseline": cpu,
        "fast_exact_workload": fast_same,
        "fast": fast,
        "fast_rms_within_1e-6": fast_1e6,
        "batch": batch,
        "warm_start": warm,
        "exact_collide": collide,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
