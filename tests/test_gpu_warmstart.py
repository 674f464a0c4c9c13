"""The learned warm start (d3force_amd.warmstart) with the committed network:
on held-out graphs, the predicted layout refined for 50 ticks from alpha =
0.5 ends closer to the energy of 300 plain fast-mode ticks than plain fast
mode squeezed into the same 50 ticks (the method without the network), and
reaches within 1% of that energy in a third of the wall-clock time of the 300
plain ticks, everything from the links to the refined layout counted (rows,
features and prediction run on the GPU)."""

import pytest

from d3force_amd import warmstart as ws
from d3force_amd.graphs import ba_graph, sbm_batch

pytestmark = pytest.mark.gpu


def test_warm_start_beats_plain_fast_mode_at_equal_ticks(gpu_core):
    net = ws.Network(core=gpu_core)
    net.load(ws.WEIGHTS)
    n, src, tgt = sbm_batch(1, seed=4242, n_min=20000, n_max=20000)[0]
    r = ws.compare(net, n, src, tgt, core=gpu_core, seed=3, short=(50,))
    run = r["runs"][0]
    print("  energy gap: warm %.3f, plain %.3f (predicted layout %.3f)"
          % (run["warm"]["gap"], run["plain_same_ticks"]["gap"], r["predicted_energy_gap"]), flush=True)
    assert run["warm"]["gap"] < run["plain_same_ticks"]["gap"]
    assert r["predicted_energy_gap"] < 1.0
    net.close()


@pytest.mark.parametrize("family", ["sbm", "ba"])
def test_warm_start_reaches_plain_energy_in_a_third_of_the_time(gpu_core, family):
    # The criterion of the learned warm start on a held-out 100k-node graph:
    # within 1% of the energy of 300 plain fast-mode ticks, in at most a
    # third of their wall-clock time (rows, features and the network's time
    # included). Stress is reported beside it.
    print("warm start criterion, %s" % family, flush=True)
    net = ws.Network(core=gpu_core)
    net.load(ws.WEIGHTS)
    if family == "sbm":
        n, src, tgt = sbm_batch(1, seed=5151, n_min=100000, n_max=100000)[0]
    else:
        n = 100000
        src, tgt = ba_graph(n, seed=6161)
    r = ws.compare(net, n, src, tgt, core=gpu_core, seed=7, short=(25, 50, 75))
    hit = next((run for run in r["runs"] if abs(run["warm"]["relative_to_plain"]) <= 0.01), None)
    for run in r["runs"]:
        print("  %d ticks: warm energy %+.4f of plain, %.3f s (stress %.3f); plain 300 ticks %.3f s (stress %.3f)"
              % (run["ticks"], run["warm"]["relative_to_plain"], run["warm"]["seconds"], run["warm"]["stress"],
                 r["plain_seconds"], r["plain_stress"]), flush=True)
    assert hit is not None
    assert r["plain_seconds"] >= 3 * hit["warm"]["seconds"]
    net.close()
