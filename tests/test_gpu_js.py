"""The Node N-API path: d3-force-wasm_amd/js/core.js over d3fgpu.node drives
the GPU core as the reference's JS facade would, and its ticks are
bit-identical to the oracle's."""

import json
import os
import shutil
import subprocess
import struct

import numpy as np
import pytest

import d3force_amd as d3
from scenarios import random_graph

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ADDON = os.path.join(ROOT, "d3-force-wasm_amd", "js", "d3fgpu.node")
needs_node = pytest.mark.skipif(shutil.which("node") is None or not os.path.exists(ADDON),
                                reason="node or the addon is missing")


@needs_node
def test_addon_loads_and_core_js_exports_the_reference_interface():
    code = ("import('%s').then(m => console.log(JSON.stringify(Object.keys(m).sort())))"
            % os.path.join(ROOT, "d3-force-wasm_amd", "js", "core.js"))
    out = subprocess.run(["node", "-e", code], capture_output=True, text=True, timeout=60)
    keys = json.loads(out.stdout.strip().splitlines()[-1])
    for k in ["CENTER", "LINK", "MANY_BODY", "Slot", "check", "context", "slot", "configure", "testing",
              "THETA2", "SEED", "F0", "SOURCE", "TARGET", "FLOATS", "SCALARS"]:
        assert k in keys, k


@needs_node
@pytest.mark.gpu
@pytest.mark.parametrize("radius", [None, 12.0])
def test_js_facade_ticks_bit_identical_to_oracle(oracle_core, tmp_path, radius):
    # With a radius, exact forceCollide runs after the link force.
    n, ticks = 5000, 12
    src, tgt = random_graph(n, 2 * n, seed=9)
    x, y = d3.phyllotaxis(n)
    path = tmp_path / "in.json"
    spec = {"x": x.tolist(), "y": y.tolist(), "source": src.tolist(), "target": tgt.tolist(), "ticks": ticks}
    if radius is not None:
        spec["radius"] = radius
    path.write_text(json.dumps(spec))
    out = subprocess.run(["node", os.path.join(ROOT, "tests", "js", "tick.mjs"), str(path)],
                         capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr[-2000:]
    got = np.array([[struct.unpack(">d", bytes.fromhex(b))[0] for b in row] for row in json.loads(out.stdout)])
    nodes = d3.Nodes(n, x=x, y=y)
    sim = d3.forceSimulation(nodes, core=oracle_core).force("charge", d3.forceManyBody()).force(
        "link", d3.forceLink((src, tgt)))
    if radius is not None:
        sim.force("collide", d3.forceCollide(radius))
    sim.force("center", d3.forceCenter())
    for _ in range(ticks):
        sim.tick()
    ref = nodes.state()
    assert np.array_equal(ref.view(np.uint64), got.view(np.uint64))


@needs_node
@pytest.mark.gpu
def test_views_kept_across_a_resize_are_detached():
    # A typed array over the library's pinned memory, kept across a call that
    # may free or move that memory (a count change, a node resize, a destroy),
    # reads as empty afterwards instead of reaching freed memory; views made
    # after the call work.
    core_js = os.path.join(ROOT, "d3-force-wasm_amd", "js", "core.js")
    code = """
import('%s').then(({context, Slot, LINK, F0}) => {
  const nodes = Array.from({length: 100}, (_, i) => ({index: i, x: i, y: 0, vx: 0, vy: 0}));
  const ctx = context(nodes);
  ctx.resize(100), ctx.views();
  const x = ctx.x;
  const link = new Slot(LINK);
  link.bind(nodes), link.count(10);
  const f = link.f64(F0, 10);
  f.fill(3);
  const before = [x.length, f.length, f[9]];
  link.count(20);
  const afterCount = [f.length, x.length];
  ctx.resize(5000);
  const afterResize = x.length;
  ctx.views();
  const g = link.f64(F0, 20);
  console.log(JSON.stringify({before, afterCount, afterResize, fresh: [ctx.x.length, g.length]}));
});
""" % core_js
    out = subprocess.run(["node", "-e", code], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr[-2000:]
    r = json.loads(out.stdout.strip().splitlines()[-1])
    assert r["before"] == [100, 10, 3]
    assert r["afterCount"] == [0, 100]  # the force's view detached, the nodes' kept
    assert r["afterResize"] == 0
    assert r["fresh"] == [5000, 20]
