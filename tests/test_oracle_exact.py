"""Pins the oracle's exact mode (quadtree, manyBody accumulate + traversal
with jiggle, Gauss-Seidel link, center, x/y/radial, integrate) bit for bit
against d3ref, an independent restatement of d3-force 3.0.0 as published,
over scenarios of the reference's parity corpus (test/parity/scenarios.js),
comparing x, y, vx, vy with Object.is semantics after every tick and the
random stream's position at the end (test/parity/parity-test.js:41-58).
The scenarios carry forceCollide where scenarios.js gives them one.
"""

import math

import pytest

from parity import SCENARIOS, build_d3ref, build_facade, d3ref_state, same
from scenarios import generator


@pytest.mark.parametrize("name", list(SCENARIOS))
def test_exact_tick_matches_d3_restatement(oracle_core, name):
    ticks = 40
    for seed in (1, 2):
        nodes, forces = SCENARIOS[name](generator(seed))
        a = build_d3ref(nodes, forces)
        b, table = build_facade(oracle_core, nodes, forces)
        assert same(d3ref_state(a), table.state()), "after initialize"
        for t in range(ticks):
            a.tick()
            b.tick()
            assert same(d3ref_state(a), table.state()), "%s seed %d tick %d" % (name, seed, t + 1)
            assert a.alpha == b.alpha()
        assert a.random() == b.random_source()(), "random stream position"


def test_alpha_decay_bits():
    # simulation.js:18-21: 1 - 0.001^(1/300) has bits 0x3f974f203e3e5400.
    import struct
    v = 1 - math.pow(0.001, 1 / 300)
    assert struct.pack(">d", v).hex() == "3f974f203e3e5400"
