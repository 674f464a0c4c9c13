"""Plants plausible mistakes in copies of the oracle (on the CPU) and checks
that the known answers and identities of tests/pins.py reject each one: a
pin that a dropped term, a wrong sign, a wrong tie rule or a swapped draw
could pass would pin nothing. Every mutant is one textual change of
oracle/oracle.cpp, compiled on its own and run in a child process (a mutant
may divide by zero or loop), which reports the checks that failed.
"""

import concurrent.futures
import json
import os
import subprocess
import sys
import tempfile

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SOURCE = os.path.join(ROOT, "oracle", "oracle.cpp")

# (name, original text, planted text, checks expected to catch it)
MUTANTS = [
    ("tie accepted (<= for <)",
     "if (w * w / p.theta2 < l) {", "if (w * w / p.theta2 <= l) {", ["exact_tie"]),
    ("far field without alpha",
     "vx += x * value * p.alpha / l;", "vx += x * value / l;", ["theta_infinite_sees_only_root"]),
    ("delta sign flipped",
     "double x = quad.x - nx, y = quad.y - ny;", "double x = nx - quad.x, y = quad.y - ny;",
     ["theta_infinite_sees_only_root", "theta_zero_is_direct_sum"]),
    ("far field ignores distanceMax",
     "      if (l < p.dmax2) {\n        if (record && (x == 0 || y == 0)) return false;",
     "      if (true) {\n        if (record && (x == 0 || y == 0)) return false;",
     ["theta_infinite_sees_only_root"]),
    ("leaf ignores distanceMax",
     "    if (l >= p.dmax2) continue;\n    uint32_t head", "    uint32_t head", ["theta_zero_is_direct_sum"]),
    ("leaf pair without alpha",
     "double wj = p.strengths[j] * p.alpha / l;", "double wj = p.strengths[j] / l;", ["theta_zero_is_direct_sum"]),
    ("distanceMin rule dropped on leaves",
     "      if (y == 0) { y = jiggle(js_random()); l += y * y; }\n      if (l < p.dmin2) l = std::sqrt(p.dmin2 * l);\n    }",
     "      if (y == 0) { y = jiggle(js_random()); l += y * y; }\n    }",
     ["theta_zero_is_direct_sum", "coincident_jiggle_order"]),
    ("charge weight without abs",
     "double w = std::fabs(q.value);", "double w = q.value;", ["cell_charges", "theta_infinite_sees_only_root"]),
    ("leaf charge of the head only",
     "strength += strengths[t.perm[k]];", "strength = strengths[t.perm[k]];", ["cell_charges"]),
    ("leaf head alone skips its jiggle",
     "if (head != i || quad.count > 1) {", "if (head != i) {", ["coincident_jiggle_order"]),
    ("jiggle draws y before x",
     "        if (x == 0) { x = jiggle(js_random()); l += x * x; }\n        if (y == 0) { y = jiggle(js_random()); l += y * y; }\n        if (l < p.dmin2) l = std::sqrt(p.dmin2 * l);\n        vx +=",
     "        if (y == 0) { y = jiggle(js_random()); l += y * y; }\n        if (x == 0) { x = jiggle(js_random()); l += x * x; }\n        if (l < p.dmin2) l = std::sqrt(p.dmin2 * l);\n        vx +=",
     []),
    ("leaf jiggle draws y before x",
     "      if (x == 0) { x = jiggle(js_random()); l += x * x; }\n      if (y == 0) { y = jiggle(js_random()); l += y * y; }\n      if (l < p.dmin2) l = std::sqrt(p.dmin2 * l);\n    }",
     "      if (y == 0) { y = jiggle(js_random()); l += y * y; }\n      if (x == 0) { x = jiggle(js_random()); l += x * x; }\n      if (l < p.dmin2) l = std::sqrt(p.dmin2 * l);\n    }",
     ["coincident_jiggle_order"]),
    ("link bias on both ends",
     "      b = 1 - b;\n      vxs[src]", "      vxs[src]", ["link_projection_identities"]),
    ("link without alpha",
     "l = (l - f.f1[i]) / l * alpha * f.f0[i];", "l = (l - f.f1[i]) / l * f.f0[i];", ["link_projection_identities"]),
    ("link source pushed the wrong way",
     "      vxs[src] += x * b;", "      vxs[src] -= x * b;",
     ["link_projection_identities", "fast_link_equals_exact_on_matching"]),
    ("center without strength",
     "sx = (sx / nf - f.scalar(S_X)) * strength;", "sx = (sx / nf - f.scalar(S_X));", ["center_identity"]),
    ("integrate moves before decaying",
     "if ((fixed & 1) == 0) { s.vx[i] *= decay; s.x[i] += s.vx[i]; }",
     "if ((fixed & 1) == 0) { s.x[i] += s.vx[i]; s.vx[i] *= decay; }", ["integrate_fixed_rule"]),
]
# "jiggle draws y before x" on accepted cells is an equivalent mutant: a cell
# is accepted only when w^2 / theta^2 < l, so l > 0 and at most one of the
# two deltas is zero; the order of the two draws cannot show. It stays in the
# list (expecting nothing) to record that.


def _build(mutant, out_dir):
    name, old, new, _ = mutant
    with open(SOURCE) as f:
        text = f.read()
    assert text.count(old) == 1, "mutation site of %r not unique in oracle.cpp" % name
    path = os.path.join(out_dir, "m%02d.cpp" % MUTANTS.index(mutant))
    with open(path, "w") as f:
        f.write(text.replace(old, new))
    lib = path[:-4] + ".so"
    subprocess.check_call(["g++", "-O1", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-fvisibility=hidden",
                           "-fPIC", "-shared", "-pthread", "-w", "-Wl,-Bsymbolic", "-o", lib, path])
    return lib


def _checks():
    import pins
    return {f[len("check_"):]: getattr(pins, f) for f in dir(pins) if f.startswith("check_")}


def _run_child(lib):
    """Child process: runs every check on the library, prints the failures."""
    sys.path[:0] = [HERE, ROOT, os.path.join(ROOT, "d3-force-wasm_amd")]
    from d3force_amd.core import Core
    core = Core(lib, prefix="oracle_")
    failed = []
    for name, check in sorted(_checks().items()):
        try:
            check(core)
        except Exception:  # a failed assertion, or a mutant's arithmetic fault
            failed.append(name)
    print(json.dumps(failed))


@pytest.fixture(scope="module")
def mutant_libs():
    with tempfile.TemporaryDirectory(prefix="oracle_mutants_") as d:
        with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as pool:
            libs = list(pool.map(lambda m: _build(m, d), MUTANTS))
        yield libs


def _failures(lib):
    r = subprocess.run([sys.executable, os.path.abspath(__file__), lib], capture_output=True, text=True, timeout=300)
    if r.returncode != 0:
        return None  # the mutant crashed its process: caught
    return set(json.loads(r.stdout.strip().splitlines()[-1]))


@pytest.mark.parametrize("k", range(len(MUTANTS)), ids=[m[0] for m in MUTANTS])
def test_pins_reject_planted_mistake(mutant_libs, k):
    name, _, _, expect = MUTANTS[k]
    failed = _failures(mutant_libs[k])
    if not expect:
        return
    assert failed is None or failed & set(expect), "%s passed every pin (failed: %s)" % (name, failed)


def test_unmutated_oracle_passes_every_pin(oracle_core):
    for name, check in _checks().items():
        check(oracle_core)


if __name__ == "__main__":
    _run_child(sys.argv[1])
