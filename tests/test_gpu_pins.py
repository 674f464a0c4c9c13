"""The known answers and identities of tests/pins.py on the MI355X library
(the same checks that pin the oracle), and error-test.js:71-81's gate on
the GPU's FP32 fast mode: RMS error against the FP64 direct sum at most 1e-3
and at most the exact mode's Barnes-Hut error."""

import pytest

import pins

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("check", [
    pins.check_node_sized_arrays_follow_node_count,
    pins.check_links_not_node_sized_and_bad_index,
    pins.check_integrate_fixed_rule,
    pins.check_fast_link_equals_exact_on_matching,
    pins.check_fast_collide_matches_exact_on_isolated_pairs,
    pins.check_center_known_answers,
    pins.check_axis_known_answers,
    pins.check_phyllotaxis_known_answer,
    pins.check_theta_zero_is_direct_sum,
    pins.check_theta_infinite_sees_only_root,
    pins.check_cell_charges,
    pins.check_exact_tie,
    pins.check_coincident_jiggle_order,
    pins.check_link_projection_identities,
    pins.check_center_identity,
], ids=lambda f: f.__name__[len("check_"):])
def test_gpu_pin(gpu_core, check):
    print("gpu pin", check.__name__, flush=True)
    check(gpu_core)


@pytest.mark.parametrize("name", list(pins.error_scenarios()))
def test_gpu_fast_error_within_barnes_hut(gpu_core, name):
    print("gpu error gate", name, flush=True)
    err_bh, err_fast = pins.error_gate(gpu_core, name, pins.error_scenarios()[name])
    print("%s: Barnes-Hut %.2e, fast %.2e" % (name, err_bh, err_fast))
    assert err_fast <= 1e-3
    assert err_fast <= err_bh
