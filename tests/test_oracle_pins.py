"""Pins the oracle's exact mode by the reference's own known answers
(src/core/tests.zig, test/d3/*.js, test/fast/error-test.js) and by
identities of the mathematics (Barnes-Hut at theta 0 and infinity, cell
charges, the exact tie, jiggle order, Gauss-Seidel link projections); see
tests/pins.py for what each check asserts and where it comes from."""

import pytest

import pins


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
def test_oracle_pin(oracle_core, check):
    check(oracle_core)


@pytest.mark.parametrize("name", list(pins.error_scenarios()))
def test_oracle_fast_error_within_barnes_hut(oracle_core, name):
    # error-test.js:71-81: fast RMS <= 1e-3 and <= d3's Barnes-Hut RMS.
    err_bh, err_fast = pins.error_gate(oracle_core, name, pins.error_scenarios()[name])
    print("%s: Barnes-Hut %.2e, fast %.2e" % (name, err_bh, err_fast))
    assert err_fast <= 1e-3
    assert err_fast <= err_bh
