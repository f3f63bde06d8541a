"""The CPU oracle against the golden restart files written by the reference
solver (tests/golden/make_golden.py): bit-for-bit after every fixture's
steps, with the same sweep count per step."""
import numpy as np
import pytest

from oracle.binding import cpu_solver
from pdac_amd import abi
from tests.golden.make_golden import CASES, load


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_matches_reference(name):
    case, ref, nits = load(name)
    s = cpu_solver(case, iter_order=abi.ORDER_LEXICOGRAPHIC)
    got_nits = []
    for _ in range(case.nsteps):
        s.step(1)
        got_nits.append(s.nit_last)
    assert got_nits == nits
    out = s.state()
    for k, v in ref.items():
        np.testing.assert_array_equal(out[k].reshape(v.shape), v, err_msg=k)


def test_golden_fixture_is_pinned_not_trivial():
    """Each fixture must actually move the state (a dead solver would pass)."""
    for name in CASES:
        case, ref, _ = load(name)
        moved = max(np.max(np.abs(ref["p"] - case.fields["p"])),
                    np.max(np.abs(ref["ug"] - case.fields["ug"])))
        assert moved > 0.0, name
