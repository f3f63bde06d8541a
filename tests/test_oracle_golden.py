"""The CPU oracle against the golden restart files written by the reference
solver (tests/golden/make_golden.py): bit-for-bit after every fixture's
steps, with the same sweep count per step."""
import numpy as np
import pytest

from oracle.binding import cpu_solver
from pdac_amd import abi
from tests.golden.make_golden import CASES, STALE_TOP_GHOST, load, top_ghost_cells


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_matches_reference(name):
    """Bit for bit, except where the reference reads past the end of ws
    for the ghost row above a free or zero-gradient lid (bdry.f:561): those
    ghost values are left out, and the rest, which that read reaches only
    through the top row's particle momentum flux, agree to 1e-12 relative
    (bit for bit where no particle reaches the lid rows)."""
    case, ref, nits = load(name)
    s = cpu_solver(case, iter_order=abi.ORDER_LEXICOGRAPHIC)
    got_nits = []
    for _ in range(case.nsteps):
        s.step(1)
        got_nits.append(s.nit_last)
    assert got_nits == nits
    out = s.state()
    stale = name in STALE_TOP_GHOST
    ghost = top_ghost_cells(case)
    for k, v in ref.items():
        o = out[k].reshape(v.shape)
        if not stale:
            np.testing.assert_array_equal(o, v, err_msg=k)
            continue
        o, v = o.copy(), v.copy()
        if k == "ws":
            o.reshape(-1, case.ntot)[:, ghost] = 0.0
            v.reshape(-1, case.ntot)[:, ghost] = 0.0
        scale = max(np.max(np.abs(v)), 1e-300)
        assert np.max(np.abs(o - v)) / scale <= 1e-12, (k, np.max(np.abs(o - v)) / scale)


def test_stale_lid_ghost_is_the_only_mismatch_without_particles_at_the_lid():
    """Where no particle reaches the lid rows the out-of-bounds read
    cannot reach a computed cell: every other value is bit for bit."""
    for name in ("mixed_2d_freetop_ord1", "mixed_2d_zerograd"):
        case, ref, _ = load(name)
        s = cpu_solver(case, iter_order=abi.ORDER_LEXICOGRAPHIC)
        s.step(case.nsteps)
        out = s.state()
        ghost = top_ghost_cells(case)
        for k, v in ref.items():
            o = out[k].reshape(v.shape).copy()
            v = v.copy()
            if k == "ws":
                o.reshape(-1, case.ntot)[:, ghost] = 0.0
                v.reshape(-1, case.ntot)[:, ghost] = 0.0
            np.testing.assert_array_equal(o, v, err_msg=f"{name} {k}")


def test_golden_fixture_is_pinned_not_trivial():
    """Each fixture must actually move the state (a dead solver would pass)."""
    for name in CASES:
        case, ref, _ = load(name)
        moved = max(np.max(np.abs(ref["p"] - case.fields["p"])),
                    np.max(np.abs(ref["ug"] - case.fields["ug"])))
        assert moved > 0.0, name


EXPLICIT = sorted(n for n in CASES if n.startswith("explicit_"))


def _explicit_run(name, iter_order=abi.ORDER_LEXICOGRAPHIC, stale_zero=1):
    case, ref, _ = load(name)
    s = cpu_solver(case, iter_order=iter_order)
    s.lib.pdac_cpu_set_explicit_stale(s.ctx, stale_zero)
    s.step(case.nsteps)
    out = s.state()
    return max(np.max(np.abs(out[k].reshape(v.shape) - v)) / max(np.max(np.abs(v)), 1e-300)
               for k, v in ref.items())


def test_explicit_fixtures_pin_the_cell_order():
    """Reading the stage's starting state everywhere (no west/south/bottom
    updates inside a stage) must miss the reference fixtures: they pin the
    serial order of the stage loop."""
    worst = max(_explicit_run(n, iter_order=abi.ORDER_STAGE) for n in EXPLICIT)
    assert worst > 1e-8


def test_explicit_fixtures_pin_the_unset_mass_flux_neighbours():
    """Taking the reflected second neighbours where calc_all_fluxes leaves
    them unset must miss the fixtures with MUSCL: they pin the zero the
    reference build reads there."""
    worst = max(_explicit_run(n, stale_zero=0) for n in EXPLICIT)
    assert worst > 1e-8


VISCOUS = sorted(n for n in CASES if "visc" in n)


@pytest.mark.parametrize("term", ["gas_viscosity", "part_viscosity", "repulsive_model", "iturb"])
def test_viscous_fixtures_pin_each_term(term):
    """With one of the gas stresses and conduction, the particle stresses
    and conduction, the Coulombic repulsive term or the Smagorinsky model
    left out of the oracle's run, the viscous fixtures that have it on must
    be missed (the oracle matches them bit for bit; the repulsive term is the
    smallest, about 5e-6 of the particle velocities over four steps): they
    pin each of the four."""
    worst, used = 0.0, 0
    for name in VISCOUS:
        case, ref, _ = load(name)
        if not getattr(case, term):
            continue
        setattr(case, term, type(getattr(case, term))(0))
        s = cpu_solver(case, iter_order=abi.ORDER_LEXICOGRAPHIC)
        s.step(case.nsteps)
        out = s.state()
        used += 1
        worst = max(worst, max(np.max(np.abs(out[k].reshape(v.shape) - v)) / max(np.max(np.abs(v)), 1e-300)
                               for k, v in ref.items()))
    assert used >= 2
    assert worst > 1e-6, worst
