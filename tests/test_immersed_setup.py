"""The immersed-boundary set-up (pdac_amd.topography, pdac_amd.immersed)
against what the reference's own set-up reported for the same decks: the
forcing points it wrote (fptx/y/z.dat: cell, interpolation type, no-slip
point), and, through the whole-step fixtures, the flags, masks and
interpolation neighbours; plus the oracle's immersed-boundary runtime, each
part of which the fixtures must pin."""
import numpy as np
import pytest

from oracle.binding import cpu_solver
from pdac_amd import abi, constants as C
from pdac_amd.topography import Dem, fortran_nint, interp_2d
from tests.golden.make_golden import load, load_extras

IB_CASES = ["cone_3d_ib", "cone_3d_ib_steep"]


FP_FIELDS = ("i", "j", "k", "int_", "di", "dj", "dk", "nsx", "nsy", "nsz")


def _restated(o, grid, topo2d, ext_cells, axis, nz, cap):
    """The oracle's restatement in place of the library's pdac_ib_forcing."""
    import pdac_amd.immersed as im
    from oracle.ib_forcing import forcing_points
    nx, ny = o.shape
    p = forcing_points(o, grid, topo2d, ext_cells, axis, nx, ny)
    return im.ForcingPoints(*(np.asarray(getattr(p, f)) for f in FP_FIELDS))


@pytest.mark.parametrize("source", ["library", "oracle"])
@pytest.mark.parametrize("name", IB_CASES)
def test_forcing_points_match_the_reference(name, source, monkeypatch):
    """The forcing points of the library (pdac_ib_forcing, host threads) and
    of the oracle's restatement against the reference's own dump."""
    if source == "oracle":
        import pdac_amd.immersed as im
        from oracle.ib_forcing import fill_cells
        monkeypatch.setattr(im, "_forcing_points", _restated)
        monkeypatch.setattr(im, "_fill_cells", fill_cells)
    case, _, _ = load(name)
    ref = load_extras(name)
    ib = case.immersed()
    for a, ax in enumerate("xyz"):
        fp = ib.fp[a]
        got = np.array([fp.i, fp.j, fp.k, fp.int_], dtype=np.int32).T
        np.testing.assert_array_equal(got, ref[f"fpt{ax}_int"], err_msg=ax)
        nsl = np.array([fp.nsx, fp.nsy, fp.nsz]).T
        # printed with F12.3
        assert np.max(np.abs(nsl - ref[f"fpt{ax}_nsl"])) <= 0.0005 + 1e-9, ax
    # the set-up exercises every kind of point and cell
    types = set(np.concatenate([np.array(f.int_) for f in ib.fp]).tolist())
    assert 0 in types and any(t >= 20 for t in types) and any(1 <= t <= 8 for t in types)
    fl = ib.flags
    for f in (C.IMMB_CELL, C.FILLED_CELL_1, C.INLET_CELL):
        assert np.any(fl == f), f


def test_library_forcing_points_equal_the_restatement(monkeypatch):
    """pdac_ib_forcing and pdac_ib_fill_cells against the oracle's
    restatements, bit for bit, on rough cones larger than the fixtures
    (every column split over threads, thousands of points of each type,
    cells closed by fill_cells)."""
    import pdac_amd.immersed as im
    from oracle.ib_forcing import fill_cells
    from pdac_amd.cases import summit_cone_3d
    real, seen = im._forcing_points, []
    real_fill, filled = im._fill_cells, []

    def both_fill(fl, tx, ty, tc, z, zb, nx, ny, nz):
        got = real_fill(fl, tx, ty, tc, z, zb, nx, ny, nz)
        want = fill_cells(fl, tx, ty, tc, z, zb, nx, ny, nz)
        for a, b, what in zip(got, want, ("flags", "bd", "vf")):
            np.testing.assert_array_equal(a, b, err_msg=what)
        filled.append(int(np.count_nonzero(got[0] != fl)))
        return got

    def both(o, grid, topo2d, ext_cells, axis, nz, cap):
        got = real(o, grid, topo2d, ext_cells, axis, nz, cap)
        want = _restated(o, grid, topo2d, ext_cells, axis, nz, cap)
        for f in FP_FIELDS:
            np.testing.assert_array_equal(getattr(got, f), getattr(want, f), err_msg=f"axis {axis} {f}")
        seen.append(np.unique(got.int_))
        return got

    monkeypatch.setattr(im, "_forcing_points", both)
    monkeypatch.setattr(im, "_fill_cells", both_fill)
    for seed, rough in ((7, 12.0), (8, 25.0)):
        summit_cone_3d(nx=64, ny=56, nz=40, d=10.0, height=250.0, sigma=150.0, rough=rough,
                       vent_half=20.0, nsteps=1, seed=seed).immersed()
    assert len(seen) == 6 and len(filled) == 2 and min(filled) > 0
    types = set(np.concatenate(seen).tolist())
    assert 0 in types and 20 in types and any(1 <= t <= 8 for t in types) and any(t > 20 for t in types)


@pytest.mark.parametrize("bit,what", [(1, "boundary forcing"), (2, "tilde forcing"),
                                      (4, "face masks and 1/vf"), (8, "momentum stencil masks")])
def test_fixtures_pin_each_part_of_the_runtime(bit, what):
    """Leaving any part of the immersed-boundary runtime out of the oracle
    must miss the reference fixtures."""
    worst = 0.0
    for name in IB_CASES:
        case, ref, _ = load(name)
        s = cpu_solver(case, iter_order=abi.ORDER_LEXICOGRAPHIC)
        s.lib.pdac_cpu_set_ib_off(s.ctx, bit)
        try:
            s.step(case.nsteps)
        except Exception:
            worst = np.inf
            continue
        out = s.state()
        worst = max(worst, max(np.max(np.abs(out[k].reshape(v.shape) - v)) / max(np.max(np.abs(v)), 1e-300)
                               for k, v in ref.items()))
    assert worst > 1e-6, what


def test_dem_file_round_trip(tmp_path):
    from pdac_amd.topography import cone_dem
    d = cone_dem(100.0, 120.0, cell=10.0, height=50.0, sigma=40.0, rough=2.0, seed=3)
    d.write(str(tmp_path / "topo.dat"))
    back = Dem.read(str(tmp_path / "topo.dat"))
    assert (back.xll, back.yll, back.cell) == (d.xll, d.yll, d.cell)
    np.testing.assert_array_equal(back.z, d.z)
    # written north to south (the reference reads rows j = ny..1)
    first = open(tmp_path / "topo.dat").read().split("\n")[6].split()
    assert float(first[0]) == d.z[-1, 0]


def test_bilinear_interpolation_closed_forms():
    """interp_2d reproduces the nodes, the average midway, and bilinear
    functions exactly (up to rounding)."""
    rng = np.random.default_rng(1)
    x1 = np.arange(-50.0, 51.0, 10.0)
    y1 = np.arange(-30.0, 41.0, 10.0)
    a, b, c, d = rng.uniform(-3, 3, 4)
    f1 = a + b * x1[:, None] + c * y1[None, :] + d * x1[:, None] * y1[None, :]
    x2 = np.array([-45.0, -40.0, 3.3, 27.1])
    y2 = np.array([-25.0, 0.0, 12.5, 33.0])
    f2 = interp_2d(x1, y1, f1, x2, y2)
    want = a + b * x2[:, None] + c * y2[None, :] + d * x2[:, None] * y2[None, :]
    np.testing.assert_allclose(f2, want, rtol=1e-12, atol=1e-10)
    assert interp_2d(x1, y1, f1, x1[1:-1], y1[1:-1]).tolist() == f1[1:-1, 1:-1].tolist()


def test_nint_rounds_halves_away_from_zero():
    assert fortran_nint([0.5, 1.5, 2.5, -0.5, -2.5, 2.4999]).tolist() == [1, 2, 3, -1, -3, 2]


def test_export_layer_equals_the_serial_loop():
    """export_layer is export_topography's serial loop (topo.f:1114-1162)
    vectorised: on random small flag fields, every column order and carry of
    the row counter, it gives what the loop gives cell by cell."""
    from types import SimpleNamespace

    import pdac_amd.constants as C
    from pdac_amd.topography import export_layer

    def serial(case, fl):
        flv = fl.reshape(case.nz, case.ny, case.nx)
        q = case.bl_row0
        for j in range(case.ny):
            for i in range(case.nx):
                col = flv[:, j, i]
                for k in range(case.nz):
                    if col[k] == C.NOSLIP_WALL:
                        q = k + 1
                    elif col[k] == C.FLUID and (k + 1) <= q + 3:
                        col[k] = C.BL_CELL
        return fl

    rng = np.random.default_rng(0)
    for _ in range(300):
        nx, ny, nz = (int(v) for v in rng.integers(1, 7, 3))
        case = SimpleNamespace(nx=nx, ny=ny, nz=nz, bl_row0=int(rng.integers(0, 4)))
        fl = rng.choice([C.FLUID, C.NOSLIP_WALL, C.SLIP_WALL, C.FREE_IO], size=nx * ny * nz,
                        p=[0.7, 0.15, 0.1, 0.05]).astype(np.int32)
        assert np.array_equal(export_layer(case, fl.copy()), serial(case, fl.copy()))
