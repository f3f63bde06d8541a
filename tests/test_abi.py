"""The C-ABI library: it is built, it loads, it exports every entry point
that include/pdacgpu.h declares, and it rejects bad arguments before it
touches a device.  No kernel runs here; the gpu-marked test creates a
context, which needs a device."""
import ctypes as ct
import os
import re

import numpy as np
import pytest

from pdac_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pdacgpu.h")


def declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(pdac_[a-z_]+)\s*\(", txt)))


@pytest.fixture(scope="module")
def lib():
    from pdac_amd import gpu
    if not os.path.exists(gpu.LIB):
        pytest.fail("libpdacgpu.so is not built (python -m pdac_amd.build)")
    return gpu.lib()


def test_header_declares_the_hot_path():
    names = declared()
    for fn in ("pdac_boundary", "pdac_fieldn", "pdac_tilde", "pdac_iter", "pdac_ygas",
               "pdac_htilde", "pdac_ftem", "pdac_step"):
        assert fn in names


def test_library_exports_every_declared_symbol(lib):
    from pdac_amd import gpu
    names = declared()
    assert sorted(gpu.EXPORTS) == names
    for fn in names:
        assert hasattr(lib, fn), fn


def test_structs_match_header_layout():
    # pdac_dims: 8 int32; pdac_params: 12 doubles + 6*4 doubles + 7 doubles
    # + 7 int32 + 12 int32
    assert abi.PDAC_MAX_SOLID == 4
    assert ct.sizeof(abi.Dims) == 32
    assert ct.sizeof(abi.Params) == 8 * (12 + 6 * 4 + 7) + 4 * (7 + 12 + 3) + 8 * 4 + 8 + 4 + 4


@pytest.mark.parametrize("dims,msg", [
    ((2, 0, 40, 1, 8, 5, 2, 0), "nsolid"),
    ((2, 0, 40, 1, 8, 1, 8, 0), "ngas"),
    ((4, 0, 40, 1, 8, 1, 2, 0), "job_type"),
    ((3, 0, 40, 2, 8, 1, 2, 0), "mesh"),
    ((2, 0, 40, 1, 8, 1, 2, 2), "precision"),
    ((3, 0, 2048, 1024, 1024, 1, 2, 0), "2^31"),
])
def test_create_rejects_bad_dims(lib, dims, msg):
    ctx = ct.c_void_p()
    rc = lib.pdac_create(ct.byref(ctx), ct.byref(abi.Dims(*dims)), 0)
    assert rc == abi.ST_BAD_ARGUMENT
    assert msg in lib.pdac_last_error().decode()
    assert not ctx.value


def test_ensemble_rejects_empty_member_list(lib):
    ens = ct.c_void_p()
    rc = lib.pdac_ens_create(ct.byref(ens), ct.POINTER(ct.c_void_p)(), 0)
    assert rc == abi.ST_BAD_ARGUMENT
    assert "ensemble" in lib.pdac_last_error().decode()
    assert not ens.value


@pytest.mark.gpu
@pytest.mark.parametrize("flag,msg", [(257, "immersed"), (512, "immersed"), (1024, "immersed"),
                                      (3, "unknown"), (0, "unknown")])
def test_set_grid_rejects_unsupported_flags(flag, msg):
    """Cell kinds the kernels do not treat are refused before any kernel
    runs: an immersed-boundary cell without forcing points would otherwise
    be advanced as plain fluid (the reference stops, bdry.f:103-120)."""
    from pdac_amd.cases import mixed_box
    from pdac_amd.gpu import gpu_solver
    from pdac_amd.solver import PdacError
    case = mixed_box(nsteps=1, bottom=3)
    flags = case.flags().copy()
    interior = int(np.flatnonzero(flags == 1)[len(np.flatnonzero(flags == 1)) // 2])
    flags[interior] = flag
    case.flags = lambda: flags
    with pytest.raises(PdacError) as e:
        gpu_solver(case, device=0)
    assert e.value.code == abi.ST_BAD_ARGUMENT
    assert msg in str(e.value)
