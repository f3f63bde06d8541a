"""The C-ABI library: it is built, it loads, it exports every entry point
that include/pdacgpu.h declares, and it rejects bad arguments before it
touches a device.  No kernel runs here."""
import ctypes as ct
import os
import re

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
    # pdac_dims: 8 int32; pdac_params: 12 doubles + 6*8 doubles + 7 doubles
    # + 7 int32 + 12 int32
    assert ct.sizeof(abi.Dims) == 32
    assert ct.sizeof(abi.Params) == 8 * (12 + 6 * 8 + 7) + 4 * (7 + 12) + 4


@pytest.mark.parametrize("dims,msg", [
    ((2, 0, 40, 1, 8, 5, 2, 0), "nsolid"),
    ((2, 0, 40, 1, 8, 1, 8, 0), "ngas"),
    ((4, 0, 40, 1, 8, 1, 2, 0), "job_type"),
    ((3, 0, 40, 2, 8, 1, 2, 0), "mesh"),
    ((2, 0, 40, 1, 8, 1, 2, 2), "precision"),
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
