"""C-ABI checks that need no GPU: the library loads, exports exactly what
include/nemo2d.h declares, the Fortran interfaces bind to declared names,
and calls without a device fail with a status code instead of crashing."""
import ctypes
import os
import re
import subprocess

import pytest

import nemo2d

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "nemo2d.h")
FORTRAN_API = os.path.join(ROOT, "nemolite2d-cuda_amd", "fortran", "nemo2d_api.f90")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nemo2d_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    for n in ("nemo2d_create", "nemo2d_continuity", "nemo2d_momentum_u", "nemo2d_momentum_v",
              "nemo2d_bc_ssh", "nemo2d_bc_solid_u", "nemo2d_bc_solid_v", "nemo2d_bc_flather_u",
              "nemo2d_bc_flather_v", "nemo2d_next", "nemo2d_step", "nemo2d_checksums",
              "nemo2d_retrieve_results", "nemo2d_destroy"):
        assert n in names
    assert sorted(nemo2d.SYMBOLS) == names


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(nemo2d.LIB_PATH)
    for n in declared():
        assert hasattr(L, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", nemo2d.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (nemo2d_\w+)", out))
    assert exported == set(declared())


def test_fortran_interfaces_bind_declared_names():
    text = open(FORTRAN_API).read()
    binds = set(re.findall(r'bind\(C,\s*name="(nemo2d_\w+)"\)', text))
    assert binds and binds <= set(declared())
    # every stage and the fused step are reachable from Fortran
    for n in ("nemo2d_continuity", "nemo2d_momentum_u", "nemo2d_momentum_v", "nemo2d_bc_ssh",
              "nemo2d_bc_solid_u", "nemo2d_bc_solid_v", "nemo2d_bc_flather_u",
              "nemo2d_bc_flather_v", "nemo2d_next", "nemo2d_step"):
        assert n in binds


def test_version_and_error_strings():
    L = nemo2d.lib()
    assert b"gfx950" in L.nemo2d_version()
    assert isinstance(L.nemo2d_last_error(), bytes)


def test_null_arguments_return_status():
    """Argument checks run on the host before any HIP call."""
    L = nemo2d.lib()
    assert L.nemo2d_create(None, 4, 4, 1.0, 1.0, 1.0, 1, 1, 1, 1.0, 0.0, 0.0) == nemo2d.ERR_ARG
    h = ctypes.c_void_p()
    assert L.nemo2d_create(ctypes.byref(h), 0, 4, 1.0, 1.0, 1.0, 1, 1, 1, 1.0, 0.0, 0.0) == nemo2d.ERR_ARG
    assert b"grid size" in L.nemo2d_last_error()
    assert L.nemo2d_create(ctypes.byref(h), 4, 4, 1.0, 1.0, 1.0, 1, 1, 1, 0.0, 0.0, 0.0) == nemo2d.ERR_ARG
    for n in ("nemo2d_continuity", "nemo2d_momentum_u", "nemo2d_momentum_v", "nemo2d_bc_solid_u",
              "nemo2d_bc_solid_v", "nemo2d_bc_flather_u", "nemo2d_bc_flather_v", "nemo2d_next"):
        assert getattr(L, n)(None) == nemo2d.ERR_ARG
    assert L.nemo2d_bc_ssh(None, 0.0) == nemo2d.ERR_ARG
    assert L.nemo2d_step(None, 1) == nemo2d.ERR_ARG
    assert L.nemo2d_set_stage_fusion(None, 1) == nemo2d.ERR_ARG
    assert L.nemo2d_set_variant(None, 0) == nemo2d.ERR_ARG
    assert L.nemo2d_destroy(None) == nemo2d.OK


def test_convolutions_reject_images_over_2_24_pixels():
    """The convolution kernels address an image with 32-bit offsets and
    range-checked buffer loads: an image of more than 2^24 pixels is refused
    on the host, before any HIP call (the pointers are never used)."""
    from nemo2d.closure import _setup_nn
    L = nemo2d.lib()
    _setup_nn(L)
    p = ctypes.c_void_p(256)
    f = ctypes.cast(p, ctypes.POINTER(ctypes.c_float))
    big = (1, 4097, 4096)  # 16,781,312 pixels
    assert L.nemo2d_nn_conv_fwd(p, *big, 32, p, f, 32, None, p, p, 1, None) == nemo2d.ERR_ARG
    assert b"too many pixels" in L.nemo2d_last_error()
    assert L.nemo2d_nn_conv_bwd_data(p, *big, 32, p, 32, f, None) == nemo2d.ERR_ARG
    assert L.nemo2d_nn_conv_bwd_data_gelu(p, *big, 32, p, 32, p, 1, p, None) == nemo2d.ERR_ARG
    assert L.nemo2d_nn_conv_bwd_weight(p, p, *big, 32, 32, f, f, f, None) == nemo2d.ERR_ARG
    assert b"too many pixels" in L.nemo2d_last_error()


def test_driver_binary_built():
    exe = os.path.join(ROOT, "nemolite2d-cuda_amd", "bin", "nemolite2d_hip")
    assert os.access(exe, os.X_OK)
