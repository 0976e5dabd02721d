"""The C-ABI library loads and exports every symbol include/eddyproj.h declares; host-side
argument validation (no GPU compute)."""
import ctypes
import os
import re

import numpy as np
import pytest

import eddyproj

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "eddyproj.h")).read()
    decl = re.compile(r"^(?:const\s+)?[a-z_]+\s*\**\s+\**(eddy_[a-z0-9_]+)\s*\(", re.M)
    return sorted(set(decl.findall(text)))


def test_library_exports_header_symbols():
    L = eddyproj.lib()
    names = declared_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n


def test_fortran_shim_exports_direct():
    path = os.path.join(ROOT, "eddy-hybrid_amd", "lib", "libdirect_hip.so")
    assert os.path.exists(path)
    f = ctypes.CDLL(path)
    assert hasattr(f, "direct_")


def _call_direct(nx=10, ny=6, nz=8, mysize=1, mp=1, lp=0, np_=1, ipres=1, nzg=None):
    L = eddyproj.lib()
    dp = np.zeros(min(nx * ny * nz, 1 << 16))  # every case here fails validation before DP is read
    m = np.ones(max(nx, nz))
    one = np.array([0.1])
    p = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))  # noqa: E731
    iv = [ctypes.c_int(v) for v in (nx, ny, nz, nz if nzg is None else nzg, mysize, 0, 0, 0, ipres, mp, lp, np_)]
    return L.eddy_direct(p(dp), p(dp), p(dp), p(one), p(one), p(m), p(m), p(m), p(m), p(m), p(m),
                         *[ctypes.byref(v) for v in iv])


@pytest.mark.parametrize("kw,status", [
    (dict(mysize=2), -5),             # EDDY_ERR_MPI
    (dict(mp=2), -1),                 # flags are 0 or 1
    (dict(ipres=2), -2),              # GENBUN branch: not wired up in direct.f
    (dict(ipres=4, mp=0), -2),        # pdc2dn branches drop the periodic couplings
    (dict(ipres=6, np_=0), -2),
    (dict(ipres=9), -2),
    (dict(nx=3), -1),                 # too small
    (dict(nzg=9), -1),                # NZ != NZG on one rank
    (dict(nx=2048, ny=1026, nz=258), -2),  # FP64 field of 4 GiB or more (32-bit offsets)
])
def test_direct_argument_validation(kw, status):
    assert _call_direct(**kw) == status
    assert eddyproj.lib().eddy_last_error()


@pytest.mark.parametrize("chunk,status", [(33, -1), (64, -1), (2, -1)])
def test_plan_rejects_xsolve_chunk_outside_kernel_templates(chunk, status):
    """The x1 solve's largest kernel template holds 32 rows per thread: a larger chunk would
    leave rows unsolved, so the plan refuses it before touching the device."""
    L = eddyproj.lib()
    nx, ny, nz = 66, 34, 18
    m = np.ones(max(nx, nz))
    p = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))  # noqa: E731
    t = eddyproj.Tuning(xsolve_chunk=chunk)
    h = ctypes.c_void_p()
    rc = L.eddy_plan_create(ctypes.byref(h), nx, ny, nz, 0.01, p(m), p(m), p(m), p(m), p(m), p(m), p(m), p(m),
                            1, 0, 1, ctypes.byref(t), None)
    assert rc == status
    assert b"xsolve_chunk" in L.eddy_last_error()
    assert not h.value


def test_version_and_error_strings():
    assert b"gfx950" in eddyproj.lib().eddy_version()
