"""ctypes wrapper of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module,
and only to compare against. It loads oracle/liboracle.so (the C restatement of the reference's
projection path, see eddy_oracle.h) and, when built, oracle/_ref/libfftpack_ref.so (the
reference's own src/fftpack.f + src/comf.f compiled by `make ref`).

All field arrays are numpy float64 arrays in Fortran order with shape (NX, NY, NZ), i.e.
the reference's DP(NX,NY,NZ) layout.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}
_REF = None

_dp = ctypes.POINTER(ctypes.c_double)


def _ptr(a):
    return a.ctypes.data_as(_dp)


def lib(extended=False):
    """liboracle.so (FP64 working precision) or, with extended=True, liboracle_ld.so (the same
    restatement carried in long double)."""
    name = "liboracle_ld.so" if extended else "liboracle.so"
    if name not in _LIBS:
        path = os.path.join(_HERE, name)
        if not os.path.exists(path):
            raise RuntimeError("oracle/%s is not built (run `make -C oracle`)" % name)
        L = ctypes.CDLL(path)
        L.orc_rfftf_many.argtypes = [ctypes.c_int, ctypes.c_long, _dp]
        L.orc_rfftb_many.argtypes = [ctypes.c_int, ctypes.c_long, _dp]
        L.orc_tridiag_eig.argtypes = [ctypes.c_int, _dp, _dp, _dp]
        L.orc_divergence.argtypes = [ctypes.c_int] * 3 + [_dp] * 9
        L.orc_correct.argtypes = [ctypes.c_int] * 3 + [_dp] * 11 + [ctypes.c_double] + [ctypes.c_int] * 3
        L.orc_laplacian.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] + [_dp] * 6 + [ctypes.c_int] * 3 + [_dp, _dp]
        L.orc_direct.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] + [_dp] * 6 + [ctypes.c_int] * 3 + [_dp]
        L.orc_direct_path.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] + [_dp] * 6 + [ctypes.c_int] * 3 + \
            [_dp, ctypes.c_int]
        L.orc_mode_solve.argtypes = [ctypes.c_int] * 4 + [ctypes.c_double] + [_dp] * 7
        L.orc_metrics_from_faces.argtypes = [ctypes.c_int, _dp, _dp, _dp]
        L.orc_philox4x32_10.argtypes = [ctypes.POINTER(ctypes.c_uint)] * 3
        L.orc_normals.argtypes = [ctypes.c_ulong, ctypes.c_uint, ctypes.c_uint, _dp]
        L.orc_perturb.argtypes = [ctypes.c_int] * 3 + [_dp] * 6 + [ctypes.c_double, ctypes.c_uint, ctypes.c_uint]
        L.orc_pressure_update.argtypes = [ctypes.c_int] * 3 + [_dp, _dp] + [ctypes.c_double] * 3 + [_dp] * 6 + \
            [ctypes.c_int] * 7
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_get_threads.restype = ctypes.c_int
        L.orc_set_threads(_THREADS[0])
        _LIBS[name] = L
    return _LIBS[name]


_THREADS = [1]


def ref_lib():
    """The reference's fftpack compiled from its own sources, or None if not built."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libfftpack_ref.so")
        if not os.path.exists(path):
            return None
        R = ctypes.CDLL(path)
        ip = ctypes.POINTER(ctypes.c_int)
        R.rffti_.argtypes = [ip, _dp]
        R.rfftf_.argtypes = [ip, _dp, _dp]
        R.rfftb_.argtypes = [ip, _dp, _dp]
        _REF = R
    return _REF


def set_threads(n):
    _THREADS[0] = int(n)
    for L in _LIBS.values():
        L.orc_set_threads(int(n))


def _f64(a):
    return np.require(a, dtype=np.float64, requirements=["F", "A", "W"])


def rfftf(x):
    """fftpack RFFTF along the last axis (copy)."""
    x = np.array(x, dtype=np.float64, order="C", copy=True)
    L = x.shape[-1]
    lib().orc_rfftf_many(L, x.size // L, _ptr(x))
    return x


def rfftb(x):
    x = np.array(x, dtype=np.float64, order="C", copy=True)
    L = x.shape[-1]
    lib().orc_rfftb_many(L, x.size // L, _ptr(x))
    return x


def ref_rfftf(x):
    """The reference's RFFTF (fftpack.f:2561) on each row, via the _ref build."""
    R = ref_lib()
    if R is None:
        raise RuntimeError("oracle/_ref/libfftpack_ref.so not built (`make -C oracle ref`)")
    x = np.array(x, dtype=np.float64, order="C", copy=True)
    L = x.shape[-1]
    n = ctypes.c_int(L)
    ws = np.zeros(2 * L + 15, dtype=np.float64)
    R.rffti_(ctypes.byref(n), _ptr(ws))
    flat = x.reshape(-1, L)
    for r in range(flat.shape[0]):
        row = np.ascontiguousarray(flat[r])
        R.rfftf_(ctypes.byref(n), _ptr(row), _ptr(ws))
        flat[r] = row
    return x


def ref_rfftb(x):
    R = ref_lib()
    if R is None:
        raise RuntimeError("oracle/_ref/libfftpack_ref.so not built (`make -C oracle ref`)")
    x = np.array(x, dtype=np.float64, order="C", copy=True)
    L = x.shape[-1]
    n = ctypes.c_int(L)
    ws = np.zeros(2 * L + 15, dtype=np.float64)
    R.rffti_(ctypes.byref(n), _ptr(ws))
    flat = x.reshape(-1, L)
    for r in range(flat.shape[0]):
        row = np.ascontiguousarray(flat[r])
        R.rfftb_(ctypes.byref(n), _ptr(row), _ptr(ws))
        flat[r] = row
    return x


def tridiag_eig(d, e):
    d = np.array(d, dtype=np.float64, copy=True)
    e = np.array(e, dtype=np.float64, copy=True)
    n = d.size
    z = np.zeros((n, n), dtype=np.float64, order="F")
    rc = lib().orc_tridiag_eig(n, _ptr(d), _ptr(e) if e.size else _ptr(np.zeros(1)), _ptr(z))
    if rc:
        raise RuntimeError("tridiag_eig failed")
    return d, z


def metrics_from_faces(nx, xu):
    xu = np.ascontiguousarray(xu, dtype=np.float64)
    assert xu.size == nx - 1
    ap = np.zeros(nx)
    au = np.zeros(nx)
    lib().orc_metrics_from_faces(nx, _ptr(xu), _ptr(ap), _ptr(au))
    return ap, au


def divergence(g, us, vs, ws):
    us, vs, ws = _f64(us), _f64(vs), _f64(ws)
    div = np.zeros(us.shape, order="F")
    lib().orc_divergence(g.nx, g.ny, g.nz, _ptr(us), _ptr(vs), _ptr(ws), _ptr(g.ap), _ptr(g.bp),
                         _ptr(g.cp), _ptr(g.rp), _ptr(g.ru), _ptr(div))
    return div


def correct(g, us, vs, ws, dp, alfxdt, uo=None, vo=None, wo=None):
    us, vs, ws, dp = _f64(us), _f64(vs), _f64(ws), _f64(dp)
    uo = np.array(us, order="F", copy=True) if uo is None else _f64(uo)
    vo = np.array(vs, order="F", copy=True) if vo is None else _f64(vo)
    wo = np.array(ws, order="F", copy=True) if wo is None else _f64(wo)
    lib().orc_correct(g.nx, g.ny, g.nz, _ptr(us), _ptr(vs), _ptr(ws), _ptr(uo), _ptr(vo), _ptr(wo),
                      _ptr(dp), _ptr(g.au), _ptr(g.bv), _ptr(g.cw), _ptr(g.rp), float(alfxdt),
                      g.mp, g.lp, g.np)
    return uo, vo, wo


def laplacian(g, phi):
    phi = _f64(phi)
    out = np.zeros(phi.shape, order="F")
    rc = lib().orc_laplacian(g.nx, g.ny, g.nz, g.delysq, _ptr(g.ap), _ptr(g.au), _ptr(g.cp), _ptr(g.cw),
                             _ptr(g.rp), _ptr(g.ru), g.mp, g.lp, g.np, _ptr(phi), _ptr(out))
    if rc:
        raise RuntimeError("orc_laplacian rc=%d" % rc)
    return out


def direct(g, rhs, extended=False, path=0):
    """DIRECT on a copy of `rhs` (interior = RHS); returns DP with ghosts as direct.f fills them.
    extended=True runs the long-double build (answer exact to well below FP64 roundoff).
    path: 0 automatic, 1 x2 eigenbasis + x1 lines, 2 x1 eigenbasis + x2 lines."""
    dp = np.array(rhs, dtype=np.float64, order="F", copy=True)
    rc = lib(extended).orc_direct_path(g.nx, g.ny, g.nz, g.delysq, _ptr(g.ap), _ptr(g.au), _ptr(g.cp),
                                       _ptr(g.cw), _ptr(g.rp), _ptr(g.ru), g.mp, g.lp, g.np, _ptr(dp), int(path))
    if rc:
        raise RuntimeError("orc_direct rc=%d" % rc)
    return dp


def mode_solve(g, jmode, rhs_kfast, extended=False):
    """One per-mode 2-D solve; rhs_kfast has shape (M, N) holding RHS1D(k + N*(i-1))."""
    r = np.array(rhs_kfast, dtype=np.float64, order="C", copy=True)
    M, N = r.shape
    rc = lib(extended).orc_mode_solve(M, N, g.ny - 2, int(jmode), g.delysq, _ptr(g.ap), _ptr(g.au), _ptr(g.cp),
                              _ptr(g.cw), _ptr(g.rp), _ptr(g.ru), _ptr(r))
    if rc:
        raise RuntimeError("orc_mode_solve rc=%d" % rc)
    return r


def philox4x32_10(ctr, key):
    """Philox4x32-10 of a 4-word counter under a 2-word key -> 4 words."""
    U4, U2 = ctypes.c_uint * 4, ctypes.c_uint * 2
    out = U4()
    lib().orc_philox4x32_10(U4(*ctr), U2(*key), out)
    return list(out)


def normals(cell, step, seed):
    """The three N(0,1) draws of `cell` (0-based DP offset) at `step` under `seed`."""
    g = np.zeros(3)
    lib().orc_normals(int(cell), int(step), int(seed), _ptr(g))
    return g


def perturb(g, u, v, w, sigma, seed, step):
    """(us, vs, ws) = (u, v, w) + sigma * N(0,1) on the faces CORRECT updates."""
    u, v, w = _f64(u), _f64(v), _f64(w)
    out = [np.zeros(u.shape, order="F") for _ in range(3)]
    lib().orc_perturb(g.nx, g.ny, g.nz, _ptr(u), _ptr(v), _ptr(w), _ptr(out[0]), _ptr(out[1]), _ptr(out[2]),
                      float(sigma), int(seed), int(step))
    return tuple(out)


def resident_steps(g, u, v, w, nsteps, sigma, seed, dt, extended=False, step0=0):
    """The resident loop on the CPU: per step u* = perturb(u), b = DIV(u*)/dt minus its
    volume-weighted mean, phi = DIRECT(b), u = CORRECT(u*, phi) (+ periodic ghost rows), and
    the (min, max) of DIV(u). Returns (u, v, w, log)."""
    vol = g.cell_volume()
    log = []
    for s in range(nsteps):
        us, vs, ws = perturb(g, u, v, w, sigma, seed, step0 + s)
        b = divergence(g, us, vs, ws)[1:-1, 1:-1, 1:-1] / dt
        b -= (vol * b).sum() / vol.sum()
        rhs = np.zeros(us.shape, order="F")
        rhs[1:-1, 1:-1, 1:-1] = b
        phi = direct(g, rhs, extended=extended)
        u, v, w = correct(g, us, vs, ws, phi, dt, uo=us.copy(order="F"), vo=vs.copy(order="F"),
                          wo=ws.copy(order="F"))
        for a in (u, v, w):
            a[:, 0, :] = a[:, -2, :]
            a[:, -1, :] = a[:, 1, :]
        d = divergence(g, u, v, w)[1:-1, 1:-1, 1:-1]
        log.append((d.min(), d.max()))
    return u, v, w, np.array(log)


def pressure_update(g, p, dp, alfxdt, ru1=0.0, rvimplx=None, rvimply=None, implx=0, imply=0, itype1=0,
                    itype5=1, itype6=1):
    """The pressure update after CORRECT (Eddy6.f:2772-2846) on a copy of p; returns the new P."""
    out = np.array(p, dtype=np.float64, order="F", copy=True)
    dp = _f64(dp)
    rx = np.ascontiguousarray(rvimplx if rvimplx is not None else np.zeros(g.nx), dtype=np.float64)
    ry = np.ascontiguousarray(rvimply if rvimply is not None else np.zeros(g.nx), dtype=np.float64)
    rc = lib().orc_pressure_update(g.nx, g.ny, g.nz, _ptr(out), _ptr(dp), float(alfxdt), float(ru1),
                                   float(g.dely), _ptr(g.ap), _ptr(g.au), _ptr(g.rp), _ptr(g.ru), _ptr(rx),
                                   _ptr(ry), int(implx), int(imply), g.mp, g.lp, int(itype1), int(itype5),
                                   int(itype6))
    if rc:
        raise RuntimeError("orc_pressure_update rc=%d" % rc)
    return out
