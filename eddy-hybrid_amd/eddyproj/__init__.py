"""Host-side binding of libeddyproj.so (the MI355X pressure projection).

This mirrors the reference's projection interface: `direct()` has DIRECT's argument meaning
(src/direct.f:2-3: DP in = RHS on the interior, DP out = pressure correction with x/y ghosts
filled and z ghosts zeroed), `Plan.divergence()` / `Plan.correct()` follow DIVERGENCE and
CORRECT (src/timeindeg.f:908, 981) and `Plan.project()` is the call sequence of
src/Eddy6.f:2712-2738 (DIVERGENCE, DP/ALFXDT, DIRECT, CORRECT).

Device fields are torch tensors of dtype float64 and shape (NZ, NY, NX) (C order), which is
exactly the reference's Fortran DP(NX, NY, NZ) layout in memory. Host fields are numpy arrays
of shape (NX, NY, NZ) in Fortran order.

There is no fallback: if libeddyproj.so is not built, or the GPU call fails, an error is
raised.
"""
import ctypes
import os

import numpy as np

from . import grid as grid  # noqa: F401
from . import configs as configs  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libeddyproj.so")

EDDY_FP64 = 0
EDDY_FP32_REFINE = 1
EDDY_ERR_NUMERIC = -4
EDDY_ERR_INCOMPATIBLE = -6

KERNELS = ("yfft_fwd", "ztrans_fwd", "xsolve", "ztrans_inv", "yfft_inv", "divergence", "correct", "other")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p
_LIB = None


class EddyError(RuntimeError):
    pass


class Tuning(ctypes.Structure):
    _fields_ = [("yfft_pencils", ctypes.c_int), ("ztrans_pencils", ctypes.c_int),
                ("xsolve_chunk", ctypes.c_int), ("use_dense_z", ctypes.c_int), ("fft_radix", ctypes.c_int),
                ("yfft_gemm", ctypes.c_int), ("slab_chunks", ctypes.c_int), ("zgemm", ctypes.c_int),
                ("yfft_iblock", ctypes.c_int), ("zgemm_lib", ctypes.c_int)]


def lib():
    """Load libeddyproj.so (raises if it is not built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise EddyError("libeddyproj.so is not built at %s (run __graft_entry__.build())" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    L.eddy_last_error.restype = ctypes.c_char_p
    L.eddy_version.restype = ctypes.c_char_p
    L.eddy_direct.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp] + [_ip] * 12
    L.eddy_plan_create.argtypes = [ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                   _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.POINTER(Tuning), _vp]
    L.eddy_plan_destroy.argtypes = [_vp]
    L.eddy_plan_set_stream.argtypes = [_vp, _vp]
    L.eddy_plan_get_tuning.argtypes = [_vp, ctypes.POINTER(Tuning)]
    L.eddy_plan_device_bytes.argtypes = [_vp]
    L.eddy_plan_device_bytes.restype = ctypes.c_size_t
    L.eddy_direct_dev.argtypes = [_vp, _vp]
    L.eddy_direct_refined_dev.argtypes = [_vp, _vp, ctypes.c_int, _dp]
    L.eddy_direct_refined_adaptive_dev.argtypes = [_vp, _vp, ctypes.c_double, ctypes.c_int, _ip, _dp]
    L.eddy_divergence_dev.argtypes = [_vp, _vp, _vp, _vp, _vp, ctypes.c_double]
    L.eddy_correct_dev.argtypes = [_vp] + [_vp] * 7 + [ctypes.c_double]
    L.eddy_project_dev.argtypes = [_vp] + [_vp] * 7 + [ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.eddy_divergence_minmax_dev.argtypes = [_vp, _vp, _vp, _vp, _dp, _dp]
    L.eddy_laplacian_dev.argtypes = [_vp, _vp, _vp]
    L.eddy_rhs_check_dev.argtypes = [_vp, _vp, _dp, ctypes.POINTER(ctypes.c_long)]
    L.eddy_graph_create.argtypes = [ctypes.POINTER(_vp), _vp] + [_vp] * 7 + [ctypes.c_double, ctypes.c_int,
                                                                         ctypes.c_int, ctypes.c_int]
    L.eddy_graph_launch.argtypes = [_vp]
    L.eddy_graph_destroy.argtypes = [_vp]
    L.eddy_plan_enable_timing.argtypes = [_vp, ctypes.c_int]
    L.eddy_plan_kernel_time.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_long), _dp]
    L.eddy_plan_reset_timing.argtypes = [_vp]
    _pp = ctypes.POINTER(_dp)
    _ppv = ctypes.POINTER(_vp)
    L.eddy_batch_create.argtypes = [ctypes.POINTER(_vp), ctypes.c_int, _ip, _ip, _ip, _dp] + [_pp] * 8 + \
        [ctypes.POINTER(Tuning), _vp]
    L.eddy_batch_bind.argtypes = [_vp] + [_ppv] * 7
    L.eddy_batch_direct_dev.argtypes = [_vp, ctypes.c_int, ctypes.c_int]
    L.eddy_batch_project_dev.argtypes = [_vp, ctypes.c_double, ctypes.c_int]
    L.eddy_batch_info.argtypes = [_vp, ctypes.POINTER(ctypes.c_long), ctypes.POINTER(ctypes.c_long)]
    L.eddy_batch_destroy.argtypes = [_vp]
    L.eddy_batch_set_mode.argtypes = [_vp, ctypes.c_int]
    L.eddy_plan_set_pressure_bc.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp, _dp,
                                            ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.eddy_pressure_update_dev.argtypes = [_vp, _vp, _vp, ctypes.c_double]
    L.eddy_perturb_dev.argtypes = [_vp] + [_vp] * 6 + [ctypes.c_double, ctypes.c_uint, _vp]
    L.eddy_resident_create.argtypes = [ctypes.POINTER(_vp), _vp] + [_vp] * 4 + \
        [ctypes.c_double, ctypes.c_double, ctypes.c_uint, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.eddy_resident_run.argtypes = [_vp, ctypes.c_int]
    L.eddy_resident_log.argtypes = [_vp, _dp, ctypes.c_int, _ip]
    L.eddy_resident_destroy.argtypes = [_vp]
    _LIB = L
    return L


def _check(rc, what):
    if rc != 0:
        raise EddyError("%s failed (status %d): %s" % (what, rc, lib().eddy_last_error().decode()))


def _p(a):
    return a.ctypes.data_as(_dp)


def _f64(a):
    return np.require(a, dtype=np.float64, requirements=["F", "A", "W"])


def direct(dp, g, ipres=1):
    """DIRECT through the Fortran-signature C entry point (host arrays, DP updated in place).

    dp: numpy float64 (NX, NY, NZ), Fortran order; g: eddyproj.grid.Grid.
    """
    L = lib()
    if not (isinstance(dp, np.ndarray) and dp.dtype == np.float64 and dp.flags.f_contiguous):
        raise EddyError("dp must be a Fortran-ordered float64 numpy array")
    dummy = np.zeros(1)
    iv = [ctypes.c_int(v) for v in (g.nx, g.ny, g.nz, g.nz, 1, 0, 0, 0, ipres, g.mp, g.lp, g.np)]
    delysq = np.array([g.delysq])
    delzsq = np.array([g.delzsq])
    rc = L.eddy_direct(_p(dp), _p(dummy), _p(dummy), _p(delysq), _p(delzsq), _p(g.ap), _p(g.au), _p(g.cp),
                       _p(g.cw), _p(g.rp), _p(g.ru), *[ctypes.byref(v) for v in iv])
    if rc == EDDY_ERR_INCOMPATIBLE:  # solved and reported, as the Fortran shim prints and continues
        import warnings
        warnings.warn(L.eddy_last_error().decode(), RuntimeWarning, stacklevel=2)
        return dp
    _check(rc, "eddy_direct")
    return dp


def release_direct_cache():
    _check(lib().eddy_direct_release(), "eddy_direct_release")


def _dptr(t):
    """Device pointer of a CUDA/HIP float64 tensor (contiguous)."""
    import torch
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
        raise EddyError("expected a contiguous float64 device tensor")
    return ctypes.c_void_p(t.data_ptr())


def to_device(a):
    """numpy (NX,NY,NZ) Fortran array -> device tensor (NZ,NY,NX)."""
    import torch
    return torch.from_numpy(np.ascontiguousarray(_f64(a).T)).to("cuda")


def to_host(t):
    """device tensor (NZ,NY,NX) -> numpy (NX,NY,NZ) Fortran array."""
    return np.asfortranarray(t.detach().cpu().numpy().T)


class Plan:
    """A device plan for DP(NX,NY,NZ) on the grid g (the SAVEd state of direct.f:121-585)."""

    def __init__(self, g, tuning=None, stream=None):
        L = lib()
        self.g = g
        self._h = _vp()
        self._stream = stream
        sp = _vp(stream.cuda_stream) if stream is not None else _vp(None)
        tun = None
        if tuning is not None:
            tun = Tuning(*[int(tuning.get(k, 0)) for k, _ in Tuning._fields_])
        rc = L.eddy_plan_create(ctypes.byref(self._h), g.nx, g.ny, g.nz, g.delysq, _p(g.ap), _p(g.au), _p(g.bp),
                                _p(g.bv), _p(g.cp), _p(g.cw), _p(g.rp), _p(g.ru), g.mp, g.lp, g.np,
                                ctypes.byref(tun) if tun is not None else None, sp)
        _check(rc, "eddy_plan_create")

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            lib().eddy_plan_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def tuning(self):
        t = Tuning()
        _check(lib().eddy_plan_get_tuning(self._h, ctypes.byref(t)), "eddy_plan_get_tuning")
        return {k: getattr(t, k) for k, _ in Tuning._fields_}

    def device_bytes(self):
        return int(lib().eddy_plan_device_bytes(self._h))

    def direct(self, dp):
        _check(lib().eddy_direct_dev(self._h, _dptr(dp)), "eddy_direct_dev")
        return dp

    def direct_refined(self, dp, sweeps=1, residual=True):
        r = ctypes.c_double(0.0)
        _check(lib().eddy_direct_refined_dev(self._h, _dptr(dp), int(sweeps), ctypes.byref(r) if residual else None),
               "eddy_direct_refined_dev")
        return r.value if residual else None

    def direct_adaptive(self, dp, tol=1e-10, max_sweeps=6):
        s = ctypes.c_int(0)
        r = ctypes.c_double(0.0)
        _check(lib().eddy_direct_refined_adaptive_dev(self._h, _dptr(dp), float(tol), int(max_sweeps),
                                                      ctypes.byref(s), ctypes.byref(r)),
               "eddy_direct_refined_adaptive_dev")
        return s.value, r.value

    def divergence(self, us, vs, ws, out, scale=1.0):
        _check(lib().eddy_divergence_dev(self._h, _dptr(us), _dptr(vs), _dptr(ws), _dptr(out), float(scale)),
               "eddy_divergence_dev")
        return out

    def correct(self, us, vs, ws, uo, vo, wo, dp, alfxdt):
        _check(lib().eddy_correct_dev(self._h, _dptr(us), _dptr(vs), _dptr(ws), _dptr(uo), _dptr(vo), _dptr(wo),
                                      _dptr(dp), float(alfxdt)), "eddy_correct_dev")

    def project(self, us, vs, ws, uo, vo, wo, dp, alfxdt, remove_mean=True, precision=EDDY_FP64, sweeps=1):
        _check(lib().eddy_project_dev(self._h, _dptr(us), _dptr(vs), _dptr(ws), _dptr(uo), _dptr(vo), _dptr(wo),
                                      _dptr(dp), float(alfxdt), int(bool(remove_mean)), int(precision),
                                      int(sweeps)), "eddy_project_dev")

    def div_minmax(self, us, vs, ws):
        lo, hi = ctypes.c_double(), ctypes.c_double()
        _check(lib().eddy_divergence_minmax_dev(self._h, _dptr(us), _dptr(vs), _dptr(ws), ctypes.byref(lo),
                                                ctypes.byref(hi)), "eddy_divergence_minmax_dev")
        return lo.value, hi.value

    def rhs_check(self, dp):
        """(|sum vol*b| / sum vol*|b|, count of non-finite values) of the RHS dp (interior)."""
        rel, nf = ctypes.c_double(), ctypes.c_long()
        _check(lib().eddy_rhs_check_dev(self._h, _dptr(dp), ctypes.byref(rel), ctypes.byref(nf)),
               "eddy_rhs_check_dev")
        return rel.value, nf.value

    def laplacian(self, phi, out):
        _check(lib().eddy_laplacian_dev(self._h, _dptr(phi), _dptr(out)), "eddy_laplacian_dev")
        return out

    def set_pressure_bc(self, implx=0, imply=0, ru1=0.0, rvimplx=None, rvimply=None, itype1=0, itype5=1,
                        itype6=1):
        """Parameters of the pressure update (Eddy6.f:2772-2846); masks are NX-long arrays."""
        rx = None if rvimplx is None else np.ascontiguousarray(rvimplx, dtype=np.float64)
        ry = None if rvimply is None else np.ascontiguousarray(rvimply, dtype=np.float64)
        _check(lib().eddy_plan_set_pressure_bc(self._h, int(implx), int(imply), float(ru1),
                                               _p(rx) if rx is not None else None,
                                               _p(ry) if ry is not None else None, int(itype1), int(itype5),
                                               int(itype6)), "eddy_plan_set_pressure_bc")
        self._press_keep = (rx, ry)

    def pressure_update(self, p, dp, alfxdt):
        """P = P + DP (+ implicit corrections) - PREF, ghost values (in place on p)."""
        _check(lib().eddy_pressure_update_dev(self._h, _dptr(p), _dptr(dp), float(alfxdt)),
               "eddy_pressure_update_dev")
        return p

    def perturb(self, u, v, w, us, vs, ws, sigma, seed, step):
        """us = u + sigma * N(0,1) (Philox4x32-10, counter {cell, step}, key {seed, 0}) on the
        faces CORRECT updates; `step` is a one-element int32 device tensor."""
        _check(lib().eddy_perturb_dev(self._h, _dptr(u), _dptr(v), _dptr(w), _dptr(us), _dptr(vs), _dptr(ws),
                                      float(sigma), int(seed), ctypes.c_void_p(step.data_ptr())),
               "eddy_perturb_dev")

    def enable_timing(self, on=True):
        _check(lib().eddy_plan_enable_timing(self._h, int(bool(on))), "eddy_plan_enable_timing")

    def reset_timing(self):
        _check(lib().eddy_plan_reset_timing(self._h), "eddy_plan_reset_timing")

    def kernel_times(self):
        """{kernel: (launches, total_ms)} for the launches recorded while timing was on."""
        out = {}
        for i, name in enumerate(KERNELS):
            n = ctypes.c_long(0)
            ms = ctypes.c_double(0.0)
            _check(lib().eddy_plan_kernel_time(self._h, i, ctypes.byref(n), ctypes.byref(ms)),
                   "eddy_plan_kernel_time")
            out[name] = (n.value, ms.value)
        return out


class Graph:
    """hipGraph of one projection step (divergence -> solve -> correction) on a plan whose
    stream is a non-default stream."""

    def __init__(self, plan, us, vs, ws, uo, vo, wo, dp, alfxdt, remove_mean=True, precision=EDDY_FP64, sweeps=1):
        self._h = _vp()
        self.plan = plan
        self._keep = (us, vs, ws, uo, vo, wo, dp)
        rc = lib().eddy_graph_create(ctypes.byref(self._h), plan.handle, _dptr(us), _dptr(vs), _dptr(ws), _dptr(uo),
                                     _dptr(vo), _dptr(wo), _dptr(dp), float(alfxdt), int(bool(remove_mean)),
                                     int(precision), int(sweeps))
        _check(rc, "eddy_graph_create")

    def launch(self):
        _check(lib().eddy_graph_launch(self._h), "eddy_graph_launch")

    def close(self):
        if self._h:
            lib().eddy_graph_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Resident:
    """The resident time loop: each step u* = u + sigma*N(0,1), u <- projection of u*, then
    the divergence range of u logged on the device; one captured hipGraph per step. u, v, w
    are updated in place; the plan needs a non-default stream."""

    def __init__(self, plan, u, v, w, dp, alfxdt, sigma, seed, precision=EDDY_FP64, sweeps=2, max_steps=1000):
        self._h = _vp()
        self.plan = plan
        self.max_steps = int(max_steps)
        self._keep = (u, v, w, dp)
        rc = lib().eddy_resident_create(ctypes.byref(self._h), plan.handle, _dptr(u), _dptr(v), _dptr(w), _dptr(dp),
                                        float(alfxdt), float(sigma), int(seed), int(precision), int(sweeps),
                                        int(max_steps))
        _check(rc, "eddy_resident_create")

    def run(self, nsteps):
        _check(lib().eddy_resident_run(self._h, int(nsteps)), "eddy_resident_run")

    def log(self):
        """(steps completed, array (steps, 2) of DIV(u) min/max after each step)."""
        buf = np.zeros(2 * self.max_steps)
        n = ctypes.c_int(0)
        _check(lib().eddy_resident_log(self._h, _p(buf), self.max_steps, ctypes.byref(n)), "eddy_resident_log")
        k = min(n.value, self.max_steps)
        return n.value, buf[:2 * k].reshape(k, 2)

    def close(self):
        if self._h:
            lib().eddy_resident_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


BATCH_BRANCHES = 0
BATCH_QUEUES = 1


class Batch:
    """Many independent problems of uneven sizes, run either as the problems' own kernels on
    parallel graph branches or as one launch per stage with a workgroup per work item of all
    problems (eddy_batch_* in include/eddyproj.h)."""

    def __init__(self, grids, tuning=None, stream=None):
        L = lib()
        self.grids = list(grids)
        n = len(self.grids)
        self._h = _vp()
        self._keep = []
        ints = lambda v: (ctypes.c_int * n)(*v)  # noqa: E731
        nx, ny, nz = ints([g.nx for g in grids]), ints([g.ny for g in grids]), ints([g.nz for g in grids])
        dely = (ctypes.c_double * n)(*[g.delysq for g in grids])

        def ptrs(attr):
            arr = (_dp * n)(*[_p(getattr(g, attr)) for g in grids])
            self._keep.append(arr)
            return arr
        tun = None
        if tuning is not None:
            tl = tuning if isinstance(tuning, (list, tuple)) else [tuning] * n
            tun = (Tuning * n)(*[Tuning(*[int(t.get(k, 0)) for k, _ in Tuning._fields_]) for t in tl])
        sp = _vp(stream.cuda_stream) if stream is not None else _vp(None)
        rc = L.eddy_batch_create(ctypes.byref(self._h), n, nx, ny, nz, dely, ptrs("ap"), ptrs("au"), ptrs("bp"),
                                 ptrs("bv"), ptrs("cp"), ptrs("cw"), ptrs("rp"), ptrs("ru"), tun, sp)
        _check(rc, "eddy_batch_create")

    def bind(self, dps, us=None, vs=None, ws=None, uo=None, vo=None, wo=None):
        n = len(self.grids)

        def arr(ts):
            if ts is None:
                return None
            a = (_vp * n)(*[_dptr(t) for t in ts])
            self._keep.append(a)
            return a
        self._fields = (dps, us, vs, ws, uo, vo, wo)
        _check(lib().eddy_batch_bind(self._h, arr(dps), arr(us), arr(vs), arr(ws), arr(uo), arr(vo), arr(wo)),
               "eddy_batch_bind")

    def direct(self, precision=EDDY_FP64, sweeps=0):
        _check(lib().eddy_batch_direct_dev(self._h, int(precision), int(sweeps)), "eddy_batch_direct_dev")

    def project(self, alfxdt, remove_mean=True):
        _check(lib().eddy_batch_project_dev(self._h, float(alfxdt), int(bool(remove_mean))), "eddy_batch_project_dev")

    def set_mode(self, mode):
        """BATCH_BRANCHES (per-problem kernels on parallel graph branches) or BATCH_QUEUES
        (one launch per stage, a workgroup per work item of all problems)."""
        _check(lib().eddy_batch_set_mode(self._h, int(mode)), "eddy_batch_set_mode")

    def info(self):
        c, i = ctypes.c_long(), ctypes.c_long()
        _check(lib().eddy_batch_info(self._h, ctypes.byref(c), ctypes.byref(i)), "eddy_batch_info")
        return c.value, i.value

    def close(self):
        if self._h:
            lib().eddy_batch_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
