"""Field files of the reference (src/io.f) through libeddyproj.so's native reader/writer.

`read_scalar` / `write_scalar` follow IOSCALAR (src/io.f:114-200) and `read_vector` /
`write_vector` IOVECTOR (src/io.f:8-108): the same record layout, the same rank slab
(`myrank`, `mysize`: NZ planes from myrank*(NZ-2)+1 on). `load_restart` reads the restart set
u_, v_, w_, p_<step>.res that src/Eddy6.f:780-796 reads. Host fields are numpy (NX, NY, NZ)
Fortran arrays; with device=True fields are float64 device tensors (NZ, NY, NX), filled
through pinned staging buffers on the current stream.

Errors (missing or damaged file, dimensions that differ from the file's) raise EddyError;
nothing is stored then.
"""
import ctypes
import os

import numpy as np

from . import EddyError, _check, _dptr, _f64, lib

_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p
_ci = ctypes.c_int
_cd = ctypes.c_double
_ready = False


class FieldInfo(ctypes.Structure):
    _fields_ = [("nx", _ci), ("ny", _ci), ("nz", _ci), ("nvar", _ci), ("per_rank", _ci), ("has_time", _ci),
                ("nstep", _ci), ("time", _cd), ("dtm1", _cd), ("grav", _cd)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def _lib():
    global _ready
    L = lib()
    if not _ready:
        pi = ctypes.POINTER(FieldInfo)
        L.eddy_field_info_read.argtypes = [ctypes.c_char_p, pi]
        L.eddy_ioscalar_read.argtypes = [ctypes.c_char_p, _dp, _ci, _ci, _ci, _ci, _ci, pi]
        L.eddy_ioscalar_read_dev.argtypes = [ctypes.c_char_p, _vp, _ci, _ci, _ci, _ci, _ci, pi, _vp]
        L.eddy_ioscalar_write.argtypes = [ctypes.c_char_p, _dp, _ci, _ci, _ci, _ci, _ci, _cd, _cd, _cd]
        L.eddy_ioscalar_write_dev.argtypes = [ctypes.c_char_p, _vp, _ci, _ci, _ci, _ci, _ci, _cd, _cd, _cd, _vp]
        L.eddy_iovector_read.argtypes = [ctypes.c_char_p, _dp, _dp, _dp, _ci, _ci, _ci, _ci, _ci, pi]
        L.eddy_iovector_read_dev.argtypes = [ctypes.c_char_p, _vp, _vp, _vp, _ci, _ci, _ci, _ci, _ci, pi, _vp]
        L.eddy_iovector_write.argtypes = [ctypes.c_char_p, _dp, _dp, _dp, _ci, _ci, _ci, _ci, _cd]
        L.eddy_iovector_write_dev.argtypes = [ctypes.c_char_p, _vp, _vp, _vp, _ci, _ci, _ci, _ci, _cd, _vp]
        _ready = True
    return L


def _path(p):
    return os.fsencode(p)


def _stream():
    import torch
    return _vp(torch.cuda.current_stream().cuda_stream)


def info(path):
    """Header (and IOSCALAR trailer) of a field file as a dict."""
    fi = FieldInfo()
    _check(_lib().eddy_field_info_read(_path(path), ctypes.byref(fi)), "eddy_field_info_read")
    return fi.as_dict()


def _dims(path, nx, ny, nz, mysize):
    if nx is None or ny is None or nz is None:
        h = info(path)
        nx = h["nx"] if nx is None else nx
        ny = h["ny"] if ny is None else ny
        if nz is None:
            if mysize != 1:
                raise EddyError("give nz (the rank's plane count) when mysize > 1")
            nz = h["nz"]
    return nx, ny, nz


def _alloc(shape, device):
    if device:
        import torch
        return torch.empty(shape[::-1], dtype=torch.float64, device="cuda")
    return np.empty(shape, dtype=np.float64, order="F")


def _ptr(a, device):
    return _dptr(a) if device else a.ctypes.data_as(_dp)


def read_scalar(path, nx=None, ny=None, nz=None, myrank=0, mysize=1, device=False):
    """IOSCALAR(NAME, P, ..., DIR=-1): returns (P, info). Dimensions default to the file's."""
    nx, ny, nz = _dims(path, nx, ny, nz, mysize)
    p = _alloc((nx, ny, nz), device)
    fi = FieldInfo()
    L = _lib()
    if device:
        rc = L.eddy_ioscalar_read_dev(_path(path), _ptr(p, True), nx, ny, nz, myrank, mysize, ctypes.byref(fi),
                                      _stream())
    else:
        rc = L.eddy_ioscalar_read(_path(path), _ptr(p, False), nx, ny, nz, myrank, mysize, ctypes.byref(fi))
    _check(rc, "eddy_ioscalar_read")
    return p, fi.as_dict()


def read_vector(path, nx=None, ny=None, nz=None, myrank=0, mysize=1, device=False):
    """IOVECTOR(NAME, UO, VO, WO, ..., DIR=-1): returns ((U, V, W), info)."""
    nx, ny, nz = _dims(path, nx, ny, nz, mysize)
    f = [_alloc((nx, ny, nz), device) for _ in range(3)]
    fi = FieldInfo()
    L = _lib()
    ptrs = [_ptr(a, device) for a in f]
    if device:
        rc = L.eddy_iovector_read_dev(_path(path), *ptrs, nx, ny, nz, myrank, mysize, ctypes.byref(fi), _stream())
    else:
        rc = L.eddy_iovector_read(_path(path), *ptrs, nx, ny, nz, myrank, mysize, ctypes.byref(fi))
    _check(rc, "eddy_iovector_read")
    return tuple(f), fi.as_dict()


def _shape(a, device):
    if device:
        nz, ny, nx = a.shape
        return nx, ny, nz
    return a.shape


def _host_field(a):
    a = _f64(a)
    if not a.flags.f_contiguous:
        a = np.asfortranarray(a)
    return a


def write_scalar(path, p, nstep=0, time=0.0, dtm1=0.0, grav=0.0, per_rank=False):
    """IOSCALAR(NAME, P, ..., DIR=1) (per_rank=False) or DIR=2 (per_rank=True)."""
    device = not isinstance(p, np.ndarray)
    if not device:
        p = _host_field(p)
    nx, ny, nz = _shape(p, device)
    L = _lib()
    if device:
        rc = L.eddy_ioscalar_write_dev(_path(path), _ptr(p, True), nx, ny, nz, int(per_rank), nstep, time, dtm1,
                                       grav, _stream())
    else:
        rc = L.eddy_ioscalar_write(_path(path), _ptr(p, False), nx, ny, nz, int(per_rank), nstep, time, dtm1, grav)
    _check(rc, "eddy_ioscalar_write")


def write_vector(path, u, v, w, time=0.0, per_rank=False):
    """IOVECTOR(NAME, UO, VO, WO, ..., DIR=1) (TIME not stored, as src/io.f:47) or DIR=2."""
    device = not isinstance(u, np.ndarray)
    f = [a if device else _host_field(a) for a in (u, v, w)]
    nx, ny, nz = _shape(f[0], device)
    if any(_shape(a, device) != (nx, ny, nz) for a in f):
        raise EddyError("U, V, W must have one shape")
    L = _lib()
    ptrs = [_ptr(a, device) for a in f]
    if device:
        rc = L.eddy_iovector_write_dev(_path(path), *ptrs, nx, ny, nz, int(per_rank), time, _stream())
    else:
        rc = L.eddy_iovector_write(_path(path), *ptrs, nx, ny, nz, int(per_rank), time)
    _check(rc, "eddy_iovector_write")


def restart_name(directory, var, step):
    """The restart file name of src/Eddy6.f:780: <dir><var>_<step, 8 digits>.res."""
    return os.path.join(directory, "%s_%08d.res" % (var, step))


def load_restart(directory, step, names=("u", "v", "w", "p"), nx=None, ny=None, nz=None, device=True):
    """The restart set of src/Eddy6.f:780-796: {name: field} and the info of the last file read
    (NSTEP, TIME, DTM1, GRAV come from every file's trailer; the reference keeps the last)."""
    out, last = {}, None
    for n in names:
        out[n], last = read_scalar(restart_name(directory, n, step), nx, ny, nz, device=device)
    return out, last


def save_restart(directory, step, fields, nstep=0, time=0.0, dtm1=0.0, grav=0.0):
    """Write {name: field} as the restart set <dir><name>_<step>.res (IOSCALAR, DIR=1)."""
    for n, a in fields.items():
        write_scalar(restart_name(directory, n, step), a, nstep=nstep, time=time, dtm1=dtm1, grav=grav)
