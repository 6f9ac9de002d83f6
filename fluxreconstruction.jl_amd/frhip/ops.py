"""Library-level operators of FluxReconstruction.jl on libfrhip (include/fr.h, "row contractions"):
``interp_face!`` (src/interpolate.jl:1-10), ``poly_derivative!`` (src/derivative.jl:1-10) and
``filter_exp`` (src/Polynomial/poly_filter.jl:11-47).

``interp_face`` and ``poly_derivative`` write their first argument in place, like the reference's
bang functions.  Arrays are either numpy arrays in the reference's column-major layout (copied to
the device and back) or torch GPU tensors whose shape is the column-major shape reversed (a
column-major (nrow, n, nbatch) array is a contiguous torch tensor of shape (nbatch, n, nrow));
device work runs on torch's current stream.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, dptr

_vp = C.c_void_p
_dp = C.POINTER(C.c_double)

_lib.SIGNATURES.update({
    "fr_filter_exp": (C.c_int, [C.c_int, C.c_int, C.c_int, _dp, C.c_int, _dp, _dp]),
    "fr_interp_face_dev": (C.c_int, [_vp, C.c_longlong, C.c_int, C.c_int, _dp, _dp, _vp, _vp]),
    "fr_poly_derivative_dev": (C.c_int, [_vp, C.c_longlong, C.c_int, C.c_int, _dp, _vp, _vp]),
})


def _bind():
    L = _lib.lib()
    for name in ("fr_filter_exp", "fr_interp_face_dev", "fr_poly_derivative_dev"):
        res, args = _lib.SIGNATURES[name]
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return L


def filter_exp(N: int, s: int, V, Nc: int = 0, invV=None) -> np.ndarray:
    """Nodal exponential filter F = V diag(sigma) inv(V) (src/Polynomial/poly_filter.jl:11-24):
    KitBase filter_exp1d's sigma for a line basis (V of N+1 rows), filter_exp2d's
    (poly_filter.jl:32-47) for a triangle basis ((N+1)(N+2)/2 rows)."""
    V = np.asfortranarray(V, dtype=np.float64)
    nv = V.shape[0]
    if V.shape != (nv, nv):
        raise ValueError("filter_exp: V must be square")
    F = np.zeros((nv, nv), order="F")
    iv = None if invV is None else np.asfortranarray(invV, dtype=np.float64)
    check(_bind().fr_filter_exp(int(N), int(s), int(Nc), dptr(V), nv, dptr(iv) if iv is not None else None, dptr(F)))
    return F


def _is_torch(a) -> bool:
    return type(a).__module__.startswith("torch")


def _shape_cm(a):
    """Column-major shape (nrow, n, nbatch) of a numpy array (Fortran layout) or a torch tensor
    (reversed shape)."""
    shp = tuple(a.shape) if not _is_torch(a) else tuple(a.shape)[::-1]
    if len(shp) == 1:
        return 1, shp[0], 1
    if len(shp) == 2:
        return shp[0], shp[1], 1
    if len(shp) == 3:
        return shp
    raise ValueError("expected a vector, a matrix or a 3-index array")


def _run(fn, out, f, nout, op_args):
    import torch

    nrow, n, nb = _shape_cm(f)
    want = (nrow, nout, nb)
    if _shape_cm(out) != want:
        raise ValueError(f"output must have the column-major shape {want}")
    L = _bind()
    stream = _vp(torch.cuda.current_stream().cuda_stream)
    if _is_torch(f):
        for a in (f, out):
            if str(a.dtype) != "torch.float64" or not a.is_cuda or not a.is_contiguous():
                raise ValueError("device arrays must be contiguous float64 torch tensors on the GPU")
        check(getattr(L, fn)(_vp(f.data_ptr()), nrow, n, nb, *op_args, _vp(out.data_ptr()), stream))
        return out
    fd = torch.from_numpy(np.asarray(f, dtype=np.float64).ravel(order="F").copy()).cuda()
    od = torch.empty(nrow * nout * nb, dtype=torch.float64, device="cuda")
    check(getattr(L, fn)(_vp(fd.data_ptr()), nrow, n, nb, *op_args, _vp(od.data_ptr()), stream))
    res = od.cpu().numpy()
    out[...] = res.reshape(out.shape, order="F")
    return out


def interp_face(ff, f, ll, lr):
    """interp_face!(fδ, f, ll, lr): fδ[i, 1] = f[i, :]·ll, fδ[i, 2] = f[i, :]·lr for every row
    (src/interpolate.jl:1-10; a vector f is one row, fδ of length 2)."""
    ll = np.ascontiguousarray(ll, dtype=np.float64)
    lr = np.ascontiguousarray(lr, dtype=np.float64)
    return _run("fr_interp_face_dev", ff, f, 2, (dptr(ll), dptr(lr)))


def poly_derivative(df, f, pdm):
    """poly_derivative!(df, f, pdm): df[i] = f·pdm[i, :] (src/derivative.jl:1-10), for every row of a
    matrix or 3-index f."""
    pdm = np.asfortranarray(pdm, dtype=np.float64)
    n = _shape_cm(f)[1]
    if pdm.shape != (n, n):
        raise ValueError("poly_derivative: pdm must be n x n with n = size(f, 2)")  # the reference's @assert
    return _run("fr_poly_derivative_dev", df, f, n, (dptr(pdm),))
