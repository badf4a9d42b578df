"""Spectral regridding between Gaussian grids / truncations (``pylbm_isca/regrid.py``).

A field moves through spectral space: native grid -> grid_to_wave ->
resize the coefficients -> wave_to_grid on the target grid.  With the
reference's normalisation a coefficient xi[m, n] means the same on every
grid, so the resize is zero-padding (finer) or dropping the n > N tail
(coarser); the coefficient copy is a HIP kernel (``lbm_spec_resize``) and
both transforms run on the device.  Winds go through vorticity and
divergence (uv_to_vordiv -> resize -> vordiv_to_uv); every other field,
including vorticity/divergence forcing fields, is a scalar.
"""

import ctypes
import types

import numpy as np

from . import _lib
from .backend import asnumpy


def detect_grid(path):
    """Native ``(nlon, nlat)`` of a big-endian Fortran sequential field file.

    The file's first 4-byte record marker is the byte count of one real*4
    field; with two longitudes per latitude (every Gaussian grid this model
    reads) the point count fixes the grid.  Behaviour of
    ``regrid.py:37-58``.
    """
    with open(path, "rb") as fh:
        head = fh.read(4)
    if len(head) < 4:
        raise ValueError(f"{path}: shorter than one record marker")
    nbytes = int.from_bytes(head, "big", signed=True)
    npts, rem = divmod(nbytes, 4)
    if nbytes <= 0 or rem:
        raise ValueError(f"{path}: first record marker {nbytes} is not a real*4 field size")
    nlat = int(np.sqrt(npts / 2.0) + 0.5)
    if 2 * nlat * nlat != npts:
        raise ValueError(f"{path}: {npts} values per field do not make an nlon = 2 nlat grid; "
                         "give the source grid instead of detecting it")
    return 2 * nlat, nlat


def resize_spec(xi, dst_nmax, batched=False):
    """Spectral ``(M, M[, K])`` complex array (``(E, M, M[, K])`` when
    ``batched``) zero-padded or truncated to ``dst_nmax`` on the device;
    returns a complex128 device tensor of the same layout (``regrid.py:66-77``)."""
    import torch
    t = torch.as_tensor(xi, device="cuda") if not isinstance(xi, torch.Tensor) else xi.to("cuda")
    t = t.to(torch.complex128)
    lead = 1 if batched else 0
    if t.dim() not in (lead + 2, lead + 3) or t.shape[lead] != t.shape[lead + 1]:
        raise ValueError(f"unexpected spectral shape {tuple(t.shape)}")
    levels = t.dim() == lead + 3
    Ms = int(t.shape[lead])
    if Ms - 1 == dst_nmax:
        return t
    B = int(t.shape[0]) if batched else 1
    K = int(t.shape[-1]) if levels else 1
    body = t.reshape(B, Ms, Ms, K).contiguous()
    Md = dst_nmax + 1
    out = torch.empty((B, Md, Md, K), dtype=torch.complex128, device="cuda")
    lib = _lib.load()
    _lib.check(lib.lbm_spec_resize(_lib.ptr(body), _lib.ptr(out), B, Ms - 1, int(dst_nmax), K,
                                   ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
    shape = ((B,) if batched else ()) + (Md, Md) + ((K,) if levels else ())
    return out.reshape(shape)


def _same(a, b):
    if _batched(a) != _batched(b):
        raise ValueError("regridding needs two batched or two single-member transforms")
    return (a.nlon, a.nlat, a.nmax) == (b.nlon, b.nlat, b.nmax)


def _batched(sp):
    return bool(getattr(sp, "batched", False))


def regrid_scalar(src_sp, dst_sp, field):
    """A scalar grid field (2-D or 3-D; leading member axis for batched
    transforms) on ``dst_sp``'s grid, as a host array (``regrid.py:84-96``)."""
    if _same(src_sp, dst_sp):
        return np.asarray(field)
    xi = resize_spec(src_sp.grid_to_wave(field), dst_sp.nmax, _batched(src_sp))
    return np.asarray(asnumpy(dst_sp.wave_to_grid(xi)))


def regrid_uv(src_sp, dst_sp, u, v):
    """Horizontal wind through vorticity/divergence (``regrid.py:99-110``)."""
    if _same(src_sp, dst_sp):
        return np.asarray(u), np.asarray(v)
    vor, div = src_sp.uv_to_vordiv(u, v)
    vor = resize_spec(vor, dst_sp.nmax, _batched(src_sp))
    div = resize_spec(div, dst_sp.nmax, _batched(src_sp))
    u2, v2 = dst_sp.vordiv_to_uv(vor, div)
    return np.asarray(asnumpy(u2)), np.asarray(asnumpy(v2))


def regrid_basic_state(bs, src_sp, dst_sp):
    """A basic state on ``dst_sp``'s grid (``regrid.py:113-124``)."""
    u, v = regrid_uv(src_sp, dst_sp, bs.u, bs.v)
    return types.SimpleNamespace(
        u=u, v=v, t=regrid_scalar(src_sp, dst_sp, bs.t), q=regrid_scalar(src_sp, dst_sp, bs.q),
        ps=regrid_scalar(src_sp, dst_sp, bs.ps), lat_deg=dst_sp.lat_deg, lon_deg=dst_sp.lon_deg,
        sig=getattr(bs, "sig", None), source=getattr(bs, "source", "") + " (regridded)")


def regrid_mjo_forcing(mjo, src_sp, dst_sp):
    """An MJOForcing moved to ``dst_sp``'s grid in place (``regrid.py:127-139``):
    all days and levels in one batched transform."""
    if _same(src_sp, dst_sp):
        return mjo
    ndays, nlat, nlon, nlev = mjo.gtf.shape
    stack = np.transpose(np.asarray(mjo.gtf, dtype=np.float64), (1, 2, 0, 3)).reshape(
        nlat, nlon, ndays * nlev)
    reg = regrid_scalar(src_sp, dst_sp, stack)
    mjo.gtf = np.transpose(reg.reshape(dst_sp.nlat, dst_sp.nlon, ndays, nlev), (2, 0, 1, 3))
    mjo.reset()
    return mjo


def regrid_steady_forcing(steady, src_sp, dst_sp):
    """A SteadyForcing moved in place; all four channels are scalars
    (``regrid.py:142-151``)."""
    if _same(src_sp, dst_sp):
        return steady
    for name in ("vor", "div", "t", "ps"):
        setattr(steady, name, regrid_scalar(src_sp, dst_sp, getattr(steady, name)))
    return steady
