"""Refractive-index volumes from spheres, generated on the GPU.

``create_dn_buffer`` and ``bpm_3d_spheres`` keep the signatures of
bpm/bpm/bpm_3d_spheres.py:11-67 (coated spheres, nearest-centre rule of
bpm/bpm/kernels/bpm_3d_spheres.cl:31-74); ``spheres_dn_device`` builds the
summed-sphere volumes of the benchmark workloads on the device (equal, bit for
bit, to the host generator ``bpm_amd.workloads.spheres_dn``).
"""
import numpy as np

from . import _lib
from .bpm3d import bpm_3d

__all__ = ["create_dn_buffer", "bpm_3d_spheres", "spheres_dn_device"]


def create_dn_buffer(size, units, points, dn_inner=.0, rad_inner=0, dn_outer=.1, rad_outer=.4):
    """(Nz,Ny,Nx) float32 GPU tensor: dn_inner inside rad_inner of the nearest
    point, dn_outer inside rad_outer, 0 elsewhere (points: [[x,y,z],...] um)."""
    Nx, Ny, Nz = size
    dx, dy, dz = units
    ps = np.array(points, np.float64).reshape(-1, 3)
    ps = ps[np.argsort(ps[:, 2], kind="stable"), :]
    pts = _lib.to_device(ps.astype(np.float32).ravel(), np.float32)
    dn = _lib.empty((Nz, Ny, Nx), np.float32)
    f = lambda v: float(np.float32(v))
    _lib.call("bpmh_fill_dn_coated", int(Nx), int(Ny), int(Nz), f(dx), f(dy), f(dz),
              _lib.ptr(pts), int(ps.shape[0]), f(dn_inner), f(rad_inner), f(dn_outer),
              f(rad_outer), _lib.ptr(dn), _lib.stream_handle())
    return dn


def bpm_3d_spheres(size, units, lam=.5, u0=None, points=None, dn_inner=0.0, rad_inner=0.,
                   dn_outer=0.05, rad_outer=2.5, return_scattering=False,
                   use_fresnel_approx=False):
    """Propagation through coated spheres at ``points`` (bpm_3d_spheres.py:42-67)."""
    dn = create_dn_buffer(size, units, points=points, dn_inner=dn_inner, rad_inner=rad_inner,
                          dn_outer=dn_outer, rad_outer=rad_outer)
    return bpm_3d(size=size, units=units, lam=lam, u0=u0, dn=dn,
                  return_scattering=return_scattering, use_fresnel_approx=use_fresnel_approx)


def spheres_dn_device(size, units, centres, radii, dns, out=None):
    """(Nz,Ny,Nx) float32 GPU tensor, sum of dn_s inside each sphere (strict);
    written into ``out`` (a contiguous (Nz,Ny,Nx) float32 GPU tensor) if given."""
    Nx, Ny, Nz = size
    dx, dy, dz = units
    tab = np.column_stack([np.asarray(centres, np.float64).reshape(-1, 3),
                           np.asarray(radii, np.float64).ravel(),
                           np.asarray(dns, np.float64).ravel()])
    t = _lib.torch()
    dev_tab = t.from_numpy(np.ascontiguousarray(tab).ravel()).to(_lib.device())
    if out is None:
        dn = _lib.empty((Nz, Ny, Nx), np.float32)
    else:
        if (tuple(out.shape) != (Nz, Ny, Nx) or out.dtype != t.float32 or not out.is_contiguous()
                or not out.is_cuda):
            raise ValueError("out must be a contiguous (Nz,Ny,Nx) float32 GPU tensor")
        dn = out
    _lib.call("bpmh_spheres_dn", int(Nx), int(Ny), int(Nz), float(dx), float(dy), float(dz),
              _lib.ptr(dev_tab), int(tab.shape[0]), _lib.ptr(dn), _lib.stream_handle())
    return dn
