"""Class API ``Bpm3d`` (bpm/bpmclass/bpm3d.py, bpm/bpmclass/_bpm3d_ocl.py) on HIP.

The host set-up (k-grid, float64 propagator, scattering weights) follows
bpm/bpmclass/bpm3d.py:64-115; the propagator is rounded to complex64 on the
device (``bpmh_propagator_table`` kind F64) and the z-loop of
``_propagate_single`` (bpm/bpmclass/_bpm3d_ocl.py:66-134) runs in
``bpmh_propagate`` with a static H and no mean correction.
"""
import numpy as np

from . import _lib
from .bpm3d import propagate_fields, propagator_table

__all__ = ["Bpm3d"]


class _Bpm3d_Base(object):
    """Split-step beam propagation through a refractive-index volume.

    model = Bpm3d(size=(128,128,128), units=(0.1,0.1,0.1), lam=0.5, n0=1.33)
    u = model.propagate(u0)
    """

    _float_type = np.float32
    _complex_type = np.complex64

    def __init__(self, size, units, lam=.5, n0=1., dn=None, use_fresnel_approx=False,
                 n_volumes=1):
        self.n_volumes = n_volumes
        # the reference forces use_fresnel_approx=False here (bpmclass/bpm3d.py:55-56)
        self._setup(size=size, units=units, lam=lam, n0=n0, use_fresnel_approx=False)
        self.set_dn(dn)
        self._setup_impl()

    def _setup(self, size, units, lam=.5, n0=1., use_fresnel_approx=False):
        """k-grid, propagator and scattering weights (bpmclass/bpm3d.py:64-115).

        Arrays are laid out (Ny, Nx); the reference's ``meshgrid(kxs,kys,'ij')``
        coincides with this only for Nx==Ny and dx==dy."""
        self.size = tuple(size)
        self.size2d = self.size[:2]
        self.units = tuple(units)
        self.n0 = n0
        self.lam = lam
        Nx, Ny, Nz = self.size
        dx, dy, dz = self.units
        self.k0 = 2. * np.pi / lam
        kxs = 2. * np.pi * np.fft.fftfreq(Nx, dx)
        kys = 2. * np.pi * np.fft.fftfreq(Ny, dy)
        self._KY, self._KX = np.meshgrid(kys, kxs, indexing="ij")
        with np.errstate(invalid="ignore"):
            self._H0 = np.sqrt(n0 ** 2 * self.k0 ** 2 - self._KX ** 2 - self._KY ** 2)
        if use_fresnel_approx:
            self._H0 = 0.j + n0 ** 2 * self.k0 - .5 * (self._KX ** 2 + self._KY ** 2)
        outside = np.isnan(self._H0)
        self._H = np.exp(-1.j * dz * self._H0)
        self._H[outside] = 0.
        self._H0[outside] = 0.
        self._cos_theta = np.real(self._H0) / self.n0 / self.k0
        self.scatter_weights = self._cos_theta
        self.gfactor_weights = self._cos_theta ** 2
        self.plain_wave_dct = Nx * Ny * np.exp(-1.j * self.k0 * n0 * np.arange(Nz) * dz).astype(np.complex64)

    def set_dn(self, dn=None):
        """dn: (Nz,Ny,Nx) float32/complex64 refractive index difference (or None)."""
        if dn is None:
            self.dn = None
            return
        if tuple(self.size) != tuple(dn.shape[::-1]):
            raise ValueError("size of dn  %s doesn't match internal size %s"
                             % (tuple(dn.shape[::-1]), self.size))
        if _lib.is_tensor(dn):
            self._is_complex_dn = bool(dn.is_complex())
            self.dn = dn
        elif np.iscomplexobj(dn):
            self._is_complex_dn = True
            self.dn = dn.astype(self._complex_type, copy=False)
        else:
            self._is_complex_dn = False
            self.dn = dn.astype(self._float_type, copy=False)
        if getattr(self, "_ready", False):
            self._upload_dn()


class _Bpm3d_HIP(_Bpm3d_Base):
    """HIP implementation (replaces _Bpm3d_OCL)."""

    def _setup_impl(self):
        Nx, Ny, Nz = self.size
        self._H_g = propagator_table(Nx, Ny, self.units, self.lam, self.n0, _lib.H_F64)
        self._ready = True
        self._upload_dn()

    def _upload_dn(self):
        self.dn_g = None
        if self.dn is not None:
            self.dn_g = _lib.to_device(self.dn, np.complex64 if self._is_complex_dn else np.float32)

    def _propagate_single(self, u0=None, return_full=True, return_intensity=False,
                          absorbing_width=0, **kwargs):
        """Propagate u0 (plane wave if None) through the whole volume.

        return_full: the (Nz,Ny,Nx) volume, else the last plane (the complex
        field, whatever return_intensity says, as _bpm3d_ocl.py:128-131);
        return_intensity: with return_full, the |u|^2 volume instead of the
        field (the reference calls an absent kernel ``fill_with_energy`` for
        it, _bpm3d_ocl.py:88-91,120-123; defined here as |u|^2)."""
        Nx, Ny, Nz = self.size
        if u0 is None:
            u0 = np.ones((Ny, Nx), np.complex64)
        if tuple(u0.shape) != (Ny, Nx):
            raise ValueError("u0 has shape %s, expected %s" % (tuple(u0.shape), (Ny, Nx)))
        plane = _lib.to_device(u0, np.complex64).reshape(1, Ny, Nx).clone()
        out_field = out_intens = None
        if return_full:
            if return_intensity:
                out_intens = _lib.empty((1, Nz, Ny, Nx), np.float32)
            else:
                out_field = _lib.empty((1, Nz, Ny, Nx), np.complex64)
        kind = _lib.DN_NONE
        if self.dn_g is not None:
            kind = _lib.DN_C64 if self._is_complex_dn else _lib.DN_F32
        propagate_fields(plane, Nz, self.units, self.lam, self.n0, dn=self.dn_g, dn_kind=kind,
                         mean_mode="ignore", H=self._H_g, absorbing_width=absorbing_width,
                         out_field=out_field, out_intens=out_intens)
        if return_full:
            return _lib.to_numpy((out_intens if return_intensity else out_field)[0])
        return _lib.to_numpy(plane[0])

    def propagate(self, u0=None, **kwargs):
        """Alias of _propagate_single (bpm/bpm/bpm_class.py:91)."""
        return self._propagate_single(u0, **kwargs)

    def __repr__(self):
        return "Bpm3d class with size %s and units %s" % (self.size, self.units)


def Bpm3d(size, units, lam=.5, n0=1., dn=None, use_fresnel_approx=False, n_volumes=1,
          backend="hip"):
    """Factory for Bpm3d objects (bpm/bpmclass/bpm3d.py:146-165).

    backend "hip" (or the reference's "opencl", accepted as an alias) selects
    the gfx950 implementation; anything else raises NotImplementedError."""
    if backend in ("hip", "opencl"):
        return _Bpm3d_HIP(size=size, units=units, lam=lam, n0=n0, dn=dn,
                          use_fresnel_approx=use_fresnel_approx, n_volumes=n_volumes)
    raise NotImplementedError


Bpm3d.__doc__ = (Bpm3d.__doc__ or "") + "\n" + (_Bpm3d_Base.__doc__ or "")
