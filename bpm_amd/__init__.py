"""bpm_amd: split-step beam propagation and vectorial Debye-Wolf focus fields
on AMD Instinct MI355X (gfx950).

Drop-in for the hot path of maweigert/bpm: ``bpm_3d`` (bpm/bpm/bpm_3d.py),
the ``Bpm3d`` class (bpm/bpmclass/), and the PSF functions ``psf``,
``psf_u0``, ``psf_lightsheet`` and ``focus_field_debye`` (bpm/psf/).  Compute
runs in the hand-written HIP kernels of libbpmhip.so (C ABI in
include/bpm_hip.h) called through ctypes; numpy arrays in, numpy arrays out,
as in the reference; torch GPU tensors are accepted as inputs.
"""
from .bpm3d import bpm_3d
from .bpmclass import Bpm3d
from .focus_fields import (psf, psf_u0, psf_lightsheet, psf_cylindrical, psf_cylindrical_u0,
                           focus_field_debye, focus_field_debye_plane, focus_field_debye_at,
                           focus_field_debye_gauss, focus_field_debye3, focus_field_cylindrical,
                           focus_field_cylindrical_plane)

__all__ = ["bpm_3d", "Bpm3d", "psf", "psf_u0", "psf_lightsheet", "psf_cylindrical",
           "psf_cylindrical_u0", "focus_field_debye", "focus_field_debye_plane",
           "focus_field_debye_at", "focus_field_debye_gauss", "focus_field_debye3",
           "focus_field_cylindrical", "focus_field_cylindrical_plane"]
__version__ = "0.1.0"
