"""Functional split-step BPM with the call signature of ``bpm.bpm_3d``.

Host logic restated from bpm/bpm/bpm_3d.py (argument handling, propagator
variants, mean modes, outputs, z-chunking); every step of the z-loop runs in
libbpmhip.so (``bpmh_propagate``).  Inputs may be numpy arrays (uploaded, as
the reference does with OCLArray.from_array) or torch tensors already on the
GPU (used in place).
"""
import atexit
import ctypes

import numpy as np

from . import _lib

__all__ = ["bpm_3d", "propagate_fields", "propagate_groups", "get_plan", "Plan"]

_MEAN = {"ignore": _lib.MEAN_IGNORE, "simple": _lib.MEAN_SIMPLE, "weighted": _lib.MEAN_WEIGHTED}
_plans = {}


def get_plan(Nx, Ny, batch=1):
    """Cached libbpmhip plan (FFT twiddles + scratch) for an (Ny, Nx) grid on
    the current device, set to the autotuned launch shapes for ``batch``
    fields (the cached choice of the nearest tuned batch size); batch None
    leaves the plan's launch shapes as they are."""
    t = _lib.torch()
    key = (int(Nx), int(Ny), t.cuda.current_device())
    p = _plans.get(key)
    if p is None:
        h = ctypes.c_void_p()
        _lib.call("bpmh_plan_create", int(Nx), int(Ny), ctypes.byref(h))
        p = _plans[key] = h
    if batch is None:
        return p
    apply_tuning(p, Nx, Ny, batch)
    return p


def apply_tuning(plan, Nx, Ny, batch):
    """Set a plan to the autotuned launch shapes for ``batch`` fields (the
    cached choice of the nearest tuned batch size), if any are cached."""
    from .tuning import lookup
    v = lookup(int(Nx), int(Ny), int(batch))
    if v is not None:  # two ints on the plan; a replayed graph is keyed by them
        _lib.call("bpmh_plan_set_variant", plan, int(v[0]), int(v[1]))


class Plan(object):
    """A plan of its own for an (Ny, Nx) grid on the current device, for
    callers that propagate on several streams at once (one plan per stream:
    a plan's scratch and captured graphs serve one stream at a time).
    Destroyed with ``close()`` or when garbage-collected."""

    def __init__(self, Nx, Ny):
        _lib.torch()
        self._h = ctypes.c_void_p()
        _lib.call("bpmh_plan_create", int(Nx), int(Ny), ctypes.byref(self._h))
        self.Nx, self.Ny = int(Nx), int(Ny)

    @property
    def _as_parameter_(self):
        return self._h

    def close(self):
        if self._h is not None and self._h.value and _lib._lib is not None:
            _lib._lib.bpmh_plan_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@atexit.register
def _destroy_plans():
    if _lib._lib is None:
        return
    for h in _plans.values():
        _lib._lib.bpmh_plan_destroy(h)
    _plans.clear()


def _dn_kind(dn, store_dn_as_half):
    """Device dtype of dn following bpm_3d.py:198-211."""
    if dn is None:
        return _lib.DN_NONE, None
    if _lib.is_tensor(dn):
        t = _lib.torch()
        if dn.is_complex():
            return _lib.DN_C64, np.complex64
        if dn.dtype == t.float16 or store_dn_as_half:
            return _lib.DN_F16, np.float16
        return _lib.DN_F32, np.float32
    if np.iscomplexobj(dn):
        return _lib.DN_C64, np.complex64
    if store_dn_as_half:
        return _lib.DN_F16, np.float16
    return _lib.DN_F32, np.float32


def propagator_table(Nx, Ny, units, lam, n0, kind=_lib.H_F32):
    """Device (Ny,Nx) complex64 propagator: float32 compute_propagator formula
    (bpm_3d_kernels.cl:408-432) or float64 host formula (bpmclass/bpm3d.py:88-107)."""
    dx, dy, dz = units
    H = _lib.empty((Ny, Nx), np.complex64)
    _lib.call("bpmh_propagator_table", int(Nx), int(Ny), int(kind), float(dx), float(dy),
              float(dz), float(n0), float(lam), _lib.ptr(H), _lib.stream_handle())
    return H


def _check_operands(plane, nz, dn, dn_kind, dn_batch_stride, mm, H, n_eff, dn0, out_field,
                    out_intens, out_sum, sw, gw, plain, sout):
    """Every tensor bpmh_propagate will touch, against what its kernels read
    and write for this batch, depth and grid (the C ABI sees bare pointers):
    raises ValueError before anything is launched."""
    import torch as t  # dtypes only; the device is checked last
    if plane.dim() != 3 or plane.dtype != t.complex64 or not plane.is_contiguous():
        raise ValueError("plane must be a contiguous (B, Ny, Nx) complex64 tensor")
    B, Ny, Nx = plane.shape
    n = Ny * Nx
    if nz < 1:
        raise ValueError("nz must be >= 1")
    if dn_batch_stride < 0:
        raise ValueError("dn_batch_stride must be >= 0")

    def need(x, name, dtype, numel):
        if x is None:
            return
        if x.dtype != dtype or not x.is_contiguous():
            raise ValueError("%s must be a contiguous %s tensor" % (name, dtype))
        if x.numel() < numel:
            raise ValueError("%s holds %d elements; this propagation reads or writes %d"
                             % (name, x.numel(), numel))
        if x.device != plane.device:
            raise ValueError("%s must be on %s" % (name, plane.device))

    if dn_kind != _lib.DN_NONE:
        if dn is None:
            raise ValueError("dn_kind needs a dn tensor")
        kind_dtype = {_lib.DN_F32: t.float32, _lib.DN_F16: t.float16, _lib.DN_C64: t.complex64}
        if dn_kind not in kind_dtype:
            raise ValueError("unknown dn kind")
        # planes 0..nz-1 of every item (slice 0 feeds the weighted mean)
        need(dn, "dn", kind_dtype[dn_kind], (B - 1) * dn_batch_stride + nz * n)
    if mm == _lib.MEAN_IGNORE:
        need(H, "H", t.complex64, n)
    if mm == _lib.MEAN_SIMPLE:
        need(n_eff, "n_eff", t.float32, nz)
        need(dn0, "dn0", t.float32, nz)
    need(out_field, "out_field", t.complex64, B * nz * n)
    need(out_intens, "out_intens", t.float32, B * nz * n)
    need(out_sum, "out_sum", t.float32, B * n)
    need(sw, "scattering weights", t.float32, n)
    need(gw, "g weights", t.float32, n)
    need(plain, "plane-wave array", t.complex64, nz)
    need(sout, "scattering output", t.float32, B * nz * 2)
    if not plane.is_cuda:
        raise ValueError("plane must be a GPU tensor")


def propagate_fields(plane, nz, units, lam, n0=1., dn=None, dn_kind=None, dn_batch_stride=0,
                     mean_mode="ignore", H=None, n_eff=None, dn0=None, absorbing_width=0,
                     out_field=None, out_intens=None, out_sum=None, scat=None, tuned=True,
                     plan=None):
    """Run nz-1 split-step steps on device for a batch of fields (no host copies).

    plane: complex64 tensor (B, Ny, Nx) on the GPU; holds u0 on entry and the
           last plane on exit.
    dn:    device tensor (nz, Ny, Nx) (shared) or (B, nz, Ny, Nx) with
           dn_batch_stride = nz*Ny*Nx; dn_kind one of _lib.DN_*.
    H:     propagator table for mean_mode "ignore"; n_eff/dn0: float32 device
           arrays (nz) for "simple".
    out_*: preallocated device tensors or None (see include/bpm_hip.h).
    scat:  None or (sw, gw, plain, out) device tensors for the scattering sums.
    tuned: apply the autotuned launch shapes for this batch size (False keeps
           the plan's current ones, as the autotuner needs).
    plan:  a plan of this grid to run on instead of the shared one of
           ``get_plan`` (a plan serves one stream at a time: callers that
           propagate on several streams at once pass a plan per stream).
    """
    B, Ny, Nx = plane.shape
    dx, dy, dz = units
    k0 = 2. * np.pi / lam
    mm = _MEAN[mean_mode] if isinstance(mean_mode, str) else int(mean_mode)
    if dn is None:
        dn_kind = _lib.DN_NONE
    sw = gw = plain = sout = None
    if scat is not None:
        sw, gw, plain, sout = scat
    _check_operands(plane, nz, dn, dn_kind, dn_batch_stride, mm, H, n_eff, dn0,
                    out_field, out_intens, out_sum, sw, gw, plain, sout)
    if plan is None:
        plan = get_plan(Nx, Ny, B if tuned else None)
    elif tuned:
        apply_tuning(plan, Nx, Ny, B)
    outs = [o.numel() for o in (out_field, out_intens) if o is not None]
    _lib.call("bpmh_propagate", plan, int(B), int(nz), _lib.ptr(plane), _lib.ptr(dn),
              int(dn_kind), int(dn_batch_stride), int(dn.numel() if dn is not None else 0),
              int(mm), _lib.ptr(H), _lib.ptr(n_eff), _lib.ptr(dn0), float(np.float32(n0)),
              float(np.float32(k0)), float(np.float32(dx)), float(np.float32(dy)),
              float(np.float32(dz)), float(np.float32(k0 * dz)), int(absorbing_width),
              _lib.ptr(out_field), _lib.ptr(out_intens), int(min(outs) if outs else 0),
              _lib.ptr(out_sum), _lib.ptr(sw), _lib.ptr(gw), _lib.ptr(plain), _lib.ptr(sout),
              _lib.stream_handle())
    return plane


def propagate_groups(plane, groups, units, lam, dn, H, dn_group_stride, out_sum=None,
                     absorbing_width=0, plan=None, done_events=None):
    """Ragged batch on the device (static propagator H, mean_mode "ignore"):
    ``groups`` = [(items, nz), ...] in order of non-increasing nz; group g's
    ``items`` fields (consecutive in ``plane``, (sum items, Ny, Nx) complex64)
    propagate through ``nz`` planes of dn volume g, at element offset
    g*dn_group_stride of the float32 tensor ``dn``.  ``out_sum``: None or
    (sum items, Ny, Nx) float32 sums of |u|^2.  ``done_events``: None or one
    torch.cuda.Event per group, recorded as soon as that group is final.
    Every item gives the bits of propagate_fields on its own volume with the
    same plan variant."""
    import torch as t
    tot = sum(int(g[0]) for g in groups)
    Bn, Ny, Nx = plane.shape
    if Bn != tot or plane.dtype != t.complex64 or not plane.is_contiguous() or not plane.is_cuda:
        raise ValueError("plane must be a contiguous (sum items, Ny, Nx) complex64 GPU tensor")
    if dn.dtype != t.float32 or not dn.is_contiguous() or dn.device != plane.device:
        raise ValueError("dn must be a contiguous float32 tensor on the plane's device")
    if out_sum is not None and (out_sum.dtype != t.float32 or not out_sum.is_contiguous()
                                or out_sum.numel() < tot * Ny * Nx or out_sum.device != plane.device):
        raise ValueError("out_sum must be a contiguous float32 tensor of (sum items, Ny, Nx)")
    if H is None or H.dtype != t.complex64 or H.numel() < Ny * Nx or H.device != plane.device:
        raise ValueError("H must be an (Ny, Nx) complex64 tensor on the plane's device")
    n = len(groups)
    items = (ctypes.c_int * n)(*[int(g[0]) for g in groups])
    nzs = (ctypes.c_int * n)(*[int(g[1]) for g in groups])
    evs = None
    if done_events is not None:
        if len(done_events) != n:
            raise ValueError("one event per group")
        for e in done_events:
            if not e.cuda_event:
                e.record()  # torch creates the event at its first record
        evs = (ctypes.c_void_p * n)(*[ctypes.c_void_p(e.cuda_event) for e in done_events])
    if plan is None:
        plan = get_plan(Nx, Ny, tot)
    k0 = 2. * np.pi / lam
    _lib.call("bpmh_propagate_groups", plan, n, items, nzs, _lib.ptr(plane), _lib.ptr(dn),
              _lib.DN_F32, int(dn_group_stride), int(dn.numel()), _lib.ptr(H),
              float(np.float32(k0 * units[2])), int(absorbing_width), _lib.ptr(out_sum), evs,
              _lib.stream_handle())
    return plane


def bpm_3d(size, units, lam=.5, u0=None, dn=None, subsample=1, n_volumes=1, n0=1.,
           return_scattering=False, return_g=False, return_full=True, return_field=True,
           absorbing_width=0, use_fresnel_approx=False, scattering_plane_ind=0,
           store_dn_as_half=False, mean_mode="weighted"):
    """Propagate a monochromatic field along z through dn (bpm/bpm/bpm_3d.py:39-101).

    size (Nx,Ny,Nz) pixels, units (dx,dy,dz) microns, lam wavelength, u0 the
    (Ny,Nx) input field (plane wave if None), dn the (Nz,Ny,Nx) index
    difference (real, float16-stored or complex), n0 the background index,
    mean_mode "weighted" (default), "simple" or "ignore".
    Returns u (volume or last plane, field or intensity)[, p[, g]].
    """
    if n_volumes == 1:
        return _bpm_3d(size, units, lam=lam, u0=u0, dn=dn, subsample=subsample, n0=n0,
                       return_scattering=return_scattering, return_full=return_full,
                       return_field=return_field, return_g=return_g,
                       absorbing_width=absorbing_width, use_fresnel_approx=use_fresnel_approx,
                       scattering_plane_ind=scattering_plane_ind,
                       store_dn_as_half=store_dn_as_half, mean_mode=mean_mode)
    return _bpm_3d_split(size, units, lam=lam, u0=u0, dn=dn, n_volumes=n_volumes,
                         subsample=subsample, n0=n0, return_scattering=return_scattering,
                         return_g=return_g, absorbing_width=absorbing_width,
                         return_full=return_full, return_field=return_field,
                         use_fresnel_approx=use_fresnel_approx,
                         store_dn_as_half=store_dn_as_half, mean_mode=mean_mode)


def _bpm_3d(size, units, lam=.5, u0=None, dn=None, subsample=1, n0=1.,
            return_scattering=False, return_g=False, return_full=True, return_field=True,
            use_fresnel_approx=False, absorbing_width=0, scattering_plane_ind=0,
            return_last_plane=False, store_dn_as_half=False, mean_mode="weighted"):
    """One unsplit propagation (bpm/bpm/bpm_3d.py:104-423)."""
    if mean_mode not in _MEAN:
        raise NotImplementedError("mean_mode  '%s' unknown" % mean_mode)
    if subsample != 1:
        raise NotImplementedError("subsample still has to be 1")
    Nx, Ny, Nz = (int(s) for s in size)
    dx, dy, dz = units
    k0 = 2. * np.pi / lam
    t = _lib.torch()

    if u0 is None:
        u0 = np.ones((Ny, Nx), np.complex64)
    if tuple(u0.shape) != (Ny, Nx):
        raise ValueError("u0 has shape %s, expected (Ny,Nx)=%s" % (tuple(u0.shape), (Ny, Nx)))
    plane = _lib.to_device(u0, np.complex64).reshape(1, Ny, Nx).clone()

    kind, ddt = _dn_kind(dn, store_dn_as_half)
    dn_g = None
    if dn is not None:
        if tuple(dn.shape) != (Nz, Ny, Nx):
            raise ValueError("dn has shape %s, expected (Nz,Ny,Nx)=%s" % (tuple(dn.shape), (Nz, Ny, Nx)))
        dn_g = _lib.to_device(dn, ddt)

    H = n_eff = dn0 = None
    if mean_mode == "ignore":
        H = propagator_table(Nx, Ny, units, lam, n0, _lib.H_F32)
    elif mean_mode == "simple":
        if dn is None:
            dn_mean = np.zeros(Nz, np.float32)
        elif _lib.is_tensor(dn):
            dn_mean = (dn.real if dn.is_complex() else dn).float().mean(dim=(1, 2)).cpu().numpy()
        else:
            dn_mean = np.mean(dn, axis=(1, 2))
            if np.iscomplexobj(dn_mean):
                dn_mean = dn_mean.real
        # n0 + dn0 formed in float64, then passed as float32 (bpm_3d.py:288-293)
        n_eff = _lib.to_device(np.array([np.float32(float(n0) + float(m)) for m in dn_mean],
                                        np.float32), np.float32)
        dn0 = _lib.to_device(np.asarray(dn_mean).astype(np.float32), np.float32)

    scat = None
    if return_scattering:
        KX, KY = np.meshgrid(2. * np.pi * np.fft.fftfreq(Nx, dx), 2. * np.pi * np.fft.fftfreq(Ny, dy),
                             indexing="xy")
        with np.errstate(invalid="ignore"):
            H0 = np.sqrt(n0 ** 2 * k0 ** 2 - KX ** 2 - KY ** 2)
        if use_fresnel_approx:
            H0 = 0.j + n0 * k0 - .5 * (KX ** 2 + KY ** 2) / n0 / k0
        H0 = np.where(np.isnan(H0), 0., H0)
        cos_theta = np.real(H0) / n0 / k0
        plain = (Nx * Ny * np.exp(-1.j * k0 * n0 * (scattering_plane_ind + np.arange(Nz)) * dz)
                 ).astype(np.complex64)
        scat = (_lib.to_device(cos_theta.astype(np.float32), np.float32),
                _lib.to_device((cos_theta ** 2).astype(np.float32), np.float32),
                _lib.to_device(plain, np.complex64),
                t.zeros((1, Nz, 2), dtype=t.float32, device=_lib.device()))

    out_field = out_intens = None
    if return_full:
        if return_field:
            out_field = _lib.empty((1, Nz, Ny, Nx), np.complex64)
        else:
            out_intens = _lib.empty((1, Nz, Ny, Nx), np.float32)

    propagate_fields(plane, Nz, units, lam, n0, dn=dn_g, dn_kind=kind, mean_mode=mean_mode,
                     H=H, n_eff=n_eff, dn0=dn0, absorbing_width=absorbing_width,
                     out_field=out_field, out_intens=out_intens, scat=scat)

    if return_full:
        u = _lib.to_numpy((out_field if return_field else out_intens)[0])
    else:
        u = _lib.to_numpy(plane[0])
        if not return_field:
            u = np.abs(u) ** 2

    if return_scattering:
        sums = _lib.to_numpy(scat[3][0])
        prefac = 1. / Nx / Ny * dx * dy
        p = np.float32(prefac) * sums[:, 0]
        if return_g:
            with np.errstate(invalid="ignore", divide="ignore"):
                g = (np.float32(prefac) * sums[:, 1]) / p
            result = (u, p, g)
        else:
            result = (u, p)
    else:
        result = u
    if return_last_plane:
        last = _lib.to_numpy(plane[0])
        result = (result + (last,)) if isinstance(result, tuple) else (result, last)
    return result


def _bpm_3d_split(size, units, lam=.5, u0=None, dn=None, n_volumes=1, n0=1., subsample=1,
                  return_scattering=False, return_full=True, return_g=False, return_field=True,
                  absorbing_width=0, use_fresnel_approx=False, store_dn_as_half=False,
                  mean_mode="weighted"):
    """z-chunked propagation (bpm/bpm/bpm_3d.py:863-992): Nz2 = Nz//n_volumes+1
    planes per chunk, chunks share their boundary plane, the last plane of a
    chunk is the next chunk's u0."""
    Nx, Ny, Nz = (int(s) for s in size)
    Nz2 = Nz // n_volumes + 1
    if (n_volumes - 1) * Nz2 >= Nz:
        raise ValueError("n_volumes=%d leaves an empty chunk for Nz=%d" % (n_volumes, Nz))
    if u0 is None:
        u0 = np.ones((subsample * Ny, subsample * Nx), np.complex64)
    u_dtype = np.complex64 if return_field else np.float32
    u = np.empty((Nz, Ny, Nx) if return_full else (Ny, Nx), u_dtype)
    p = np.empty(Nz, np.float32)
    g = np.empty(Nz, np.float32)
    kw = dict(units=units, lam=lam, n0=n0, subsample=subsample, return_full=return_full,
              return_field=return_field, return_g=return_g, absorbing_width=absorbing_width,
              return_scattering=return_scattering, use_fresnel_approx=use_fresnel_approx,
              store_dn_as_half=store_dn_as_half, mean_mode=mean_mode)
    for i in range(n_volumes):
        i1, i2 = i * Nz2, int(np.clip((i + 1) * Nz2, 0, Nz))
        if i < n_volumes - 1:
            dn_part = None if dn is None else dn[i1:i2 + 1]
            res = _bpm_3d((Nx, Ny, i2 - i1 + 1), u0=u0, dn=dn_part, return_last_plane=True,
                          scattering_plane_ind=Nz2 * i, **kw)
            if return_scattering:
                if return_g:
                    u_part, p_part, g_part, u0 = res
                    g[i1:i2] = g_part[1:]
                else:
                    u_part, p_part, u0 = res
                p[i1:i2] = p_part[1:]
            else:
                u_part, u0 = res
            if return_full:
                u[i1:i2] = u_part[:-1]
        else:
            dn_part = None if dn is None else dn[i1:i2]
            res = _bpm_3d((Nx, Ny, i2 - i1), u0=u0, dn=dn_part, scattering_plane_ind=Nz2 * i, **kw)
            if return_scattering:
                if return_g:
                    u_part, p_part, g_part = res
                    g[i1:i2] = g_part
                    g[i1] = g[i1 - 1]
                else:
                    u_part, p_part = res
                p[i1:i2] = p_part
                p[i1] = p[i1 - 1]
            else:
                u_part = res
            if return_full:
                u[i1:i2] = u_part
            else:
                u = u_part
    if return_scattering:
        return (u, p, g) if return_g else (u, p)
    return u
