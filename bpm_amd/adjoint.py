"""Differentiable split-step BPM: d(loss)/d(dn) through a hand-written adjoint.

``propagate(dn, u0, units, lam, n0, checkpoint=16)`` returns the exit fields
of a batch of inputs (e.g. illumination angles) propagated through a dn
volume -- real float32, or complex64 for an absorbing medium (the
``mult_dn_complex`` screen of bpm/bpmclass/kernels/bpm_3d_kernels.cl, no
taper) -- on the static-propagator path of ``Bpm3d`` (no mean correction);
it is a ``torch.autograd.Function``:

* forward: the z-loop runs in ``bpmh_propagate`` segment by segment, keeping
  only the plane at every ``checkpoint``-th slice;
* backward: for each segment, last to first, the forward planes are
  recomputed from the segment's checkpoint and ``bpmh_adjoint_segment``
  runs the adjoint steps  G_{z-1} = F^-1 conj(H) F (conj(s_z) G_z)  while
  accumulating  dL/ddn_z = k0 dz Im(conj(G_z) u_z)  (G = torch's complex
  gradient, 2 dL/du*); for complex dn also dL/dIm dn_z = k0 dz Re(conj(G_z)
  u_z), returned as the complex gradient dL/dRe + i dL/dIm (torch's
  convention for complex leaves).

With ``mean_mode="weighted"`` (the default of ``bpm_3d``,
bpm/bpm/bpm_3d.py:149-159,302-311) the propagator of each step is H(n0+dn0)
with dn0 = sum|U| d / sum|U| over the k-space plane U of that step (d = dn
slice 0 at the same flat index), and the screen uses dn - dn0.  dn0 depends
on the field and on slice 0, so the adjoint (``bpmh_mean_segment``)
carries two more terms per step: dL/ddn0 = -sum(screen gradient) + the
propagator's derivative, summed over k-space, and through it a k-space term
in the field gradient and the gradient of slice 0.  With
``mean_mode="simple"`` dn0 = mean(dn[z]) of the step's target plane, so
dL/ddn0 / N is added to every voxel of that plane.  Real dn only.

Memory: checkpoints B*ceil(nz/K) planes + one segment of B*(K+1) planes
(weighted: about 2*B*(K+2) planes of recorded forward state).
"""
import ctypes

import numpy as np

from . import _lib
from . import bpm3d as B

__all__ = ["propagate"]


def _segments(nz, K):
    return [(z0, min(z0 + K, nz - 1)) for z0 in range(0, nz - 1, K)]


def _make_fn():
    t = _lib.torch()

    class _Fn(t.autograd.Function):
        @staticmethod
        def forward(ctx, dn, u0, units, lam, n0, absorbing_width, checkpoint):
            nz, ny, nx = dn.shape
            kind = _lib.DN_C64 if dn.is_complex() else _lib.DN_F32
            H = B.propagator_table(nx, ny, units, lam, n0, _lib.H_F64)
            plane = u0.detach().to(dtype=t.complex64).contiguous().clone()
            segs = _segments(nz, checkpoint)
            ckpts = []
            dnc = dn.detach().contiguous()
            for z0, z1 in segs:
                ckpts.append(plane.clone())
                B.propagate_fields(plane, z1 - z0 + 1, units, lam, n0, dn=dnc[z0:],
                                   dn_kind=kind, mean_mode="ignore", H=H,
                                   absorbing_width=absorbing_width)
            ctx.save_for_backward(dnc, *ckpts)
            ctx.meta = (units, lam, n0, absorbing_width, segs, H, kind)
            return plane

        @staticmethod
        def backward(ctx, grad_out):
            dnc, *ckpts = ctx.saved_tensors
            units, lam, n0, absorb, segs, H, kind = ctx.meta
            nz, ny, nx = dnc.shape
            Bn = grad_out.shape[0]
            g = grad_out.to(dtype=t.complex64).contiguous().clone()
            grad_dn = t.zeros_like(dnc)
            gitem = t.empty((Bn, ny, nx), dtype=dnc.dtype, device=dnc.device)
            k0dz = float(np.float32(2. * np.pi / lam * units[2]))
            plan = B.get_plan(nx, ny)
            for (z0, z1), ck in zip(reversed(segs), reversed(ckpts)):
                useg = t.empty((Bn, z1 - z0 + 1, ny, nx), dtype=t.complex64, device=dnc.device)
                p = ck.clone()
                B.propagate_fields(p, z1 - z0 + 1, units, lam, n0, dn=dnc[z0:],
                                   dn_kind=kind, mean_mode="ignore", H=H,
                                   absorbing_width=absorb, out_field=useg)
                _lib.call("bpmh_adjoint_segment", plan, int(Bn), int(z0), int(z1), _lib.ptr(g),
                          _lib.ptr(useg), int(useg.numel()), _lib.ptr(dnc), int(kind),
                          int(dnc.numel()), _lib.ptr(H), k0dz, int(absorb), _lib.ptr(grad_dn),
                          _lib.ptr(gitem), _lib.stream_handle())
            return grad_dn, g, None, None, None, None, None

    return _Fn


def _mean_args(units, lam, n0):
    dx, dy, dz = units
    k0 = 2. * np.pi / lam
    f = lambda v: float(np.float32(v))
    return f(n0), f(k0), f(dx), f(dy), f(dz), f(k0 * dz)


def _mean_call(mode, plane, z0, z1, dnc, means, consts, absorb, g, grad_dn, work):
    n0, k0, dx, dy, dz, unit_k = consts
    n_eff, dn0 = means if means is not None else (None, None)
    _lib.call("bpmh_mean_segment", B.get_plan(dnc.shape[2], dnc.shape[1]), int(mode),
              int(plane.shape[0]), int(z0), int(z1), _lib.ptr(plane), _lib.ptr(dnc),
              int(dnc.numel()), _lib.ptr(n_eff), _lib.ptr(dn0), n0, k0, dx, dy, dz, unit_k,
              int(absorb), _lib.ptr(g) if g is not None else None,
              _lib.ptr(grad_dn) if grad_dn is not None else None, _lib.ptr(work),
              int(work.numel()), _lib.stream_handle())


def _work(nx, ny, batch, z0, z1, adjoint):
    t = _lib.torch()
    n = ctypes.c_longlong()
    _lib.call("bpmh_mean_segment_work", B.get_plan(nx, ny), int(batch), int(z0), int(z1),
              int(adjoint), ctypes.byref(n))
    return t.empty(int(n.value), dtype=t.float32, device=_lib.device())


def _simple_means(dnc, n0):
    """Per-plane means of dn and n0 + mean as bpm_3d's simple mode forms them
    (float32 mean over the plane, n0 + mean in float64, both stored as
    float32; bpm/bpm/bpm_3d.py:216-217,288-293)."""
    dn_mean = dnc.float().mean(dim=(1, 2)).cpu().numpy()
    n_eff = np.array([np.float32(float(n0) + float(m)) for m in dn_mean], np.float32)
    return (_lib.to_device(n_eff, np.float32),
            _lib.to_device(np.asarray(dn_mean).astype(np.float32), np.float32))


def _make_mean_fn(mode):
    t = _lib.torch()

    class _MFn(t.autograd.Function):
        @staticmethod
        def forward(ctx, dn, u0, units, lam, n0, absorbing_width, checkpoint):
            nz, ny, nx = dn.shape
            Bn = u0.shape[0]
            consts = _mean_args(units, lam, n0)
            plane = u0.detach().to(dtype=t.complex64).contiguous().clone()
            dnc = dn.detach().contiguous()
            means = _simple_means(dnc, n0) if mode == _lib.MEAN_SIMPLE else None
            segs = _segments(nz, checkpoint)
            work = _work(nx, ny, Bn, 0, 0, False)
            ckpts = []
            for z0, z1 in segs:
                ckpts.append(plane.clone())
                _mean_call(mode, plane, z0, z1, dnc, means, consts, absorbing_width, None, None,
                           work)
            ctx.save_for_backward(dnc, *ckpts)
            ctx.meta = (consts, absorbing_width, segs, means)
            return plane

        @staticmethod
        def backward(ctx, grad_out):
            dnc, *ckpts = ctx.saved_tensors
            consts, absorb, segs, means = ctx.meta
            nz, ny, nx = dnc.shape
            Bn = grad_out.shape[0]
            g = grad_out.to(dtype=t.complex64).contiguous().clone()
            grad_dn = t.zeros_like(dnc)
            longest = max(z1 - z0 for z0, z1 in segs)
            work = _work(nx, ny, Bn, 0, longest, True)
            for (z0, z1), ck in zip(reversed(segs), reversed(ckpts)):
                _mean_call(mode, ck, z0, z1, dnc, means, consts, absorb, g, grad_dn, work)
            return grad_dn, g, None, None, None, None, None

    return _MFn


_FN = None
_MFN = {}


def propagate(dn, u0, units=(.1, .1, .1), lam=.5, n0=1., absorbing_width=0, checkpoint=16,
              mean_mode="ignore"):
    """Exit fields (B, Ny, Nx) complex64 of u0 (B, Ny, Nx) through dn
    (Nz, Ny, Nx), float32 or complex64 (both GPU tensors); differentiable in
    dn and u0.  mean_mode "ignore" is the Bpm3d path (static propagator);
    "weighted" (bpm_3d's default) and "simple" are bpm_3d's mean corrections
    (float32 dn)."""
    global _FN
    t = _lib.torch()
    if dn.dtype not in (t.float32, t.complex64):
        raise ValueError("dn must be float32 or complex64")
    if dn.dim() != 3 or u0.dim() != 3 or tuple(u0.shape[1:]) != tuple(dn.shape[1:]):
        raise ValueError("dn must be (Nz,Ny,Nx) and u0 (B,Ny,Nx)")
    if checkpoint < 1:
        raise ValueError("checkpoint must be >= 1")
    if mean_mode in ("weighted", "simple"):
        if dn.dtype != t.float32:
            raise NotImplementedError("the mean-corrected adjoints take float32 dn")
        mode = _lib.MEAN_WEIGHTED if mean_mode == "weighted" else _lib.MEAN_SIMPLE
        if mode not in _MFN:
            _MFN[mode] = _make_mean_fn(mode)
        return _MFN[mode].apply(dn, u0, tuple(units), lam, n0, int(absorbing_width),
                                int(checkpoint))
    if mean_mode != "ignore":
        raise NotImplementedError("mean_mode must be 'ignore', 'simple' or 'weighted'")
    if _FN is None:
        _FN = _make_fn()
    return _FN.apply(dn, u0, tuple(units), lam, n0, int(absorbing_width), int(checkpoint))

