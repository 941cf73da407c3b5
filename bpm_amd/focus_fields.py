"""Vectorial Debye-Wolf focus fields with the signatures of ``bpm.psf``.

Host logic restated from bpm/psf/psf_functions.py and
bpm/psf/_focus_fields_debye.py; the theta quadrature (bpmh_debye_wolf*) and
the octant-to-grid mirror (bpmh_debye_mirror) run in libbpmhip.so.

Reference: M. R. Foreman, P. Toeroek, Computational methods in vectorial
imaging, J. Mod. Opt. 58(5-6), 339 (2011).
"""
import numpy as np

from . import _lib

__all__ = ["psf", "psf_u0", "psf_lightsheet", "focus_field_debye", "focus_field_debye_plane",
           "focus_field_debye_at", "focus_field_debye_gauss", "focus_field_debye3",
           "focus_field_debye_device", "bessel_jn", "psf_cylindrical", "psf_cylindrical_u0",
           "focus_field_cylindrical", "focus_field_cylindrical_plane"]


def _alphas(NA, n0):
    if np.isscalar(NA):
        NA = [0., NA]
    alphas = np.arcsin(np.array(NA) / n0)
    assert len(alphas) % 2 == 0
    return alphas


def _alpha_arg(alphas):
    a = np.ascontiguousarray(np.asarray(alphas, np.float32))
    return a, a.ctypes.data


def _grid_call(Nx, Ny, Nz, x1, x2, y1, y2, z1, z2, lam, alphas, steps, variant, sig=0.,
               outputs=(True, True, True, True), i_mode=0):
    t = _lib.torch()
    ex = _lib.empty((Nz, Ny, Nx), np.complex64) if outputs[0] else None
    ey = _lib.empty((Nz, Ny, Nx), np.complex64) if outputs[1] else None
    ez = _lib.empty((Nz, Ny, Nx), np.complex64) if outputs[2] else None
    u = _lib.empty((Nz, Ny, Nx), np.float32) if outputs[3] else None
    a, ap = _alpha_arg(alphas)
    f = lambda v: float(np.float32(v))
    _lib.call("bpmh_debye_wolf", int(Nx), int(Ny), int(Nz), f(x1), f(x2), f(y1), f(y2), f(z1),
              f(z2), f(lam), ap, len(a), int(steps), 1.0, 0.0, int(variant), f(sig), int(i_mode),
              _lib.ptr(ex), _lib.ptr(ey), _lib.ptr(ez), _lib.ptr(u), _lib.stream_handle())
    return u, ex, ey, ez


def focus_field_debye_device(shape, units, lam, NA, n0=1., n_integration_steps=200):
    """focus_field_debye with the result left on the GPU as torch tensors
    (u, ex, ey, ez).  The four are views of one allocation (Ex, Ey, Ez, then
    I; each 16-byte aligned), which saves three allocator calls per call."""
    Nx0, Ny0, Nz0 = shape
    dx, dy, dz = units
    if not Nx0 % 2 + Ny0 % 2 + Nz0 % 2 == 0:
        raise NotImplementedError("odd shapes not supported yet")
    alphas = _alphas(NA, n0)
    n = Nz0 * Ny0 * Nx0  # even: every axis is
    buf = _lib.empty((3 * n + n // 2,), np.complex64)
    ex, ey, ez = (buf[i * n:(i + 1) * n].view(Nz0, Ny0, Nx0) for i in range(3))
    u = buf[3 * n:].view(_lib.torch().float32).view(Nz0, Ny0, Nx0)
    a, ap = _alpha_arg(alphas)
    # the octant's quadrature with each value written to its mirror positions
    _lib.call("bpmh_focus_field_debye", int(Nx0), int(Ny0), int(Nz0), float(dx), float(dy),
              float(dz), float(np.float32(1. * lam / n0)), ap, len(a), int(n_integration_steps),
              _lib.ptr(ex), _lib.ptr(ey), _lib.ptr(ez), _lib.ptr(u), _lib.stream_handle())
    return u, ex, ey, ez


def focus_field_debye(shape, units, lam, NA, n0=1., n_integration_steps=200):
    """Focus field of an aberration-free objective via the vectorial Debye
    integral (bpm/psf/_focus_fields_debye.py:48-146): returns u, ex, ey, ez
    on the (Nz,Ny,Nx) grid centred at the focus.  NA is a number or an even
    list of annulus limits, e.g. [.1,.2,.5,.6]."""
    return tuple(_lib.to_numpy(v) for v in
                 focus_field_debye_device(shape, units, lam, NA, n0, n_integration_steps))


def focus_field_debye_plane(shape=(128, 128), units=(.1, .1), z=0., lam=.5, NA=.6, n0=1.,
                            ex_g=None, n_integration_steps=200):
    """Ex on the plane z (bpm/psf/_focus_fields_debye.py:149-213, sin-weighted
    variant).  If ex_g is a (Ny,Nx) complex64 GPU tensor it is filled and None
    returned, else a numpy array is returned."""
    Nx, Ny = shape
    dx, dy = units
    alphas = _alphas(NA, n0)
    t = _lib.torch()
    if ex_g is not None:
        if tuple(ex_g.shape) != (Ny, Nx):
            raise AssertionError("ex_g must have shape (Ny,Nx)")
        # the kernel writes Ny*Nx complex64 values through the bare pointer
        if ex_g.dtype != t.complex64 or not ex_g.is_contiguous() or not ex_g.is_cuda:
            raise ValueError("ex_g must be a contiguous complex64 GPU tensor")
    out = ex_g if ex_g is not None else _lib.empty((Ny, Nx), np.complex64)
    a, ap = _alpha_arg(alphas)
    f = lambda v: float(np.float32(v))
    _lib.call("bpmh_debye_wolf", int(Nx), int(Ny), 1, f(-(Nx // 2) * dx), f((Nx - Nx // 2) * dx),
              f(-(Ny // 2) * dy), f((Ny - Ny // 2) * dy), f(z), f(z), f(lam / n0), ap, len(a),
              int(n_integration_steps), 1.0, 0.0, _lib.DW_SIN, 0.0, 0, _lib.ptr(out), None, None,
              None, _lib.stream_handle())
    if ex_g is None:
        return _lib.to_numpy(out)
    return None


def focus_field_debye_at(x, y, z, lam, NA, n0=1., n_integration_steps=200):
    """Fields at arbitrary points (bpm/psf/_focus_fields_debye.py:315-358);
    u is |Ex|^2 as in the reference kernel (psf_debye.cl:267)."""
    alphas = _alphas(NA, n0)
    x = np.asarray(x)
    if not (x.shape == np.shape(y) == np.shape(z)):
        raise AssertionError("x, y, z must have the same shape")
    dshape = x.shape
    N = int(np.prod(dshape))
    xg = _lib.to_device(np.asarray(x).ravel(), np.float32)
    yg = _lib.to_device(np.asarray(y).ravel(), np.float32)
    zg = _lib.to_device(np.asarray(z).ravel(), np.float32)
    u = _lib.empty((N,), np.float32)
    ex = _lib.empty((N,), np.complex64)
    ey = _lib.empty((N,), np.complex64)
    ez = _lib.empty((N,), np.complex64)
    a, ap = _alpha_arg(alphas)
    _lib.call("bpmh_debye_wolf_points", N, _lib.ptr(xg), _lib.ptr(yg), _lib.ptr(zg),
              float(np.float32(lam / n0)), ap, len(a), int(n_integration_steps), 1.0, 0.0,
              _lib.DW_SIN, 0.0, 1, _lib.ptr(ex), _lib.ptr(ey), _lib.ptr(ez), _lib.ptr(u),
              _lib.stream_handle())
    return tuple(_lib.to_numpy(v).reshape(dshape) for v in (u, ex, ey, ez))


def _mirror_half_offset(u, ex, ey, ez, shape, conj_high_z):
    """Host spread used by focus_field_debye3/_gauss (_focus_fields_debye.py:292-304, :430-439)."""
    import itertools
    Nx0, Ny0, Nz0 = shape
    Nz, Ny, Nx = u.shape
    sx = [slice(0, Nx), slice(Nx0 - Nx, Nx0)]
    sy = [slice(0, Ny), slice(Ny0 - Ny, Ny0)]
    sz = [slice(0, Nz), slice(Nz0 - Nz, Nz0)]
    outs = [np.empty((Nz0, Ny0, Nx0), a.dtype) for a in (u, ex, ey, ez)]
    for i, j, k in itertools.product([0, 1], [0, 1], [0, 1]):
        sl = (sz[1 - i], sy[1 - j], sx[1 - k])
        rv = (slice(None, None, (-1) ** i), slice(None, None, (-1) ** j),
              slice(None, None, (-1) ** k))
        outs[0][sl] = u[rv]
        for o, f in zip(outs[1:], (ex, ey, ez)):
            v = f[rv]
            if conj_high_z and i == 0:
                v = np.conjugate(v)
            o[sl] = v
    return tuple(outs)


def focus_field_debye3(shape, units, lam, NA, n0=1., n_integration_steps=200):
    """(N+1)/2 octant at half-voxel offsets (bpm/psf/_focus_fields_debye.py:218-310)."""
    Nx0, Ny0, Nz0 = shape
    dx, dy, dz = units
    alphas = _alphas(NA, n0)
    Nx, Ny, Nz = (Nx0 + 1) // 2, (Ny0 + 1) // 2, (Nz0 + 1) // 2
    res = _grid_call(Nx, Ny, Nz, .5 * dx, dx * .5 * (Nx0 - 1), .5 * dy, dy * .5 * (Ny0 - 1),
                     .5 * dz, dz * .5 * (Nz0 - 1), lam / n0, alphas, n_integration_steps,
                     _lib.DW_INDEX)
    u, ex, ey, ez = (_lib.to_numpy(v) for v in res)
    return _mirror_half_offset(u, ex, ey, ez, shape, conj_high_z=True)


def focus_field_debye_gauss(shape, units, lam, NAs, sig=1. / np.sqrt(2), n_integration_steps=200):
    """Gaussian-apodised pupil (bpm/psf/_focus_fields_debye.py:361-441); as in
    the reference, lam (not lam/n0) and arcsin(NAs) are used."""
    assert sig > 0
    Nx0, Ny0, Nz0 = shape
    dx, dy, dz = units
    alphas = np.arcsin(np.array(NAs))
    Nx, Ny, Nz = (Nx0 + 1) // 2, (Ny0 + 1) // 2, (Nz0 + 1) // 2
    res = _grid_call(Nx, Ny, Nz, 0., dx * Nx, 0., dy * Ny, 0., dz * Nz, lam, alphas,
                     n_integration_steps, _lib.DW_GAUSS, sig)
    u, ex, ey, ez = (_lib.to_numpy(v) for v in res)
    return _mirror_half_offset(u, ex, ey, ez, shape, conj_high_z=False)


def psf(shape, units, lam, NA, n0=1., n_integration_steps=200, return_field=False):
    """3-D PSF of an aberration-free system via the vectorial Debye integral
    (bpm/psf/psf_functions.py:18-71): u, or (u, ex, ey, ez) if return_field."""
    u, ex, ey, ez = focus_field_debye(shape=shape, units=units, lam=lam, NA=NA, n0=n0,
                                      n_integration_steps=n_integration_steps)
    if return_field:
        return u, ex, ey, ez
    return u


def psf_lightsheet(shape, units, lam_illum, NA_illum, lam_detect, NA_detect, n0=1.,
                   n_integration_steps=200, return_field=False):
    """Product of detection and (transposed) illumination PSFs (psf_functions.py:75-98)."""
    u_detect = psf(shape=shape, units=units, lam=lam_detect, NA=NA_detect, n0=n0,
                   n_integration_steps=n_integration_steps)
    u_illum = psf(shape=shape[::-1], units=units[::-1], lam=lam_illum, NA=NA_illum, n0=n0,
                  n_integration_steps=n_integration_steps)
    u_illum = u_illum.transpose((2, 1, 0))
    return u_detect * u_illum


def psf_u0(shape, units, zfoc=0, NA=.4, lam=.5, n0=1., n_integration_steps=200):
    """Initial plane (Ny,Nx) of a beam focused at zfoc (psf_functions.py:104-120):
    Ex of psf((Nx,Ny,4),(dx,dy,zfoc/2)) at its first plane, z = -zfoc."""
    Nx, Ny = shape
    dx, dy = units
    u, ex, ey, ez = psf((Nx, Ny, 4), (dx, dy, zfoc / 2.), n0=n0, lam=lam, NA=NA,
                        n_integration_steps=n_integration_steps, return_field=True)
    return ex[0, ...]


def bessel_jn(n, x):
    """J_n(x), n in {0,1,2}, of float32 arguments on the GPU with the
    evaluation the Debye tables use (float64 series below 14, float32
    asymptotic expansion above), returned as float32 (the bessel_fill
    self-test of bpm/psf/kernels/bessel.cl:120-129)."""
    xg = _lib.to_device(np.asarray(x, np.float32).ravel(), np.float32)
    out = _lib.empty(xg.shape, np.float32)
    _lib.call("bpmh_bessel_fill", int(n), _lib.ptr(xg), _lib.ptr(out), int(xg.numel()),
              _lib.stream_handle())
    return _lib.to_numpy(out).reshape(np.shape(x))


def _cyl_call(Ny, Nz, y1, y2, z1, z2, lam, alpha, steps, plane_variant, outputs):
    ex = _lib.empty((Nz, Ny), np.complex64) if outputs[0] else None
    ey = _lib.empty((Nz, Ny), np.complex64) if outputs[1] else None
    ez = _lib.empty((Nz, Ny), np.complex64) if outputs[2] else None
    u = _lib.empty((Nz, Ny), np.float32) if outputs[3] else None
    f = lambda v: float(np.float32(v))
    _lib.call("bpmh_cylindrical", int(Ny), int(Nz), f(y1), f(y2), f(z1), f(z2), f(lam), f(alpha),
              int(steps), int(plane_variant), _lib.ptr(ex), _lib.ptr(ey), _lib.ptr(ez),
              _lib.ptr(u), _lib.stream_handle())
    return u, ex, ey, ez


def focus_field_cylindrical(shape, units, lam=.5, NA=.3, n0=1., n_integration_steps=100):
    """Focus field of a cylindrical lens (bpm/psf/_focus_fields_cylindrical.py:32-79;
    C. J. R. Sheppard, Appl. Opt. 52, 538 (2013)): the (y, z) field repeated
    along x; returns u, ex, ey, ez of shape (Nz, Ny, Nx)."""
    Nx, Ny, Nz = shape
    dx, dy, dz = units
    alpha = np.arcsin(NA / n0)
    res = _cyl_call(Ny, Nz, -dy * (Ny - 1) / 2., dy * (Ny - 1) / 2., -dz * (Nz - 1) / 2.,
                    dz * (Nz - 1) / 2., lam / n0, alpha, n_integration_steps, 0,
                    (True, True, True, True))
    return tuple(np.repeat(_lib.to_numpy(v)[..., np.newaxis], Nx, axis=-1) for v in res)


def focus_field_cylindrical_plane(shape=(128, 128), units=(.1, .1), z=0., lam=.5, NA=.6, n0=1.,
                                  ex_g=None, n_integration_steps=200):
    """Ex of a cylindrical lens on the plane z (_focus_fields_cylindrical.py:83-134),
    (Ny, Nx); fills the GPU tensor ex_g if given, else returns numpy."""
    Nx, Ny = shape
    dx, dy = units
    alpha = np.arcsin(NA / n0)
    if ex_g is not None and tuple(ex_g.shape) != (Ny, Nx):
        raise AssertionError("ex_g must have shape (Ny,Nx)")
    _, ex, _, _ = _cyl_call(Ny, 1, -dy * (Ny - 1) / 2., dy * (Ny - 1) / 2., z, z, lam / n0, alpha,
                            n_integration_steps, 1, (True, False, False, False))
    full = ex.reshape(Ny, 1).expand(Ny, Nx)
    if ex_g is not None:
        ex_g.copy_(full)
        return None
    return _lib.to_numpy(full.contiguous())


def psf_cylindrical(shape, units, lam, NA, n0=1., return_field=False, n_integration_steps=100):
    """PSF of a cylindrical lens (bpm/psf/psf_functions.py:122-137): u, or
    (u, ex) if return_field.  (The reference unpacks the four return values of
    focus_field_cylindrical into two names and raises; this returns what its
    docstring and psf_cylindrical_u0 expect.)"""
    u, ex, _, _ = focus_field_cylindrical(shape=shape, units=units, lam=lam, NA=NA, n0=n0,
                                          n_integration_steps=n_integration_steps)
    if return_field:
        return u, ex
    return u


def psf_cylindrical_u0(shape, units, zfoc, lam, NA, n0=1., n_integration_steps=200):
    """Initial plane of a cylindrical-lens beam focused at zfoc
    (psf_functions.py:140-156): conj(Ex) of psf_cylindrical((Nx,Ny,4),
    (dx,dy,2 zfoc/3)) at its first plane."""
    Nx, Ny = shape
    dx, dy = units
    u, ex = psf_cylindrical(shape=(Nx, Ny, 4), units=(dx, dy, 2. * zfoc / 3.), lam=lam, NA=NA,
                            n0=n0, n_integration_steps=n_integration_steps, return_field=True)
    return ex[0, ...].conjugate()
