"""CPU (numpy) restatement of the reference split-step beam propagation.

TEST INFRASTRUCTURE ONLY -- imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py``; never by the product.

Two precisions share one code path:

* ``precision="c64"`` restates the reference float32 OpenCL path operation by
  operation (float32 scalars, complex64 planes, numpy's single-precision FFT).
  Transcendentals (cos/sin) are evaluated in float64 on the float32 argument
  and rounded to float32 (the reference's OpenCL ``cos``/``sin`` are <= 4 ulp;
  their exact ulp behaviour is not reproducible without the original device).
* ``precision="f64"`` is the float64 split-step: the same algorithm with every
  quantity in float64 / complex128.  It is the accuracy yardstick.

Function -> reference -> what pins it (tests/test_oracle_bpm.py unless noted)
------------------------------------------------------------------------------
fftfreq_f32        FFTFREQ macro  bpm/bpm/kernels/bpm_3d_kernels.cl:6
                   pinned: equals numpy.fft.fftfreq to float32 rounding for
                   even and odd N.
propagator_f32     compute_propagator / mult_propagator[_buff]
                   bpm_3d_kernels.cl:408-493
                   pinned: closed form of an on-grid plane wave (exact phase
                   exp(-i z kz)); evanescent cut-off; agreement with
                   propagator_f64 to float32 rounding.
propagator_f64     bpm/bpm/bpm_3d.py:163-181, bpm/bpmclass/bpm3d.py:88-107
                   pinned: same closed forms.
taper_f32          absorb term of mult_dn*  bpm_3d_kernels.cl:27-31
                   pinned: independent sin^2 form of the Hann ramp, brute force
                   distance-to-edge on small grids.
screen             mult_dn / mult_dn_half / mult_dn_complex /
                   mult_dn_mean[_half][_complex][_buff]  bpm_3d_kernels.cl:20-283
                   pinned: homogeneous-medium closed form (phase k0 dz c per
                   step), absorbing complex dn closed form (decay
                   exp(-k0 dz Im dn) per step), dn[i+1] indexing via a
                   z-varying homogeneous medium.
bpm_3d             _bpm_3d  bpm/bpm/bpm_3d.py:104-423
                   pinned: plane-wave closed forms in all three mean modes,
                   weighted-mean k-space/slice-0 quirk via a brute-force case,
                   Parseval energy conservation, paraxial Gaussian beam,
                   field==intensity identity (reference tests/test_field_intens.py:29-43),
                   scattering reduction vs a brute-force DFT on a tiny grid,
                   c64 vs f64 agreement.
bpm_3d_split       _bpm_3d_split  bpm/bpm/bpm_3d.py:863-992
                   pinned: equals one unsplit run in "ignore" mode (chunks share
                   boundary planes), chunk-local slice-0 dn0 in weighted mode.
propagate_single   _Bpm3d_OCL._propagate_single  bpm/bpmclass/_bpm3d_ocl.py:66-134
                   pinned: same plane-wave / energy closed forms; equals
                   bpm_3d(mean_mode="ignore") up to the H variant; its streamed
                   return_sum equals the z-sum of the intensity volume.
fill_dn_coated     fill_dn  bpm/bpm/kernels/bpm_3d_spheres.cl:31-74
                   pinned: brute-force nearest centre over all points.
"""
import numpy as np

F32 = np.float32
# ``#define M_PI 3.14159265358979f`` (bpm_3d_kernels.cl:3)
PI_F32 = F32(3.14159265358979)

__all__ = ["fftfreq_f32", "propagator_f32", "propagator_f64", "taper_f32",
           "screen", "bpm_3d", "bpm_3d_split", "propagate_single",
           "intensity_f32", "rel_l2", "fill_dn_coated"]


def rel_l2(a, b):
    """Relative L2 distance ||a-b|| / ||b|| (float64 accumulation)."""
    a = np.asarray(a).astype(np.complex128)
    b = np.asarray(b).astype(np.complex128)
    nb = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (nb if nb > 0 else 1.0))


def _cos32(x):
    return np.cos(np.asarray(x, np.float64)).astype(F32)


def _sin32(x):
    return np.sin(np.asarray(x, np.float64)).astype(F32)


def fftfreq_f32(N, d):
    """FFTFREQ(i,N,dx) of bpm_3d_kernels.cl:6 in float32 for i = 0..N-1."""
    i = np.arange(N)
    half = (N - 1) // 2 + 1
    num = np.where(i < half, i, (i - half) - N // 2).astype(F32)
    return (num / F32(N)) / F32(d)


def propagator_f32(Nx, Ny, dx, dy, dz, n0, k0):
    """Device float32 propagator H[ky,kx] (compute_propagator, .cl:408-432).

    ``n0`` is the (possibly mean-corrected) index as the float32 kernel
    argument; mult_propagator (.cl:437-462) passes ``n0+dn0`` and
    mult_propagator_buff (.cl:465-493) forms ``n00 + dn_sum/u_sum`` on device.
    """
    n0 = F32(n0)
    k0 = F32(k0)
    dz = F32(dz)
    twopi = F32(2.0) * PI_F32
    kx = twopi * fftfreq_f32(Nx, dx)
    ky = twopi * fftfreq_f32(Ny, dy)
    a = ((n0 * n0) * k0) * k0
    tmp = (a - (kx * kx)[None, :]) - (ky * ky)[:, None]
    h0 = np.where(tmp < 0, F32(0), np.sqrt(np.maximum(tmp, F32(0))))
    ph = (-dz) * h0
    ok = (tmp >= 0).astype(F32)
    H = np.empty((Ny, Nx), np.complex64)
    H.real = ok * _cos32(ph)
    H.imag = ok * _sin32(ph)
    return H


def _kgrid_f64(Nx, Ny, dx, dy):
    kxs = 2. * np.pi * np.fft.fftfreq(Nx, dx)
    kys = 2. * np.pi * np.fft.fftfreq(Ny, dy)
    KY, KX = np.meshgrid(kys, kxs, indexing="ij")
    return KX, KY


def propagator_f64(Nx, Ny, dx, dy, dz, n0, lam, use_fresnel_approx=False):
    """Host float64 propagator (bpm_3d.py:163-181; bpmclass/bpm3d.py:88-107).

    Returns (H, H0) as complex128 / float64 (or complex128 H0 with Fresnel),
    laid out (Ny, Nx).  H is zero on the evanescent part.
    """
    k0 = 2. * np.pi / lam
    KX, KY = _kgrid_f64(Nx, Ny, dx, dy)
    with np.errstate(invalid="ignore"):
        H0 = np.sqrt(n0 ** 2 * k0 ** 2 - KX ** 2 - KY ** 2)
    if use_fresnel_approx:
        H0 = 0.j + n0 * k0 - .5 * (KX ** 2 + KY ** 2) / n0 / k0
    outside = np.isnan(H0)
    H = np.exp(-1.j * dz * H0)
    H[outside] = 0.
    H0 = np.array(H0)
    H0[outside] = 0.
    return H, H0


def propagator_exact_f64(Nx, Ny, dx, dy, dz, n):
    """float64 H for index n (the f64 oracle's propagator): exp(-i dz sqrt(n^2k0^2-k^2)), 0 if evanescent.

    ``n`` here already includes k0: it is the wavenumber n*k0."""
    KX, KY = _kgrid_f64(Nx, Ny, dx, dy)
    tmp = n * n - KX ** 2 - KY ** 2
    h0 = np.sqrt(np.maximum(tmp, 0.))
    return np.where(tmp >= 0, np.exp(-1.j * dz * h0), 0.)


def taper_f32(Nx, Ny, absorb):
    """Absorbing boundary factor of mult_dn* (bpm_3d_kernels.cl:27-31), float32 (Ny,Nx)."""
    i = np.arange(Nx)
    j = np.arange(Ny)
    distx = np.minimum(Nx - i - 1, i)
    disty = np.minimum(Ny - j - 1, j)
    dist = np.minimum(distx[None, :], disty[:, None])
    if absorb <= 0:
        return np.ones((Ny, Nx), F32)
    arg = (PI_F32 * dist.astype(F32)) / F32(absorb)
    val = F32(0.5) * (F32(1.0) - _cos32(arg))
    return np.where(dist < absorb, val, F32(1.0)).astype(F32)


def _cmul(a, b):
    """complex64 product with the float32 component formula of cfloat_mul."""
    out = np.empty(np.broadcast(a, b).shape, np.complex64)
    out.real = a.real * b.real - a.imag * b.imag
    out.imag = a.real * b.imag + a.imag * b.real
    return out


def screen(plane, dn_slice, unit_k, dn0=0.0, absorb=0, precision="c64"):
    """Phase screen + taper applied after the inverse FFT (mult_dn*, .cl:20-283).

    real dn  : u *= exp(-i unit_k (dn - dn0)) * taper        (.cl:20-49, 198-229)
    complex  : u *= cexp(-i unit_k (dn - dn0)), no taper     (.cl:84-101, 265-283)
    float16  : dn read with vload_half, otherwise as real     (.cl:52-81)
    """
    if precision == "f64":
        d = np.asarray(dn_slice, np.complex128 if np.iscomplexobj(dn_slice) else np.float64) - dn0
        if np.iscomplexobj(d):
            return plane * np.exp(-1.j * unit_k * d)
        ph = -unit_k * d
        out = plane * np.exp(1.j * ph)
        if absorb > 0:
            out = out * taper_f32(plane.shape[1], plane.shape[0], absorb).astype(np.float64)
        return out
    unit_k = F32(unit_k)
    if np.iscomplexobj(dn_slice):
        d = dn_slice.astype(np.complex64) - np.complex64(dn0)
        # cfloat_mul((0,-k), d) = (k*d.im, -k*d.re); cfloat_exp(z) = e^re (cos im, sin im)
        re = unit_k * d.imag
        im = (-unit_k) * d.real
        mag = np.exp(re.astype(np.float64)).astype(F32)
        ph = np.empty(d.shape, np.complex64)
        ph.real = mag * _cos32(im)
        ph.imag = mag * _sin32(im)
        return _cmul(plane, ph)
    dnv = np.asarray(dn_slice).astype(F32)
    dnDiff = (-unit_k) * (dnv - F32(dn0))
    ph = np.empty(dnv.shape, np.complex64)
    ph.real = _cos32(dnDiff)
    ph.imag = _sin32(dnDiff)
    res = _cmul(plane, ph)
    T = taper_f32(plane.shape[1], plane.shape[0], absorb)
    out = np.empty_like(res)
    out.real = res.real * T
    out.imag = res.imag * T
    return out


def intensity_f32(plane):
    """|u|^2 as copy_intens (.cl:396-404): hypot, then square, float32."""
    a = np.abs(plane.astype(np.complex64))
    return (a * a).astype(F32)


_FFT_WORKERS = 1


def set_fft_workers(n):
    """Threads for the 2-D FFTs (1: numpy.fft; more: scipy.fft with that many
    workers, the same pocketfft transforms).  Only the CPU-baseline timing
    uses more than one."""
    global _FFT_WORKERS
    _FFT_WORKERS = max(1, int(n))


def _fft2(a):
    if _FFT_WORKERS > 1:
        import scipy.fft
        return scipy.fft.fft2(a, workers=_FFT_WORKERS)
    return np.fft.fft2(a)


def _ifft2(a):
    if _FFT_WORKERS > 1:
        import scipy.fft
        return scipy.fft.ifft2(a, workers=_FFT_WORKERS)
    return np.fft.ifft2(a)


def bpm_3d(size, units, lam=.5, u0=None, dn=None, n0=1.,
           return_scattering=False, return_g=False, return_full=True,
           return_field=True, use_fresnel_approx=False, absorbing_width=0,
           scattering_plane_ind=0, return_last_plane=False,
           store_dn_as_half=False, mean_mode="weighted", precision="c64",
           sum_intensity=False):
    """Restatement of ``_bpm_3d`` (bpm/bpm/bpm_3d.py:104-423).

    Extra keyword ``sum_intensity``: additionally return the float32
    sequential z-sum of |u_z|^2 over all Nz planes (the summed-intensity
    output of the plain-BPM workloads); it is appended last.

    Semantics where the reference kernels are type-inconsistent (complex dn
    with a mean mode, weighted mode with float16 dn): dn0 is real, computed
    from the real part of dn, and subtracted from dn before the complex
    exponential; half dn is converted to float32 first.
    """
    if mean_mode not in ("ignore", "simple", "weighted"):
        raise NotImplementedError("mean_mode  '%s' unknown" % mean_mode)
    f64 = precision == "f64"
    cdt = np.complex128 if f64 else np.complex64
    Nx, Ny, Nz = size
    dx, dy, dz = units
    k0 = 2. * np.pi / lam

    H64, H0 = propagator_f64(Nx, Ny, dx, dy, dz, n0, lam, use_fresnel_approx)
    if f64:
        H_static = propagator_exact_f64(Nx, Ny, dx, dy, dz, n0 * k0)
    else:
        # the device kernel overwrites the host H (bpm_3d.py:190-196)
        H_static = propagator_f32(Nx, Ny, dx, dy, dz, n0, k0)

    if u0 is None:
        u0 = np.ones((Ny, Nx), np.complex64)
    plane = np.asarray(u0).astype(np.complex64).astype(cdt)

    is_complex = dn is not None and np.iscomplexobj(dn)
    if dn is not None:
        if is_complex:
            dn_v = np.asarray(dn).astype(np.complex128 if f64 else np.complex64)
        elif store_dn_as_half and not f64:
            dn_v = np.asarray(dn).astype(np.float16).astype(F32)
        else:
            dn_v = np.asarray(dn).astype(np.float64 if f64 else F32)
        dn_mean = np.mean(dn, axis=(1, 2))
        if np.iscomplexobj(dn_mean):
            dn_mean = dn_mean.real
        dn_flat0 = (dn_v[0].real if is_complex else dn_v[0]).ravel()
    else:
        dn_v = None
        dn_mean = np.zeros(Nz, F32)
        dn_flat0 = np.zeros(Nx * Ny, np.float64 if f64 else F32)

    if return_scattering:
        cos_theta = np.real(H0) / n0 / k0
        sw = cos_theta.astype(np.float64 if f64 else F32).ravel()
        gw = (cos_theta ** 2).astype(np.float64 if f64 else F32).ravel()
        pw = Nx * Ny * np.exp(-1.j * k0 * n0 * (scattering_plane_ind + np.arange(Nz)) * dz)
        pw = pw.astype(cdt)
        scat = np.zeros(Nz, np.float64 if f64 else F32)
        gfac = np.zeros(Nz, np.float64 if f64 else F32)

    def inten(p):
        return (np.abs(p) ** 2) if f64 else intensity_f32(p)

    if return_full:
        u_out = np.empty((Nz, Ny, Nx), cdt if return_field else (np.float64 if f64 else F32))
        u_out[0] = plane if return_field else inten(plane)
    if sum_intensity:
        acc = inten(plane).astype(np.float64 if f64 else F32)

    unit_k = k0 * dz if f64 else F32(k0 * dz)

    for i in range(Nz - 1):
        U = _fft2(plane)
        dn0 = 0.0
        if mean_mode == "simple":
            dn0 = dn_mean[i + 1]
            n_eff = n0 + float(dn0)
            H = (propagator_exact_f64(Nx, Ny, dx, dy, dz, n_eff * k0) if f64 else
                 propagator_f32(Nx, Ny, dx, dy, dz, F32(n_eff), k0))
            if not f64:
                dn0 = F32(dn0)
        elif mean_mode == "weighted":
            a = np.abs(U).ravel()
            if f64:
                dn_sum = np.sum(a * dn_flat0)
                u_sum = np.sum(a)
                dn0 = dn_sum / u_sum
                H = propagator_exact_f64(Nx, Ny, dx, dy, dz, (n0 + dn0) * k0)
            else:
                dn_sum = F32(np.sum((a * dn_flat0.astype(F32)).astype(np.float64)))
                u_sum = F32(np.sum(a.astype(np.float64)))
                dn0 = dn_sum / u_sum
                n_eff = F32(n0) + dn0
                H = propagator_f32(Nx, Ny, dx, dy, dz, n_eff, k0)
        else:
            H = H_static
        U = (U * H) if f64 else _cmul(U, H)

        if return_scattering:
            f = U.ravel().copy()
            f[0] = f[0] - pw[i + 1]
            a = np.abs(f)
            if f64:
                scat[i + 1] = np.sum(sw * a * a)
                gfac[i + 1] = np.sum(gw * a * a)
            else:
                scat[i + 1] = F32(np.sum((sw * a * a).astype(np.float64)))
                gfac[i + 1] = F32(np.sum((gw * a * a).astype(np.float64)))

        plane = _ifft2(U)
        if dn_v is not None:
            plane = screen(plane, dn_v[i + 1], unit_k, dn0, absorbing_width, precision)
        if return_full:
            u_out[i + 1] = plane if return_field else inten(plane)
        if sum_intensity:
            acc = acc + inten(plane)

    if return_full:
        u = u_out
    else:
        u = plane.copy()
        if not return_field:
            u = inten(u)

    if return_scattering:
        prefac = 1. / Nx / Ny * dx * dy
        p = (prefac * scat) if f64 else (F32(prefac) * scat)
        if return_g:
            with np.errstate(invalid="ignore", divide="ignore"):
                g = (prefac * gfac / p) if f64 else ((F32(prefac) * gfac) / p)
            result = (u, p, g)
        else:
            result = (u, p)
    else:
        result = u
    if return_last_plane:
        result = (result + (plane.copy(),)) if isinstance(result, tuple) else (result, plane.copy())
    if sum_intensity:
        result = (result + (acc,)) if isinstance(result, tuple) else (result, acc)
    return result


def bpm_3d_split(size, units, lam=.5, u0=None, dn=None, n_volumes=1, n0=1.,
                 return_scattering=False, return_full=True, return_g=False,
                 return_field=True, absorbing_width=0,
                 use_fresnel_approx=False, store_dn_as_half=False,
                 mean_mode="weighted", precision="c64"):
    """Restatement of ``_bpm_3d_split`` (bpm/bpm/bpm_3d.py:863-992).

    Nz2 = Nz//n_volumes + 1; chunk i spans planes i*Nz2 .. min((i+1)*Nz2, Nz)
    inclusive and hands its last plane on as the next chunk's u0.
    """
    Nx, Ny, Nz = size
    Nz2 = Nz // n_volumes + 1
    if u0 is None:
        u0 = np.ones((Ny, Nx), np.complex64)
    f64 = precision == "f64"
    u_dtype = (np.complex128 if f64 else np.complex64) if return_field else (np.float64 if f64 else F32)
    u = np.empty((Nz, Ny, Nx) if return_full else (Ny, Nx), u_dtype)
    p = np.empty(Nz, np.float64 if f64 else F32)
    g = np.empty(Nz, np.float64 if f64 else F32)
    kw = dict(units=units, lam=lam, n0=n0, return_full=return_full,
              return_field=return_field, return_g=return_g,
              absorbing_width=absorbing_width,
              return_scattering=return_scattering,
              use_fresnel_approx=use_fresnel_approx,
              store_dn_as_half=store_dn_as_half, mean_mode=mean_mode,
              precision=precision)
    for i in range(n_volumes):
        i1, i2 = i * Nz2, int(np.clip((i + 1) * Nz2, 0, Nz))
        if i < n_volumes - 1:
            dn_part = None if dn is None else dn[i1:i2 + 1]
            res = bpm_3d((Nx, Ny, i2 - i1 + 1), u0=u0, dn=dn_part,
                         return_last_plane=True, scattering_plane_ind=Nz2 * i, **kw)
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
            res = bpm_3d((Nx, Ny, i2 - i1), u0=u0, dn=dn_part,
                         scattering_plane_ind=Nz2 * i, **kw)
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


def propagate_single(size, units, lam=.5, n0=1., dn=None, u0=None,
                     return_full=True, return_intensity=False,
                     absorbing_width=0, precision="c64", return_sum=False):
    """Restatement of the class path ``_Bpm3d_OCL._propagate_single``
    (bpm/bpmclass/_bpm3d_ocl.py:66-134) with the fp64-host -> complex64 H of
    bpm/bpmclass/bpm3d.py:88-107 (laid out (Ny,Nx)).

    ``return_intensity`` (the reference's missing ``fill_with_energy``) is
    defined as |u|^2 of the returned planes.  ``return_sum`` (with
    return_full=False) returns (last plane, sum_z |u_z|^2) accumulated plane
    by plane, in float32 for "c64" (as the intensity volume summed in z
    order) and float64 for "f64".
    """
    f64 = precision == "f64"
    Nx, Ny, Nz = size
    dx, dy, dz = units
    k0 = 2. * np.pi / lam
    H, _ = propagator_f64(Nx, Ny, dx, dy, dz, n0, lam)
    if not f64:
        H = H.astype(np.complex64)
    cdt = np.complex128 if f64 else np.complex64
    if u0 is None:
        u0 = np.ones((Ny, Nx), np.complex64)
    plane = np.asarray(u0).astype(np.complex64).astype(cdt)
    if dn is not None:
        if np.iscomplexobj(dn):
            dn_v = np.asarray(dn).astype(np.complex128 if f64 else np.complex64)
        else:
            dn_v = np.asarray(dn).astype(np.float64 if f64 else F32)
    unit_k = k0 * dz if f64 else F32(k0 * dz)

    def inten(p):
        return (np.abs(p) ** 2) if f64 else intensity_f32(p)

    if return_full:
        out = np.empty((Nz, Ny, Nx), (np.float64 if f64 else F32) if return_intensity else cdt)
        out[0] = inten(plane) if return_intensity else plane
    acc = inten(plane) if return_sum else None
    for i in range(Nz - 1):
        U = _fft2(plane)
        U = (U * H) if f64 else _cmul(U, H)
        plane = _ifft2(U)
        if dn is not None:
            plane = screen(plane, dn_v[i + 1], unit_k, 0.0, absorbing_width, precision)
        if return_full:
            out[i + 1] = inten(plane) if return_intensity else plane
        if return_sum:
            acc = acc + inten(plane)
    if return_full:
        return out
    last = inten(plane) if return_intensity else plane
    return (last, acc) if return_sum else last


def fill_dn_coated(size, units, points, dn_inner=.0, rad_inner=0., dn_outer=.1, rad_outer=.4):
    """Coated-sphere volume of fill_dn (bpm/bpm/kernels/bpm_3d_spheres.cl:31-74)
    with create_dn_buffer's z-sort (bpm/bpm/bpm_3d_spheres.py:24-26), float32.

    Pinned (tests/test_oracle_bpm.py): equals the brute-force nearest-centre
    rule over all points (the z window only drops centres farther than rad2)."""
    Nx, Ny, Nz = size
    dx, dy, dz = (F32(u) for u in units)
    ps = np.array(points, np.float64).reshape(-1, 3)
    ps = ps[np.argsort(ps[:, 2], kind="stable")].astype(F32)
    r2, r1 = F32(rad_outer), F32(rad_inner)
    x = dx * np.arange(Nx).astype(F32)
    y = dy * np.arange(Ny).astype(F32)
    out = np.zeros((Nz, Ny, Nx), F32)
    Np = ps.shape[0]
    for k in range(Nz):
        z = dz * F32(k)
        ind1 = 0
        while ps[ind1, 2] < z - r2 and ind1 < Np - 1:
            ind1 += 1
        ind2 = Np - 1
        while ps[ind2, 2] > z + r2 and ind2 > 0:
            ind2 -= 1
        dmin = np.full((Ny, Nx), F32(1000000000.) * r2, F32)
        for n in range(ind1, ind2 + 1):
            ex = x[None, :] - ps[n, 0]
            ey = y[:, None] - ps[n, 1]
            ez = z - ps[n, 2]
            d = np.sqrt((ex * ex + ey * ey) + ez * ez).astype(F32)
            dmin = np.minimum(dmin, d)
        out[k] = np.where(dmin < r2, np.where(dmin < r1, F32(dn_inner), F32(dn_outer)), F32(0))
    return out
