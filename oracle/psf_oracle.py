"""CPU (numpy/scipy) restatement of the reference vectorial Debye-Wolf focus field.

TEST INFRASTRUCTURE ONLY -- imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py``; never by the product.

``precision="f32"`` follows the float32 operation order of the OpenCL kernels
(bpm/psf/kernels/psf_debye.cl); cos/sin/atan2 are evaluated in float64 on the
float32 argument and rounded; Bessel functions are ``scipy.special.jv`` in
float64 rounded to float32 (the reference's own J_n, bessel.cl:14-115, is a
<=1e-7-class rational approximation whose exact rounding is not reproducible
without the original device).  ``precision="f64"`` evaluates the same
quadrature entirely in float64.

Function -> reference -> what pins it (tests/test_oracle_psf.py)
-----------------------------------------------------------------
debye_wolf_grid   debye_wolf / debye_wolf_plane / debye_wolf_gauss kernels
                  psf_debye.cl:24-113, :117-183, :273-362
                  pinned: (1) the sin-weighted variant against an independent
                  adaptive quadrature (scipy.integrate.quad) of the textbook
                  Richards-Wolf integrals at points (the reference's own
                  check, tests/test_psf_vectorial.py:15-42), to trapezoid
                  accuracy; (2) scalar low-NA limits, Airy J1(v)^2/v^2
                  laterally and sinc^2(u/4) axially (tests/test_psf.py:17-52);
                  (3) the quadrature-rule quirks (S+1 samples, overshoot to
                  alpha2+dt, index vs float-equality end weights) by explicit
                  single-sample brute force; (4) Ex/Ey/Ez symmetry identities
                  (Ey=0 and Ez=0 on axes, Ey antisymmetric in phi).
debye_wolf_points debye_wolf_at  psf_debye.cl:187-269 -- pinned as (1),(3).
mirror_octant     focus_field_debye host mirror  _focus_fields_debye.py:116-143
                  pinned: brute-force |x|,|y|,|z| index map on a tiny grid and
                  the TIFF profile fixtures of the reference
                  (tests/data/psf_NA_0.8_dx_0.1_n_*.tif, central lines).
focus_field_debye _focus_fields_debye.py:48-146 -- composition of the above.
focus_field_debye_gauss  debye_wolf_gauss psf_debye.cl:273-362, host
                  _focus_fields_debye.py:361-441
                  pinned: after removing the quirk samples one by one, equal
                  to scipy quad of the sin-weighted integrand times the pupil
                  exp(-t^2/(2 amax^2 sig^2)) per annulus, with the residual
                  falling as dt^2 (the trapezoid rule's own error); the host
                  entry's lam, arcsin(NAs) and octant coordinates.
_mirror_half_offset  focus_field_debye3 / _gauss mirrors
                  _focus_fields_debye.py:282-304, :425-439
                  pinned: element-by-element index map from the per-axis rule
                  (low half reversed, high half in order, overlap written by
                  the reversed copy; conjugation on the high z half for
                  debye3 only) on even, odd and mixed grids, and for even
                  grids the physical |x| map.
focus_field_debye3  _focus_fields_debye.py:218-310 -- composition of the
                  index quadrature (pinned above) and _mirror_half_offset.
cylindrical       psf_cylindrical[_plane] kernels psf_cylindrical.cl:74-185
                  pinned: trapezoid on [-alpha, alpha+dt] vs scipy quad; the
                  plane variant's float-equality end weights against the
                  index variant plus single-sample corrections (both the
                  case where t_{S-1} lands on alpha and the case where it
                  misses).
"""
import numpy as np
from scipy.special import jv

F32 = np.float32
# bessel.cl:11 ``#define M_PI 3.141592653589793`` is the one in scope of
# psf_debye.cl (bessel.cl is included first); it is a double literal.
PI_D = 3.141592653589793

__all__ = ["debye_wolf_grid", "debye_wolf_points", "mirror_octant",
           "focus_field_debye", "focus_field_debye_plane",
           "focus_field_debye_at", "focus_field_debye_gauss",
           "focus_field_debye3", "bessel_jn", "cylindrical"]


def bessel_jn(n, x):
    """J_n(x) in float64 (scipy), the Bessel oracle (stand-in for bessel.cl:68-115)."""
    return jv(n, np.asarray(x, np.float64))


def _f(x, f64):
    return np.asarray(x, np.float64) if f64 else np.asarray(x, F32)


def _c(x, f64):
    return np.cos(np.asarray(x, np.float64)) if f64 else np.cos(np.asarray(x, np.float64)).astype(F32)


def _s(x, f64):
    return np.sin(np.asarray(x, np.float64)) if f64 else np.sin(np.asarray(x, np.float64)).astype(F32)


def _axis(n, a1, a2, f64):
    """``x1 + i*(x2-x1)/(N-1.f)`` (psf_debye.cl:46-48) in float32."""
    i = np.arange(n)
    if n == 1:  # single sample: the start coordinate (used for one z plane)
        return np.array([a1], np.float64 if f64 else F32)
    if f64:
        return a1 + i * (a2 - a1) / (n - 1.)
    a1 = F32(a1)
    a2 = F32(a2)
    return a1 + (i.astype(F32) * (a2 - a1)) / F32(n - 1.)


def _quadrature(x, y, z, lam, alphas, int_steps, Ex0, Ey0, variant, sig, f64):
    """Core of every debye_wolf* kernel on broadcastable x,y (lateral) and z.

    x, y: arrays of identical shape L (lateral points); z: shape Zs.
    Returns Ex, Ey, Ez complex (Zs + L) and I.
    variant: "index"  -- debye_wolf: end weights by sample index, no sin(t)
             "sin"    -- debye_wolf_plane/_at: end weights by float equality
                         t==alpha1||t==alpha2, weight includes sin(t)
             "gauss"  -- debye_wolf_gauss: as "sin" times exp(t^2 * prefac)
    """
    S = int(int_steps)
    x = _f(x, f64)
    y = _f(y, f64)
    z = _f(z, f64)
    lam = float(lam) if f64 else F32(lam)
    if f64:
        r = np.sqrt(x * x + y * y)
        kr = 2. * PI_D / lam * r
        kz = 2. * PI_D / lam * z
        phi = np.arctan2(y, x)
    else:
        r = np.sqrt(x * x + y * y).astype(F32)  # f32 sqrt of f32 sum
        kr = ((2. * PI_D / np.float64(lam)) * r.astype(np.float64)).astype(F32)
        kz = ((2. * PI_D / np.float64(lam)) * z.astype(np.float64)).astype(F32)
        phi = np.arctan2(y.astype(np.float64), x.astype(np.float64)).astype(F32)
    L = x.shape
    Zs = z.shape
    zdim = (Ellipsis,) + (None,) * len(L)
    rdt = np.float64 if f64 else F32
    I0r = np.zeros(Zs + L, rdt); I0i = np.zeros(Zs + L, rdt)
    I1r = np.zeros(Zs + L, rdt); I1i = np.zeros(Zs + L, rdt)
    I2r = np.zeros(Zs + L, rdt); I2i = np.zeros(Zs + L, rdt)
    alphas = np.asarray(alphas, np.float64 if f64 else F32)
    one = rdt(1.0)
    half = rdt(0.5)
    if variant == "gauss":
        amax = alphas[-1]
        sg = rdt(sig)
        exp_prefac = (((rdt(-.5) / amax) / amax) / sg) / sg
    for i_n in range(len(alphas) // 2):
        a1 = alphas[2 * i_n]
        a2 = alphas[2 * i_n + 1]
        dt = (a2 - a1) / rdt(S - 1.)
        for i_t in range(S + 1):
            t = a1 + rdt(i_t) * dt
            co = _c(t, f64)
            si = _s(t, f64)
            arg = kz * co
            pr = _c(arg, f64)
            pi_ = _s(arg, f64)
            if variant == "index":
                w = half if (i_t == 0 or i_t == S) else one
                sq = np.sqrt(co) if f64 else F32(np.sqrt(co))
                w = w * (dt * sq)
            else:
                w = half if (t == a1 or t == a2) else one
                sq = np.sqrt(co) if f64 else F32(np.sqrt(co))
                rhs = (dt * sq) * si
                if variant == "gauss":
                    e = np.exp(np.float64((t * t) * exp_prefac))
                    rhs = rhs * (e if f64 else F32(e))
                w = w * rhs
            ks = kr * si
            J0 = bessel_jn(0, ks)
            J1 = bessel_jn(1, ks)
            J2 = bessel_jn(2, ks)
            if not f64:
                J0 = J0.astype(F32); J1 = J1.astype(F32); J2 = J2.astype(F32)
            r0 = (w * (co + one)) * J0
            r1 = (w * si) * J1
            r2 = (w * (co - one)) * J2
            prz = pr[zdim]
            piz = pi_[zdim]
            I0r += r0 * prz; I0i += r0 * piz
            I1r += r1 * prz; I1i += r1 * piz
            I2r += r2 * prz; I2i += r2 * piz
    c2 = _c(rdt(2.) * phi, f64)
    s2 = _s(rdt(2.) * phi, f64)
    cp = _c(phi, f64)
    sp = _s(phi, f64)
    Ex0 = rdt(Ex0); Ey0 = rdt(Ey0)
    # ex = Ex0*(I0 + I2 cos2phi) + (Ey0 sin2phi) I2          (psf_debye.cl:88-90)
    exr = Ex0 * (I0r + I2r * c2) + (Ey0 * s2) * I2r
    exi = Ex0 * (I0i + I2i * c2) + (Ey0 * s2) * I2i
    # ey = Ey0*(I0 - I2 cos2phi) + (Ex0 sin2phi) I2          (:92-94)
    eyr = Ey0 * (I0r - I2r * c2) + (Ex0 * s2) * I2r
    eyi = Ey0 * (I0i - I2i * c2) + (Ex0 * s2) * I2i
    # ez = (0,-2)*I1 * (Ex0 cos phi + Ey0 sin phi)           (:96-97)
    m = Ex0 * cp + Ey0 * sp
    ezr = (rdt(2.) * I1i) * m
    ezi = (rdt(-2.) * I1r) * m
    cdt = np.complex128 if f64 else np.complex64
    ex = (exr + 1j * exi).astype(cdt)
    ey = (eyr + 1j * eyi).astype(cdt)
    ez = (ezr + 1j * ezi).astype(cdt)
    if f64:
        vx2 = exr ** 2 + exi ** 2; vy2 = eyr ** 2 + eyi ** 2; vz2 = ezr ** 2 + ezi ** 2
    else:
        vx = np.abs(ex); vy = np.abs(ey); vz = np.abs(ez)
        vx2 = vx * vx; vy2 = vy * vy; vz2 = vz * vz
    I = (vx2 + vy2 + vz2).astype(rdt)
    return ex, ey, ez, I, vx2.astype(rdt)


def debye_wolf_grid(Nx, Ny, Nz, x1, x2, y1, y2, z1, z2, lam, alphas, int_steps,
                    Ex0=1., Ey0=0., variant="index", sig=None, precision="f32"):
    """Fields on the regular grid of the debye_wolf (variant "index") or
    debye_wolf_gauss ("gauss") kernels; returns (Ex, Ey, Ez, I), each (Nz,Ny,Nx)."""
    f64 = precision == "f64"
    xs = _axis(Nx, x1, x2, f64)
    ys = _axis(Ny, y1, y2, f64)
    zs = _axis(Nz, z1, z2, f64)
    Y, X = np.meshgrid(ys, xs, indexing="ij")
    ex, ey, ez, I, _ = _quadrature(X, Y, zs, lam, alphas, int_steps, Ex0, Ey0,
                                   variant, sig, f64)
    return ex, ey, ez, I


def debye_wolf_points(x, y, z, lam, alphas, int_steps, Ex0=1., Ey0=0., precision="f32"):
    """debye_wolf_at (psf_debye.cl:187-269): per-point fields, I = |Ex|^2 only (:267)."""
    f64 = precision == "f64"
    x = np.asarray(x).ravel(); y = np.asarray(y).ravel(); z = np.asarray(z).ravel()
    out = [[], [], [], []]
    for k in range(x.size):
        ex, ey, ez, _, vx2 = _quadrature(np.array([x[k]]), np.array([y[k]]), np.array([z[k]]),
                                         lam, alphas, int_steps, Ex0, Ey0, "sin", None, f64)
        for o, v in zip(out, (ex, ey, ez, vx2)):
            o.append(v[0, 0])
    return tuple(np.array(o) for o in out)


def mirror_octant(u, ex, ey, ez, shape):
    """Spread the computed octant to the full grid (_focus_fields_debye.py:116-143).

    Full index f along each axis maps to octant index |f - N0/2|; the field
    is conjugated for f_z <= Nz0/2 (z <= 0) and copied otherwise.
    """
    Nx0, Ny0, Nz0 = shape
    ix = np.abs(np.arange(Nx0) - Nx0 // 2)
    iy = np.abs(np.arange(Ny0) - Ny0 // 2)
    iz = np.abs(np.arange(Nz0) - Nz0 // 2)
    sel = np.ix_(iz, iy, ix)
    conj = (np.arange(Nz0) <= Nz0 // 2)[:, None, None]
    outs = [u[sel]]
    for f in (ex, ey, ez):
        v = f[sel]
        outs.append(np.where(conj, np.conjugate(v), v).astype(f.dtype))
    return tuple(outs)


def _alphas(NA, n0):
    if np.isscalar(NA):
        NA = [0., NA]
    a = np.arcsin(np.array(NA) / n0)
    assert len(a) % 2 == 0
    return a


def focus_field_debye(shape, units, lam, NA, n0=1., n_integration_steps=200, precision="f32"):
    """_focus_fields_debye.py:48-146: octant via debye_wolf, then mirror."""
    Nx0, Ny0, Nz0 = shape
    dx, dy, dz = units
    if not Nx0 % 2 + Ny0 % 2 + Nz0 % 2 == 0:
        raise NotImplementedError("odd shapes not supported yet")
    alphas = _alphas(NA, n0)
    Nx, Ny, Nz = Nx0 // 2 + 1, Ny0 // 2 + 1, Nz0 // 2 + 1
    ex, ey, ez, u = debye_wolf_grid(Nx, Ny, Nz, 0., dx * (Nx - 1.), 0., dy * (Ny - 1.),
                                    0., dz * (Nz - 1.), 1. * lam / n0, alphas,
                                    n_integration_steps, 1., 0., "index", None, precision)
    return mirror_octant(u, ex, ey, ez, shape)


def focus_field_debye_plane(shape=(128, 128), units=(.1, .1), z=0., lam=.5, NA=.6, n0=1.,
                            n_integration_steps=200, precision="f32"):
    """_focus_fields_debye.py:149-213 / debye_wolf_plane: Ex on one z plane (Ny,Nx)."""
    Nx, Ny = shape
    dx, dy = units
    alphas = _alphas(NA, n0)
    ex, _, _, _ = debye_wolf_grid(Nx, Ny, 2, -(Nx // 2) * dx, (Nx - Nx // 2) * dx,
                                  -(Ny // 2) * dy, (Ny - Ny // 2) * dy, z, z,
                                  lam / n0, alphas, n_integration_steps, 1., 0., "sin",
                                  None, precision)
    return ex[0]


def focus_field_debye_at(x, y, z, lam, NA, n0=1., n_integration_steps=200, precision="f32"):
    """_focus_fields_debye.py:315-358: fields at arbitrary points, I=|Ex|^2."""
    alphas = _alphas(NA, n0)
    shp = np.asarray(x).shape
    ex, ey, ez, u = debye_wolf_points(x, y, z, lam / n0, alphas, n_integration_steps,
                                      precision=precision)
    return u.reshape(shp), ex.reshape(shp), ey.reshape(shp), ez.reshape(shp)


def focus_field_debye3(shape, units, lam, NA, n0=1., n_integration_steps=200, precision="f32"):
    """_focus_fields_debye.py:218-310: (N+1)/2 octant at half-voxel offsets.

    Mirror: octant reversed into the low half (not conjugated for the
    ``i==1`` z-half), conjugated in the high z half, as at :292-304."""
    Nx0, Ny0, Nz0 = shape
    dx, dy, dz = units
    alphas = _alphas(NA, n0)
    Nx, Ny, Nz = (Nx0 + 1) // 2, (Ny0 + 1) // 2, (Nz0 + 1) // 2
    ex, ey, ez, u = debye_wolf_grid(Nx, Ny, Nz, .5 * dx, dx * .5 * (Nx0 - 1), .5 * dy,
                                    dy * .5 * (Ny0 - 1), .5 * dz, dz * .5 * (Nz0 - 1),
                                    lam / n0, alphas, n_integration_steps, 1., 0., "index",
                                    None, precision)
    return _mirror_half_offset(u, ex, ey, ez, shape, conj_high_z=True)


def focus_field_debye_gauss(shape, units, lam, NAs, sig=1. / np.sqrt(2),
                            n_integration_steps=200, precision="f32"):
    """_focus_fields_debye.py:361-441: Gaussian pupil; lam (not lam/n0) and
    arcsin(NAs) without /n0 are passed, as the reference does (:389, :410)."""
    Nx0, Ny0, Nz0 = shape
    dx, dy, dz = units
    alphas = np.arcsin(np.array(NAs))
    Nx, Ny, Nz = (Nx0 + 1) // 2, (Ny0 + 1) // 2, (Nz0 + 1) // 2
    ex, ey, ez, u = debye_wolf_grid(Nx, Ny, Nz, 0., dx * Nx, 0., dy * Ny, 0., dz * Nz,
                                    lam, alphas, n_integration_steps, 1., 0., "gauss",
                                    sig, precision)
    return _mirror_half_offset(u, ex, ey, ez, shape, conj_high_z=None)


def _mirror_half_offset(u, ex, ey, ez, shape, conj_high_z):
    """Mirror used by focus_field_debye3/_gauss: out[s[1-i],s[1-j],s[1-k]] =
    oct[::(-1)**i, ...] with s=[slice(0,N),slice(N0-N,N0)], in product order
    (later octants overwrite earlier ones where the slices overlap)."""
    import itertools
    Nx0, Ny0, Nz0 = shape
    Nz, Ny, Nx = u.shape
    sx = [slice(0, Nx), slice(Nx0 - Nx, Nx0)]
    sy = [slice(0, Ny), slice(Ny0 - Ny, Ny0)]
    sz = [slice(0, Nz), slice(Nz0 - Nz, Nz0)]
    outs = [np.empty((Nz0, Ny0, Nx0), a.dtype) for a in (u, ex, ey, ez)]
    for i, j, k in itertools.product([0, 1], [0, 1], [0, 1]):
        sl = (sz[1 - i], sy[1 - j], sx[1 - k])
        rv = (slice(None, None, (-1) ** i), slice(None, None, (-1) ** j), slice(None, None, (-1) ** k))
        outs[0][sl] = u[rv]
        for o, f in zip(outs[1:], (ex, ey, ez)):
            v = f[rv]
            if conj_high_z is True and i == 0:
                v = np.conjugate(v)
            o[sl] = v
    return tuple(outs)


def cylindrical(Ny, Nz, y1, y2, z1, z2, lam, alpha, int_steps, plane_variant=False,
                precision="f32"):
    """psf_cylindrical / psf_cylindrical_plane kernels (bpm/psf/kernels/psf_cylindrical.cl:74-185)
    on an (Nz, Ny) grid; returns ex, ey, ez, I.  M_PI is the float literal of
    that file (:16).  Pinned (tests/test_oracle_psf.py): trapezoid rule on
    [-alpha, alpha + dt] against scipy.integrate.quad of the same integrand."""
    f64 = precision == "f64"
    rdt = np.float64 if f64 else F32
    S = int(int_steps)
    y = _axis(Ny, y1, y2, f64)
    z = _axis(Nz, z1, z2, f64)
    if f64:
        kf = 2. * np.pi / lam
    else:
        kf = (F32(2.) * F32(3.141592653589793)) / F32(lam)
    ky = (kf * y).astype(rdt)[None, :]
    kz = (kf * z).astype(rdt)[:, None]
    alpha = rdt(alpha)
    dt = (rdt(2.) * alpha) / rdt(S - 1.)
    acc = [np.zeros((Nz, Ny), rdt) for _ in range(6)]
    for it in range(S + 1):
        t = -alpha + rdt(it) * dt
        co, si = _c(t, f64), _s(t, f64)
        if plane_variant:
            w = rdt(.5) if (t == -alpha or t == alpha) else rdt(1.)
        else:
            w = rdt(.5) if (it == 0 or it == S) else rdt(1.)
        sq = np.sqrt(co) if f64 else F32(np.sqrt(co))
        w = w * (dt * sq)
        arg = ky * si + kz * co
        c, s = _c(arg, f64), _s(arg, f64)
        for k, ww in enumerate((w, w * co, w * si)):
            acc[2 * k] += ww * c
            acc[2 * k + 1] += ww * s
    cdt = np.complex128 if f64 else np.complex64
    ex, ey, ez = ((acc[2 * k] + 1j * acc[2 * k + 1]).astype(cdt) for k in range(3))
    I = (acc[0] ** 2 + acc[1] ** 2 + acc[2] ** 2 + acc[3] ** 2 + acc[4] ** 2 + acc[5] ** 2).astype(rdt)
    return ex, ey, ez, I
