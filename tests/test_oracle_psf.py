"""Pin the numpy Debye-Wolf oracle (oracle/psf_oracle.py).  CPU only.

* debye_wolf (variant "index": no sin(t), end weights by index, S+1 samples)
  is exactly the trapezoid rule on [alpha1, alpha2 + dt]; it is checked
  against scipy.integrate.quad of the same integrands on that interval.
* the sin-weighted variants are the trapezoid rule on [alpha1, alpha2] plus
  the quirk samples (overshoot at alpha2+dt with weight 1, alpha2 weight 1
  when (S-1)*dt misses alpha2), checked against quad on [alpha1, alpha2]
  after removing the quirk samples computed one by one.
* scalar low-NA limits (Airy laterally, sinc^2 axially; reference
  tests/test_psf.py:17-52), field symmetries, the literal slicing loop of
  the reference mirror, and the reference's PSF fixture profiles.
"""
import itertools
import os

import numpy as np
import pytest
from scipy.integrate import quad
from scipy.special import jv

from oracle import psf_oracle as P

HERE = os.path.dirname(os.path.abspath(__file__))


def integrals(x, y, z, lam, a1, a2, with_sin):
    """Richards-Wolf I0, I1, I2 at a point by adaptive quadrature (float64)."""
    k = 2 * np.pi / lam
    kr = k * np.hypot(x, y)
    kz = k * z

    def f(t, n, part):
        s, c = np.sin(t), np.cos(t)
        w = np.sqrt(c) * (s if with_sin else 1.)
        amp = {0: (c + 1), 1: s, 2: (c - 1)}[n] * jv(n, kr * s) * w
        ph = kz * c
        return amp * (np.cos(ph) if part == 0 else np.sin(ph))

    out = []
    for n in range(3):
        re = quad(f, a1, a2, args=(n, 0), limit=400, epsabs=1e-13, epsrel=1e-12)[0]
        im = quad(f, a1, a2, args=(n, 1), limit=400, epsabs=1e-13, epsrel=1e-12)[0]
        out.append(re + 1j * im)
    return out


def fields(I0, I1, I2, x, y):
    phi = np.arctan2(y, x)
    ex = I0 + I2 * np.cos(2 * phi)
    ey = I2 * np.sin(2 * phi)
    ez = -2j * I1 * np.cos(phi)
    return np.array([ex, ey, ez])


POINTS = [(0.0, 0.0, 0.0), (0.21, 0.0, 0.1), (0.13, 0.17, -0.35), (0.0, 0.4, 0.6), (0.5, -0.3, 0.2)]


@pytest.mark.parametrize("NA,n0", [(0.8, 1.0), (1.1, 1.33)])
def test_debye_wolf_is_trapezoid_on_extended_interval(NA, n0):
    lam = .5 / n0
    S = 300
    a1, a2 = 0.0, np.arcsin(NA / n0)
    dt = (a2 - a1) / (S - 1)
    for (x, y, z) in POINTS:
        ex, ey, ez, _ = P.debye_wolf_grid(1, 1, 1, x, x, y, y, z, z, lam, [a1, a2], S,
                                          variant="index", precision="f64")
        got = np.array([ex[0, 0, 0], ey[0, 0, 0], ez[0, 0, 0]])
        ref = fields(*integrals(x, y, z, lam, a1, a2 + dt, False), x, y)
        scale = np.abs(ref).max() + 1e-3
        assert np.abs(got - ref).max() / scale < 2e-5, (x, y, z)


def test_sin_variant_quirks_and_trapezoid():
    lam, n0, NA, S = .5, 1., .9, 601
    a1, a2 = 0.0, np.arcsin(NA / n0)
    dt = (a2 - a1) / (S - 1)
    for (x, y, z) in POINTS[1:]:
        ex, ey, ez, _ = P.debye_wolf_grid(1, 1, 1, x, x, y, y, z, z, lam / n0, [a1, a2], S,
                                          variant="sin", precision="f64")
        got = np.array([ex[0, 0, 0], ey[0, 0, 0], ez[0, 0, 0]])
        k = 2 * np.pi / (lam / n0)
        kr, kz = k * np.hypot(x, y), k * z

        def sample(t):
            s, c = np.sin(t), np.cos(t)
            w = np.sqrt(c) * s
            e = np.exp(1j * kz * c)
            return [w * (c + 1) * jv(0, kr * s) * e, w * s * jv(1, kr * s) * e,
                    w * (c - 1) * jv(2, kr * s) * e]

        extra = [dt * v for v in sample(a1 + S * dt)]          # overshoot, weight 1
        t_last = a1 + (S - 1) * dt
        if t_last != a2:                                         # float-equality miss
            extra = [e + .5 * dt * v for e, v in zip(extra, sample(t_last))]
        corrected = got - fields(*extra, x, y)
        ref = fields(*integrals(x, y, z, lam / n0, a1, a2, True), x, y)
        assert np.abs(corrected - ref).max() / (np.abs(ref).max() + 1e-3) < 2e-5


def test_scalar_low_na_limits():
    # Airy (2 J1(v)/v)^2 laterally, sinc^2(u/4) axially (tests/test_psf.py:17-52)
    lam, NA, n0 = .5, .12, 1.
    a = [0., np.arcsin(NA)]
    r = np.linspace(0.05, 4.0, 40)
    ex, ey, ez, _ = P.debye_wolf_grid(40, 1, 1, r[0], r[-1], 0, 0, 0, 0, lam, a, 400, variant="sin",
                                      precision="f64")
    I = (np.abs(ex) ** 2 + np.abs(ey) ** 2 + np.abs(ez) ** 2)[0, 0]
    ex0, _, _, _ = P.debye_wolf_grid(1, 1, 1, 0, 0, 0, 0, 0, 0, lam, a, 400, variant="sin",
                                     precision="f64")
    v = 2 * np.pi * NA / lam * r
    airy = (2 * jv(1, v) / v) ** 2
    assert np.abs(I / np.abs(ex0[0, 0, 0]) ** 2 - airy).max() < 0.01
    z = np.linspace(0.2, 40, 30)
    exz, _, _, _ = P.debye_wolf_grid(1, 1, 30, 0, 0, 0, 0, z[0], z[-1], lam, a, 400, variant="sin",
                                     precision="f64")
    Iz = np.abs(exz[:, 0, 0]) ** 2 / np.abs(ex0[0, 0, 0]) ** 2
    u = 2 * np.pi / lam * NA ** 2 * z
    sinc2 = (np.sin(u / 4) / (u / 4)) ** 2
    assert np.abs(Iz - sinc2).max() < 0.01


def test_field_symmetries():
    lam, n0 = .5, 1.33
    a = P._alphas(1.1, n0)
    ex, ey, ez, I = P.debye_wolf_grid(9, 9, 3, 0., .4, 0., .4, -.2, .2, lam / n0, a, 100,
                                      variant="index", precision="f64")
    # y = 0 row (phi = 0) and x = 0 column (phi = pi/2): sin(2 phi) = 0 -> Ey = 0
    assert np.abs(ey[:, 0, :]).max() < 1e-12 and np.abs(ey[:, :, 0]).max() < 1e-12
    # x = 0 (phi = pi/2): cos(phi) = 0 -> Ez = 0
    assert np.abs(ez[:, :, 0]).max() < 1e-9
    assert np.allclose(I, np.abs(ex) ** 2 + np.abs(ey) ** 2 + np.abs(ez) ** 2)
    # z -> -z conjugates the fields (phase e^{i kz cos t})
    assert np.allclose(ex[0], np.conj(ex[2]), atol=1e-12)


def test_mirror_matches_reference_slicing_loop():
    rng = np.random.default_rng(0)
    shape = (8, 6, 10)
    Nx0, Ny0, Nz0 = shape
    Nx, Ny, Nz = Nx0 // 2 + 1, Ny0 // 2 + 1, Nz0 // 2 + 1
    u = rng.uniform(0, 1, (Nz, Ny, Nx)).astype(np.float32)
    ex, ey, ez = (rng.normal(size=(Nz, Ny, Nx)) + 1j * rng.normal(size=(Nz, Ny, Nx))
                  for _ in range(3))
    got = P.mirror_octant(u, ex, ey, ez, shape)
    # the slicing loop of bpm/psf/_focus_fields_debye.py:124-143
    ref = [np.empty((Nz0, Ny0, Nx0), a.dtype) for a in (u, ex, ey, ez)]
    sx = [slice(0, Nx), slice(Nx, Nx0)]
    sy = [slice(0, Ny), slice(Ny, Ny0)]
    sz = [slice(0, Nz), slice(Nz, Nz0)]
    for i, j, k in itertools.product([0, 1], [0, 1], [0, 1]):
        src = (slice(1 - i, Nz - 1 + i), slice(1 - j, Ny - 1 + j), slice(1 - k, Nx - 1 + k))
        rev = (slice(None, None, (-1) ** i), slice(None, None, (-1) ** j), slice(None, None, (-1) ** k))
        dst = (sz[1 - i], sy[1 - j], sx[1 - k])
        ref[0][dst] = u[src][rev]
        for r, f in zip(ref[1:], (ex, ey, ez)):
            v = f[src][rev]
            r[dst] = v if i == 0 else np.conjugate(v)
    for g, r in zip(got, ref):
        assert np.array_equal(g, r)


def test_reference_psf_fixture_lateral_profile():
    # tests/data/psf_NA_0.8_dx_0.1_n_*.tif central lines (dx = 0.05 um, as the
    # reference test passes); the textbook (sin-weighted) integral matches the
    # lateral profile, the reference's psf (no sin(t)) is broader.
    g = np.load(os.path.join(HERE, "golden", "psf_tiff_profiles.npz"))
    for n0, key in ((1.0, "1p0"), (1.5, "1p5")):
        a = P._alphas(.8, n0)
        N = 33
        res = {}
        for variant in ("sin", "index"):
            ex, ey, ez, u = P.debye_wolf_grid(N, N, 1, 0., .05 * (N - 1), 0., .05 * (N - 1), 0., 0.,
                                              .5 / n0, a, 200, variant=variant, precision="f64")
            line = u[0, 0, :] / u[0, 0, 0]
            res[variant] = np.abs(line[:32] - g["x_" + key][32:64]).max()
        assert res["sin"] < 0.08
        assert res["index"] > res["sin"]


def test_bessel_oracle_known_values():
    assert abs(P.bessel_jn(0, 2.404825557695773)) < 1e-15
    assert abs(P.bessel_jn(1, 3.8317059702075125)) < 1e-15
    assert abs(P.bessel_jn(2, 5.135622301840683)) < 1e-15
    assert abs(P.bessel_jn(0, 1.0) - 0.7651976865579666) < 1e-15
    x = np.linspace(0.1, 50, 200)
    assert np.allclose(P.bessel_jn(0, x) + P.bessel_jn(2, x), 2 / x * P.bessel_jn(1, x), atol=1e-13)


def test_cylindrical_is_trapezoid_on_extended_interval():
    lam, alpha, S = .5, np.arcsin(.6), 800
    dt = 2 * alpha / (S - 1)
    ex, ey, ez, I = P.cylindrical(5, 4, -.6, .6, -1., .8, lam, alpha, S, precision="f64")
    y = np.linspace(-.6, .6, 5)
    z = np.linspace(-1., .8, 4)
    k = 2 * np.pi / lam
    for (j, zz) in enumerate(z):
        for (i, yy) in enumerate(y):
            for m, f in enumerate((lambda t: 1., np.cos, np.sin)):
                re = quad(lambda t: np.sqrt(np.cos(t)) * f(t) * np.cos(k * (yy * np.sin(t) + zz * np.cos(t))),
                          -alpha, alpha + dt, limit=200)[0]
                im = quad(lambda t: np.sqrt(np.cos(t)) * f(t) * np.sin(k * (yy * np.sin(t) + zz * np.cos(t))),
                          -alpha, alpha + dt, limit=200)[0]
                got = (ex, ey, ez)[m][j, i]
                assert abs(got - (re + 1j * im)) < 2e-5
    assert np.allclose(I, np.abs(ex) ** 2 + np.abs(ey) ** 2 + np.abs(ez) ** 2)


def integrals_gauss(x, y, z, lam, a1, a2, amax, sig):
    """I0, I1, I2 of the sin-weighted integrand apodised by the Gaussian
    pupil exp(-t^2 / (2 amax^2 sig^2)) (psf_debye.cl:307-308,327), by
    adaptive quadrature in float64."""
    k = 2 * np.pi / lam
    kr = k * np.hypot(x, y)
    kz = k * z

    def f(t, n, part):
        s, c = np.sin(t), np.cos(t)
        w = np.sqrt(c) * s * np.exp(-t * t / (2. * amax * amax * sig * sig))
        amp = {0: (c + 1), 1: s, 2: (c - 1)}[n] * jv(n, kr * s) * w
        ph = kz * c
        return amp * (np.cos(ph) if part == 0 else np.sin(ph))

    out = []
    for n in range(3):
        re = quad(f, a1, a2, args=(n, 0), limit=400, epsabs=1e-13, epsrel=1e-12)[0]
        im = quad(f, a1, a2, args=(n, 1), limit=400, epsabs=1e-13, epsrel=1e-12)[0]
        out.append(re + 1j * im)
    return out


@pytest.mark.parametrize("NAs,sig", [([0.0, 0.9], 1. / np.sqrt(2)), ([0.2, 0.5, 0.6, 0.85], 0.4)])
def test_gauss_variant_is_apodised_trapezoid(NAs, sig):
    """debye_wolf_gauss (psf_debye.cl:273-362): per annulus the sin-weighted
    quadrature with its quirk samples (overshoot at alpha2 + dt with weight 1;
    alpha2 itself with weight 1 when the float sum misses it), every sample
    times exp(t^2 * (-1/2)/amax/amax/sig/sig), amax the last aperture angle.
    After removing the quirk samples one by one the result must equal quad of
    the apodised integrand on each [alpha1, alpha2]; a dropped pupil factor,
    a wrong amax or sig, or a missing sin t each fail it."""
    lam = .5
    alphas = np.arcsin(np.array(NAs))
    amax = alphas[-1]

    def residual(x, y, z, S):
        ex, ey, ez, _ = P.debye_wolf_grid(1, 1, 1, x, x, y, y, z, z, lam, alphas, S,
                                          variant="gauss", sig=sig, precision="f64")
        got = np.array([ex[0, 0, 0], ey[0, 0, 0], ez[0, 0, 0]])
        k = 2 * np.pi / lam
        kr, kz = k * np.hypot(x, y), k * z

        def sample(t):
            s, c = np.sin(t), np.cos(t)
            w = np.sqrt(c) * s * np.exp(-t * t / (2. * amax * amax * sig * sig))
            e = np.exp(1j * kz * c)
            return [w * (c + 1) * jv(0, kr * s) * e, w * s * jv(1, kr * s) * e,
                    w * (c - 1) * jv(2, kr * s) * e]

        corrected = got.copy()
        ref = np.zeros(3, complex)
        for a1, a2 in zip(alphas[0::2], alphas[1::2]):
            dt = (a2 - a1) / (S - 1)
            extra = [dt * v for v in sample(a1 + S * dt)]          # overshoot, weight 1
            t_last = a1 + (S - 1) * dt
            if t_last != a2:                                         # float-equality miss
                extra = [e + .5 * dt * v for e, v in zip(extra, sample(t_last))]
            corrected = corrected - fields(*extra, x, y)
            ref = ref + fields(*integrals_gauss(x, y, z, lam, a1, a2, amax, sig), x, y)
        return np.abs(corrected - ref).max() / np.abs(ref).max()

    for (x, y, z) in POINTS[1:]:
        e1, e2 = residual(x, y, z, 401), residual(x, y, z, 801)
        assert e2 < 2e-5
        # what is left is the trapezoid rule's own error: it falls as dt^2
        assert 3. < e1 / e2 < 5.
    # and the host entry passes lam (not lam/n0) and arcsin(NAs) (no /n0),
    # on the (N+1)/2 octant at x = i * dx*N/(N-1) (_focus_fields_debye.py:389-413)
    shape, units = (6, 4, 5), (.1, .12, .15)
    u, ex, ey, ez = P.focus_field_debye_gauss(shape, units, lam, NAs, sig=sig,
                                              n_integration_steps=201, precision="f64")
    Nx, Ny, Nz = 3, 2, 3
    oex, _, _, _ = P.debye_wolf_grid(Nx, Ny, Nz, 0., .1 * Nx, 0., .12 * Ny, 0., .15 * Nz, lam,
                                     alphas, 201, variant="gauss", sig=sig, precision="f64")
    # octant sample (k, j, i) lands at full index (Nz0-Nz+k, Ny0-Ny+j, Nx0-Nx+i)
    assert np.array_equal(ex[2:, 2:, 3:], oex)


def _half_offset_source(f, N0, N):
    """Octant index feeding full index f along one axis of the half-offset
    mirror, and whether it comes from the high half: the low half [0, N)
    holds the octant reversed, the high half [N0-N, N0) the octant in order;
    where the halves overlap (odd N0, f = N-1) the reversed copy is the one
    written last."""
    if f < N:
        return N - 1 - f, False
    return f - (N0 - N), True


@pytest.mark.parametrize("shape", [(6, 4, 8), (5, 7, 3), (4, 5, 6), (1, 2, 3)])
@pytest.mark.parametrize("conj_high_z", [True, None])
def test_mirror_half_offset_index_map(shape, conj_high_z):
    """The mirror of focus_field_debye3 / _gauss (_focus_fields_debye.py:
    282-304, 425-439) against an element-by-element index map built from the
    per-axis rule, with every octant value distinct, on even, odd and mixed
    grids; conjugation on the high z half for debye3, never for gauss.  For
    even N0 the rule is the physical |x| map: octant sample m sits at
    (m + 1/2) dx and full index f at (f - (N0-1)/2) dx."""
    Nx0, Ny0, Nz0 = shape
    Nx, Ny, Nz = (Nx0 + 1) // 2, (Ny0 + 1) // 2, (Nz0 + 1) // 2
    n = Nx * Ny * Nz
    u = np.arange(n, dtype=np.float32).reshape(Nz, Ny, Nx)
    ex = (np.arange(n) + 1j * (np.arange(n) + .5)).reshape(Nz, Ny, Nx)
    ey, ez = 2 * ex, -3 * ex
    got = P._mirror_half_offset(u, ex, ey, ez, shape, conj_high_z)
    for fz in range(Nz0):
        kz, high = _half_offset_source(fz, Nz0, Nz)
        for fy in range(Ny0):
            ky, _ = _half_offset_source(fy, Ny0, Ny)
            for fx in range(Nx0):
                kx, _ = _half_offset_source(fx, Nx0, Nx)
                assert got[0][fz, fy, fx] == u[kz, ky, kx]
                for g, f in zip(got[1:], (ex, ey, ez)):
                    v = f[kz, ky, kx]
                    assert g[fz, fy, fx] == (np.conjugate(v) if (high and conj_high_z) else v)
    for N0 in (2, 4, 6, 8):
        N = N0 // 2
        for f in range(N0):
            m, _ = _half_offset_source(f, N0, N)
            assert (m + .5) == abs(f - (N0 - 1) / 2.)


def test_cylindrical_plane_variant_end_weights():
    """psf_cylindrical_plane (psf_cylindrical.cl:140-185) differs from
    psf_cylindrical only in its end weights, which test t == -alpha or
    t == alpha in float32 instead of the sample index: so the plane variant
    equals the index variant plus half a sample at t_S = alpha + dt (full
    weight there instead of half) minus half a sample at t_{S-1} whenever
    that float32 sum lands exactly on alpha.  Each single sample is evaluated
    on its own; S values are picked so that both cases occur."""
    F32 = np.float32
    lam, alpha = .5, F32(np.arcsin(.6))
    y1, y2, z1, z2, Ny, Nz = -.6, .6, -1., .8, 5, 4
    cases = {}
    for S in range(300, 420):
        dt = (F32(2.) * alpha) / F32(S - 1.)
        t_last = F32(-alpha + F32(S - 1) * dt)
        cases.setdefault(bool(t_last == alpha), S)
    assert set(cases) == {True, False}
    kf = (F32(2.) * F32(3.141592653589793)) / F32(lam)
    y = P._axis(Ny, y1, y2, False)
    z = P._axis(Nz, z1, z2, False)
    ky = (kf * y).astype(np.float64)[None, :]
    kz = (kf * z).astype(np.float64)[:, None]
    for hit, S in cases.items():
        dt = (F32(2.) * alpha) / F32(S - 1.)

        def sample(i):
            t = np.float64(F32(-alpha + F32(i) * dt))
            w = float(dt) * np.sqrt(np.cos(t))
            e = np.exp(1j * (ky * np.sin(t) + kz * np.cos(t)))
            return [w * e, w * np.cos(t) * e, w * np.sin(t) * e]

        idx = P.cylindrical(Ny, Nz, y1, y2, z1, z2, lam, alpha, S, plane_variant=False)
        pln = P.cylindrical(Ny, Nz, y1, y2, z1, z2, lam, alpha, S, plane_variant=True)
        extra = [.5 * v for v in sample(S)]
        if hit:
            extra = [e - .5 * v for e, v in zip(extra, sample(S - 1))]
        for m in range(3):
            d = pln[m].astype(np.complex128) - idx[m].astype(np.complex128) - extra[m]
            scale = np.abs(idx[m]).max()
            # float32 accumulation leaves ~1e-7; the corrections are 1e-4..4e-3
            assert np.abs(d).max() < 1e-6 * scale, (hit, S, m)
            assert np.abs(extra[m]).max() > 2e-5 * scale, (hit, S, m)
