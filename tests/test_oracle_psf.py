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
