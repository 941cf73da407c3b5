"""Pin the numpy BPM oracle (oracle/bpm_oracle.py) against closed forms,
invariants and brute force.  CPU only.

Each check is chosen so that a plausible slip in the restatement (sign of the
propagator or screen, dn[i] vs dn[i+1], slice-0 / k-space weighting of the
weighted mean, taper ramp, normalisation) fails at least one of them.
"""
import numpy as np
import pytest

from oracle import bpm_oracle as O

LAM, N0 = .5, 1.33
K0 = 2 * np.pi / LAM


def plane(Nx, Ny, dx, dy, mx, my):
    """On-grid plane wave exp(-i(kx x + ky y)) and its (kx, ky)."""
    kx = 2 * np.pi * mx / (Nx * dx)
    ky = 2 * np.pi * my / (Ny * dy)
    x = dx * np.arange(Nx)
    y = dy * np.arange(Ny)
    # rounded to complex64 as bpm_3d does with u0 (bpm_3d.py:189)
    u0 = np.exp(-1j * (kx * x[None, :] + ky * y[:, None])).astype(np.complex64)
    return u0.astype(np.complex128), kx, ky


def layered(Nz, Ny, Nx, c):
    return np.broadcast_to(np.asarray(c, np.float64)[:, None, None], (Nz, Ny, Nx)).copy()


# f64: the reference rounds u0 to complex64; that ~3e-8 noise is not a plane wave
TOL = {"c64": 2e-5, "f64": 1e-7}


def test_fftfreq_f32_matches_numpy():
    for N in (16, 17, 64, 255, 256):
        for d in (.1, .07):
            ref = np.fft.fftfreq(N, d)
            got = O.fftfreq_f32(N, d).astype(np.float64)
            assert np.allclose(got, ref, rtol=2e-7, atol=1e-12)


def test_propagator_f32_vs_f64_and_cutoff():
    Nx, Ny, d = 64, 32, (.1, .13, .1)
    Hf = O.propagator_f32(Nx, Ny, d[0], d[1], d[2], N0, K0)
    Hd, H0 = O.propagator_f64(Nx, Ny, d[0], d[1], d[2], N0, LAM)
    ev = H0 == 0
    assert np.all(Hf[ev] == 0) and np.all(Hd[ev] == 0)
    # float32 cancellation in n^2k0^2 - k^2 near the cut-off costs a few ulp of phase
    assert np.abs(Hf - Hd).max() < 1e-5
    assert np.allclose(np.abs(Hf[~ev]), 1, atol=2e-7)


@pytest.mark.parametrize("precision", ["c64", "f64"])
def test_plane_wave_ignore_mode_layered_medium(precision):
    # u_z = u0 exp(-i sum_{j=1..z} (dz kz + k0 dz c_j)): pins H sign, screen
    # sign and the dn[i+1] indexing of the screen.
    Nx, Ny, Nz, dx, dz = 32, 16, 12, .1, .1
    u0, kx, ky = plane(Nx, Ny, dx, dx, 3, -2)
    c = np.linspace(0.0, 0.06, Nz) ** 1.5
    dn = layered(Nz, Ny, Nx, c)
    u = O.bpm_3d((Nx, Ny, Nz), (dx, dx, dz), LAM, u0, dn, N0, mean_mode="ignore",
                 precision=precision)
    kz = np.sqrt(N0 ** 2 * K0 ** 2 - kx ** 2 - ky ** 2)
    ph = np.concatenate([[0], np.cumsum(dz * kz + K0 * dz * c[1:])])
    exact = u0[None] * np.exp(-1j * ph)[:, None, None]
    assert O.rel_l2(u, exact) < TOL[precision]


@pytest.mark.parametrize("precision", ["c64", "f64"])
def test_plane_wave_simple_mode(precision):
    # simple: H(n0 + mean dn[z]) and screen dn - mean = 0
    Nx, Ny, Nz, dx, dz = 32, 32, 10, .1, .1
    u0, kx, ky = plane(Nx, Ny, dx, dx, 1, 4)
    c = np.linspace(0.01, 0.05, Nz)
    dn = layered(Nz, Ny, Nx, c).astype(np.float32)
    u = O.bpm_3d((Nx, Ny, Nz), (dx, dx, dz), LAM, u0, dn, N0, mean_mode="simple",
                 precision=precision)
    cz = dn[:, 0, 0].astype(np.float64)
    kz = np.sqrt((N0 + cz) ** 2 * K0 ** 2 - kx ** 2 - ky ** 2)
    ph = np.concatenate([[0], np.cumsum(dz * kz[1:])])
    exact = u0[None] * np.exp(-1j * ph)[:, None, None]
    assert O.rel_l2(u, exact) < TOL[precision]


@pytest.mark.parametrize("precision", ["c64", "f64"])
def test_weighted_mode_uses_slice0_for_every_step(precision):
    # weighted dn0 = sum|U| dn_flat[slice 0] / sum|U| (bpm_3d.py:149-159,302):
    # with a layered medium dn0 = c_0 at every step, so
    # phase/step = dz sqrt((n0+c0)^2 k0^2 - k^2) + k0 dz (c_j - c0)
    Nx, Ny, Nz, dx, dz = 16, 32, 10, .1, .1
    u0, kx, ky = plane(Nx, Ny, dx, dx, -2, 5)
    c = np.linspace(0.04, 0.0, Nz).astype(np.float32)
    dn = layered(Nz, Ny, Nx, c).astype(np.float32)
    u = O.bpm_3d((Nx, Ny, Nz), (dx, dx, dz), LAM, u0, dn, N0, mean_mode="weighted",
                 precision=precision)
    c = c.astype(np.float64)
    kz = np.sqrt((N0 + c[0]) ** 2 * K0 ** 2 - kx ** 2 - ky ** 2)
    ph = np.concatenate([[0], np.cumsum(dz * kz + K0 * dz * (c[1:] - c[0]))])
    exact = u0[None] * np.exp(-1j * ph)[:, None, None]
    assert O.rel_l2(u, exact) < TOL[precision]


@pytest.mark.parametrize("precision", ["c64", "f64"])
def test_weighted_mode_kspace_index_quirk(precision):
    # a plane wave occupies a single k-space index p*; the weight dn_flat is
    # indexed by that k-space index in slice 0, so dn0 = dn[0].flat[p*].
    Nx, Ny, Nz, dx, dz = 16, 16, 6, .1, .1
    mx, my = 3, 2
    u0, kx, ky = plane(Nx, Ny, dx, dx, mx, my)
    rng = np.random.default_rng(1)
    dn = np.zeros((Nz, Ny, Nx), np.float32)
    dn[0] = rng.uniform(0, .05, (Ny, Nx))
    dn[1:] = .02
    u = O.bpm_3d((Nx, Ny, Nz), (dx, dx, dz), LAM, u0, dn, N0, mean_mode="weighted",
                 precision=precision)
    p = ((-my) % Ny) * Nx + ((-mx) % Nx)
    d0 = float(dn[0].ravel()[p])
    kz = np.sqrt((N0 + d0) ** 2 * K0 ** 2 - kx ** 2 - ky ** 2)
    ph = np.concatenate([[0], np.cumsum(np.full(Nz - 1, dz * kz + K0 * dz * (.02 - d0)))])
    exact = u0[None] * np.exp(-1j * ph)[:, None, None]
    assert O.rel_l2(u, exact) < (2e-4 if precision == "c64" else 1e-7)


def test_taper_matches_hann_ramp():
    # one step of a uniform field through dn = 0: u1 = exp(-i n0 k0 dz) T(x,y)
    # with T = sin^2(pi d / 2w) for d < w (the Hann ramp), d = distance to edge
    Nx, Ny, w = 24, 20, 5
    dn = np.zeros((2, Ny, Nx), np.float32)
    u = O.bpm_3d((Nx, Ny, 2), (.1, .1, .1), LAM, None, dn, N0, mean_mode="ignore",
                 absorbing_width=w)
    T = np.ones((Ny, Nx))
    for j in range(Ny):
        for i in range(Nx):
            d = min(i, j, Nx - 1 - i, Ny - 1 - j)
            if d < w:
                T[j, i] = np.sin(np.pi * d / (2 * w)) ** 2
    exact = np.exp(-1j * N0 * K0 * .1) * T
    assert np.abs(u[1] - exact).max() < 2e-6
    # no taper without dn (bpm_3d.py:345) and no taper for complex dn
    u_none = O.bpm_3d((Nx, Ny, 2), (.1, .1, .1), LAM, None, None, N0, mean_mode="ignore",
                      absorbing_width=w)
    assert np.allclose(np.abs(u_none[1]), 1, atol=1e-6)
    u_c = O.bpm_3d((Nx, Ny, 2), (.1, .1, .1), LAM, None, dn.astype(np.complex64), N0,
                   mean_mode="ignore", absorbing_width=w)
    assert np.allclose(np.abs(u_c[1]), 1, atol=1e-6)


@pytest.mark.parametrize("precision", ["c64", "f64"])
def test_absorbing_complex_medium(precision):
    # dn = c + i b: per step exp(-i k0 dz c) exp(+k0 dz b) (cfloat_exp of
    # (0,-k0dz)*dn); b < 0 absorbs (reference tests/test_complex.py)
    Nx, Ny, Nz, dx, dz = 16, 16, 9, .05, .05
    c, b = .03, -.01
    u0, kx, ky = plane(Nx, Ny, dx, dx, 1, 0)
    dn = np.full((Nz, Ny, Nx), c + 1j * b, np.complex64)
    u = O.bpm_3d((Nx, Ny, Nz), (dx, dx, dz), LAM, u0, dn, N0, mean_mode="ignore",
                 precision=precision)
    kz = np.sqrt(N0 ** 2 * K0 ** 2 - kx ** 2 - ky ** 2)
    z = np.arange(Nz)
    exact = u0[None] * (np.exp(-1j * z * (dz * kz + K0 * dz * c)) * np.exp(K0 * dz * b * z))[:, None, None]
    assert O.rel_l2(u, exact) < TOL[precision]
    # the same through the class path
    v = O.propagate_single((Nx, Ny, Nz), (dx, dx, dz), LAM, N0, dn, u0, precision=precision)
    assert O.rel_l2(v, exact) < TOL[precision]


@pytest.mark.parametrize("precision", ["c64", "f64"])
def test_energy_conserved_on_all_propagating_grid(precision):
    # dx = 0.3 um puts every FFT frequency inside n0 k0, so |H| = 1 and the
    # step is unitary: ||u_z|| is constant (Parseval), for any real dn.
    Nx, Ny, Nz, dx = 32, 32, 40, .3
    rng = np.random.default_rng(3)
    dn = rng.uniform(-.05, .05, (Nz, Ny, Nx)).astype(np.float32)
    u0 = np.exp(1j * rng.uniform(0, 6, (Ny, Nx))).astype(np.complex64)
    u = O.bpm_3d((Nx, Ny, Nz), (dx, dx, .1), LAM, u0, dn, N0, mean_mode="weighted",
                 precision=precision)
    e = np.sum(np.abs(u.astype(np.complex128)) ** 2, axis=(1, 2))
    assert np.abs(e / e[0] - 1).max() < (2e-5 if precision == "c64" else 1e-12)


def test_gaussian_beam_paraxial():
    # free propagation of a paraxial Gaussian beam (reference tests/test_gaussian.py)
    from bpm_amd.workloads import gaussian_beam
    Nx = Ny = 128
    dx, dz, n0, w0 = .2, .2, 1., 2.
    Nz = 60
    z0 = -6.
    u0 = gaussian_beam(Nx, Ny, dx, dx, LAM, n0, w0, z0)
    u = O.bpm_3d((Nx, Ny, Nz), (dx, dx, dz), LAM, u0, None, n0, mean_mode="ignore",
                 precision="f64")
    err = [O.rel_l2(u[k], gaussian_beam(Nx, Ny, dx, dx, LAM, n0, w0, z0 + k * dz))
           for k in (0, Nz // 2, Nz - 1)]
    assert err[0] < 1e-6 and max(err) < 0.02


def test_field_intensity_identity():
    # reference tests/test_field_intens.py:29-43
    Nx, Ny, Nz = 32, 32, 12
    rng = np.random.default_rng(5)
    dn = (.05 * rng.uniform(0, 1, (Nz, Ny, Nx))).astype(np.float32)
    uf = O.bpm_3d((Nx, Ny, Nz), (.1, .1, .1), LAM, None, dn, N0, return_field=True)
    ui = O.bpm_3d((Nx, Ny, Nz), (.1, .1, .1), LAM, None, dn, N0, return_field=False)
    assert np.allclose(np.abs(uf) ** 2, ui, rtol=1e-6, atol=1e-7)


def test_scattering_sums_brute_force_dft():
    Nx, Ny, Nz, dx, dz = 8, 8, 4, .1, .1
    rng = np.random.default_rng(7)
    dn = (.05 * rng.uniform(0, 1, (Nz, Ny, Nx)))
    u0 = np.exp(1j * rng.uniform(0, 1, (Ny, Nx)))
    u, p, g = O.bpm_3d((Nx, Ny, Nz), (dx, dx, dz), LAM, u0, dn, N0, mean_mode="ignore",
                       return_scattering=True, return_g=True, precision="f64")
    # brute-force 2-D DFT matrices
    Wx = np.exp(-2j * np.pi * np.outer(np.arange(Nx), np.arange(Nx)) / Nx)
    Wy = np.exp(-2j * np.pi * np.outer(np.arange(Ny), np.arange(Ny)) / Ny)
    kx = 2 * np.pi * np.fft.fftfreq(Nx, dx)
    k2 = kx[None, :] ** 2 + kx[:, None] ** 2
    h0 = np.sqrt(np.maximum(N0 ** 2 * K0 ** 2 - k2, 0))
    H = np.where(N0 ** 2 * K0 ** 2 - k2 >= 0, np.exp(-1j * dz * h0), 0)
    cos_t = h0 / N0 / K0
    for i in range(Nz - 1):
        U = Wy @ u[i] @ Wx.T * H
        U[0, 0] -= Nx * Ny * np.exp(-1j * K0 * N0 * (i + 1) * dz)
        s = np.sum(cos_t * np.abs(U) ** 2)
        gs = np.sum(cos_t ** 2 * np.abs(U) ** 2)
        pref = dx * dx / Nx / Ny
        assert np.isclose(p[i + 1], pref * s, rtol=1e-9)
        assert np.isclose(g[i + 1], gs / s, rtol=1e-9)
    assert p[0] == 0 and np.isnan(g[0])


@pytest.mark.parametrize("n_volumes", [2, 3])
def test_split_equals_unsplit_in_ignore_mode(n_volumes):
    Nx, Ny, Nz = 32, 32, 17
    rng = np.random.default_rng(9)
    dn = (.05 * rng.uniform(0, 1, (Nz, Ny, Nx))).astype(np.float32)
    a = O.bpm_3d((Nx, Ny, Nz), (.1, .1, .1), LAM, None, dn, N0, mean_mode="ignore")
    b = O.bpm_3d_split((Nx, Ny, Nz), (.1, .1, .1), LAM, None, dn, n_volumes, N0,
                       mean_mode="ignore")
    assert np.array_equal(a, b)


def test_split_weighted_uses_chunk_slice0():
    Nx, Ny, Nz, dx, dz = 16, 16, 10, .1, .1
    u0, kx, ky = plane(Nx, Ny, dx, dx, 2, 1)
    c = np.linspace(0.0, 0.05, Nz).astype(np.float32)
    dn = layered(Nz, Ny, Nx, c).astype(np.float32)
    u = O.bpm_3d_split((Nx, Ny, Nz), (dx, dx, dz), LAM, u0, dn, 2, N0, mean_mode="weighted",
                       precision="f64")
    c = c.astype(np.float64)
    Nz2 = Nz // 2 + 1
    ph = [0.]
    for z in range(1, Nz):
        c0 = c[0] if z <= Nz2 else c[Nz2]
        kz = np.sqrt((N0 + c0) ** 2 * K0 ** 2 - kx ** 2 - ky ** 2)
        ph.append(ph[-1] + dz * kz + K0 * dz * (c[z] - c0))
    exact = u0[None] * np.exp(-1j * np.array(ph))[:, None, None]
    assert O.rel_l2(u, exact) < 1e-7


def test_c64_tracks_f64_in_random_medium():
    from bpm_amd.workloads import plain_small_workload
    u0, dn = plain_small_workload(seed=0, size=(64, 64, 64))
    for mode in ("ignore", "weighted"):
        a = O.bpm_3d((64, 64, 64), (.1, .1, .1), LAM, u0, dn, N0, mean_mode=mode,
                     return_full=False)
        b = O.bpm_3d((64, 64, 64), (.1, .1, .1), LAM, u0, dn, N0, mean_mode=mode,
                     return_full=False, precision="f64")
        assert O.rel_l2(a, b) < 2e-5


def test_class_path_close_to_functional_ignore():
    from bpm_amd.workloads import plain_small_workload
    u0, dn = plain_small_workload(seed=1, size=(32, 32, 40))
    a = O.propagate_single((32, 32, 40), (.1, .1, .1), LAM, N0, dn, u0)
    b = O.bpm_3d((32, 32, 40), (.1, .1, .1), LAM, u0, dn, N0, mean_mode="ignore")
    assert O.rel_l2(a, b) < 2e-5


def test_sum_intensity_is_z_sum_of_intensity_volume():
    Nx, Ny, Nz = 16, 16, 9
    rng = np.random.default_rng(2)
    dn = (.05 * rng.uniform(0, 1, (Nz, Ny, Nx))).astype(np.float32)
    ui, acc = O.bpm_3d((Nx, Ny, Nz), (.1, .1, .1), LAM, None, dn, N0, return_field=False,
                       sum_intensity=True)
    assert np.allclose(acc, ui.astype(np.float64).sum(0), rtol=1e-5)


def test_fill_dn_coated_equals_brute_force_nearest_centre():
    rng = np.random.default_rng(4)
    size, units = (24, 20, 16), (.1, .12, .15)
    pts = rng.uniform(0, 2.4, (9, 3))
    got = O.fill_dn_coated(size, units, pts, dn_inner=.04, rad_inner=.3, dn_outer=.02, rad_outer=.5)
    Nx, Ny, Nz = size
    Z, Y, X = np.meshgrid(np.arange(Nz) * units[2], np.arange(Ny) * units[1],
                          np.arange(Nx) * units[0], indexing="ij")
    d = np.min([np.sqrt((X - p[0]) ** 2 + (Y - p[1]) ** 2 + (Z - p[2]) ** 2) for p in pts], axis=0)
    ref = np.where(d < .5, np.where(d < .3, .04, .02), 0.)
    # voxels within float32 rounding of a sphere surface may differ
    edge = (np.abs(d - .5) < 1e-5) | (np.abs(d - .3) < 1e-5)
    assert np.array_equal(got[~edge], ref[~edge].astype(np.float32))
    assert (got > 0).sum() > 50


def test_propagate_single_streamed_sum_equals_volume_sum():
    # return_sum accumulates |u_z|^2 plane by plane; it must equal the z-sum
    # (in z order) of the return_full intensity volume and leave the exit
    # plane unchanged
    rng = np.random.default_rng(4)
    dn = rng.uniform(0, .05, (9, 32, 48)).astype(np.float32)
    u0 = np.exp(1j * rng.uniform(0, 3, (32, 48))).astype(np.complex64)
    for prec in ("c64", "f64"):
        kw = dict(size=(48, 32, 9), units=(.1, .1, .1), lam=.5, n0=1.33, dn=dn, u0=u0, precision=prec)
        vol = O.propagate_single(return_full=True, return_intensity=True, **kw)
        last, acc = O.propagate_single(return_full=False, return_sum=True, **kw)
        ref = vol[0].copy()
        for z in range(1, 9):
            ref = ref + vol[z]
        assert np.array_equal(acc, ref)
        assert np.array_equal(last, O.propagate_single(return_full=False, **kw))
