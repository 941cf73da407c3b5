"""Physics checks of the HIP path on the GPU, in the spirit of the
reference's own scripts (tests/test_single_focus.py, tests/test_gaussian.py):
a focused field from the Debye-Wolf kernels propagated by the split-step
kernels must reproduce the Debye field at every plane, and a paraxial
Gaussian beam must follow its closed form."""
import sys

import numpy as np
import pytest

from oracle import bpm_oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def bpm():
    import bpm_amd
    from bpm_amd import _lib
    _lib.torch()
    return bpm_amd


@pytest.mark.parametrize("NA,n0", [(.3, 1.), (.6, 1.33)])
def test_bpm_of_psf_u0_reproduces_debye_focus(bpm, NA, n0):
    # psf_u0 is Ex of the Debye field at z = -zfoc; the psf grid's planes are
    # z = (k - Nz/2) dz, so with zfoc = Nz dz / 2 free split-step propagation
    # of u0 must give the Debye Ex volume plane by plane: both solve the same
    # homogeneous problem, up to the Debye approximation and the sampling of
    # the aperture.  The Debye field carries its phase as exp(+i k z) while a
    # split step multiplies a plane wave by exp(-i k dz), so the comparison
    # propagates conj(u0) and compares with conj(Ex) (psf_functions.py:104
    # returns the unconjugated plane; its cylindrical sibling at
    # psf_functions.py:156 conjugates for exactly this reason).  The float64
    # oracle gives 2.6e-2 .. 5.5e-2 at these planes for both cases.
    size, units, lam = (64, 64, 64), (.1, .1, .1), .5
    Nx, Ny, Nz = size
    say("case focus NA", NA, "n0", n0)
    _, ex, _, _ = bpm.psf(size, units, lam, NA, n0=n0, return_field=True)
    u0 = bpm.psf_u0((Nx, Ny), units[:2], zfoc=.5 * units[2] * Nz, NA=NA, lam=lam, n0=n0)
    assert O.rel_l2(u0, ex[0]) < 1e-5  # the same plane from two calls
    u = bpm.bpm_3d(size, units, lam=lam, n0=n0, u0=np.conj(u0), absorbing_width=0)
    err = [O.rel_l2(u[k], np.conj(ex[k])) for k in (Nz // 4, Nz // 2, Nz - 1)]
    say("   rel-L2 bpm vs Debye at z = -zfoc/2, 0, +zfoc: %.3e %.3e %.3e" % tuple(err))
    assert max(err) < 8e-2
    # without the conjugation the two disagree completely
    u = bpm.bpm_3d(size, units, lam=lam, n0=n0, u0=u0, absorbing_width=0)
    assert O.rel_l2(u[Nz // 2], ex[Nz // 2]) > .5


def test_gaussian_beam_follows_closed_form(bpm):
    from bpm_amd.workloads import gaussian_beam
    Nx = Ny = 256
    dx, dz, n0, w0, lam = .2, .2, 1.33, 2., .5
    Nz, z0 = 120, -12.
    u0 = gaussian_beam(Nx, Ny, dx, dx, lam, n0, w0, z0)
    say("case Gaussian beam")
    u = bpm.bpm_3d((Nx, Ny, Nz), (dx, dx, dz), lam=lam, n0=n0, u0=u0, mean_mode="ignore")
    err = [O.rel_l2(u[k], gaussian_beam(Nx, Ny, dx, dx, lam, n0, w0, z0 + k * dz))
           for k in (0, Nz // 2, Nz - 1)]
    say("   rel-L2 vs paraxial closed form at z = %g, %g, %g: %.2e %.2e %.2e"
        % (z0, z0 + Nz // 2 * dz, z0 + (Nz - 1) * dz, *err))
    assert err[0] < 1e-6 and max(err) < 0.02


def test_refraction_reference_script_vs_oracle(bpm):
    """The reference's refraction script (tests/test_refraction.py:6-44): a
    focused NA 0.3 beam (psf_u0, zfoc at the volume's centre) through an
    interface tilted by w = 0.8 rad into dn0 = 0.1, 256^3 at 0.1 um, n0 = 1,
    bpm_3d defaults (weighted mean).  The script only plots; here the whole
    field volume is checked against the complex64 oracle (bar 1e-4) given
    the same u0."""
    Nx = Nz = 256
    dx, NA, lam, w, dn0 = .1, .3, .5, .8, .1
    say("case refraction script")
    x = dx * (np.arange(-Nx / 2, Nx / 2) + .5)
    z = dx * (np.arange(-Nz / 2, Nz / 2) + .5)
    Z, Y, X = np.meshgrid(z, x, x, indexing="ij")
    Zr = np.cos(w) * Z - np.sin(w) * X
    dn = (dn0 * (Zr > 0)).astype(np.float32)
    u0 = bpm.psf_u0((Nx, Nx), units=(dx,) * 2, zfoc=dx * (Nz - 1.) / 2., lam=lam, NA=NA)
    u = bpm.bpm_3d((Nx, Nx, Nz), units=(dx,) * 3, lam=lam, dn=dn, u0=u0)
    ref = O.bpm_3d((Nx, Nx, Nz), (dx,) * 3, lam, u0, dn, 1.)
    e = O.rel_l2(u, ref)
    say("   field volume rel-L2 vs c64 oracle %.2e" % e)
    assert u.shape == (Nz, Nx, Nx) and e < 1e-4


def test_refraction_follows_snell(bpm):
    """A Gaussian beam (w0 = 3 um) at normal incidence crosses a plane
    interface tilted by w = 0.4 rad into n2 = n0 + dn0 (n0 = 1.33, dn0 =
    0.05).  Behind the interface its centroid moves on a straight line whose
    slope dx/dz is measured over 25 um.  Phase matching of the split step
    (thin phase screen, propagator in n0) gives the transverse wavenumber
    kx = -k0 dn0 tan w, so dx/dz = kx / sqrt(n0^2 k0^2 - kx^2): checked to
    1 %.  Snell's law, n0 sin w = n2 sin(w - delta), bends the beam towards
    the normal (dx/dz < 0) by delta; the split step agrees with it to first
    order in dn0 (their ratio is n2/n0 = 1.038 here): checked to 5 %."""
    from bpm_amd.workloads import gaussian_beam
    Nx, Nz = 256, 512
    dx, lam, n0, dn0, w = .1, .5, 1.33, .05, .4
    say("case Snell")
    x = dx * (np.arange(Nx) - Nx / 2)
    z = dx * np.arange(Nz)
    Z, Y, X = np.meshgrid(z, x, x, indexing="ij")
    dn = (dn0 * (np.cos(w) * (Z - 10.) - np.sin(w) * X > 0)).astype(np.float32)
    u0 = gaussian_beam(Nx, Nx, dx, dx, lam, n0, 3., 0.)
    I = bpm.bpm_3d((Nx, Nx, Nz), (dx,) * 3, lam=lam, u0=u0, dn=dn, n0=n0, mean_mode="ignore",
                   return_field=False, absorbing_width=16)
    ks = np.arange(250, Nz)  # z from 25 um on: behind the interface
    prof = I[ks].sum(axis=1).astype(np.float64)  # (planes, x)
    xc = (prof * x[None, :]).sum(1) / prof.sum(1)
    slope = np.polyfit(z[ks], xc, 1)[0]
    k0 = 2 * np.pi / lam
    kx = -k0 * dn0 * np.tan(w)
    pred = kx / np.sqrt((n0 * k0) ** 2 - kx ** 2)
    delta = w - np.arcsin(n0 * np.sin(w) / (n0 + dn0))
    say("   centroid slope %.5f; split-step phase matching %.5f; Snell %.5f"
        % (slope, pred, -np.tan(delta)))
    assert slope < 0
    assert abs(slope / pred - 1) < .01
    assert abs(slope / -np.tan(delta) - 1) < .05
