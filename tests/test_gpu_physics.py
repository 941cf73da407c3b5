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
