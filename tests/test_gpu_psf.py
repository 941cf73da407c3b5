"""GPU parity of the HIP Debye-Wolf kernels (through the C ABI) against the oracle."""
import sys

import numpy as np
import pytest

from oracle import bpm_oracle as O
from oracle import psf_oracle as P

pytestmark = pytest.mark.gpu

TOL = 1e-4


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def bpm():
    import bpm_amd
    from bpm_amd import _lib
    _lib.torch()
    return bpm_amd


@pytest.mark.parametrize("NA,n0,steps", [(1.1, 1.33, 200), ([.3, .5, .7, .9], 1., 64), (.4, 1., 100)])
def test_focus_field_debye_vs_oracle(bpm, NA, n0, steps):
    shape, units, lam = (32, 24, 40), (.05, .06, .05), .5
    say("case focus_field_debye", NA, n0, steps)
    got = bpm.focus_field_debye(shape, units, lam, NA, n0=n0, n_integration_steps=steps)
    ref = P.focus_field_debye(shape, units, lam, NA, n0=n0, n_integration_steps=steps)
    ref64 = P.focus_field_debye(shape, units, lam, NA, n0=n0, n_integration_steps=steps,
                                precision="f64")
    for name, g, r, r64 in zip(("u", "ex", "ey", "ez"), got, ref, ref64):
        e, e64 = O.rel_l2(g, r), O.rel_l2(g, r64)
        say("   %s rel-L2 vs f32 oracle %.2e, vs f64 %.2e" % (name, e, e64))
        assert g.shape == (shape[2], shape[1], shape[0])
        assert e < TOL


def test_psf_and_psf_u0(bpm):
    say("case psf, psf_u0, lightsheet")
    u = bpm.psf((16, 16, 16), (.1, .1, .1), .5, .8)
    r = P.focus_field_debye((16, 16, 16), (.1, .1, .1), .5, .8)[0]
    assert O.rel_l2(u, r) < TOL
    u0 = bpm.psf_u0((64, 48), (.1, .1), zfoc=3., NA=.4, lam=.5, n0=1.33)
    r0 = P.focus_field_debye((64, 48, 4), (.1, .1, 1.5), .5, .4, n0=1.33)[1][0]
    assert u0.shape == (48, 64) and O.rel_l2(u0, r0) < TOL
    ls = bpm.psf_lightsheet((16, 16, 16), (.1, .1, .1), .4, .3, .5, .8)
    a = P.focus_field_debye((16, 16, 16), (.1, .1, .1), .5, .8)[0]
    b = P.focus_field_debye((16, 16, 16), (.1, .1, .1), .4, .3)[0].transpose((2, 1, 0))
    assert O.rel_l2(ls, a * b) < TOL
    with pytest.raises(NotImplementedError):
        bpm.psf((15, 16, 16), (.1, .1, .1), .5, .8)


def test_secondary_variants(bpm):
    say("case plane / at / gauss / debye3")
    ex = bpm.focus_field_debye_plane((40, 32), (.1, .1), z=-1.3, lam=.5, NA=.7, n0=1.2,
                                     n_integration_steps=80)
    rx = P.focus_field_debye_plane((40, 32), (.1, .1), z=-1.3, lam=.5, NA=.7, n0=1.2,
                                   n_integration_steps=80)
    assert O.rel_l2(ex, rx) < TOL
    rng = np.random.default_rng(0)
    x, y, z = (rng.uniform(-1, 1, (5, 7)) for _ in range(3))
    got = bpm.focus_field_debye_at(x, y, z, .5, .9, n0=1.1, n_integration_steps=90)
    ref = P.focus_field_debye_at(x, y, z, .5, .9, n0=1.1, n_integration_steps=90)
    for g, r in zip(got, ref):
        assert g.shape == (5, 7) and O.rel_l2(g, r) < TOL
    got = bpm.focus_field_debye_gauss((20, 18, 16), (.1, .1, .1), .5, [0., .8], sig=.6,
                                      n_integration_steps=60)
    ref = P.focus_field_debye_gauss((20, 18, 16), (.1, .1, .1), .5, [0., .8], sig=.6,
                                    n_integration_steps=60)
    for g, r in zip(got, ref):
        assert O.rel_l2(g, r) < TOL
    got = bpm.focus_field_debye3((17, 16, 15), (.1, .1, .1), .5, .7, n_integration_steps=60)
    ref = P.focus_field_debye3((17, 16, 15), (.1, .1, .1), .5, .7, n_integration_steps=60)
    for g, r in zip(got, ref):
        assert O.rel_l2(g, r) < TOL


def test_bessel_fill_vs_scipy(bpm):
    from scipy.special import jv
    from bpm_amd import focus_fields
    x = np.linspace(0, 130, 100001).astype(np.float32)
    for n in range(3):
        got = focus_fields.bessel_jn(n, x)
        ref = jv(n, x.astype(np.float64))
        assert np.abs(got - ref).max() < 1e-7


def test_full_size_debye_planes(bpm):
    # the 256^3, NA 1.1, n0 1.33, 512-step workload; planes of the octant
    # checked against the oracle evaluated on those planes only
    from bpm_amd import focus_fields as Q
    shape, units, lam, NA, n0, S = (256, 256, 256), (.05,) * 3, .5, 1.1, 1.33, 512
    say("case 256^3 Debye")
    u, ex, ey, ez = (t.cpu().numpy() for t in Q.focus_field_debye_device(shape, units, lam, NA, n0, S))
    N = 129
    a = P._alphas(NA, n0)
    zmax = np.float32(.05 * (N - 1.))
    for k in (0, 37, 128):
        zk = np.float32(0) + (np.float32(k) * (zmax - np.float32(0))) / np.float32(N - 1.)
        rex, rey, rez, ru = P.debye_wolf_grid(N, N, 1, 0., .05 * (N - 1.), 0., .05 * (N - 1.), zk, zk,
                                              lam / n0, a, S)
        # full-grid plane of octant plane k: 128+k for z > 0; plane 128 (z=0)
        # and plane 0 (z=-6.4, octant 128) are conjugated
        fz = 128 + k if 0 < k < 128 else (128 if k == 0 else 0)
        if k in (0, 128):
            rex, rey, rez = np.conj(rex), np.conj(rey), np.conj(rez)
        q = (0, slice(0, 128), slice(0, 128))
        for g, r in ((u[fz, 128:, 128:], ru[q]), (ex[fz, 128:, 128:], rex[q]),
                     (ey[fz, 128:, 128:], rey[q]), (ez[fz, 128:, 128:], rez[q])):
            e = O.rel_l2(g, r)
            say("   plane %d rel-L2 %.2e" % (k, e))
            assert e < TOL


def test_cylindrical_lens(bpm):
    say("case cylindrical")
    shape, units = (16, 40, 24), (.1, .1, .1)
    got = bpm.focus_field_cylindrical(shape, units, lam=.5, NA=.5, n0=1.2, n_integration_steps=90)
    ref = P.cylindrical(40, 24, -.1 * 39 / 2., .1 * 39 / 2., -.1 * 23 / 2., .1 * 23 / 2., .5 / 1.2,
                        np.arcsin(.5 / 1.2), 90)
    for g, r in zip(got, (ref[3], ref[0], ref[1], ref[2])):
        assert g.shape == (24, 40, 16)
        assert O.rel_l2(g, np.repeat(r[..., None], 16, -1)) < TOL
    u, ex = bpm.psf_cylindrical(shape, units, .5, .5, n0=1.2, return_field=True, n_integration_steps=90)
    assert O.rel_l2(u, got[0]) < 1e-7
    u0 = bpm.psf_cylindrical_u0((16, 40), (.1, .1), 2., .5, .5, n0=1.2, n_integration_steps=90)
    r0 = P.cylindrical(40, 4, -.1 * 39 / 2., .1 * 39 / 2., -2., 2., .5 / 1.2, np.arcsin(.5 / 1.2), 90)[0]
    assert O.rel_l2(u0, np.repeat(np.conj(r0[0])[:, None], 16, -1)) < TOL
    pl = bpm.focus_field_cylindrical_plane((16, 40), (.1, .1), z=-.7, lam=.5, NA=.5, n0=1.2,
                                           n_integration_steps=90)
    rp = P.cylindrical(40, 1, -.1 * 39 / 2., .1 * 39 / 2., -.7, -.7, .5 / 1.2, np.arcsin(.5 / 1.2), 90,
                       plane_variant=True)[0]
    assert O.rel_l2(pl, np.repeat(rp[0][:, None], 16, -1)) < TOL


def test_mirror_vector_and_unaligned_paths_agree(bpm):
    """bpmh_debye_mirror writes two x-neighbours per lane with 16-byte stores
    when the outputs are aligned, and one voxel per lane otherwise; both must
    equal the oracle's mirror (pure data movement: bit-exact)."""
    import torch
    from bpm_amd import _lib
    say("case mirror aligned / unaligned")
    rng = np.random.default_rng(7)
    shape = (12, 10, 8)
    nx, ny, nz = (s // 2 + 1 for s in shape)
    u = rng.standard_normal((nz, ny, nx)).astype(np.float32)
    e = [(rng.standard_normal((nz, ny, nx)) + 1j * rng.standard_normal((nz, ny, nx))).astype(np.complex64)
         for _ in range(3)]
    ref = P.mirror_octant(u, *e, shape)
    dev = [_lib.to_device(a, a.dtype.type) for a in [u] + e]
    n = shape[0] * shape[1] * shape[2]
    for offset in (0, 1):  # offset 1 element: 8-byte aligned complex outputs -> scalar path
        outs = [torch.zeros(n + 1, dtype=torch.float32, device="cuda")] + \
               [torch.zeros(n + 1, dtype=torch.complex64, device="cuda") for _ in range(3)]
        views = [o[offset:offset + n] for o in outs]
        _lib.call("bpmh_debye_mirror", shape[0], shape[1], shape[2], *[_lib.ptr(d) for d in dev],
                  *[_lib.ptr(v) for v in views], _lib.stream_handle())
        torch.cuda.synchronize()
        for got, r in zip(views, ref):
            assert np.array_equal(_lib.to_numpy(got).reshape(r.shape), r)


def test_debye_symmetric_sample_table_is_bit_identical(bpm):
    """With identical x and y axes bpmh_debye_wolf fills its sample table once
    per pair of columns (i, j), i <= j.  An (n, n+1) grid whose first n y
    coordinates equal the x coordinates exactly (power-of-two spacing) runs
    the general kernel; its first n rows must equal the symmetric run bit for
    bit, and the oracle to the usual tolerance."""
    from bpm_amd import focus_fields as Q
    from bpm_amd import _lib
    n, nz, d, lam = 17, 6, 0.0625, .5 / 1.33
    alphas = Q._alphas(1.1, 1.33)
    say("case Debye symmetric table, n", n)
    sym = Q._grid_call(n, n, nz, 0., (n - 1) * d, 0., (n - 1) * d, 0., (nz - 1) * d, lam, alphas,
                       64, _lib.DW_INDEX)
    gen = Q._grid_call(n, n + 1, nz, 0., (n - 1) * d, 0., n * d, 0., (nz - 1) * d, lam, alphas,
                       64, _lib.DW_INDEX)
    for a, b in zip(sym, gen):
        a, b = _lib.to_numpy(a), _lib.to_numpy(b)
        assert np.array_equal(a, b[:, :n, :])
    ex_ref = P.debye_wolf_grid(n, n, nz, 0., (n - 1) * d, 0., (n - 1) * d, 0., (nz - 1) * d,
                               lam, P._alphas(1.1, 1.33), 64)[0]
    assert O.rel_l2(_lib.to_numpy(sym[1]), ex_ref) < TOL


def test_plane_variant_fills_a_provided_tensor(bpm):
    """focus_field_debye_plane(ex_g=...) writes the same bits into a given
    (Ny, Nx) complex64 GPU tensor as it returns without one, and refuses a
    tensor of another dtype before launching (the kernel writes complex64
    through the bare pointer)."""
    import torch
    from bpm_amd import focus_fields as Q
    say("case plane into ex_g")
    ref = Q.focus_field_debye_plane((32, 24), (.1, .1), z=.2, NA=.6, n0=1.)
    g = torch.zeros((24, 32), dtype=torch.complex64, device="cuda")
    assert Q.focus_field_debye_plane((32, 24), (.1, .1), z=.2, NA=.6, n0=1., ex_g=g) is None
    assert np.array_equal(g.cpu().numpy(), ref)
    with pytest.raises(ValueError):
        Q.focus_field_debye_plane((32, 24), (.1, .1), ex_g=torch.zeros((24, 32), device="cuda"))


@pytest.mark.parametrize("shape", [(64, 64, 64), (48, 32, 40), (256, 256, 256), (128, 128, 200),
                                   (96, 96, 512), (2048, 2048, 4), (2, 2, 2), (4, 4, 6)])
def test_fused_focus_field_equals_octant_plus_mirror(bpm, shape):
    """bpmh_focus_field_debye (the product's epilogue writes every octant
    value to its mirror positions) against bpmh_debye_wolf on the octant
    followed by bpmh_debye_mirror: the same values bit for bit, for the
    pair-table path (square x/y axes) and the general one, and at the 256^3
    workload size, and square grids whose octant depth (101, 257 planes) is
    not a whole number of the product's plane groups; 2048^2 octant rows
    (1025 points) take the spread's one-row work-groups; 2^3 and 4^2 x 6 are
    the smallest grids (one lane per full-grid row pair)."""
    import torch
    from bpm_amd import focus_fields as Q
    from bpm_amd import _lib
    say("case fused Debye", shape)
    Nx0, Ny0, Nz0 = shape
    units, lam, NA, n0, S = (.05, .06, .05), .5, 1.1, 1.33, 128
    if Nx0 == Ny0:
        units = (.05,) * 3
    dx, dy, dz = units
    alphas = Q._alphas(NA, n0)
    Nx, Ny, Nz = Nx0 // 2 + 1, Ny0 // 2 + 1, Nz0 // 2 + 1
    oct_ = Q._grid_call(Nx, Ny, Nz, 0., dx * (Nx - 1.), 0., dy * (Ny - 1.), 0., dz * (Nz - 1.),
                        1. * lam / n0, alphas, S, _lib.DW_INDEX)
    ref = [_lib.empty((Nz0, Ny0, Nx0), np.float32)] + \
          [_lib.empty((Nz0, Ny0, Nx0), np.complex64) for _ in range(3)]
    _lib.call("bpmh_debye_mirror", Nx0, Ny0, Nz0, *[_lib.ptr(o) for o in oct_],
              *[_lib.ptr(r) for r in ref], _lib.stream_handle())
    got = Q.focus_field_debye_device(shape, units, lam, NA, n0, S)
    torch.cuda.synchronize()
    for g, r in zip(got, ref):
        assert torch.equal(g, r)


def test_back_to_back_calls_equal_synchronised_calls(bpm):
    """The grid path's calls share one library scratch buffer, ordered by an
    event (calls on another stream wait for its last use; larger calls grow
    it).  Calls issued back to back (fused full-grid calls of different
    sizes, NA and steps, an octant-only call, a call on another stream) must
    give the same bits as the same calls each synchronised."""
    import torch
    from bpm_amd import focus_fields as Q
    from bpm_amd import _lib
    say("case Debye back to back")
    cfgs = [((64, 64, 64), .05, 1.1, 128), ((96, 96, 48), .04, .8, 64),
            ((128, 128, 32), .06, 1.2, 96), ((64, 64, 64), .05, .5, 200)]

    def full(c):
        shape, d, NA, S = c
        return Q.focus_field_debye_device(shape, (d,) * 3, .5, NA, 1.33, S)

    def octant():
        return Q._grid_call(33, 33, 17, 0., 1.6, 0., 1.6, 0., .8, .5 / 1.33, Q._alphas(.9, 1.33),
                            80, _lib.DW_INDEX)

    ref = []
    for c in cfgs:
        ref.append([t.clone() for t in full(c)])
        torch.cuda.synchronize()
    ref_oct = [t.clone() for t in octant()]
    torch.cuda.synchronize()
    got = []
    for rep in range(2):
        for c in cfgs:
            got.append(full(c))
        got.append(octant())
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        other = full(cfgs[2])
    s.synchronize()
    torch.cuda.synchronize()
    want = (ref + [ref_oct]) * 2
    for g, r in zip(got, want):
        for a, b in zip(g, r):
            assert torch.equal(a, b)
    for a, b in zip(other, ref[2]):
        assert torch.equal(a, b)

