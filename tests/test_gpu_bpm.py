"""GPU parity of the HIP split-step BPM (through the C ABI) against the oracle.

Bar: complex64 fields and intensities within relative L2 <= 1e-4 of the
complex64 restatement of the reference float32 path; the
error against the float64 oracle is printed alongside.  Full-size cases are
checked through size-independent properties (closed form, Parseval,
batch == independent runs).
"""
import sys

import numpy as np
import pytest

from oracle import bpm_oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-4
LAM, N0 = .5, 1.33


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def bpm():
    import bpm_amd
    from bpm_amd import _lib
    _lib.torch()  # raises without a GPU or without libbpmhip.so
    return bpm_amd


def rng_dn(shape, kind, seed=0, amp=.05):
    rng = np.random.default_rng(seed)
    Nz, Ny, Nx = shape
    from bpm_amd.workloads import random_spheres, spheres_dn
    size = (Nx, Ny, Nz)
    c, r, d = random_spheres(rng, size, (.1, .1, .1), 12, (.3, 1.), (.01, amp))
    dn = spheres_dn(size, (.1, .1, .1), c, r, d)
    if kind == "c64":
        dn = (dn - 1j * 0.2 * dn).astype(np.complex64)
    return dn


@pytest.mark.parametrize("mean_mode", ["ignore", "simple", "weighted"])
@pytest.mark.parametrize("kind", ["none", "f32", "f16", "c64"])
@pytest.mark.parametrize("absorb", [0, 6])
def test_bpm_3d_modes_vs_oracle(bpm, mean_mode, kind, absorb):
    Nx, Ny, Nz = 64, 32, 24
    say("case bpm_3d", mean_mode, kind, absorb, (Nx, Ny, Nz))
    dn = None if kind == "none" else rng_dn((Nz, Ny, Nx), kind, seed=1)
    half = kind == "f16"
    rng = np.random.default_rng(2)
    u0 = np.exp(1j * rng.uniform(0, .5, (Ny, Nx))).astype(np.complex64)
    got = bpm.bpm_3d((Nx, Ny, Nz), (.1, .1, .1), LAM, u0=u0, dn=dn, n0=N0, mean_mode=mean_mode,
                     absorbing_width=absorb, store_dn_as_half=half)
    ref = O.bpm_3d((Nx, Ny, Nz), (.1, .1, .1), LAM, u0, dn, N0, mean_mode=mean_mode,
                   absorbing_width=absorb, store_dn_as_half=half)
    e = O.rel_l2(got, ref)
    say("   rel-L2 vs c64 oracle %.2e" % e)
    assert got.shape == (Nz, Ny, Nx) and got.dtype == np.complex64
    assert e < TOL


@pytest.mark.parametrize("nx,ny", [(16, 16), (32, 64), (128, 16), (256, 256), (512, 128),
                                   (1024, 64), (2048, 32), (64, 2048), (4096, 16), (16, 4096)])
def test_all_grid_sizes(bpm, nx, ny):
    Nz = 5
    say("case sizes", nx, ny)
    dn = rng_dn((Nz, ny, nx), "f32", seed=nx + ny)
    rng = np.random.default_rng(3)
    u0 = np.exp(1j * rng.uniform(0, 1, (ny, nx))).astype(np.complex64)
    got = bpm.bpm_3d((nx, ny, Nz), (.1, .1, .1), LAM, u0=u0, dn=dn, n0=N0, mean_mode="weighted",
                     return_full=False)
    ref = O.bpm_3d((nx, ny, Nz), (.1, .1, .1), LAM, u0, dn, N0, mean_mode="weighted",
                   return_full=False)
    e = O.rel_l2(got, ref)
    say("   rel-L2 %.2e" % e)
    assert e < TOL


def test_outputs_intensity_scattering_split(bpm):
    Nx, Ny, Nz = 64, 64, 21
    say("case outputs")
    dn = rng_dn((Nz, Ny, Nx), "f32", seed=4)
    kw = dict(n0=N0, mean_mode="ignore")
    # intensity volume, last plane field / intensity
    ui = bpm.bpm_3d((Nx, Ny, Nz), (.1, .1, .1), LAM, dn=dn, return_field=False, **kw)
    ri = O.bpm_3d((Nx, Ny, Nz), (.1, .1, .1), LAM, None, dn, return_field=False, **kw)
    assert ui.dtype == np.float32 and O.rel_l2(ui, ri) < TOL
    ul = bpm.bpm_3d((Nx, Ny, Nz), (.1, .1, .1), LAM, dn=dn, return_full=False, **kw)
    rl = O.bpm_3d((Nx, Ny, Nz), (.1, .1, .1), LAM, None, dn, return_full=False, **kw)
    assert ul.shape == (Ny, Nx) and O.rel_l2(ul, rl) < TOL
    # scattering cross-section and g factor
    u, p, g = bpm.bpm_3d((Nx, Ny, Nz), (.1, .1, .1), LAM, dn=dn, return_scattering=True,
                         return_g=True, **kw)
    ru, rp, rg = O.bpm_3d((Nx, Ny, Nz), (.1, .1, .1), LAM, None, dn, return_scattering=True,
                          return_g=True, **kw)
    say("   p rel %.2e g rel %.2e" % (O.rel_l2(p[1:], rp[1:]), O.rel_l2(g[1:], rg[1:])))
    assert p[0] == 0 and np.isnan(g[0])
    assert O.rel_l2(p[1:], rp[1:]) < 1e-3 and O.rel_l2(g[1:], rg[1:]) < 1e-3
    # z-chunked, weighted, with scattering
    for nv in (2, 3):
        us, ps = bpm.bpm_3d((Nx, Ny, Nz), (.1, .1, .1), LAM, dn=dn, n_volumes=nv,
                            return_scattering=True, n0=N0)
        rs, rps = O.bpm_3d_split((Nx, Ny, Nz), (.1, .1, .1), LAM, None, dn, nv, N0,
                                 return_scattering=True)
        assert O.rel_l2(us, rs) < TOL and O.rel_l2(ps[1:], rps[1:]) < 1e-3


def test_field_intensity_identity_reference_case(bpm):
    # reference tests/test_field_intens.py:14-43 (its only assert), Nz reduced
    Nx, Ny, Nz = 256, 256, 128
    say("case field==intensity")
    dx, dz = .05, .05
    x = dx * np.arange(-Nx / 2, Nx / 2)
    z = dz * np.arange(-Nz / 2, Nz / 2)
    Z, Y, X = np.meshgrid(z, x, x, indexing="ij")
    dn = .05 * (np.sqrt(X ** 2 + Y ** 2 + Z ** 2) < 1.)
    u1 = bpm.bpm_3d((Nx, Ny, Nz), (dx, dx, dz), LAM, dn=dn, n_volumes=2, return_field=True)
    u2 = bpm.bpm_3d((Nx, Ny, Nz), (dx, dx, dz), LAM, dn=dn, n_volumes=2, return_field=False)
    assert np.allclose(np.abs(u1) ** 2, u2)


def test_class_path_vs_oracle(bpm):
    Nx, Ny, Nz = 128, 64, 33
    say("case Bpm3d")
    for kind in ("f32", "c64"):
        dn = rng_dn((Nz, Ny, Nx), kind, seed=5)
        m = bpm.Bpm3d((Nx, Ny, Nz), (.1, .1, .1), lam=LAM, n0=N0, dn=dn)
        for absorb in (0, 4):
            got = m._propagate_single(absorbing_width=absorb)
            ref = O.propagate_single((Nx, Ny, Nz), (.1, .1, .1), LAM, N0, dn, None,
                                     absorbing_width=absorb)
            e = O.rel_l2(got, ref)
            say("   %s absorb=%d rel-L2 %.2e" % (kind, absorb, e))
            assert e < TOL
        # return_full=False gives the last plane's field even with
        # return_intensity (_bpm3d_ocl.py:128-131); the intensity volume
        # is |u|^2 of the field volume
        gl = m.propagate(return_full=False, return_intensity=True, absorbing_width=4)
        assert gl.shape == (Ny, Nx) and gl.dtype == np.complex64
        assert O.rel_l2(gl, ref[-1]) < TOL
        gi = m.propagate(return_full=True, return_intensity=True, absorbing_width=4)
        assert gi.dtype == np.float32 and O.rel_l2(gi, np.abs(ref) ** 2) < TOL
    with pytest.raises(ValueError):
        m.set_dn(np.zeros((Nz, Ny, Nx + 1), np.float32))
    with pytest.raises(NotImplementedError):
        bpm.Bpm3d((Nx, Ny, Nz), (.1, .1, .1), backend="cuda")


def test_errors_map_to_reference_exceptions(bpm):
    with pytest.raises(NotImplementedError):
        bpm.bpm_3d((32, 32, 4), (.1, .1, .1), mean_mode="median")
    with pytest.raises(NotImplementedError):
        bpm.bpm_3d((32, 32, 4), (.1, .1, .1), subsample=2)
    with pytest.raises(NotImplementedError):
        bpm.bpm_3d((48, 32, 4), (.1, .1, .1))
    with pytest.raises(ValueError):
        bpm.bpm_3d((32, 32, 4), (.1, .1, .1), dn=np.zeros((4, 32, 16), np.float32))


def test_full_size_plane_wave_closed_form(bpm):
    # 1024^2 grid, 256 steps through a layered medium: exact solution known
    Nx = Ny = 1024
    Nz, dx, dz = 256, .1, .1
    say("case 1024^2 plane wave")
    mx, my = 7, -3
    kx = 2 * np.pi * mx / (Nx * dx)
    ky = 2 * np.pi * my / (Ny * dx)
    x = dx * np.arange(Nx)
    u0 = np.exp(-1j * (kx * x[None, :] + ky * x[:, None])).astype(np.complex64)
    c = (0.03 * np.sin(np.arange(Nz) / 17.)).astype(np.float32)
    dn = np.broadcast_to(c[:, None, None], (Nz, Ny, Nx)).copy()
    got = bpm.bpm_3d((Nx, Ny, Nz), (dx, dx, dz), LAM, u0=u0, dn=dn, n0=N0, mean_mode="ignore",
                     return_full=False)
    kz = np.sqrt(N0 ** 2 * (2 * np.pi / LAM) ** 2 - kx ** 2 - ky ** 2)
    ph = (Nz - 1) * dz * kz + 2 * np.pi / LAM * dz * c[1:].astype(np.float64).sum()
    exact = u0.astype(np.complex128) * np.exp(-1j * ph)
    e = O.rel_l2(got, exact)
    say("   rel-L2 vs closed form %.2e" % e)
    assert e < 1e-4


def test_full_size_energy_conservation_and_batch(bpm):
    # dx = 0.3 um: every FFT frequency propagates, so ||u_z|| is conserved.
    import torch
    from bpm_amd import bpm3d as B
    from bpm_amd import _lib
    Nx = Ny = 1024
    Nz, B_ = 64, 3
    say("case 1024^2 energy + batch")
    rng = np.random.default_rng(11)
    dn = torch.from_numpy(rng.uniform(-.05, .05, (Nz, Ny, Nx)).astype(np.float32)).cuda()
    u0 = np.exp(1j * rng.uniform(0, 6, (B_, Ny, Nx))).astype(np.complex64)
    units = (.3, .3, .1)
    H = B.propagator_table(Nx, Ny, units, LAM, N0)
    plane = torch.from_numpy(u0).cuda()
    acc = torch.empty((B_, Ny, Nx), dtype=torch.float32, device="cuda")
    B.propagate_fields(plane, Nz, units, LAM, N0, dn=dn, dn_kind=_lib.DN_F32, mean_mode="ignore",
                       H=H, out_sum=acc)
    out = plane.cpu().numpy()
    e0 = np.sum(np.abs(u0.astype(np.complex128)) ** 2, axis=(1, 2))
    e1 = np.sum(np.abs(out.astype(np.complex128)) ** 2, axis=(1, 2))
    drift = np.abs(e1 / e0 - 1).max()
    say("   energy drift over %d steps: %.2e" % (Nz - 1, drift))
    assert drift < 1e-5
    # summed intensity of a unitary propagation: total = Nz * energy
    tot = acc.double().sum(dim=(1, 2)).cpu().numpy()
    assert np.allclose(tot, Nz * e0, rtol=1e-5)
    # each batch item equals its own single run, bit for bit
    for b in range(B_):
        p1 = torch.from_numpy(u0[b:b + 1]).cuda()
        B.propagate_fields(p1, Nz, units, LAM, N0, dn=dn, dn_kind=_lib.DN_F32,
                           mean_mode="ignore", H=H)
        assert torch.equal(p1[0], plane[b])


def test_graph_replay_matches_eager(bpm):
    import torch
    from bpm_amd import bpm3d as B
    from bpm_amd import _lib
    Nx, Ny, Nz = 256, 128, 40
    say("case graph replay")
    rng = np.random.default_rng(12)
    dn = torch.from_numpy(rng.uniform(0, .05, (Nz, Ny, Nx)).astype(np.float32)).cuda()
    u0 = torch.from_numpy(np.exp(1j * rng.uniform(0, 3, (1, Ny, Nx))).astype(np.complex64)).cuda()
    units = (.1, .1, .1)
    H = B.propagator_table(Nx, Ny, units, LAM, N0, _lib.H_F64)
    plan = B.get_plan(Nx, Ny)
    lib = _lib.load()
    plane = torch.empty_like(u0)
    acc = torch.empty((1, Ny, Nx), dtype=torch.float32, device="cuda")

    def run(mode):
        plane.copy_(u0)
        B.propagate_fields(plane, Nz, units, LAM, N0, dn=dn, dn_kind=_lib.DN_F32,
                           mean_mode=mode, H=H, out_sum=acc)
        torch.cuda.synchronize()
        return plane.clone(), acc.clone()

    for mode in ("ignore", "weighted"):
        _lib.check(lib.bpmh_plan_set_graph(plan, 0))
        ref_p, ref_a = run(mode)
        _lib.check(lib.bpmh_plan_set_graph(plan, 1))
        try:
            for r in range(4):
                got_p, got_a = run(mode)
                assert torch.equal(got_p, ref_p) and torch.equal(got_a, ref_a), (mode, r)
                # unrelated eager work between replays
                x = torch.randn(1 << 20, device="cuda")
                (x * 2).sum().item()
        finally:
            _lib.check(lib.bpmh_plan_set_graph(plan, 0))


def test_device_sphere_volumes(bpm):
    from bpm_amd import spheres as S
    from bpm_amd.workloads import random_spheres, spheres_dn
    say("case sphere volumes")
    rng = np.random.default_rng(21)
    size, units = (96, 80, 64), (.1, .1, .1)
    pts = rng.uniform(0, 8, (15, 3))
    got = S.create_dn_buffer(size, units, pts, dn_inner=.04, rad_inner=.6, dn_outer=.02,
                             rad_outer=1.1).cpu().numpy()
    ref = O.fill_dn_coated(size, units, pts, dn_inner=.04, rad_inner=.6, dn_outer=.02, rad_outer=1.1)
    assert np.array_equal(got, ref)
    c, r, d = random_spheres(rng, size, units, 40, (.3, 1.5), (.01, .05))
    got = S.spheres_dn_device(size, units, c, r, d).cpu().numpy()
    assert np.array_equal(got, spheres_dn(size, units, c, r, d))
    # propagation through a coated-sphere volume (bpm_3d_spheres)
    size = (64, 32, 40)
    ref = O.fill_dn_coated(size, units, pts, dn_inner=.04, rad_inner=.6, dn_outer=.02, rad_outer=1.1)
    u = S.bpm_3d_spheres(size, units, points=pts, dn_inner=.04, rad_inner=.6, dn_outer=.02,
                         rad_outer=1.1)
    uref = O.bpm_3d(size, units, .5, None, ref, 1.)
    assert O.rel_l2(u, uref) < TOL


@pytest.mark.parametrize("nx,ny", [(1024, 1024), (256, 512), (128, 256), (16, 256), (4096, 16),
                                   (16, 2048)])
def test_launch_variants_agree(bpm, nx, ny):
    import torch
    from bpm_amd import bpm3d as B
    from bpm_amd import _lib
    say("case variants", nx, ny)
    Nz = 6
    rng = np.random.default_rng(7)
    dn = torch.from_numpy(rng.uniform(0, .05, (Nz, ny, nx)).astype(np.float32)).cuda()
    u0 = torch.from_numpy(np.exp(1j * rng.uniform(0, 3, (2, ny, nx))).astype(np.complex64)).cuda()
    units = (.1, .1, .1)
    H = B.propagator_table(nx, ny, units, LAM, N0, _lib.H_F64)
    plan = B.get_plan(nx, ny, None)
    lib = _lib.load()
    outs = {}
    try:
        for rv in range(6):
            for cv in range(10):
                _lib.check(lib.bpmh_plan_set_variant(plan, rv, cv))
                for mode in ("ignore", "weighted"):
                    p = u0.clone()
                    B.propagate_fields(p, Nz, units, LAM, N0, dn=dn, dn_kind=_lib.DN_F32,
                                       mean_mode=mode, H=H, tuned=False)
                    outs[(rv, cv, mode)] = p
        # per-line arithmetic and the per-column reductions do not depend on
        # the launch shape, so every variant of one pair of radix schedules
        # (rv // 3, cv // 5) gives the same bits; the schedules differ only
        # in rounding
        for (rv, cv, mode), p in outs.items():
            ref = outs[(3 * (rv // 3) + 1, 5 * (cv // 5) + 1, mode)]
            assert torch.equal(p, ref), (rv, cv, mode)
            d = ((p - outs[(1, 1, mode)]).abs().norm() / outs[(1, 1, mode)].abs().norm()).item()
            assert d < 1e-5, (rv, cv, mode, d)
    finally:
        _lib.check(lib.bpmh_plan_set_variant(plan, 1, 1))


@pytest.mark.parametrize("nx,ny", [(512, 512), (64, 1024)])
def test_paired_column_variant_on_every_column_path(bpm, nx, ny):
    """Column variants 3 (two adjacent columns per thread, 16-byte accesses)
    and 4 (the same with each column's accesses separate) against variant 0
    (the same work-group, one column per thread) on the column pass's other
    paths: the simple mean (propagator formula with n_eff per
    plane), the scattering sums, and a plane that is only 8-byte aligned (a
    view at an odd element offset, which the library runs as variant 0).
    Same bits throughout."""
    import torch
    from bpm_amd import bpm3d as B
    from bpm_amd import _lib
    say("case paired columns", nx, ny)
    Nz, nb = 6, 2
    rng = np.random.default_rng(11)
    dn = torch.from_numpy(rng.uniform(0, .05, (Nz, ny, nx)).astype(np.float32)).cuda()
    u0 = torch.from_numpy(np.exp(1j * rng.uniform(0, 3, (nb, ny, nx))).astype(np.complex64)).cuda()
    units = (.1, .1, .1)
    H = B.propagator_table(nx, ny, units, LAM, N0, _lib.H_F64)
    dn0 = dn.mean(dim=(1, 2)).contiguous()
    n_eff = (dn0 + N0).contiguous()
    sw = torch.from_numpy(rng.uniform(0, 1, (ny, nx)).astype(np.float32)).cuda()
    gw = torch.from_numpy(rng.uniform(-1, 1, (ny, nx)).astype(np.float32)).cuda()
    plain = torch.from_numpy(np.exp(1j * rng.uniform(0, 3, Nz)).astype(np.complex64)).cuda()
    plan = B.get_plan(nx, ny, None)
    lib = _lib.load()
    outs = {}
    try:
        for cv in (0, 3, 4):
            _lib.check(lib.bpmh_plan_set_variant(plan, 1, cv))
            p = u0.clone()
            B.propagate_fields(p, Nz, units, LAM, N0, dn=dn, dn_kind=_lib.DN_F32,
                               mean_mode="simple", n_eff=n_eff, dn0=dn0, tuned=False)
            outs[(cv, "simple")] = p
            p = u0.clone()
            sout = torch.zeros(nb * Nz * 2, dtype=torch.float32, device="cuda")
            B.propagate_fields(p, Nz, units, LAM, N0, dn=dn, dn_kind=_lib.DN_F32,
                               mean_mode="ignore", H=H, scat=(sw, gw, plain, sout), tuned=False)
            outs[(cv, "scat")] = p
            outs[(cv, "sums")] = sout
            buf = torch.zeros(nb * ny * nx + 1, dtype=torch.complex64, device="cuda")
            p = buf[1:].view(nb, ny, nx)
            assert p.data_ptr() % 16 == 8
            p.copy_(u0)
            B.propagate_fields(p, Nz, units, LAM, N0, dn=dn, dn_kind=_lib.DN_F32,
                               mean_mode="weighted", tuned=False)
            outs[(cv, "unaligned")] = p.clone()
        for kind in ("simple", "scat", "sums", "unaligned"):
            assert torch.equal(outs[(3, kind)], outs[(0, kind)]), kind
            assert torch.equal(outs[(4, kind)], outs[(0, kind)]), kind
        assert outs[(0, "sums")].abs().sum() > 0
    finally:
        _lib.check(lib.bpmh_plan_set_variant(plan, 1, 1))


@pytest.mark.parametrize("config", ["256x256x256", "1024x1024x1024"])
def test_plain_workloads_vs_c64_and_f64_oracles(bpm, config):
    """The plain workloads of the benchmark (Bpm3d path: float64 propagator
    rounded to complex64, real dn, exit field + sum_z |u_z|^2) in full:
    256^2 x 256 (21 spheres, plane wave) and 1024^2 x 1024 (400 spheres,
    focused Gaussian beam), against the complex64 oracle (bar 1e-4) and the
    float64 oracle (error stated; the oracle runs about a minute here)."""
    import torch
    from bpm_amd import _lib
    from bpm_amd import bpm3d as B
    from bpm_amd.workloads import plain_large_workload, plain_small_workload
    say("case config", config)
    units = (.1, .1, .1)
    if config == "256x256x256":
        u0, dn = plain_small_workload(0)
    else:
        u0, dn = plain_large_workload(0)
    nz, ny, nx = dn.shape
    dn_g = _lib.to_device(dn, np.float32)
    plane = _lib.to_device(u0[None], np.complex64)
    acc = torch.empty((1, ny, nx), dtype=torch.float32, device="cuda")
    H = B.propagator_table(nx, ny, units, LAM, N0, _lib.H_F64)
    B.propagate_fields(plane, nz, units, LAM, N0, dn=dn_g, dn_kind=_lib.DN_F32,
                       mean_mode="ignore", H=H, out_sum=acc)
    got_u, got_s = _lib.to_numpy(plane)[0], _lib.to_numpy(acc)[0]
    O.set_fft_workers(min(16, len(__import__("os").sched_getaffinity(0))))
    try:
        errs = {}
        for prec in ("c64", "f64"):
            t0 = __import__("time").perf_counter()
            ref_u, ref_s = O.propagate_single((nx, ny, nz), units, LAM, N0, dn, u0,
                                              return_full=False, precision=prec, return_sum=True)
            errs[prec] = (O.rel_l2(got_u, ref_u), O.rel_l2(got_s, ref_s))
            say("   %s oracle %.1f s" % (prec, __import__("time").perf_counter() - t0))
    finally:
        O.set_fft_workers(1)
    say("   exit field rel-L2 vs c64 oracle %.2e, vs f64 oracle %.2e" % (errs["c64"][0], errs["f64"][0]))
    say("   sum |u|^2  rel-L2 vs c64 oracle %.2e, vs f64 oracle %.2e" % (errs["c64"][1], errs["f64"][1]))
    assert errs["c64"][0] < TOL and errs["c64"][1] < TOL
    assert errs["f64"][0] < 1e-3 and errs["f64"][1] < 1e-3


@pytest.mark.parametrize("mode", ["ignore", "weighted"])
def test_distinct_volumes_in_one_batch(bpm, mode):
    """B fields through B different dn volumes in one launch sequence
    (dn_batch_stride = Nz*Ny*Nx): every item equals its own batch-1 run bit
    for bit (exit plane and summed intensity), and one item matches the
    complex64 oracle of bpm_3d.  In weighted mode each item's dn0 comes from
    its own volume's slice 0."""
    import torch
    from bpm_amd import bpm3d as B
    from bpm_amd import _lib
    Nx, Ny, Nz, nb = 256, 128, 40, 3
    say("case distinct volumes", mode)
    dns = np.stack([rng_dn((Nz, Ny, Nx), "f32", seed=30 + b) for b in range(nb)])
    rng = np.random.default_rng(31)
    u0 = np.exp(1j * rng.uniform(0, 2, (nb, Ny, Nx))).astype(np.complex64)
    units = (.1, .1, .1)
    dn_g = torch.from_numpy(dns).cuda()
    H = B.propagator_table(Nx, Ny, units, LAM, N0, _lib.H_F32) if mode == "ignore" else None
    plane = torch.from_numpy(u0).cuda()
    acc = torch.empty((nb, Ny, Nx), dtype=torch.float32, device="cuda")
    B.propagate_fields(plane, Nz, units, LAM, N0, dn=dn_g, dn_kind=_lib.DN_F32,
                       dn_batch_stride=Nz * Ny * Nx, mean_mode=mode, H=H, out_sum=acc)
    for b in range(nb):
        p1 = torch.from_numpy(u0[b:b + 1]).cuda()
        a1 = torch.empty((1, Ny, Nx), dtype=torch.float32, device="cuda")
        B.propagate_fields(p1, Nz, units, LAM, N0, dn=dn_g[b].contiguous(), dn_kind=_lib.DN_F32,
                           mean_mode=mode, H=H, out_sum=a1)
        assert torch.equal(p1[0], plane[b]) and torch.equal(a1[0], acc[b]), b
    b = nb - 1
    ref = O.bpm_3d((Nx, Ny, Nz), units, LAM, u0[b], dns[b], N0, mean_mode=mode,
                   return_full=False)
    e = O.rel_l2(plane[b].cpu().numpy(), ref)
    say("   item %d rel-L2 vs c64 oracle %.2e" % (b, e))
    assert e < TOL
    # the items really differ (distinct volumes were read)
    assert not torch.equal(plane[0], plane[1])


def test_weighted_mode_small_workload_vs_c64_and_f64_oracles(bpm):
    """bpm_3d's default mean_mode="weighted" on the 256^2 x 256 plain workload
    (21 spheres, plane wave, n0 1.33): the exit field (through the public
    bpm_3d) and the summed intensity against the complex64 oracle (bar 1e-4)
    and the float64 oracle (error stated)."""
    import torch
    from bpm_amd import bpm3d as B
    from bpm_amd import _lib
    from bpm_amd.workloads import plain_small_workload
    say("case weighted 256^3")
    units = (.1, .1, .1)
    u0, dn = plain_small_workload(0)
    nz, ny, nx = dn.shape
    last = bpm.bpm_3d((nx, ny, nz), units, LAM, u0=u0, dn=dn, n0=N0, return_full=False)
    plane = _lib.to_device(u0[None], np.complex64)
    acc = torch.empty((1, ny, nx), dtype=torch.float32, device="cuda")
    B.propagate_fields(plane, nz, units, LAM, N0, dn=_lib.to_device(dn, np.float32),
                       dn_kind=_lib.DN_F32, mean_mode="weighted", out_sum=acc)
    assert np.array_equal(plane[0].cpu().numpy(), last)
    got_s = acc[0].cpu().numpy()
    errs = {}
    for prec in ("c64", "f64"):
        ru, rs = O.bpm_3d((nx, ny, nz), units, LAM, u0, dn, N0, mean_mode="weighted",
                          return_full=False, precision=prec, sum_intensity=True)
        errs[prec] = (O.rel_l2(last, ru), O.rel_l2(got_s, rs))
    say("   exit field rel-L2 vs c64 oracle %.2e, vs f64 oracle %.2e" % (errs["c64"][0], errs["f64"][0]))
    say("   sum |u|^2  rel-L2 vs c64 oracle %.2e, vs f64 oracle %.2e" % (errs["c64"][1], errs["f64"][1]))
    assert errs["c64"][0] < TOL and errs["c64"][1] < TOL
    assert errs["f64"][0] < 1e-3 and errs["f64"][1] < 1e-3


def test_ragged_batch_equals_per_group_runs(bpm):
    """bpmh_propagate_groups: three groups of 3, 1 and 2 fields with their own
    dn volumes and depths 40, 25 and 25 in one launch sequence; every field's
    exit plane and summed intensity equal its group's own propagate_fields
    run bit for bit (same plan variant), one field matches the complex64
    oracle of the class path, and each group's event is recorded."""
    import torch
    from bpm_amd import bpm3d as B
    from bpm_amd import _lib
    say("case ragged batch")
    Nx, Ny, units = 256, 128, (.1, .1, .1)
    groups = [(3, 40), (1, 25), (2, 25)]
    nzmax = 40
    rng = np.random.default_rng(41)
    vols = [rng_dn((nzmax, Ny, Nx), "f32", seed=50 + g) for g in range(len(groups))]
    dn = torch.from_numpy(np.stack(vols)).cuda()
    tot = sum(g[0] for g in groups)
    u0 = np.exp(1j * rng.uniform(0, 2, (tot, Ny, Nx))).astype(np.complex64)
    H = B.propagator_table(Nx, Ny, units, LAM, N0, _lib.H_F64)
    plan = B.Plan(Nx, Ny)
    try:
        _lib.check(_lib.load().bpmh_plan_set_variant(plan, 1, 1))
        plane = torch.from_numpy(u0).cuda()
        acc = torch.empty((tot, Ny, Nx), dtype=torch.float32, device="cuda")
        evs = [torch.cuda.Event() for _ in groups]
        B.propagate_groups(plane, groups, units, LAM, dn, H, nzmax * Ny * Nx, out_sum=acc,
                           plan=plan, done_events=evs)
        torch.cuda.synchronize()
        assert all(e.query() for e in evs)
        i0 = 0
        for g, (items, nz) in enumerate(groups):
            p1 = torch.from_numpy(u0[i0:i0 + items]).cuda()
            a1 = torch.empty((items, Ny, Nx), dtype=torch.float32, device="cuda")
            B.propagate_fields(p1, nz, units, LAM, N0, dn=dn[g, :nz].contiguous(),
                               dn_kind=_lib.DN_F32, mean_mode="ignore", H=H, out_sum=a1,
                               plan=plan, tuned=False)
            assert torch.equal(p1, plane[i0:i0 + items]), g
            assert torch.equal(a1, acc[i0:i0 + items]), g
            i0 += items
        ref = O.propagate_single((Nx, Ny, 25), units, LAM, N0, vols[1][:25], u0[3],
                                 return_full=False)
        e = O.rel_l2(plane[3].cpu().numpy(), ref)
        say("   group 1 vs c64 oracle rel-L2 %.2e" % e)
        assert e < TOL
    finally:
        plan.close()
