"""GPU checks of the hand-written adjoint (bpm_amd.adjoint) against a float64
torch autograd reference of the same split-step map and finite differences."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LAM, N0 = .5, 1.33


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def reference_f64(dn, u0, units, lam, n0, absorbing_width=0):
    """u_z = exp(-i k0 dz dn_z) T * IFFT2(H FFT2(u_{z-1})) in complex128 (torch
    autograd); complex dn (absorbing, Im dn < 0 decays) has no taper, as in
    the forward pass (bpm/bpmclass/kernels/bpm_3d_kernels.cl mult_dn_complex)."""
    import torch as t
    nz, ny, nx = dn.shape
    dx, dy, dz = units
    k0 = 2. * np.pi / lam
    kx = 2. * np.pi * np.fft.fftfreq(nx, dx)
    ky = 2. * np.pi * np.fft.fftfreq(ny, dy)
    tmp = n0 ** 2 * k0 ** 2 - kx[None, :] ** 2 - ky[:, None] ** 2
    H = t.from_numpy(np.where(tmp >= 0, np.exp(-1j * dz * np.sqrt(np.maximum(tmp, 0))), 0)).to(dn.device)
    T = None
    if absorbing_width > 0 and not dn.is_complex():
        i, j = np.arange(nx), np.arange(ny)
        d = np.minimum(np.minimum(nx - 1 - i, i)[None, :], np.minimum(ny - 1 - j, j)[:, None])
        T = t.from_numpy(np.where(d < absorbing_width,
                                  .5 * (1 - np.cos(np.pi * d / absorbing_width)), 1.)).to(dn.device)
    u = u0.to(t.complex128)
    for z in range(1, nz):
        u = t.fft.ifft2(t.fft.fft2(u) * H)
        u = u * t.exp(-1j * k0 * dz * dn[z])
        if T is not None:
            u = u * T
    return u


def setup(nz, ny, nx, Bn, seed=0, units=(.1, .1, .1)):
    """dn and target dn with the 256^2 workload's sphere recipe (a centred
    r = 2 um sphere plus 20 random ones), and Bn tilted plane waves."""
    import torch as t
    from bpm_amd.spheres import spheres_dn_device
    from bpm_amd.workloads import small_workload_spheres
    rng = np.random.default_rng(seed)
    size = (nx, ny, nz)
    dn = spheres_dn_device(size, units, *small_workload_spheres(rng, size, units))
    dn2 = spheres_dn_device(size, units, *small_workload_spheres(rng, size, units))
    th = rng.uniform(-.2, .2, (Bn, 2))
    x = np.arange(nx) * units[0]
    y = np.arange(ny) * units[1]
    k = 2 * np.pi * N0 / LAM
    u0 = np.exp(-1j * k * (np.sin(th[:, 0])[:, None, None] * x[None, None, :] +
                           np.sin(th[:, 1])[:, None, None] * y[None, :, None])).astype(np.complex64)
    return dn, dn2, t.from_numpy(u0).cuda()


@pytest.mark.parametrize("absorb", [0, 5])
def test_adjoint_vs_float64_autograd(absorb):
    import torch as t
    from bpm_amd import adjoint
    nz, ny, nx, Bn = 24, 32, 64, 3
    say("case adjoint vs f64 autograd absorb=%d" % absorb)
    dn, dn2, u0 = setup(nz, ny, nx, Bn, seed=1)
    target = adjoint.propagate(dn2, u0, lam=LAM, n0=N0, absorbing_width=absorb).detach()
    dn_a = dn.clone().requires_grad_(True)
    u0_a = u0.clone().requires_grad_(True)
    out = adjoint.propagate(dn_a, u0_a, lam=LAM, n0=N0, absorbing_width=absorb, checkpoint=5)
    loss = (out - target).abs().pow(2).sum()
    loss.backward()
    dn_r = dn.double().clone().requires_grad_(True)
    u0_r = u0.to(t.complex128).clone().requires_grad_(True)
    ref = reference_f64(dn_r, u0_r, (.1, .1, .1), LAM, N0, absorb)
    loss_r = (ref - target.to(t.complex128)).abs().pow(2).sum()
    loss_r.backward()
    g, gr = dn_a.grad.double(), dn_r.grad
    e = ((g - gr).norm() / gr.norm()).item()
    eu = ((u0_a.grad.to(t.complex128) - u0_r.grad).norm() / u0_r.grad.norm()).item()
    say("   loss %.6e vs f64 %.6e; grad_dn rel-L2 %.2e; grad_u0 rel-L2 %.2e"
        % (loss.item(), loss_r.item(), e, eu))
    assert abs(loss.item() - loss_r.item()) / loss_r.item() < 1e-4
    assert e < 1e-4 and eu < 1e-4
    assert g[0].abs().max().item() == 0  # dn[0] does not enter the propagation


def test_adjoint_complex_dn_vs_float64_autograd():
    """Absorbing (complex) dn: the gradient in Re dn and Im dn (returned as
    dL/dRe + i dL/dIm, torch's convention for complex leaves) and in u0,
    against float64 autograd of the same map.  The loss compares with
    targets through a second, real dn."""
    import torch as t
    from bpm_amd import adjoint
    nz, ny, nx, Bn = 20, 32, 64, 2
    say("case adjoint complex dn vs f64 autograd")
    dn, dn2, u0 = setup(nz, ny, nx, Bn, seed=4)
    dnc = (dn - 0.3j * dn).to(t.complex64).contiguous()
    target = adjoint.propagate(dn2, u0, lam=LAM, n0=N0).detach()
    dn_a = dnc.clone().requires_grad_(True)
    u0_a = u0.clone().requires_grad_(True)
    out = adjoint.propagate(dn_a, u0_a, lam=LAM, n0=N0, checkpoint=6)
    loss = (out - target).abs().pow(2).sum()
    loss.backward()
    dn_r = dnc.to(t.complex128).clone().requires_grad_(True)
    u0_r = u0.to(t.complex128).clone().requires_grad_(True)
    ref = reference_f64(dn_r, u0_r, (.1, .1, .1), LAM, N0)
    loss_r = (ref - target.to(t.complex128)).abs().pow(2).sum()
    loss_r.backward()
    g, gr = dn_a.grad.to(t.complex128), dn_r.grad
    e = ((g - gr).norm() / gr.norm()).item()
    e_im = ((g.imag - gr.imag).norm() / gr.imag.norm()).item()
    eu = ((u0_a.grad.to(t.complex128) - u0_r.grad).norm() / u0_r.grad.norm()).item()
    say("   loss %.6e vs f64 %.6e; grad_dn rel-L2 %.2e (Im part %.2e); grad_u0 rel-L2 %.2e"
        % (loss.item(), loss_r.item(), e, e_im, eu))
    assert dn_a.grad.dtype == t.complex64
    assert abs(loss.item() - loss_r.item()) / loss_r.item() < 1e-4
    assert e < 1e-4 and e_im < 1e-4 and eu < 1e-4
    assert gr.imag.abs().max().item() > 0


def test_checkpoint_interval_does_not_change_gradient():
    from bpm_amd import adjoint
    nz, ny, nx, Bn = 30, 32, 32, 2
    dn, dn2, u0 = setup(nz, ny, nx, Bn, seed=2)
    grads = []
    for K in (1, 7, 29, 64):
        d = dn.clone().requires_grad_(True)
        out = adjoint.propagate(d, u0, lam=LAM, n0=N0, checkpoint=K)
        (out - u0).abs().pow(2).sum().backward()
        grads.append(d.grad.clone())
    for g in grads[1:]:
        assert (g == grads[0]).all()


def test_adjoint_full_size_vs_finite_differences():
    # the differentiable workload: 512^2 x 256 slices, 16 angles, loss = L2
    # between exit fields and targets from a second seeded dn, checkpoint 16;
    # float64 central differences of the float64 reference on 32 voxels
    import torch as t
    from bpm_amd import adjoint
    nz, ny, nx, Bn = 256, 512, 512, 16
    say("case adjoint 512^2 x 256 x 16")
    dn, dn2, u0 = setup(nz, ny, nx, Bn, seed=3)
    target = adjoint.propagate(dn2, u0, lam=LAM, n0=N0).detach().to(t.complex128)
    d = dn.clone().requires_grad_(True)
    out = adjoint.propagate(d, u0, lam=LAM, n0=N0, checkpoint=16)
    ((out.to(t.complex128) - target).abs().pow(2).sum()).backward()
    g = d.grad.double()
    rng = np.random.default_rng(5)
    # voxels inside spheres and outside, planes 1..nz-1
    nzv = t.nonzero(dn[1:] > 0)
    pick = nzv[t.from_numpy(rng.choice(nzv.shape[0], 16, replace=False)).cuda()]
    pick[:, 0] += 1
    rand = t.from_numpy(np.stack([rng.integers(1, nz, 16), rng.integers(0, ny, 16),
                                  rng.integers(0, nx, 16)], 1)).cuda()
    vox = t.cat([pick, rand])
    base = dn.double()
    eps = 1e-4
    errs = []
    with t.no_grad():
        for (z, y, x) in vox.tolist():
            vals = []
            for sgn in (1, -1):
                p = base.clone()
                p[z, y, x] += sgn * eps
                r = reference_f64(p, u0, (.1, .1, .1), LAM, N0)
                vals.append((r - target).abs().pow(2).sum().item())
            fd = (vals[0] - vals[1]) / (2 * eps)
            errs.append((g[z, y, x].item(), fd))
    a = np.array(errs)
    scale = np.abs(a[:, 1]).max()
    err = np.abs(a[:, 0] - a[:, 1]).max() / scale
    say("   max |adjoint - FD| / max|FD| over 32 voxels: %.2e" % err)
    assert err < 1e-3


def reference_mean_f64(dn, u0, units, lam, n0, absorbing_width=0, mode="weighted"):
    """bpm_3d's mean corrections in complex128 with autograd.  Weighted (the
    default): per step U = FFT2(u), dn0 = sum|U| d / sum|U| with d = dn[0]
    read at U's flat index (bpm/bpm/bpm_3d.py:149-159,302-311); simple:
    dn0 = mean(dn_z) of the step's target plane (bpm_3d.py:216-217,288-293).
    Then u = exp(-i k0 dz (dn_z - dn0)) T * IFFT2(H(n0 + dn0) U) with
    H(n) = exp(-i dz sqrt(n^2 k0^2 - kx^2 - ky^2)) (0 where negative)."""
    import torch as t
    nz, ny, nx = dn.shape
    dx, dy, dz = units
    k0 = 2. * np.pi / lam
    kx = 2. * np.pi * np.fft.fftfreq(nx, dx)
    ky = 2. * np.pi * np.fft.fftfreq(ny, dy)
    k2 = t.from_numpy(kx[None, :] ** 2 + ky[:, None] ** 2).to(dn.device)
    T = None
    if absorbing_width > 0:
        i, j = np.arange(nx), np.arange(ny)
        d = np.minimum(np.minimum(nx - 1 - i, i)[None, :], np.minimum(ny - 1 - j, j)[:, None])
        T = t.from_numpy(np.where(d < absorbing_width,
                                  .5 * (1 - np.cos(np.pi * d / absorbing_width)), 1.)).to(dn.device)
    u = u0.to(t.complex128)
    for z in range(1, nz):
        U = t.fft.fft2(u)
        if mode == "weighted":
            a = U.abs()
            dn0 = (a * dn[0]).sum((-2, -1)) / a.sum((-2, -1))
        else:
            dn0 = dn[z].mean().expand(u.shape[0])
        tmp = ((n0 + dn0) ** 2 * k0 ** 2)[:, None, None] - k2
        beta = t.sqrt(t.clamp(tmp, min=1e-30))
        H = t.where(tmp > 0, t.exp(-1j * dz * beta), t.zeros_like(beta, dtype=t.complex128))
        u = t.fft.ifft2(H * U) * t.exp(-1j * k0 * dz * (dn[z] - dn0[:, None, None]))
        if T is not None:
            u = u * T
    return u


@pytest.mark.parametrize("absorb", [0, 5])
def test_weighted_adjoint_vs_float64_autograd(absorb):
    """The weighted-mean adjoint (dn0 from the field's k-space magnitudes and
    dn slice 0): loss, dL/ddn (every plane, slice 0 included: it enters only
    through the mean's weights) and dL/du0 against float64 autograd of the
    same map; the checkpointed forward equals bpm_3d's weighted propagation
    through the whole volume bit for bit."""
    import torch as t
    from bpm_amd import adjoint, bpm3d as B, _lib
    nz, ny, nx, Bn = 20, 32, 64, 3
    say("case weighted adjoint vs f64 autograd absorb=%d" % absorb)
    dn, dn2, u0 = setup(nz, ny, nx, Bn, seed=7)
    target = adjoint.propagate(dn2, u0, lam=LAM, n0=N0, absorbing_width=absorb,
                               mean_mode="weighted").detach()
    # forward: checkpointed segments == one weighted propagation of the volume
    plane = u0.clone()
    B.propagate_fields(plane, nz, (.1, .1, .1), LAM, N0, dn=dn, dn_kind=_lib.DN_F32,
                       mean_mode="weighted", absorbing_width=absorb)
    out0 = adjoint.propagate(dn, u0, lam=LAM, n0=N0, absorbing_width=absorb, checkpoint=6,
                             mean_mode="weighted")
    assert t.equal(out0, plane), "checkpointed weighted forward differs from bpmh_propagate"
    dn_a = dn.clone().requires_grad_(True)
    u0_a = u0.clone().requires_grad_(True)
    out = adjoint.propagate(dn_a, u0_a, lam=LAM, n0=N0, absorbing_width=absorb, checkpoint=6,
                            mean_mode="weighted")
    loss = (out - target).abs().pow(2).sum()
    loss.backward()
    dn_r = dn.double().clone().requires_grad_(True)
    u0_r = u0.to(t.complex128).clone().requires_grad_(True)
    ref = reference_mean_f64(dn_r, u0_r, (.1, .1, .1), LAM, N0, absorb)
    loss_r = (ref - target.to(t.complex128)).abs().pow(2).sum()
    loss_r.backward()
    g, gr = dn_a.grad.double(), dn_r.grad
    e = ((g - gr).norm() / gr.norm()).item()
    e0 = ((g[0] - gr[0]).norm() / gr[0].norm()).item()
    eu = ((u0_a.grad.to(t.complex128) - u0_r.grad).norm() / u0_r.grad.norm()).item()
    say("   loss %.6e vs f64 %.6e; grad_dn rel-L2 %.2e (slice 0: %.2e, |g0|/|g| %.2e); "
        "grad_u0 rel-L2 %.2e" % (loss.item(), loss_r.item(), e, e0,
                                 (gr[0].norm() / gr.norm()).item(), eu))
    assert abs(loss.item() - loss_r.item()) / loss_r.item() < 1e-4
    assert e < 1e-4 and eu < 1e-4
    assert e0 < 1e-3
    # the same gradient with every checkpoint (segments of one step): planes
    # 1.. bit for bit; slice 0 sums its per-step terms in another grouping
    dn_b = dn.clone().requires_grad_(True)
    out_b = adjoint.propagate(dn_b, u0, lam=LAM, n0=N0, absorbing_width=absorb, checkpoint=1,
                              mean_mode="weighted")
    (out_b - target).abs().pow(2).sum().backward()
    assert t.equal(dn_b.grad[1:], dn_a.grad[1:])
    assert ((dn_b.grad[0] - dn_a.grad[0]).norm() / dn_a.grad[0].norm()).item() < 1e-5


def test_simple_mean_adjoint_vs_float64_autograd():
    """mean_mode "simple" (dn0 = mean of the step's target plane): the
    checkpointed forward equals bpm_3d's simple propagation bit for bit, and
    loss, dL/ddn (the mean's share included: dL/ddn0 / N on every voxel of
    the plane) and dL/du0 match float64 autograd of the same map."""
    import torch as t
    from bpm_amd import adjoint, bpm3d as B, _lib
    nz, ny, nx, Bn = 18, 32, 64, 2
    say("case simple-mean adjoint vs f64 autograd")
    dn, dn2, u0 = setup(nz, ny, nx, Bn, seed=9)
    target = adjoint.propagate(dn2, u0, lam=LAM, n0=N0, mean_mode="simple").detach()
    # forward == bpmh_propagate with the same per-plane arrays
    n_eff, dn0 = adjoint._simple_means(dn, N0)
    plane = u0.clone()
    B.propagate_fields(plane, nz, (.1, .1, .1), LAM, N0, dn=dn, dn_kind=_lib.DN_F32,
                       mean_mode="simple", n_eff=n_eff, dn0=dn0)
    out0 = adjoint.propagate(dn, u0, lam=LAM, n0=N0, checkpoint=5, mean_mode="simple")
    assert t.equal(out0, plane), "checkpointed simple forward differs from bpmh_propagate"
    dn_a = dn.clone().requires_grad_(True)
    u0_a = u0.clone().requires_grad_(True)
    out = adjoint.propagate(dn_a, u0_a, lam=LAM, n0=N0, checkpoint=5, mean_mode="simple")
    loss = (out - target).abs().pow(2).sum()
    loss.backward()
    dn_r = dn.double().clone().requires_grad_(True)
    u0_r = u0.to(t.complex128).clone().requires_grad_(True)
    ref = reference_mean_f64(dn_r, u0_r, (.1, .1, .1), LAM, N0, mode="simple")
    loss_r = (ref - target.to(t.complex128)).abs().pow(2).sum()
    loss_r.backward()
    g, gr = dn_a.grad.double(), dn_r.grad
    e = ((g - gr).norm() / gr.norm()).item()
    eu = ((u0_a.grad.to(t.complex128) - u0_r.grad).norm() / u0_r.grad.norm()).item()
    # the mean's share: the plane-average of the gradient, against autograd's
    em = ((g.mean(dim=(1, 2)) - gr.mean(dim=(1, 2))).norm() / gr.mean(dim=(1, 2)).norm()).item()
    say("   loss %.6e vs f64 %.6e; grad_dn rel-L2 %.2e (plane means %.2e); grad_u0 rel-L2 %.2e"
        % (loss.item(), loss_r.item(), e, em, eu))
    assert abs(loss.item() - loss_r.item()) / loss_r.item() < 1e-4
    assert e < 1e-4 and eu < 1e-4 and em < 1e-3
    assert g[0].abs().max().item() == 0  # dn[0] enters no step
