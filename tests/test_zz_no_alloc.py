"""bpmh_propagate_noalloc through the C ABI: its step scratch is a fixed
device array instead of plan memory, so its results must equal
bpmh_propagate's bit for bit (the same kernels).  Two calls in a row, the
second on a smaller grid and batch, with the weighted mean and scattering
sums (the paths that use the scratch); a call whose scratch would exceed the
fixed array is refused before anything is launched."""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LAM, N0, UNITS = .5, 1.33, (.1, .1, .1)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _call(lib, name, plan, plane, nz, dn, mode, scat):
    from bpm_amd import _lib
    B, ny, nx = plane.shape
    k0 = 2 * np.pi / LAM
    f = lambda v: float(np.float32(v))
    sw, gw, plain, sout = scat
    return getattr(lib, name)(plan, B, nz, _lib.ptr(plane), _lib.ptr(dn), _lib.DN_F32,
                              nz * ny * nx, dn.numel(), mode, None, None, None, f(N0), f(k0),
                              f(UNITS[0]), f(UNITS[1]), f(UNITS[2]), f(k0 * UNITS[2]), 4, None,
                              None, 0, None, _lib.ptr(sw), _lib.ptr(gw), _lib.ptr(plain),
                              _lib.ptr(sout), _lib.stream_handle())


def _inputs(torch, B, ny, nx, nz, seed):
    rng = np.random.default_rng(seed)
    dn = torch.from_numpy(rng.uniform(0, .05, (B, nz, ny, nx)).astype(np.float32)).cuda()
    u0 = torch.from_numpy(np.exp(1j * rng.uniform(0, 3, (B, ny, nx))).astype(np.complex64)).cuda()
    sw = torch.from_numpy(rng.uniform(0, 1, (ny, nx)).astype(np.float32)).cuda()
    gw = torch.from_numpy(rng.uniform(-1, 1, (ny, nx)).astype(np.float32)).cuda()
    plain = torch.from_numpy(np.exp(1j * rng.uniform(0, 3, nz)).astype(np.complex64)).cuda()
    return dn, u0, sw, gw, plain


def test_noalloc_equals_propagate_twice_then_smaller():
    import torch
    from bpm_amd import _lib
    from bpm_amd import bpm3d as Bm
    _lib.torch()
    lib = _lib.load()
    for (B, ny, nx, nz, seed) in ((3, 256, 512, 24, 1), (2, 128, 128, 12, 2)):
        say("case noalloc", B, ny, nx, nz)
        dn, u0, sw, gw, plain = _inputs(torch, B, ny, nx, nz, seed)
        plan = Bm.Plan(nx, ny)
        try:
            outs = {}
            for name in ("bpmh_propagate", "bpmh_propagate_noalloc"):
                p = u0.clone()
                sout = torch.zeros((B, nz, 2), dtype=torch.float32, device="cuda")
                _lib.check(_call(lib, name, plan, p, nz, dn, _lib.MEAN_WEIGHTED,
                                 (sw, gw, plain, sout)))
                torch.cuda.synchronize()
                outs[name] = (p, sout)
            a, b = outs["bpmh_propagate"], outs["bpmh_propagate_noalloc"]
            assert torch.equal(a[0], b[0]) and torch.equal(a[1], b[1])
            assert a[1].abs().sum() > 0
        finally:
            plan.close()


def test_noalloc_refuses_scratch_beyond_its_array():
    import torch
    from bpm_amd import _lib
    from bpm_amd import bpm3d as Bm
    _lib.torch()
    lib = _lib.load()
    # scattering sums need 8 * batch * nz * nx bytes: 1 x 2100 x 1024 > 16 MiB
    B, ny, nx, nz = 1, 64, 1024, 2100
    dn = torch.zeros((nz, ny, nx), dtype=torch.float32, device="cuda")
    plane = torch.zeros((B, ny, nx), dtype=torch.complex64, device="cuda")
    sw = torch.zeros((ny, nx), dtype=torch.float32, device="cuda")
    plain = torch.zeros(nz, dtype=torch.complex64, device="cuda")
    sout = torch.zeros((B, nz, 2), dtype=torch.float32, device="cuda")
    plan = Bm.Plan(nx, ny)
    try:
        rc = _call(lib, "bpmh_propagate_noalloc", plan, plane, nz, dn, _lib.MEAN_WEIGHTED,
                   (sw, sw, plain, sout))
        assert rc == _lib.ERR_INVALID and b"fixed scratch holds" in lib.bpmh_last_error()
    finally:
        plan.close()
