"""C ABI of libbpmhip.so on the CPU: it loads, exports every function declared
in include/bpm_hip.h, rejects bad arguments before touching a device, and its
Bessel code (the one the Debye kernels run) matches scipy."""
import ctypes
import os
import re

import numpy as np
import pytest
from scipy.special import jv

from bpm_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "bpm_hip.h")


def header_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(bpmh_[a-z_0-9]+)\s*\(", txt)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libbpmhip.so is not built (run python -m bpm_amd.build)")
    return _lib.load()


def test_every_declared_symbol_is_exported_and_bound(lib):
    names = header_functions()
    assert len(names) >= 10
    for n in names:
        assert hasattr(lib, n), n
        assert n in _lib.SIGNATURES, "no ctypes signature for " + n
    assert lib.bpmh_version() == 201


def test_plan_rejects_unsupported_sizes_without_gpu(lib):
    h = ctypes.c_void_p()
    for nx, ny in ((100, 64), (64, 8), (8192, 64), (0, 16)):
        rc = lib.bpmh_plan_create(nx, ny, ctypes.byref(h))
        assert rc == _lib.ERR_UNSUPPORTED
        assert b"powers of two" in lib.bpmh_last_error()
    with pytest.raises(NotImplementedError):
        _lib.check(lib.bpmh_plan_create(48, 64, ctypes.byref(h)))


def test_argument_validation_paths(lib):
    # invalid before any launch: null plan / bad Debye arguments / bad Bessel order
    rc = lib.bpmh_propagate(None, 1, 2, None, None, 0, 0, 0, 0, None, None, None, 1., 1., .1,
                            .1, .1, 1., 0, None, None, 0, None, None, None, None, None, None)
    assert rc == _lib.ERR_INVALID
    al = np.array([0., .5, .6], np.float32)
    rc = lib.bpmh_debye_wolf(8, 8, 8, 0, 1, 0, 1, 0, 1, .5, al.ctypes.data, 3, 100, 1., 0., 0,
                             0., 0, None, None, None, None, None)
    assert rc == _lib.ERR_INVALID
    # sample counts beyond the launch grids' y extent, including an
    # int_steps whose (int_steps + 1) * pairs overflows 32 bits, and too many
    # planes for the product grid are rejected before any allocation
    al = np.array([0., .5], np.float32)
    for steps, nz in ((65000, 8), (2 ** 31 - 1, 8), (100, 1 << 24)):
        rc = lib.bpmh_debye_wolf(8, 8, nz, 0, 1, 0, 1, 0, 1, .5, al.ctypes.data, 2, steps, 1.,
                                 0., 0, 0., 0, None, None, None, None, None)
        assert rc == _lib.ERR_INVALID, (steps, nz)
        assert b"too large" in lib.bpmh_last_error()
    rc = lib.bpmh_debye_mirror(9, 8, 8, None, None, None, None, None, None, None, None, None)
    assert rc == _lib.ERR_UNSUPPORTED
    # the one-call full-grid entry: odd shapes (as _focus_fields_debye.py:81-82),
    # a grid beyond the launch limits and missing aperture angles are refused
    # before any device call
    al = np.array([0., .5], np.float32)
    rc = lib.bpmh_focus_field_debye(9, 8, 8, .1, .1, .1, .5, al.ctypes.data, 2, 64,
                                    None, None, None, None, None)
    assert rc == _lib.ERR_UNSUPPORTED and b"odd" in lib.bpmh_last_error()
    rc = lib.bpmh_focus_field_debye(8, 8, 1 << 17, .1, .1, .1, .5, al.ctypes.data, 2, 64,
                                    None, None, None, None, None)
    assert rc == _lib.ERR_INVALID and b"too large" in lib.bpmh_last_error()
    rc = lib.bpmh_focus_field_debye(8, 8, 8, .1, .1, .1, .5, None, 2, 64,
                                    None, None, None, None, None)
    assert rc == _lib.ERR_INVALID
    rc = lib.bpmh_bessel_fill(3, None, None, 4, None)
    assert rc == _lib.ERR_UNSUPPORTED
    with pytest.raises(ValueError):
        _lib.check(_lib.ERR_INVALID)


def test_bessel_code_matches_scipy_on_host(lib):
    x = np.concatenate([np.linspace(0, 140, 140001), np.linspace(13.99, 14.01, 101),
                        [1e-20, 1e-6, -3.5]])
    for n in range(3):
        out = np.empty_like(x)
        assert lib.bpmh_bessel_host(n, x.ctypes.data, out.ctypes.data, x.size) == 0
        ref = jv(n, x)
        assert np.abs(out - ref).max() < 1e-10
        small = (np.abs(x) < 1) & (x != 0)
        assert np.allclose(out[small], ref[small], rtol=1e-13, atol=0)


def test_propagate_fields_checks_operand_extents_before_launch():
    """propagate_fields hands bare pointers to the library, so it checks every
    tensor against what the passes read and write for this batch, depth and
    grid first: a dn with fewer planes than nz, a short output volume or a
    propagator of the wrong type raise ValueError before any device work
    (CPU tensors here; the GPU check comes last)."""
    import torch
    from bpm_amd import bpm3d as B
    plane = torch.zeros((2, 16, 16), dtype=torch.complex64)
    H = torch.zeros((16, 16), dtype=torch.complex64)
    dn = torch.zeros((3, 16, 16), dtype=torch.float32)
    cases = [
        (dict(dn=dn, dn_kind=_lib.DN_F32, H=H), 129, "dn holds"),  # 3 planes, 129 steps
        (dict(dn=torch.zeros((2, 5, 16, 16)), dn_kind=_lib.DN_F32, dn_batch_stride=5 * 256, H=H), 6,
         "dn holds"),
        (dict(dn=dn.half(), dn_kind=_lib.DN_F32, H=H), 3, "dn must be"),
        (dict(H=H.real.contiguous()), 3, "H must be"),
        (dict(H=H, out_field=torch.zeros((2, 2, 16, 16), dtype=torch.complex64)), 3, "out_field holds"),
        (dict(H=H, out_sum=torch.zeros((1, 16, 16))), 3, "out_sum holds"),
        (dict(H=H, dn=dn, dn_kind=_lib.DN_F32, dn_batch_stride=-1), 3, "dn_batch_stride"),
        (dict(H=H, dn=dn, dn_kind=_lib.DN_F32), 3, "GPU tensor"),  # all extents fine
    ]
    for kw, nz, msg in cases:
        with pytest.raises(ValueError, match=msg):
            B.propagate_fields(plane, nz, (.1, .1, .1), .5, 1.33, mean_mode="ignore", **kw)


def test_library_checks_extents_before_launch(lib):
    """The library itself rejects a dn volume, output volume or recomputed
    segment shorter than the call reads or writes (BPMH_ERR_INVALID, nothing
    launched): the fault class of a 129-step propagation through a 3-plane
    dn.  Plan creation touches no device, so this runs on the CPU; the
    pointers are never dereferenced."""
    h = ctypes.c_void_p()
    assert lib.bpmh_plan_create(64, 32, ctypes.byref(h)) == 0
    try:
        n = 64 * 32
        fake = ctypes.c_void_p(4096)

        def prop(batch, nz, dn_numel, stride=0, out=None, out_numel=0, kind=_lib.DN_F32):
            return lib.bpmh_propagate(h, batch, nz, fake, fake, kind, stride, dn_numel,
                                      _lib.MEAN_IGNORE, fake, None, None, 1.33, 12.5, .1, .1, .1,
                                      1.25, 0, out, None, out_numel, None, None, None, None,
                                      None, None)

        cases = [
            (lambda: prop(1, 129, 3 * n), b"dn holds"),
            (lambda: prop(2, 5, 9 * n, stride=5 * n), b"dn holds"),  # item 1 ends at 10 planes
            (lambda: prop(1, 4, 4 * n, out=fake, out_numel=4 * n - 1), b"output volume holds"),
            (lambda: prop(2, 4, 4 * n, out=fake, out_numel=4 * n), b"output volume holds"),
        ]
        for call, msg in cases:
            rc = call()
            assert rc == _lib.ERR_INVALID and msg in lib.bpmh_last_error(), (rc, msg)
        with pytest.raises(ValueError, match="dn holds"):
            _lib.check(prop(1, 129, 3 * n))
        # adjoint: dn / grad_dn up to plane z_hi, and the recomputed planes
        rc = lib.bpmh_adjoint_segment(h, 2, 0, 16, fake, fake, 2 * 17 * n, fake, _lib.DN_F32,
                                      16 * n, fake, 1.25, 0, fake, fake, None)
        assert rc == _lib.ERR_INVALID and b"dn and grad_dn hold" in lib.bpmh_last_error()
        rc = lib.bpmh_adjoint_segment(h, 2, 0, 16, fake, fake, 17 * n, fake, _lib.DN_F32,
                                      17 * n, fake, 1.25, 0, fake, fake, None)
        assert rc == _lib.ERR_INVALID and b"u_seg holds" in lib.bpmh_last_error()
        rc = lib.bpmh_adjoint_segment(h, 1, 0, 4, fake, fake, 5 * n, fake, _lib.DN_F16,
                                      5 * n, fake, 1.25, 0, fake, fake, None)
        assert rc == _lib.ERR_UNSUPPORTED
    finally:
        assert lib.bpmh_plan_destroy(h) == 0


def test_ragged_batch_arguments_checked_before_launch(lib):
    """bpmh_propagate_groups refuses groups out of depth order, a dn volume
    too short for some group, empty groups and more than 16 groups, before
    touching a device (plan creation is host-only here)."""
    h = ctypes.c_void_p()
    assert lib.bpmh_plan_create(32, 32, ctypes.byref(h)) == 0
    try:
        n = 32 * 32
        fake = ctypes.c_void_p(4096)

        def call(items, nzs, stride, numel):
            k = len(items)
            it = (ctypes.c_int * k)(*items)
            nz = (ctypes.c_int * k)(*nzs)
            return lib.bpmh_propagate_groups(h, k, it, nz, fake, fake, _lib.DN_F32, stride, numel,
                                             fake, 1.25, 0, None, None, None)

        cases = [
            (lambda: call([2, 3], [10, 20], 20 * n, 40 * n), b"non-increasing depth"),
            (lambda: call([2, 3], [20, 10], 20 * n, 29 * n), b"group 1 reads"),
            (lambda: call([2, 0], [20, 10], 20 * n, 40 * n), b">= 1 item"),
            (lambda: call([1] * 17, [5] * 17, 5 * n, 17 * 5 * n), b"1 to 16 groups"),
        ]
        for fn, msg in cases:
            rc = fn()
            assert rc == _lib.ERR_INVALID and msg in lib.bpmh_last_error(), (rc, msg)
    finally:
        assert lib.bpmh_plan_destroy(h) == 0


def test_mean_segment_and_floor_checked_before_launch(lib):
    """bpmh_mean_segment refuses a dn volume short of plane z_hi, a work
    buffer below what bpmh_mean_segment_work reports, an adjoint call
    without grad_dn, the simple mean without its arrays and an unknown mean
    mode; bpmh_floor_steps refuses a dn of fewer than nz planes -- all before
    touching a device."""
    h = ctypes.c_void_p()
    assert lib.bpmh_plan_create(64, 32, ctypes.byref(h)) == 0
    try:
        n = 64 * 32
        fake = ctypes.c_void_p(4096)
        need = ctypes.c_longlong()
        assert lib.bpmh_mean_segment_work(h, 2, 0, 8, 0, ctypes.byref(need)) == 0
        assert need.value == 256 + 64  # 2*batch*nx partials + batch dn0, 256-byte aligned parts
        assert lib.bpmh_mean_segment_work(h, 2, 0, 8, 1, ctypes.byref(need)) == 0
        adj = need.value
        # recorded planes (9), k-space planes (8), the forward copy (c64) and two f32 planes
        assert adj >= 2 * 2 * n * (9 + 8 + 1) + 2 * 2 * n

        def seg(z_hi, dn_numel, g, grad, work_numel, mode=_lib.MEAN_WEIGHTED, arrays=None):
            return lib.bpmh_mean_segment(h, mode, 2, 0, z_hi, fake, fake, dn_numel, arrays, arrays,
                                         1.33, 12.5, .1, .1, .1, 1.25, 0, g, grad, fake,
                                         work_numel, None)

        cases = [
            (lambda: seg(8, 8 * n, None, None, 10 ** 9), b"dn holds"),
            (lambda: seg(8, 9 * n, fake, fake, adj - 1), b"work holds"),
            (lambda: seg(8, 9 * n, fake, None, adj), b"needs grad_dn"),
            (lambda: seg(8, 9 * n, None, None, adj, mode=_lib.MEAN_SIMPLE), b"needs n_eff"),
            (lambda: seg(8, 9 * n, None, None, adj, mode=_lib.MEAN_IGNORE, arrays=fake),
             b"simple or weighted"),
            (lambda: lib.bpmh_floor_steps(h, 1, 4, fake, fake, 3 * n, None), b"fewer than nz"),
        ]
        for fn, msg in cases:
            rc = fn()
            assert rc == _lib.ERR_INVALID and msg in lib.bpmh_last_error(), (rc, msg)
    finally:
        assert lib.bpmh_plan_destroy(h) == 0
