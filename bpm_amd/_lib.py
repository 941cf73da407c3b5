"""ctypes binding of libbpmhip.so and the device-memory plumbing around it.

The library is loaded from this package directory (built in-tree by
``python -m bpm_amd.build``).  There is no fallback: if the library is
missing, or no ROCm GPU is visible, every compute call raises.
Error codes of the C ABI map onto the exception types the reference raises
(ValueError, NotImplementedError) or RuntimeError/MemoryError.
"""
import ctypes
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libbpmhip.so")

OK, ERR_INVALID, ERR_UNSUPPORTED, ERR_HIP, ERR_NOMEM = 0, -1, -2, -3, -4
DN_NONE, DN_F32, DN_F16, DN_C64 = 0, 1, 2, 3
MEAN_IGNORE, MEAN_SIMPLE, MEAN_WEIGHTED = 0, 1, 2
H_F32, H_F64 = 0, 1
DW_INDEX, DW_SIN, DW_GAUSS = 0, 1, 2

_P = ctypes.c_void_p
_I = ctypes.c_int
_F = ctypes.c_float
_D = ctypes.c_double
_LL = ctypes.c_longlong

# name -> (restype, argtypes); every symbol declared in include/bpm_hip.h
SIGNATURES = {
    "bpmh_version": (_I, []),
    "bpmh_last_error": (ctypes.c_char_p, []),
    "bpmh_plan_create": (_I, [_I, _I, ctypes.POINTER(_P)]),
    "bpmh_plan_destroy": (_I, [_P]),
    "bpmh_plan_set_timing": (_I, [_P, _I]),
    "bpmh_plan_kernel_times": (_I, [_P, _P, _P]),
    "bpmh_plan_set_graph": (_I, [_P, _I]),
    "bpmh_plan_set_variant": (_I, [_P, _I, _I]),
    "bpmh_plan_get_variant": (_I, [_P, _P, _P, _P, _P, _P, _P]),
    "bpmh_propagator_table": (_I, [_I, _I, _I, _D, _D, _D, _D, _D, _P, _P]),
    "bpmh_propagate": (_I, [_P, _I, _I, _P, _P, _I, _LL, _LL, _I, _P, _P, _P, _F, _F, _F, _F,
                            _F, _F, _I, _P, _P, _LL, _P, _P, _P, _P, _P, _P]),
    "bpmh_propagate_noalloc": (_I, [_P, _I, _I, _P, _P, _I, _LL, _LL, _I, _P, _P, _P, _F, _F,
                                    _F, _F, _F, _F, _I, _P, _P, _LL, _P, _P, _P, _P, _P, _P]),
    "bpmh_propagate_groups": (_I, [_P, _I, _P, _P, _P, _P, _I, _LL, _LL, _P, _F, _I, _P, _P,
                                   _P]),
    "bpmh_adjoint_segment": (_I, [_P, _I, _I, _I, _P, _P, _LL, _P, _I, _LL, _P, _F, _I, _P, _P,
                                  _P]),
    "bpmh_mean_segment": (_I, [_P, _I, _I, _I, _I, _P, _P, _LL, _P, _P, _F, _F, _F, _F, _F, _F,
                               _I, _P, _P, _P, _LL, _P]),
    "bpmh_mean_segment_work": (_I, [_P, _I, _I, _I, _I, _P]),
    "bpmh_floor_steps": (_I, [_P, _I, _I, _P, _P, _LL, _P]),
    "bpmh_debye_wolf": (_I, [_I, _I, _I, _F, _F, _F, _F, _F, _F, _F, _P, _I, _I, _F, _F, _I,
                             _F, _I, _P, _P, _P, _P, _P]),
    "bpmh_debye_wolf_points": (_I, [_I, _P, _P, _P, _F, _P, _I, _I, _F, _F, _I, _F, _I, _P,
                                    _P, _P, _P, _P]),
    "bpmh_debye_mirror": (_I, [_I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "bpmh_focus_field_debye": (_I, [_I, _I, _I, _D, _D, _D, _F, _P, _I, _I, _P, _P, _P, _P, _P]),
    "bpmh_debye_set_timing": (_I, [_I]),
    "bpmh_debye_kernel_times": (_I, [_P, _P]),
    "bpmh_bessel_fill": (_I, [_I, _P, _P, _I, _P]),
    "bpmh_cylindrical": (_I, [_I, _I, _F, _F, _F, _F, _F, _F, _I, _I, _P, _P, _P, _P, _P]),
    "bpmh_fill_dn_coated": (_I, [_I, _I, _I, _F, _F, _F, _P, _I, _F, _F, _F, _F, _P, _P]),
    "bpmh_spheres_dn": (_I, [_I, _I, _I, _D, _D, _D, _P, _I, _P, _P]),
    "bpmh_bessel_host": (_I, [_I, _P, _P, _I]),
}

_lib = None


def load():
    """Load libbpmhip.so (once) and declare its signatures; raises if absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libbpmhip.so not found at %s: build it with "
                              "`python -m bpm_amd.build`" % LIB_PATH)
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(rc):
    """Raise the Python exception that corresponds to a C-ABI return code."""
    if rc == OK:
        return
    msg = load().bpmh_last_error().decode(errors="replace")
    if rc == ERR_INVALID:
        raise ValueError(msg)
    if rc == ERR_UNSUPPORTED:
        raise NotImplementedError(msg)
    if rc == ERR_NOMEM:
        raise MemoryError(msg)
    raise RuntimeError("libbpmhip error %d: %s" % (rc, msg))


def call(name, *args):
    check(getattr(load(), name)(*args))


# ----------------------------------------------------------------------------
# device memory (torch is only the allocator / stream provider)
# ----------------------------------------------------------------------------
_torch = None


def torch():
    global _torch
    if _torch is None:
        import torch as t
        if not t.cuda.is_available():
            raise RuntimeError("bpm_amd needs a ROCm GPU: torch.cuda.is_available() is False")
        load()
        _torch = t
    return _torch


def device():
    return torch().device("cuda", torch().cuda.current_device())


def stream_handle():
    return ctypes.c_void_p(torch().cuda.current_stream().cuda_stream)


_TDT = {np.dtype(np.float32): "float32", np.dtype(np.float16): "float16",
        np.dtype(np.complex64): "complex64"}


def to_device(x, np_dtype):
    """numpy array or torch tensor -> contiguous torch tensor on the GPU with dtype np_dtype."""
    t = torch()
    tdt = getattr(t, _TDT[np.dtype(np_dtype)])
    if isinstance(x, t.Tensor):
        return x.to(device=device(), dtype=tdt).contiguous()
    a = np.ascontiguousarray(np.asarray(x).astype(np_dtype, copy=False))
    if not a.flags.writeable:  # torch.from_numpy wants a writable array
        a = a.copy()
    return t.from_numpy(a).to(device())


def empty(shape, np_dtype):
    t = torch()
    return t.empty(tuple(shape), dtype=getattr(t, _TDT[np.dtype(np_dtype)]), device=device())


def ptr(x):
    """Device pointer of a tensor (or None -> NULL)."""
    if x is None:
        return None
    return ctypes.c_void_p(x.data_ptr())


def is_tensor(x):
    try:
        import torch as t
    except ImportError:
        return False
    return isinstance(x, t.Tensor)


def to_numpy(x):
    return x.cpu().numpy()
