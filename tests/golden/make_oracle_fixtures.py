"""Write the small golden fixtures of the oracle (tests/golden/oracle_fixtures.npz).

Inputs are seeded; outputs come from the numpy oracle (oracle/), which is
pinned by tests/test_oracle_*.py.  The fixtures freeze those outputs so that
a later change of the oracle shows up (tests/test_golden.py) and so that the
GPU tests can compare the HIP path against stored vectors.
Run from the repository root:  python tests/golden/make_oracle_fixtures.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import bpm_oracle as O  # noqa: E402
from oracle import psf_oracle as P  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "oracle_fixtures.npz")
SIZE = (32, 16, 12)
UNITS = (.1, .1, .1)


def inputs(seed=0):
    rng = np.random.default_rng(seed)
    Nx, Ny, Nz = SIZE
    u0 = np.exp(1j * rng.uniform(0, 1, (Ny, Nx))).astype(np.complex64)
    dn = (0.04 * rng.uniform(0, 1, (Nz, Ny, Nx))).astype(np.float32)
    return u0, dn


def cases():
    """(name, kwargs of oracle.bpm_3d)"""
    out = []
    for mode in ("ignore", "simple", "weighted"):
        for kind in ("none", "f32", "f16", "c64"):
            out.append(("bpm_%s_%s" % (mode, kind), dict(mean_mode=mode, kind=kind, absorb=3)))
    return out


def dn_for(kind, dn):
    if kind == "none":
        return None
    if kind == "c64":
        return (dn - 0.3j * dn).astype(np.complex64)
    return dn


def main():
    u0, dn = inputs()
    d = {"u0": u0, "dn": dn}
    for name, c in cases():
        d[name] = O.bpm_3d(SIZE, UNITS, .5, u0, dn_for(c["kind"], dn), 1.33,
                           mean_mode=c["mean_mode"], absorbing_width=c["absorb"],
                           store_dn_as_half=c["kind"] == "f16", return_full=False)
    d["class_f32"] = O.propagate_single(SIZE, UNITS, .5, 1.33, dn, u0, return_full=False,
                                        absorbing_width=2)
    u, ex, ey, ez = P.focus_field_debye((16, 16, 16), (.05, .05, .05), .5, 1.1, n0=1.33,
                                        n_integration_steps=64)
    d.update(debye_u=u, debye_ex=ex, debye_ey=ey, debye_ez=ez)
    d["debye_plane_ex"] = P.focus_field_debye_plane((16, 12), (.1, .1), z=-.7, lam=.5, NA=.8,
                                                    n0=1.1, n_integration_steps=50)
    np.savez_compressed(OUT, **d)
    print("wrote", OUT, len(d), "arrays")


if __name__ == "__main__":
    main()
