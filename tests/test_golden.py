"""The oracle still reproduces its committed golden fixtures (CPU), and the
HIP path matches them (GPU)."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_oracle_fixtures as G  # noqa: E402

from oracle import bpm_oracle as O  # noqa: E402
from oracle import psf_oracle as P  # noqa: E402

FIX = np.load(os.path.join(HERE, "golden", "oracle_fixtures.npz"))


def test_oracle_reproduces_fixtures():
    u0, dn = FIX["u0"], FIX["dn"]
    for name, c in G.cases():
        got = O.bpm_3d(G.SIZE, G.UNITS, .5, u0, G.dn_for(c["kind"], dn), 1.33,
                       mean_mode=c["mean_mode"], absorbing_width=c["absorb"],
                       store_dn_as_half=c["kind"] == "f16", return_full=False)
        assert O.rel_l2(got, FIX[name]) < 1e-6, name  # numpy FFT SIMD paths vary by CPU
    got = P.focus_field_debye((16, 16, 16), (.05, .05, .05), .5, 1.1, n0=1.33,
                              n_integration_steps=64)
    for k, g in zip(("debye_u", "debye_ex", "debye_ey", "debye_ez"), got):
        assert O.rel_l2(g, FIX[k]) < 1e-6, k


@pytest.mark.gpu
def test_hip_path_matches_fixtures():
    import bpm_amd
    u0, dn = FIX["u0"], FIX["dn"]
    for name, c in G.cases():
        got = bpm_amd.bpm_3d(G.SIZE, G.UNITS, .5, u0=u0, dn=G.dn_for(c["kind"], dn), n0=1.33,
                             mean_mode=c["mean_mode"], absorbing_width=c["absorb"],
                             store_dn_as_half=c["kind"] == "f16", return_full=False)
        assert O.rel_l2(got, FIX[name]) < 1e-5, name
    m = bpm_amd.Bpm3d(G.SIZE, G.UNITS, lam=.5, n0=1.33, dn=dn)
    assert O.rel_l2(m.propagate(u0, return_full=False, absorbing_width=2), FIX["class_f32"]) < 1e-5
    got = bpm_amd.focus_field_debye((16, 16, 16), (.05, .05, .05), .5, 1.1, n0=1.33,
                                    n_integration_steps=64)
    for k, g in zip(("debye_u", "debye_ex", "debye_ey", "debye_ez"), got):
        assert O.rel_l2(g, FIX[k]) < 1e-5, k
    ex = bpm_amd.focus_field_debye_plane((16, 12), (.1, .1), z=-.7, lam=.5, NA=.8, n0=1.1,
                                         n_integration_steps=50)
    assert O.rel_l2(ex, FIX["debye_plane_ex"]) < 1e-5
