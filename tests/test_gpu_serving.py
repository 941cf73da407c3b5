"""GPU tests of the continuous-batching server (bpm_amd.serving).

Every job's results (B exit fields and summed intensities) must be the same
bits whether the job is served alone or in a mix of overlapping and fused
jobs (the server keeps one launch shape and radix schedule per grid size),
and agree with the per-request loop that propagates each angle alone (which
uses its own batch-1 tuning, possibly another radix schedule) to the parity
bar, 1e-4 rel-L2 (measured up to 1.3e-5 over ~500 planes); one angle per grid size is checked against the complex64 oracle (bar
1e-4 rel-L2) and the float64 oracle (error stated).  The per-request entry each job reproduces is
bpm/bpm/bpm_3d.py:39-101 (with the class path's static propagator,
bpm/bpmclass/_bpm3d_ocl.py:66-134).
"""
import os
import sys
import time

import numpy as np
import pytest

from oracle import bpm_oracle as O

pytestmark = pytest.mark.gpu

UNITS, LAM, N0 = (.1, .1, .1), .5, 1.33


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def srv_mod():
    from bpm_amd import _lib
    _lib.torch()  # raises without a GPU or without libbpmhip.so
    from bpm_amd import serving
    return serving


def rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def test_served_results_equal_alone_and_in_mix(srv_mod):
    # small jobs: continuous batching with more jobs than slots, fused and
    # unfused; each job's fields equal to its run alone bit for bit and to
    # the per-request baseline (each angle alone) to the parity bar
    S = srv_mod
    say("case serving small")
    jobs = S.poisson_jobs(seed=3, n_jobs=8, rate_hz=200., grids=(64, 128, 256), depth=(16, 48),
                          angles=(2, 6))
    budget = 3 * max(S.job_footprint(j) for j in jobs)
    out = {}
    for fuse in (True, False):
        srv = S.BpmServer(hbm_budget=budget, slots=3, fuse=fuse)
        try:
            r = srv.run(jobs, realtime=True, keep_results=True)
            alone = {j.jid: srv.run([j], realtime=False, keep_results=True)["results"][j.jid]
                     for j in jobs}
        finally:
            srv.close()
        say("   fuse=%s: launches %s, p50 %.1f ms p99 %.1f ms"
            % (fuse, r["launch_sizes"], r["p50_s"] * 1e3, r["p99_s"] * 1e3))
        for j in jobs:
            assert all(np.array_equal(a, b) for a, b in zip(r["results"][j.jid], alone[j.jid])), j.jid
        out[fuse] = r
        assert r["p99_s"] >= r["p50_s"] > 0
    assert max(out[True]["launch_sizes"]) > 1  # some jobs were fused
    p = S.plain_loop(jobs, keep_results=True)
    for j in jobs:
        for fuse in (True, False):
            gu, gi = out[fuse]["results"][j.jid]
            pu, pi = p["results"][j.jid]
            assert rel(gu, pu) < 1e-4 and rel(gi, pi) < 1e-4, (j.jid, fuse)
    say("   p50 %.1f ms p99 %.1f ms" % (r["p50_s"] * 1e3, r["p99_s"] * 1e3))
    assert r["p99_s"] >= r["p50_s"] > 0
    # a job that does not fit the budget is refused up front
    with pytest.raises(MemoryError):
        S.BpmServer(hbm_budget=1000, slots=1).run(jobs[:1])


@pytest.mark.timeout(900)
def test_c4_served_stream_alone_mix_and_oracle(srv_mod):
    """The served workload's shape: a seeded Poisson stream of jobs with grids
    {128, 256, 512, 1024}^2, depth U{64..512}, 8-64 plane-wave angles
    (|tilt| <= 0.3 rad) and random-sphere dn; 4 slots.  The shared plans of
    every grid size are left in graph mode (as the benchmark leaves them)
    while the server runs on its own per-slot plans."""
    import torch
    S = srv_mod
    from bpm_amd import _lib
    from bpm_amd import bpm3d as B
    from bpm_amd.workloads import spheres_dn
    jobs = S.poisson_jobs(seed=21, n_jobs=10, rate_hz=40.)
    grids = sorted({j.N for j in jobs})
    say("case C4 stream: %d jobs, grids %s, depths %s, angles %s"
        % (len(jobs), grids, [j.nz for j in jobs], [len(j.angles) for j in jobs]))
    assert grids == [128, 256, 512, 1024]
    lib = _lib.load()
    for N in grids:
        _lib.check(lib.bpmh_plan_set_graph(B.get_plan(N, N), 1))
    budget = 24 << 30
    srv = S.BpmServer(hbm_budget=budget, slots=4)
    try:
        t0 = time.perf_counter()
        mix = srv.run(jobs, realtime=True, keep_results=True)
        say("   mix: wall %.2f s, p50 %.1f ms, p99 %.1f ms"
            % (time.perf_counter() - t0, mix["p50_s"] * 1e3, mix["p99_s"] * 1e3))
        alone = {}
        for j in jobs:
            alone[j.jid] = srv.run([j], realtime=False, keep_results=True)["results"][j.jid]
        say("   alone runs done")
    finally:
        srv.close()
        for N in grids:
            _lib.check(lib.bpmh_plan_set_graph(B.get_plan(N, N), 0))
    for j in jobs:
        mu, mi = mix["results"][j.jid]
        au, ai = alone[j.jid]
        assert mu.shape == (len(j.angles), j.N, j.N) and mi.shape == mu.shape
        assert np.array_equal(mu, au) and np.array_equal(mi, ai), j.jid
    say("   every job: mix == alone bit for bit; launches %s" % mix["launch_sizes"])
    plain = S.plain_loop(jobs, keep_results=True)
    worst = 0.
    for j in jobs:
        pu, pi = plain["results"][j.jid]
        mu, mi = mix["results"][j.jid]
        worst = max(worst, rel(mu, pu), rel(mi, pi))
    say("   served vs per-request loop (each angle alone): max rel-L2 %.2e" % worst)
    assert worst < 1e-4
    # one angle of one job per grid size against the oracles
    O.set_fft_workers(min(16, len(os.sched_getaffinity(0))))
    try:
        for N in grids:
            j = min((j for j in jobs if j.N == N), key=lambda j: j.nz)
            b = len(j.angles) // 2
            dn = spheres_dn((N, N, j.nz), UNITS, *S.job_spheres(j, UNITS))
            u0 = S.job_u0(j, UNITS, LAM, N0)[b].cpu().numpy()  # the served input
            mu, mi = mix["results"][j.jid]
            t0 = time.perf_counter()
            errs = {}
            for prec in ("c64", "f64"):
                ru, ri = O.propagate_single((N, N, j.nz), UNITS, LAM, N0, dn, u0,
                                            return_full=False, precision=prec, return_sum=True)
                errs[prec] = (O.rel_l2(mu[b], ru), O.rel_l2(mi[b], ri))
            say("   job %d (%d^2 x %d, angle %d/%d): exit field rel-L2 c64 %.2e f64 %.2e; "
                "sum|u|^2 c64 %.2e f64 %.2e (oracle %.1f s)"
                % (j.jid, N, j.nz, b, len(j.angles), errs["c64"][0], errs["f64"][0],
                   errs["c64"][1], errs["f64"][1], time.perf_counter() - t0))
            assert errs["c64"][0] < 1e-4 and errs["c64"][1] < 1e-4
            assert errs["f64"][0] < 1e-3 and errs["f64"][1] < 1e-3
    finally:
        O.set_fft_workers(1)
