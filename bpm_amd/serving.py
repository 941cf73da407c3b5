"""Device-resident continuous-batching server for split-step BPM jobs.

A job is one dn volume (N x N x depth, random spheres generated on the GPU)
illuminated by B tilted plane waves; its result is the B exit fields and
summed intensities, copied back to the host.  The server

* admits queued jobs (FIFO) whenever an execution slot is free and the job's
  HBM footprint fits the remaining budget;
* runs each admitted job on its slot's HIP stream with the slot's own plan
  (a plan's scratch and captured graphs serve one stream at a time), its B
  angles batched into one launch sequence (``bpmh_propagate`` with batch=B
  sharing dn), so jobs that arrive while others are running overlap with
  them on the device;
* retires jobs as their completion events fire, returning their memory to
  the budget and recording latency = host-result-ready - arrival.

``plain_loop`` is the per-request baseline: requests one after another, each
angle propagated on its own (batch 1), the way the reference's bpm_3d is
called.  Both produce identical fields for a job (same kernels, same inputs).
"""
import collections
import math
import time

import numpy as np

from . import _lib
from . import bpm3d as B
from .spheres import spheres_dn_device
from .workloads import plane_waves, small_workload_spheres

__all__ = ["Job", "BpmServer", "poisson_jobs", "plain_loop", "job_footprint", "job_spheres",
           "job_u0"]

Job = collections.namedtuple("Job", "jid arrival N nz angles seed")


def poisson_jobs(seed=0, n_jobs=24, rate_hz=40., grids=(128, 256, 512, 1024), depth=(64, 512),
                 angles=(8, 64), max_tilt=.3):
    """Seeded Poisson stream: exponential inter-arrival times, grid uniform in
    ``grids``, depth U{depth}, angle count U{angles}, tilts |theta| <= max_tilt."""
    rng = np.random.default_rng(seed)
    t = 0.
    jobs = []
    for j in range(n_jobs):
        t += rng.exponential(1. / rate_hz)
        N = int(rng.choice(grids))
        nz = int(rng.integers(depth[0], depth[1] + 1))
        B_ = int(rng.integers(angles[0], angles[1] + 1))
        th = rng.uniform(-max_tilt, max_tilt, (B_, 2)) / math.sqrt(2.)
        jobs.append(Job(j, t, N, nz, th, int(rng.integers(1 << 30))))
    return jobs


def job_footprint(job):
    """HBM bytes a job holds while admitted: dn + planes + u0 + intensity sums."""
    n2 = job.N * job.N
    B_ = len(job.angles)
    return 4 * n2 * job.nz + B_ * n2 * (8 + 8 + 4)


class _Ctx:
    """Per-server constants: propagators per grid size and a plan per (grid
    size, slot).  The plans are the server's own (eager launches, their own
    launch-shape settings), never the shared ones of ``bpm3d.get_plan``."""

    def __init__(self, units, lam, n0, slots):
        self.units, self.lam, self.n0 = units, lam, n0
        self.slots = slots
        self.H = {}
        self.plans = {}

    def prepare(self, N):
        if N not in self.H:
            self.H[N] = B.propagator_table(N, N, self.units, self.lam, self.n0, _lib.H_F64)
            for s in range(self.slots):
                self.plans[(N, s)] = B.Plan(N, N)
            _lib.torch().cuda.synchronize()
        return self.H[N]

    def plan(self, N, slot):
        return self.plans[(N, slot)]

    def close(self):
        for p in self.plans.values():
            p.close()
        self.plans.clear()


def job_spheres(job, units):
    """The job's sphere list (centres, radii, dns), the 256^2 workload's recipe
    on the job's grid, from the job's seed."""
    rng = np.random.default_rng(job.seed)
    return small_workload_spheres(rng, (job.N, job.N, job.nz), units)


def job_u0(job, units, lam, n0):
    """The job's B plane-wave inputs (B, N, N) complex64 on the host."""
    return plane_waves(job.N, job.N, units[0], units[1], lam, n0, job.angles)


def _job_inputs(job, ctx):
    """dn volume (device sphere generator) and the B plane-wave inputs."""
    t = _lib.torch()
    dn = spheres_dn_device((job.N, job.N, job.nz), ctx.units, *job_spheres(job, ctx.units))
    u0 = t.from_numpy(job_u0(job, ctx.units, ctx.lam, ctx.n0)).to(_lib.device())
    return dn, u0


class BpmServer:
    def __init__(self, hbm_budget, slots=4, units=(.1, .1, .1), lam=.5, n0=1.33):
        self.budget = int(hbm_budget)
        self.slots = int(slots)
        self.ctx = _Ctx(units, lam, n0, self.slots)

    def close(self):
        """Release the server's plans."""
        self.ctx.close()

    def _launch(self, job, slot, stream):
        t = _lib.torch()
        with t.cuda.stream(stream):
            H = self.ctx.H[job.N]
            dn, plane = _job_inputs(job, self.ctx)
            acc = t.empty((len(job.angles), job.N, job.N), dtype=t.float32, device=_lib.device())
            B.propagate_fields(plane, job.nz, self.ctx.units, self.ctx.lam, self.ctx.n0, dn=dn,
                               dn_kind=_lib.DN_F32, mean_mode="ignore", H=H, out_sum=acc,
                               plan=self.ctx.plan(job.N, slot))
            out_u = t.empty(plane.shape, dtype=plane.dtype, pin_memory=True)
            out_i = t.empty(acc.shape, dtype=acc.dtype, pin_memory=True)
            out_u.copy_(plane, non_blocking=True)
            out_i.copy_(acc, non_blocking=True)
            ev = t.cuda.Event()
            ev.record(stream)
        # keep device tensors alive until the event fires
        return {"ev": ev, "keep": (dn, plane, acc), "out": (out_u, out_i)}

    def run(self, jobs, realtime=True, keep_results=False):
        """Serve ``jobs`` (arrival offsets in seconds from the start).

        Returns dict with per-job latency (s), p50/p99, throughput in
        voxel-steps/s and, if keep_results, the host results per job id."""
        t = _lib.torch()
        for N in sorted({j.N for j in jobs}):
            self.ctx.prepare(N)
        for j in jobs:
            if job_footprint(j) > self.budget:
                raise MemoryError("job %d needs %d B, budget %d B" % (j.jid, job_footprint(j), self.budget))
        streams = [t.cuda.Stream() for _ in range(self.slots)]
        free = list(range(self.slots))
        pending = sorted(jobs, key=lambda j: j.arrival)
        queue = collections.deque()
        inflight = {}
        used = 0
        lat = {}
        results = {}
        t.cuda.synchronize()
        t0 = time.perf_counter()
        clock = (lambda: time.perf_counter() - t0) if realtime else None
        vsteps = 0
        while pending or queue or inflight:
            now = clock() if realtime else float("inf")
            while pending and pending[0].arrival <= now:
                queue.append(pending.pop(0))
            # admit: FIFO while a slot is free and the footprint fits
            while queue and free and used + job_footprint(queue[0]) <= self.budget:
                j = queue.popleft()
                s = free.pop()
                used += job_footprint(j)
                inflight[j.jid] = (j, s, self._launch(j, s, streams[s]))
            # retire finished jobs
            done = [jid for jid, (_, _, h) in inflight.items() if h["ev"].query()]
            for jid in done:
                j, s, h = inflight.pop(jid)
                tdone = time.perf_counter() - t0
                lat[jid] = tdone - (j.arrival if realtime else 0.)
                vsteps += j.N * j.N * (j.nz - 1) * len(j.angles)
                used -= job_footprint(j)
                free.append(s)
                if keep_results:
                    results[jid] = tuple(o.numpy().copy() for o in h["out"])
            if not done and not (queue and free):
                if inflight:
                    time.sleep(0.0002)
                elif pending and realtime:
                    time.sleep(max(0., min(0.002, pending[0].arrival - clock())))
        t.cuda.synchronize()
        wall = time.perf_counter() - t0
        ls = np.array([lat[j.jid] for j in jobs])
        return {"latency_s": lat, "p50_s": float(np.percentile(ls, 50)),
                "p99_s": float(np.percentile(ls, 99)), "wall_s": wall,
                "voxel_steps": vsteps, "throughput": vsteps / wall, "results": results}


def plain_loop(jobs, units=(.1, .1, .1), lam=.5, n0=1.33, keep_results=False):
    """Per-request baseline: jobs in arrival order, one at a time, each angle
    propagated alone (batch 1); returns wall time and throughput."""
    t = _lib.torch()
    ctx = _Ctx(units, lam, n0, 1)
    for N in sorted({j.N for j in jobs}):
        ctx.prepare(N)
    try:
        return _plain_loop(jobs, ctx, units, lam, n0, keep_results)
    finally:
        ctx.close()


def _plain_loop(jobs, ctx, units, lam, n0, keep_results):
    t = _lib.torch()
    t.cuda.synchronize()
    t0 = time.perf_counter()
    vsteps = 0
    results = {}
    for j in sorted(jobs, key=lambda j: j.arrival):
        dn, u0 = _job_inputs(j, ctx)
        outs_u, outs_i = [], []
        for b in range(len(j.angles)):
            plane = u0[b:b + 1].clone()
            acc = t.empty((1, j.N, j.N), dtype=t.float32, device=_lib.device())
            B.propagate_fields(plane, j.nz, units, lam, n0, dn=dn, dn_kind=_lib.DN_F32,
                               mean_mode="ignore", H=ctx.H[j.N], out_sum=acc,
                               plan=ctx.plan(j.N, 0))
            outs_u.append(plane.cpu())
            outs_i.append(acc.cpu())
        vsteps += j.N * j.N * (j.nz - 1) * len(j.angles)
        if keep_results:
            results[j.jid] = (t.cat(outs_u).numpy(), t.cat(outs_i).numpy())
    t.cuda.synchronize()
    wall = time.perf_counter() - t0
    return {"wall_s": wall, "voxel_steps": vsteps, "throughput": vsteps / wall, "results": results}
