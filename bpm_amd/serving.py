"""Device-resident continuous-batching server for split-step BPM jobs.

A job is one dn volume (N x N x depth, random spheres generated on the GPU)
illuminated by B tilted plane waves; its result is the B exit fields and
summed intensities, copied back to the host.  The server

* admits queued jobs whenever an execution slot is free and the job's HBM
  footprint fits the remaining budget -- in arrival order (``policy="fifo"``),
  or smallest work first (``"sjf"``, work = voxel-steps) except that a job
  waiting longer than ``max_wait`` seconds goes first, in arrival order;
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
from .workloads import small_workload_spheres

__all__ = ["Job", "BpmServer", "poisson_jobs", "plain_loop", "job_footprint", "job_work",
           "job_spheres", "job_u0"]

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


def job_work(job):
    """Voxel-steps of a job: N^2 x (depth - 1) x angles."""
    return job.N * job.N * (job.nz - 1) * len(job.angles)


def job_footprint(job):
    """HBM bytes a job holds while admitted: dn + planes + u0 + intensity sums."""
    return _group_footprint([job])


def job_spheres(job, units):
    """The job's sphere list (centres, radii, dns), the 256^2 workload's recipe
    on the job's grid, from the job's seed."""
    rng = np.random.default_rng(job.seed)
    return small_workload_spheres(rng, (job.N, job.N, job.nz), units)


def job_u0(job, units, lam, n0):
    """The job's B plane-wave inputs exp(-i k0 n0 (x sin tx + y sin ty)) as a
    (B, N, N) complex64 GPU tensor, built on the device from its separable
    factors: exp(-i k sx x) and exp(-i k sy y) per angle in float64, rounded
    to complex64, multiplied in complex64 (one element-wise kernel)."""
    t = _lib.torch()
    dev = _lib.device()
    N = job.N
    k = 2. * np.pi * n0 / lam
    s = t.from_numpy(np.sin(np.asarray(job.angles, np.float64)).reshape(-1, 2)).to(dev)
    x = t.arange(N, dtype=t.float64, device=dev) * units[0]
    y = t.arange(N, dtype=t.float64, device=dev) * units[1]
    one = t.ones((), dtype=t.float64, device=dev)
    ex = t.polar(one, -k * s[:, 0, None] * x[None, :])
    ey = t.polar(one, -k * s[:, 1, None] * y[None, :])
    return ey.to(t.complex64)[:, :, None] * ex.to(t.complex64)[:, None, :]


def _job_inputs(job, ctx):
    """dn volume (device sphere generator) and the B plane-wave inputs."""
    dn = spheres_dn_device((job.N, job.N, job.nz), ctx.units, *job_spheres(job, ctx.units))
    return dn, job_u0(job, ctx.units, ctx.lam, ctx.n0)


class _Ctx:
    """Per-server constants: propagators per grid size and a plan per (grid
    size, slot).  The plans are the server's own (eager launches), never the
    shared ones of ``bpm3d.get_plan``, and keep one launch shape and radix
    schedule per grid size (the autotuned choice for ``serve_batch`` fields)
    whatever the batch of a launch, so a job's bits do not depend on what it
    is batched with."""

    def __init__(self, units, lam, n0, slots, serve_batch=16):
        self.units, self.lam, self.n0 = units, lam, n0
        self.slots = slots
        self.serve_batch = serve_batch
        self.H = {}
        self.plans = {}

    def prepare(self, N):
        if N not in self.H:
            self.H[N] = B.propagator_table(N, N, self.units, self.lam, self.n0, _lib.H_F64)
            for s in range(self.slots):
                p = B.Plan(N, N)
                B.apply_tuning(p, N, N, self.serve_batch)
                self.plans[(N, s)] = p
            _lib.torch().cuda.synchronize()
        return self.H[N]

    def plan(self, N, slot):
        return self.plans[(N, slot)]

    def close(self):
        for p in self.plans.values():
            p.close()
        self.plans.clear()


def _group_footprint(js):
    """HBM bytes of one fused launch: the dn volumes at the deepest job's
    depth plus every angle's working plane, input plane and intensity sum."""
    n2 = js[0].N * js[0].N
    items = sum(len(j.angles) for j in js)
    return 4 * n2 * len(js) * max(j.nz for j in js) + items * n2 * (8 + 8 + 4)


class BpmServer:
    """Continuous-batching server (see the module docstring).  With ``fuse``
    the jobs of one grid size that are admitted together (up to
    ``max_group_jobs`` jobs and ``max_group_items`` fields) run as one ragged
    launch sequence (``bpmh_propagate_groups``): deeper jobs keep
    propagating after shallower ones have finished and left the launch grid,
    and each job's results are copied back on the slot's copy stream as soon
    as its own planes are final."""

    def __init__(self, hbm_budget, slots=4, units=(.1, .1, .1), lam=.5, n0=1.33, fuse=True,
                 max_group_jobs=16, max_group_items=256, serve_batch=16, policy="fifo",
                 max_wait=1.0):
        if policy not in ("fifo", "sjf"):
            raise ValueError("policy must be 'fifo' or 'sjf'")
        self.budget = int(hbm_budget)
        self.slots = int(slots)
        self.fuse = bool(fuse)
        self.policy = policy
        self.max_wait = float(max_wait)
        self.max_group_jobs = max(1, min(16, int(max_group_jobs)))
        self.max_group_items = int(max_group_items)
        self.ctx = _Ctx(units, lam, n0, self.slots, serve_batch)

    def close(self):
        """Release the server's plans."""
        self.ctx.close()

    def _launch(self, js, slot, stream, copy_stream):
        """One launch sequence for jobs js (one grid size) on `stream`; returns
        per-job handles (completion event, device tensors to keep, host
        outputs)."""
        t = _lib.torch()
        N = js[0].N
        js = sorted(js, key=lambda j: -j.nz)  # non-increasing depth
        nzmax = js[0].nz
        with t.cuda.stream(stream):
            H = self.ctx.H[N]
            dn = t.empty((len(js), nzmax, N, N), dtype=t.float32, device=_lib.device())
            u0s = []
            for g, j in enumerate(js):
                spheres_dn_device((N, N, j.nz), self.ctx.units, *job_spheres(j, self.ctx.units),
                                  out=dn[g, :j.nz])
                u0s.append(job_u0(j, self.ctx.units, self.ctx.lam, self.ctx.n0))
            plane = t.cat(u0s) if len(u0s) > 1 else u0s[0]
            acc = t.empty(plane.shape, dtype=t.float32, device=_lib.device())
            evs = [t.cuda.Event() for _ in js]
            B.propagate_groups(plane, [(len(j.angles), j.nz) for j in js], self.ctx.units,
                               self.ctx.lam, dn, H, nzmax * N * N, out_sum=acc,
                               plan=self.ctx.plan(N, slot), done_events=evs)
        handles = []
        i0 = 0
        for j, ev in zip(js, evs):
            nb = len(j.angles)
            out_u = t.empty((nb, N, N), dtype=plane.dtype, pin_memory=True)
            out_i = t.empty((nb, N, N), dtype=acc.dtype, pin_memory=True)
            with t.cuda.stream(copy_stream):
                copy_stream.wait_event(ev)
                out_u.copy_(plane[i0:i0 + nb], non_blocking=True)
                out_i.copy_(acc[i0:i0 + nb], non_blocking=True)
                done = t.cuda.Event()
                done.record(copy_stream)
            handles.append((j, {"ev": done, "out": (out_u, out_i)}))
            i0 += nb
        # the device tensors stay alive until every job of the launch is retired
        return handles, (dn, plane, acc, u0s)

    def _order(self, queue, now):
        """Queued jobs in admission order under the server's policy."""
        if self.policy == "fifo":
            return list(queue)
        late = [j for j in queue if now - j.arrival > self.max_wait]
        rest = sorted((j for j in queue if now - j.arrival <= self.max_wait),
                      key=lambda j: (job_work(j), j.arrival))
        return late + rest

    def _take_group(self, queue, used, now=0.):
        """The first job in admission order plus (when fusing) the later
        queued jobs of its grid size that fit the group limits and the HBM
        budget; None when that job does not fit yet."""
        order = self._order(queue, now)
        head = order[0]
        if used + _group_footprint([head]) > self.budget:
            return None
        js = [head]
        if self.fuse:
            items = len(head.angles)
            for j in order[1:]:
                if len(js) >= self.max_group_jobs:
                    break
                if j.N != head.N or items + len(j.angles) > self.max_group_items:
                    continue
                if used + _group_footprint(js + [j]) > self.budget:
                    continue
                js.append(j)
                items += len(j.angles)
        for j in js:
            queue.remove(j)
        return js

    def run(self, jobs, realtime=True, keep_results=False):
        """Serve ``jobs`` (arrival offsets in seconds from the start).

        Returns dict with per-job latency (s), p50/p99, throughput in
        voxel-steps/s, the launches (jobs per fused launch) and, if
        keep_results, the host results per job id."""
        t = _lib.torch()
        for N in sorted({j.N for j in jobs}):
            self.ctx.prepare(N)
        for j in jobs:
            if _group_footprint([j]) > self.budget:
                raise MemoryError("job %d needs %d B, budget %d B" % (j.jid, _group_footprint([j]), self.budget))
        streams = [t.cuda.Stream() for _ in range(self.slots)]
        copies = [t.cuda.Stream() for _ in range(self.slots)]
        free = list(range(self.slots))
        pending = sorted(jobs, key=lambda j: j.arrival)
        queue = collections.deque()
        launches = {}   # slot -> (remaining job ids, footprint, device tensors)
        inflight = {}   # jid -> (job, slot, handle)
        used = 0
        lat = {}
        results = {}
        sizes = []
        t.cuda.synchronize()
        t0 = time.perf_counter()
        clock = (lambda: time.perf_counter() - t0) if realtime else None
        vsteps = 0
        while pending or queue or inflight:
            now = clock() if realtime else float("inf")
            while pending and pending[0].arrival <= now:
                queue.append(pending.pop(0))
            # admit: the first job in policy order (and same-grid followers)
            # while a slot is free
            while queue and free:
                js = self._take_group(queue, used, clock() if realtime else 0.)
                if js is None:
                    break
                s = free.pop()
                fp = _group_footprint(js)
                used += fp
                handles, keep = self._launch(js, s, streams[s], copies[s])
                launches[s] = ({j.jid for j in js}, fp, keep)
                sizes.append(len(js))
                for j, h in handles:
                    inflight[j.jid] = (j, s, h)
            # retire finished jobs; a slot is free once all its jobs are
            done = [jid for jid, (_, _, h) in inflight.items() if h["ev"].query()]
            for jid in done:
                j, s, h = inflight.pop(jid)
                tdone = time.perf_counter() - t0
                lat[jid] = tdone - (j.arrival if realtime else 0.)
                vsteps += j.N * j.N * (j.nz - 1) * len(j.angles)
                if keep_results:
                    results[jid] = tuple(o.numpy().copy() for o in h["out"])
                rem, fp, keep = launches[s]
                rem.discard(jid)
                if not rem:
                    del launches[s]
                    used -= fp
                    free.append(s)
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
                "voxel_steps": vsteps, "throughput": vsteps / wall, "results": results,
                "launch_sizes": sizes}


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
