"""Benchmark: split-step BPM throughput on one MI355X.

Workload (one "step" = one full propagation): 1024 x 1024 grid x 1024 planes
(1023 z-steps), voxel 0.1 um, lam 0.5 um, n0 1.33, complex64 paraxial
Gaussian beam (w0 3 um) focused 20 um into the volume, float32 dn of 400
seeded spheres (r ~ U(1,4) um, dn ~ U(.01,.05), overlaps summed); Bpm3d
propagate path (static float64->complex64 propagator, phase screen per
plane); outputs the exit field and the summed intensity sum_z |u_z|^2.
Inputs are resident in HBM before the timed region; each step copies u0
into the working plane and runs the whole z-loop.

Prints one JSON line.  value = voxel-steps/s = Nx*Ny*(Nz-1) per step / time.
roofline: the dominant kernel's algorithmic bytes per launch (row pass:
plane read+write 16 B + dn 4 B per voxel; column pass: plane read+write
16 B per voxel; 36 B per voxel-step in all) divided by its average launch
duration.  That duration comes from two HIP-event measurements on the stream
the kernels run on: events around the graph-replayed timed region give the
total, and an eager repeat of the same steps with an event pair around every
launch gives each kernel's share of it.
traffic: HBM bytes per launch of that kernel from rocprofv3 FETCH_SIZE and
WRITE_SIZE runs, read from the committed counter summary when one exists for
the same kernel and launch shape.
cpu_baseline: the numpy complex64 restatement of the same loop (the oracle)
timed on a bounded sample of the workload, on all usable host cores (2-D FFTs
through scipy.fft workers) and on one core.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
FP32_PEAK_TFLOPS = 157.3  # MI355X dense FP32 (matrix, and vector with packed math)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nx", type=int, default=1024)
    ap.add_argument("--nz", type=int, default=1024)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-sample-planes", type=int, default=129,
                    help="planes of the workload the all-cores CPU baseline propagates")
    ap.add_argument("--cpu-sample-planes-1core", type=int, default=65,
                    help="planes of the workload the one-core CPU baseline propagates")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-debye", action="store_true")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--no-differentiable", action="store_true")
    ap.add_argument("--no-small", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-floor", action="store_true")
    ap.add_argument("--no-weighted", action="store_true")
    ap.add_argument("--copy-reference", type=int, default=0, metavar="MIB",
                    help="only run a MIB-sized device copy (counter scale reference)")
    ap.add_argument("--debye-only", action="store_true",
                    help="only time the Debye-Wolf workload and print its JSON object")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch the z-loop kernel by kernel instead of replaying a HIP graph")
    return ap.parse_args()


def host_cores():
    """Cores this process may use (affinity), capped by OMP_NUM_THREADS when set."""
    n = len(os.sched_getaffinity(0))
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        n = min(n, int(os.environ["OMP_NUM_THREADS"]))
    return max(1, n)


def cpu_baseline(u0, dn, units, lam, n0, planes, planes_1core):
    """Time the oracle's complex64 restatement of the Bpm3d loop on the host:
    on all usable cores (2-D FFTs through scipy.fft workers; the element-wise
    steps are numpy, single-threaded) and on one core (numpy.fft)."""
    from oracle import bpm_oracle as O
    Ny, Nx = u0.shape
    out = {}
    for cores, pl in ((host_cores(), planes), (1, planes_1core)):
        sub = np.ascontiguousarray(dn[:pl])
        O.set_fft_workers(cores)
        # untimed warm-up: imports, FFT plans and thread pools
        O.propagate_single((Nx, Ny, 2), units, lam, n0, sub[:2], u0, return_full=False)
        t0 = time.perf_counter()
        O.propagate_single((Nx, Ny, pl), units, lam, n0, sub, u0, return_full=False)
        dt = time.perf_counter() - t0
        O.set_fft_workers(1)
        out[cores] = (Nx * Ny * (pl - 1) / dt, dt, pl)
    n = host_cores()
    vs, dt, pl = out[n]
    vs1, dt1, pl1 = out[1]
    try:
        model = [l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("model name")][0]
    except (OSError, IndexError):
        model = "unknown"
    return {"value": vs, "unit": "voxel-steps/s", "cores": n, "kind": "port",
            "sample": "%dx%d grid x %d planes (%d steps) of the same workload, the oracle's "
                      "complex64 restatement (oracle.bpm_oracle.propagate_single) with "
                      "scipy.fft on %d workers, %.1f s; CPU %s"
                      % (Nx, Ny, pl, pl - 1, n, dt, model),
            "single_core_value": vs1,
            "single_core_sample": "%d planes, numpy.fft, %.1f s" % (pl1, dt1)}


def debye_extra(torch):
    """The 256^3 Debye-Wolf workload (NA 1.1, n0 1.33, 512 theta steps): the
    whole call (octant quadrature written straight to the full grid) --
    calls issued back to back (wall time and one event pair over all of
    them, per call) and single calls each waited for (host included) --
    then one call with HIP events around each of its kernels for the
    product's MFMA rate."""
    import ctypes
    from bpm_amd import focus_fields as Q
    from bpm_amd import _lib
    shape, units, lam, NA, n0, S = (256, 256, 256), (.05,) * 3, .5, 1.1, 1.33, 512
    for _ in range(2):
        Q.focus_field_debye_device(shape, units, lam, NA, n0, S)
    torch.cuda.synchronize()
    reps = 20
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    for r in range(reps):
        Q.focus_field_debye_device(shape, units, lam, NA, n0, S)
    e1.record()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    gpu_ms = e0.elapsed_time(e1) / reps
    single = []
    for r in range(5):
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        Q.focus_field_debye_device(shape, units, lam, NA, n0, S)
        torch.cuda.synchronize()
        single.append(time.perf_counter() - t1)
    lib = _lib.load()
    _lib.check(lib.bpmh_debye_set_timing(1))
    Q.focus_field_debye_device(shape, units, lam, NA, n0, S)
    ms = (ctypes.c_float * 5)()
    cnt = (ctypes.c_int * 5)()
    _lib.check(lib.bpmh_debye_kernel_times(ms, cnt))
    _lib.check(lib.bpmh_debye_set_timing(0))
    # the product: for every (plane, table column, sample) 6 real
    # multiply-adds (Re and Im of I0, I1, I2), 12 flop; the table runs over
    # the 129*130/2 column pairs of the 129^2 octant plane and is padded to
    # whole MFMA tiles (planes to 144, columns to 8400, samples to 528)
    nz, npairs, total = 129, 129 * 130 // 2, S + 1
    zp, cp, tp = 144, (npairs + 15) // 16 * 16, (total + 15) // 16 * 16
    prod_s = ms[1] * 1e-3
    mfma_tf = 12. * zp * cp * tp / prod_s / 1e12
    useful_tf = 12. * nz * npairs * total / prod_s / 1e12
    samples = 129 ** 3 * (S + 1)
    return {"workload": "Debye-Wolf 256^3, NA 1.1, n0 1.33, 512 theta steps, x-polarised",
            "ms": dt * 1e3, "ms_gpu_span": gpu_ms, "ms_single_call": float(np.median(single)) * 1e3,
            "theta_samples_per_s": samples / dt,
            "kernel_us": {"tables": ms[0] * 1e3, "product": ms[1] * 1e3,
                          "full_grid_spread": ms[2] * 1e3},
            "product_mfma_tflops": mfma_tf, "product_mfma_frac": mfma_tf / FP32_PEAK_TFLOPS,
            "product_useful_tflops": useful_tf,
            "timing": "ms = wall time of %d calls issued back to back / %d (GPU-bound: the host "
                      "part of a call is ~30 us); ms_gpu_span = one event pair around them / %d; "
                      "ms_single_call = median of 5 calls each synchronised before and after "
                      "(host launch and completion wait included)" % (reps, reps, reps),
            "note": "MFMA fraction = flops the MFMA units execute over the padded "
                    "144 x 8400 x 528 product (12 per plane, column pair and sample) / the "
                    "product kernel's time (event pair) / 157.3 TFLOP/s dense FP32 matrix "
                    "peak; useful = unpadded 129 x 8385 x 513"}


def served_extra(seed, n_jobs=200, loads=(0.8, 0.95), plain_jobs=48, runs=2):
    """Mixed-size served stream (grids 128^2..1024^2, depth U{64..512}, 8-64
    plane-wave angles per job, spheres as the 256^2 workload) through the
    continuous-batching server.  Capacity = the n_jobs all queued at t = 0;
    then, for each load in `loads`, `runs` seeded Poisson streams of the same
    jobs at load x that capacity (jobs/s) give latency percentiles and served
    throughput, with the spread over the runs (served throughput at a load
    is bounded by load x capacity: the arrivals set it).  The per-request
    loop (each angle alone, one job after another) is timed on the first
    `plain_jobs` jobs; throughputs are compared as voxel-steps/s."""
    from bpm_amd.serving import BpmServer, plain_loop, poisson_jobs
    jobs = poisson_jobs(seed=seed, n_jobs=n_jobs, rate_hz=1.)
    srv = BpmServer(hbm_budget=48 << 30, slots=4)
    unf = BpmServer(hbm_budget=48 << 30, slots=4, fuse=False)
    streams = []
    try:
        srv.run(jobs[:4], realtime=False)  # warm-up (plans, propagators)
        cap = srv.run(jobs, realtime=False)
        unf.run(jobs[:4], realtime=False)
        cap_unfused = unf.run(jobs, realtime=False)
        for load in loads:
            rate = load * n_jobs / cap["wall_s"]
            rs = []
            for r_i in range(runs):
                js = poisson_jobs(seed=seed, n_jobs=n_jobs, rate_hz=rate)
                rs.append(srv.run(js, realtime=True))
                print("served load %.2f run %d: %.1f s, p50 %.1f ms, p99 %.1f ms"
                      % (load, r_i, rs[-1]["wall_s"], rs[-1]["p50_s"] * 1e3,
                         rs[-1]["p99_s"] * 1e3), file=sys.stderr, flush=True)
            streams.append((load, rate, rs))
    finally:
        srv.close()
        unf.close()
    p = plain_loop(jobs[:plain_jobs])
    poisson = []
    for load, rate, rs in streams:
        thr = [r["throughput"] for r in rs]
        poisson.append({
            "load_of_capacity": load, "rate_jobs_per_s": rate,
            "p50_ms": [r["p50_s"] * 1e3 for r in rs], "p99_ms": [r["p99_s"] * 1e3 for r in rs],
            "served_voxel_steps_per_s": thr, "served_wall_s": [r["wall_s"] for r in rs],
            "served_jobs_per_launch": [float(np.mean(r["launch_sizes"])) for r in rs],
            "served_over_plain": [v / p["throughput"] for v in thr]})
    return {"workload": "%d jobs, grids {128,256,512,1024}^2, depth U{64..512}, 8-64 angles "
                        "(tilt <= 0.3 rad), spheres as the 256^2 workload; 4 slots, 48 GiB "
                        "HBM budget; latency = arrival -> results on host" % n_jobs,
            "capacity_voxel_steps_per_s": cap["throughput"], "capacity_wall_s": cap["wall_s"],
            "capacity_jobs_per_launch": float(np.mean(cap["launch_sizes"])),
            "capacity_unfused_voxel_steps_per_s": cap_unfused["throughput"],
            "plain_loop_voxel_steps_per_s": p["throughput"], "plain_loop_jobs": plain_jobs,
            "plain_wall_s": p["wall_s"],
            "capacity_over_plain": cap["throughput"] / p["throughput"],
            "poisson": poisson}


def plan_shape(lib, plan, kernel):
    """Launch shape the plan uses: rows per row work-group or columns per
    column tile, as "lines=<n>" / "cols=<n>"."""
    import ctypes
    v = [ctypes.c_int(-1) for _ in range(6)]
    if lib.bpmh_plan_get_variant(plan, *[ctypes.byref(x) for x in v]) != 0:
        return "?"
    if kernel == "row_pass":
        return "lines=%d" % v[2].value + ("" if v[4].value == 4 else " e=%d" % v[4].value)
    return "cols=%d" % v[3].value + ("" if v[5].value == 8 else " e=%d" % v[5].value)


def measured_traffic(kernel, nx, ny, shape):
    """HBM bytes per launch from the committed counter summary, if it covers
    this kernel, grid and launch shape.  Its lines read
    `traffic <kernel> <nx>x<ny> <shape> <bytes>`."""
    path = os.path.join(ROOT, "profiles", "bpm_1024x1024_pass_traffic.md")
    if not os.path.exists(path):
        return None
    key = "traffic %s %dx%d %s " % (kernel, nx, ny, shape)
    with open(path) as f:
        for line in f:
            line = line.strip().strip("`")
            if line.startswith(key):
                return {"bytes": float(line[len(key):].split()[0]),
                        "source": "profiles/bpm_1024x1024_pass_traffic.md"}
    return None


def small_extra(torch, seed):
    """The 256^2 x 256 plain workload (20 spheres + a centred r = 2 um sphere,
    plane wave at normal incidence, n0 1.33), Bpm3d path, exit field +
    summed intensity, graph-replayed; batch 1 and a batch of 16 fields
    through the same dn."""
    from bpm_amd import _lib
    from bpm_amd import bpm3d as B
    from bpm_amd.workloads import plain_small_workload
    units, lam, n0 = (.1, .1, .1), .5, 1.33
    u0, dn = plain_small_workload(seed)
    nz, ny, nx = dn.shape
    dn_g = _lib.to_device(dn, np.float32)
    H = B.propagator_table(nx, ny, units, lam, n0, _lib.H_F64)
    _lib.check(_lib.load().bpmh_plan_set_graph(B.get_plan(nx, ny), 1))
    out = {"workload": "256^2 x 256, 21 spheres, plane wave, Bpm3d path, exit field + summed intensity"}
    for batch in (1, 16):
        u0_g = _lib.to_device(np.broadcast_to(u0, (batch, ny, nx)), np.complex64)
        plane = torch.empty_like(u0_g)
        acc = torch.empty((batch, ny, nx), dtype=torch.float32, device="cuda")

        def step():
            plane.copy_(u0_g)
            B.propagate_fields(plane, nz, units, lam, n0, dn=dn_g, dn_kind=_lib.DN_F32,
                               mean_mode="ignore", H=H, out_sum=acc)

        for _ in range(3):
            step()
        torch.cuda.synchronize()
        reps = 20
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        ev0.record()
        for _ in range(reps):
            step()
        ev1.record()
        torch.cuda.synchronize()
        ms = ev0.elapsed_time(ev1) / reps
        vs = nx * ny * (nz - 1) * batch / (ms * 1e-3)
        out["batch%d" % batch] = {"ms_per_propagation": ms, "voxel_steps_per_s": vs,
                                  "effective_GBps_36B": 36. * vs / 1e9,
                                  "effective_GBps_44B": 44. * vs / 1e9}
    # bpm_3d's default mean_mode "weighted" (k-space weighted dn0 each step:
    # three passes per z-step), batch 1, same outputs
    u0_g = _lib.to_device(u0[None], np.complex64)
    plane = torch.empty_like(u0_g)
    acc = torch.empty((1, ny, nx), dtype=torch.float32, device="cuda")

    def wstep():
        plane.copy_(u0_g)
        B.propagate_fields(plane, nz, units, lam, n0, dn=dn_g, dn_kind=_lib.DN_F32,
                           mean_mode="weighted", out_sum=acc)

    for _ in range(3):
        wstep()
    torch.cuda.synchronize()
    reps = 20
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(reps):
        wstep()
    ev1.record()
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / reps
    _lib.check(_lib.load().bpmh_plan_set_graph(B.get_plan(nx, ny), 0))
    vs = nx * ny * (nz - 1) / (ms * 1e-3)
    # three passes per step: 52 B per voxel-step
    out["weighted_batch1"] = {"ms_per_propagation": ms, "voxel_steps_per_s": vs,
                              "effective_GBps_52B": 52. * vs / 1e9,
                              "hbm_frac_52B": 52. * vs / 1e9 / HBM_PEAK_GBS}
    return out


def batched_extra(torch, seed, dn_g, units, lam, n0, batches=(16, 64)):
    """Batched form at 1024^2: B tilted plane waves through the whole
    benchmark dn (all its planes) in one launch sequence per pass,
    graph-replayed; exit fields + summed intensities."""
    from bpm_amd import _lib
    from bpm_amd import bpm3d as B
    from bpm_amd.workloads import plane_waves
    nz, ny, nx = dn_g.shape
    H = B.propagator_table(nx, ny, units, lam, n0, _lib.H_F64)
    rng = np.random.default_rng(seed)
    out = {"workload": "%d^2 x %d planes of the benchmark dn, B tilted plane waves "
                       "(|theta| <= 0.2 rad), exit fields + summed intensities" % (nx, nz)}
    for batch in batches:
        u0_g = torch.from_numpy(plane_waves(nx, ny, units[0], units[1], lam, n0,
                                            rng.uniform(-.2, .2, (batch, 2)))).cuda()
        plane = torch.empty_like(u0_g)
        acc = torch.empty((batch, ny, nx), dtype=torch.float32, device="cuda")
        plan = B.get_plan(nx, ny, batch)
        _lib.check(_lib.load().bpmh_plan_set_graph(plan, 1))

        def step():
            plane.copy_(u0_g)
            B.propagate_fields(plane, nz, units, lam, n0, dn=dn_g, dn_kind=_lib.DN_F32,
                               mean_mode="ignore", H=H, out_sum=acc)

        step()
        torch.cuda.synchronize()
        reps = 2
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        ev0.record()
        for _ in range(reps):
            step()
        ev1.record()
        torch.cuda.synchronize()
        _lib.check(_lib.load().bpmh_plan_set_graph(plan, 0))
        ms = ev0.elapsed_time(ev1) / reps
        vs = nx * ny * (nz - 1) * batch / (ms * 1e-3)
        out["batch%d" % batch] = {"ms_per_propagation": ms,
                                  "us_per_field_step": ms * 1e3 / (nz - 1) / batch,
                                  "voxel_steps_per_s": vs,
                                  "effective_GBps_36B": 36. * vs / 1e9,
                                  "hbm_frac_36B": 36. * vs / 1e9 / HBM_PEAK_GBS,
                                  "effective_GBps_44B": 44. * vs / 1e9,
                                  "hbm_frac_44B": 44. * vs / 1e9 / HBM_PEAK_GBS}
        del plane, acc, u0_g
    return out


def weighted_extra(torch, dn_g, u0_g, units, lam, n0, reps=3):
    """bpm_3d's default mean_mode "weighted" on the benchmark workload (1024^2
    x all planes of its dn, same focused beam, batch 1), graph-replayed: per
    step column pass A (FFT_y, k-space partial sums, dn0 formed once by the
    last work-group), column pass B (H(n0+dn0), IFFT_y) and the row pass,
    52 B per voxel-step."""
    from bpm_amd import _lib
    from bpm_amd import bpm3d as B
    nz, ny, nx = dn_g.shape
    plane = torch.empty_like(u0_g)
    acc = torch.empty((1, ny, nx), dtype=torch.float32, device="cuda")
    plan = B.get_plan(nx, ny)
    _lib.check(_lib.load().bpmh_plan_set_graph(plan, 1))

    def step():
        plane.copy_(u0_g)
        B.propagate_fields(plane, nz, units, lam, n0, dn=dn_g, dn_kind=_lib.DN_F32,
                           mean_mode="weighted", out_sum=acc)

    step()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(reps):
        step()
    ev1.record()
    torch.cuda.synchronize()
    _lib.check(_lib.load().bpmh_plan_set_graph(plan, 0))
    ms = ev0.elapsed_time(ev1) / reps
    vs = nx * ny * (nz - 1) / (ms * 1e-3)
    return {"workload": "%d^2 x %d planes, benchmark dn and focused beam, mean_mode weighted "
                        "(bpm_3d default), exit field + summed intensity, batch 1" % (nx, nz),
            "ms_per_propagation": ms, "us_per_step": ms * 1e3 / (nz - 1),
            "voxel_steps_per_s": vs, "effective_GBps_52B": 52. * vs / 1e9,
            "hbm_frac_52B": 52. * vs / 1e9 / HBM_PEAK_GBS}


def floor_extra(torch, dn_g):
    """Batch-1 floors of a two-pass step over all planes of the benchmark dn,
    each captured in one graph and replayed like the z-loop:
    * "copy_pair": two plane copies per step (plane -> tmp -> plane, device
      copies), 32 B per voxel-step;
    * "elementwise_pair": a plane copy (16 B/voxel) and a plane update with
      the dn slice (plane read and write + dn read, 20 B/voxel), 36 B per
      voxel-step, as torch element-wise kernels;
    * "pass_pattern_batch1/16": the passes' own access patterns and launch
      shapes without the transforms (pass_pattern_floor)."""
    nz, ny, nx = dn_g.shape
    plane = torch.zeros((ny, nx), dtype=torch.complex64, device="cuda")
    tmp = torch.empty_like(plane)

    def copy_pair(z):
        tmp.copy_(plane)
        plane.copy_(tmp)

    def elementwise_pair(z):
        torch.mul(plane, 1.0, out=tmp)
        torch.add(tmp, dn_g[z], out=plane)

    out = {"workload": "%d^2 x %d planes, batch 1, graph-replayed" % (nx, nz)}
    for name, fn, nbytes in (("copy_pair", copy_pair, 32.), ("elementwise_pair", elementwise_pair, 36.)):
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            for z in range(1, 3):  # warm-up outside the capture
                fn(z)
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            for z in range(1, nz):
                fn(z)
        g.replay()
        torch.cuda.synchronize()
        reps = 5
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        ev0.record()
        for _ in range(reps):
            g.replay()
        ev1.record()
        torch.cuda.synchronize()
        ms = ev0.elapsed_time(ev1) / reps
        vs = nx * ny * (nz - 1) / (ms * 1e-3)
        del g
        out[name] = {"bytes_per_voxel_step": nbytes, "ms_per_propagation": ms,
                     "us_per_step": ms * 1e3 / (nz - 1), "voxel_steps_per_s": vs,
                     "effective_GBps": nbytes * vs / 1e9,
                     "hbm_frac": nbytes * vs / 1e9 / HBM_PEAK_GBS}
    out.update(pass_pattern_floor(torch, dn_g))
    return out


def pass_pattern_floor(torch, dn_g, batches=(1, 16), nz_batched=129):
    """bpmh_floor_steps: per step the column pass's and the row pass's global
    loads and stores in the autotuned launch shapes of each batch size, with
    no transforms (36 B per voxel-step), graph-replayed like the z-loop.  The
    FFT passes' step time over this one's is the cost of the transforms on
    top of the step's data movement (its cross-XCD exchange included)."""
    from bpm_amd import _lib
    from bpm_amd.bpm3d import Plan, apply_tuning
    nz, ny, nx = dn_g.shape
    res = {}
    for B in batches:
        depth = nz if B == 1 else min(nz, nz_batched)
        plan = Plan(nx, ny)
        apply_tuning(plan, nx, ny, B)
        plane = torch.zeros((B, ny, nx), dtype=torch.complex64, device="cuda")

        def run():
            _lib.call("bpmh_floor_steps", plan, int(B), int(depth), _lib.ptr(plane),
                      _lib.ptr(dn_g), int(dn_g.numel()), _lib.stream_handle())

        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            run()
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            run()
        g.replay()
        torch.cuda.synchronize()
        reps = 5 if B == 1 else 3
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        ev0.record()
        for _ in range(reps):
            g.replay()
        ev1.record()
        torch.cuda.synchronize()
        ms = ev0.elapsed_time(ev1) / reps
        del g
        plan.close()
        vs = B * nx * ny * (depth - 1) / (ms * 1e-3)
        res["pass_pattern_batch%d" % B] = {
            "bytes_per_voxel_step": 36.0, "planes": depth,
            "us_per_field_step": ms * 1e3 / (depth - 1) / B, "voxel_steps_per_s": vs,
            "effective_GBps": 36. * vs / 1e9, "hbm_frac": 36. * vs / 1e9 / HBM_PEAK_GBS}
    return res


def differentiable_extra(seed):
    """The differentiable workload: 512^2 x 256 planes, 16 plane-wave angles,
    loss = sum |u_exit - target|^2 with targets through a second seeded dn,
    gradient in dn by the hand-written adjoint with checkpoints every 16
    planes.  Times forward + loss + backward."""
    import torch
    from bpm_amd import adjoint
    from bpm_amd.spheres import spheres_dn_device
    from bpm_amd.workloads import small_workload_spheres
    nz, ny, nx, nb = 256, 512, 512, 16
    units, lam, n0 = (.1, .1, .1), .5, 1.33
    rng = np.random.default_rng(seed)
    size = (nx, ny, nz)
    # the 256^2 workload's sphere recipe: a centred r = 2 um sphere (dn .05)
    # plus 20 spheres r ~ U(.5, 1.5) um, dn ~ U(.01, .03)
    dn = spheres_dn_device(size, units, *small_workload_spheres(rng, size, units))
    dn2 = spheres_dn_device(size, units, *small_workload_spheres(rng, size, units))
    th = rng.uniform(-.2, .2, (nb, 2))
    k = 2 * np.pi * n0 / lam
    x = np.arange(nx) * units[0]
    y = np.arange(ny) * units[1]
    u0 = torch.from_numpy(np.exp(-1j * k * (np.sin(th[:, 0])[:, None, None] * x[None, None, :] +
                                            np.sin(th[:, 1])[:, None, None] * y[None, :, None]))
                          .astype(np.complex64)).cuda()
    target = adjoint.propagate(dn2, u0, units, lam, n0).detach()

    def step():
        d = dn.clone().requires_grad_(True)
        out = adjoint.propagate(d, u0, units, lam, n0, checkpoint=16)
        (out - target).abs().pow(2).sum().backward()
        return d.grad

    step()
    torch.cuda.synchronize()
    reps = 3
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(reps):
        step()
    ev1.record()
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / reps
    vsteps = nx * ny * (nz - 1) * nb
    return {"workload": "512^2 x 256 planes, 16 angles, spheres as the 256^2 workload, L2 loss "
                        "vs second-dn targets, gradient in dn, checkpoint every 16 planes",
            "ms_forward_backward": ms, "forward_voxel_steps": vsteps,
            "voxel_steps_per_s_forward_equivalent": vsteps / (ms * 1e-3)}


def copy_reference(mib):
    """A plain device-to-device copy of known size (the scale for the HBM
    byte counters FETCH_SIZE / WRITE_SIZE read under rocprofv3 --pmc)."""
    import torch
    n = mib << 18  # float32 elements
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_reference_bytes": n * 4, "dispatches": 3}), flush=True)


def main():
    a = parse()
    if a.copy_reference:
        copy_reference(a.copy_reference)
        return
    if a.debye_only:
        import torch
        print(json.dumps({"debye": debye_extra(torch)}), flush=True)
        return
    if a.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(independent jobs would shard by job with no communication); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    import ctypes
    import torch
    from bpm_amd import _lib
    from bpm_amd import bpm3d as B
    from bpm_amd.spheres import spheres_dn_device
    from bpm_amd.workloads import gaussian_beam, plain_large_spheres

    Nx = Ny = a.nx
    Nz = a.nz
    units, lam, n0 = (.1, .1, .1), .5, 1.33
    t0 = time.perf_counter()
    sph = plain_large_spheres(seed=a.seed, size=(Nx, Ny, Nz), units=units)
    u0 = gaussian_beam(Nx, Ny, units[0], units[1], lam, n0, 3., -20.)
    dev = torch.device("cuda", 0)
    dn_g = spheres_dn_device((Nx, Ny, Nz), units, *sph)
    torch.cuda.synchronize()
    gen_s = time.perf_counter() - t0
    print("workload generated on device in %.1f s" % gen_s, file=sys.stderr, flush=True)
    u0_g = torch.from_numpy(u0).to(dev).reshape(1, Ny, Nx)
    plane = torch.empty_like(u0_g)
    acc = torch.empty((1, Ny, Nx), dtype=torch.float32, device=dev)
    H = B.propagator_table(Nx, Ny, units, lam, n0, _lib.H_F64)
    plan = B.get_plan(Nx, Ny)
    lib = _lib.load()

    def step():
        plane.copy_(u0_g)
        B.propagate_fields(plane, Nz, units, lam, n0, dn=dn_g, dn_kind=_lib.DN_F32,
                           mean_mode="ignore", H=H, out_sum=acc)

    _lib.check(lib.bpmh_plan_set_graph(plan, 0 if a.no_graph else 1))
    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()

    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record()
    for _ in range(a.steps):
        step()
    ev1.record()
    torch.cuda.synchronize()
    el_wall = time.perf_counter() - t0
    el = ev0.elapsed_time(ev1) * 1e-3

    # per-kernel durations: the same steps again, eagerly launched with a HIP
    # event pair around every kernel (on the stream the kernels run on)
    _lib.check(lib.bpmh_plan_set_timing(plan, 1))
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(a.steps):
        step()
    torch.cuda.synchronize()
    el_events = time.perf_counter() - t1
    ms = (ctypes.c_float * 4)()
    cnt = (ctypes.c_int * 4)()
    _lib.check(lib.bpmh_plan_kernel_times(plan, ms, cnt))
    _lib.check(lib.bpmh_plan_set_timing(plan, 0))
    # the shared plan goes back to eager launches (its captured graphs are
    # released) before the other workloads run
    _lib.check(lib.bpmh_plan_set_graph(plan, 0))

    ms_step = el / a.steps * 1e3
    vsteps = Nx * Ny * (Nz - 1)
    value = vsteps / (ms_step * 1e-3)
    # algorithmic bytes per launch (36 B per voxel-step: each pass reads and
    # writes the plane, 16 B per voxel, and the row pass reads dn, 4 B; the
    # z = 0 row pass has no screen, so no dn read).  The summed-intensity
    # accumulator (read + write, 8 B per voxel per row pass) is not credited;
    # the bytes moved with it are reported beside.
    row_bytes = (16. + 20. * (Nz - 1)) / Nz * Nx * Ny
    row_bytes_sum = (24. + 28. * (Nz - 1)) / Nz * Nx * Ny
    col_bytes = 16. * Nx * Ny
    ev_total = sum(ms[:4])
    kern = {}
    for i, (name, nbytes) in enumerate((("row_pass", row_bytes), ("col_pass", col_bytes))):
        launches = cnt[i] / a.steps  # per step
        share = ms[i] / ev_total
        avg_ms = share * ms_step / launches
        kern[name] = {"avg_us": avg_ms * 1e3, "launches_per_step": launches,
                      "event_avg_us": ms[i] / cnt[i] * 1e3, "share": share,
                      "alg_bytes": nbytes, "GBps": nbytes / (avg_ms * 1e-3) / 1e9}
    kern["row_pass"]["bytes_with_sum"] = row_bytes_sum
    kern["row_pass"]["GBps_with_sum"] = row_bytes_sum / (kern["row_pass"]["avg_us"] * 1e-6) / 1e9
    dom = max(kern, key=lambda k: kern[k]["share"])
    ach = kern[dom]["GBps"]
    step_gbs = 36. * vsteps / (ms_step * 1e-3) / 1e9
    step_gbs44 = 44. * vsteps / (ms_step * 1e-3) / 1e9
    traffic = measured_traffic(dom, Nx, Ny, plan_shape(lib, plan, dom))

    out = {
        "metric": "voxel-steps/s (split-step BPM, 1024^2 grid)",
        "value": value,
        "unit": "voxel-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic",
        "config": {"workload": "plain BPM %dx%dx%d, voxel 0.1um, lam 0.5um, n0 1.33, "
                               "Gaussian beam w0 3um focused 20um deep, 400 seeded spheres, "
                               "Bpm3d propagate path; exit field + summed intensity"
                               % (Nx, Ny, Nz),
                   "nx": Nx, "ny": Ny, "nz": Nz, "batch": 1, "seed": a.seed},
        "hbm_frac_step": step_gbs / HBM_PEAK_GBS,
        "step_effective_GBps": step_gbs,
        "step_bytes_per_voxel_step": 36,
        "hbm_frac_step_44B": step_gbs44 / HBM_PEAK_GBS,
        "note_44B": "44 B per voxel-step also counts the summed-intensity read and write",
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": ach, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                     "traffic": traffic["bytes"] if traffic else None},
        "traffic_source": traffic["source"] if traffic else None,
        "kernels": kern,
        "launch": "eager" if a.no_graph else "hip-graph",
        "ms_per_step_wall": el_wall / a.steps * 1e3,
        "ms_per_step_eager_with_events": el_events / a.steps * 1e3,
    }
    if not a.no_debye:
        out["debye"] = debye_extra(torch)
    if not a.no_served:
        out["served"] = served_extra(a.seed)
    if not a.no_differentiable:
        out["differentiable"] = differentiable_extra(a.seed)
    if not a.no_small:
        out["small_256"] = small_extra(torch, a.seed)
    if not a.no_batched:
        out["batched_1024"] = batched_extra(torch, a.seed, dn_g, units, lam, n0)
    if not a.no_floor:
        out["floor_batch1"] = floor_extra(torch, dn_g)
    if not a.no_weighted:
        out["weighted_1024"] = weighted_extra(torch, dn_g, u0_g, units, lam, n0)
    if not a.no_cpu_baseline:
        from bpm_amd.workloads import spheres_dn
        planes = min(a.cpu_sample_planes, Nz)
        planes1 = min(a.cpu_sample_planes_1core, Nz)
        dn_s = spheres_dn((Nx, Ny, max(planes, planes1)), units, *sph)
        out["cpu_baseline"] = cpu_baseline(u0, dn_s, units, lam, n0, planes, planes1)
        out["gpu_over_cpu"] = value / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
