"""Autotuner for the launch shapes and radix schedules of the split-step passes.

For a grid (Nx, Ny) and batch size it times every (row, column) variant of
``bpmh_plan_set_variant`` -- work-group shapes, paired or split column
accesses, and for each pass both radix schedules (4 or 8 elements per
thread, e.g. a 1024-point line as 4.4.4.4.4 or 2.8.8.8) -- on the device
(HIP events on the current stream around a short propagation) and keeps the
fastest.  Results are cached in
``tuning_cache.json`` next to the library, keyed by "gfx:NxxNy:batch";
``get_plan`` applies a cached choice automatically.
"""
import ctypes
import json
import os

import numpy as np

from . import _lib

CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning_cache.json")
_cache = None


def _key(nx, ny, batch):
    return "gfx950:%dx%d:%d" % (nx, ny, batch)


def load_cache():
    global _cache
    if _cache is None:
        _cache = {}
        if os.path.exists(CACHE):
            with open(CACHE) as f:
                _cache = json.load(f)
    return _cache


def lookup(nx, ny, batch=1):
    """Cached (row_variant, col_variant) for this grid and batch, else for
    the nearest cached batch of the same grid, else None."""
    c = load_cache()
    k = _key(nx, ny, batch)
    if k in c:
        return tuple(c[k]["best"])
    same = [(abs(int(v.split(":")[2]) - batch), v) for v in c if v.split(":")[1] == "%dx%d" % (nx, ny)]
    if same:
        return tuple(c[min(same)[1]]["best"])
    return None


ROW_VARIANTS, COL_VARIANTS = 6, 10  # V + 3*s, V + 5*s (s: radix schedule)


def shape_of(plan):
    """(row variant, column variant, lines per row work-group, columns per
    column tile, row elements per thread, column elements per thread)."""
    v = [ctypes.c_int() for _ in range(6)]
    _lib.call("bpmh_plan_get_variant", plan, *[ctypes.byref(x) for x in v])
    return tuple(x.value for x in v)


def autotune(nx, ny, batch=1, nz=65, reps=3, units=(.1, .1, .1), lam=.5, n0=1.33, save=None):
    """Time all 6x10 (row, column) variants for one grid/batch (variants that
    map to an identical kernel pair are timed once); returns (best, table)."""
    from . import bpm3d as B
    t = _lib.torch()
    plan = B.get_plan(nx, ny, None)
    rng = np.random.default_rng(0)
    dn = t.from_numpy(rng.uniform(0, .05, (nz, ny, nx)).astype(np.float32)).to(_lib.device())
    u0 = t.ones((batch, ny, nx), dtype=t.complex64, device=_lib.device())
    H = B.propagator_table(nx, ny, units, lam, n0, _lib.H_F64)
    acc = t.empty((batch, ny, nx), dtype=t.float32, device=_lib.device())
    plane = t.empty_like(u0)
    table = {}
    for rv in range(ROW_VARIANTS):
        for cv in range(COL_VARIANTS):
            _lib.call("bpmh_plan_set_variant", plan, rv, cv)
            shp = shape_of(plan)
            key = [shp[2], shp[3], shp[4], shp[5]]
            if cv % 5 in (3, 4) and 256 <= ny <= 1024:
                key.append(cv % 5)  # paired / split columns: kernels of their own
            if key in [v["shape"] for v in table.values()]:
                continue  # clamped to an identical launch shape
            def run():
                plane.copy_(u0)
                B.propagate_fields(plane, nz, units, lam, n0, dn=dn, dn_kind=_lib.DN_F32,
                                   mean_mode="ignore", H=H, out_sum=acc, tuned=False)
            run()
            t.cuda.synchronize()
            e0, e1 = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                run()
            e1.record()
            t.cuda.synchronize()
            ms = e0.elapsed_time(e1) / reps
            table["%d,%d" % (rv, cv)] = {"ms": ms, "shape": key,
                                         "us_per_step": ms * 1e3 / (nz - 1)}
    best = min(table, key=lambda k: table[k]["ms"])
    bv = tuple(int(x) for x in best.split(","))
    _lib.call("bpmh_plan_set_variant", plan, *bv)
    c = load_cache()
    c[_key(nx, ny, batch)] = {"best": list(bv), "table": table}
    if save:
        with open(save, "w") as f:
            json.dump(c, f, indent=1, sort_keys=True)
    return bv, table


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="autotune the split-step launch shapes")
    ap.add_argument("--grids", default="1024x1024,512x512,256x256,128x128")
    ap.add_argument("--batches", default="1")
    ap.add_argument("--nz", type=int, default=65)
    ap.add_argument("--out", default=CACHE)
    a = ap.parse_args(argv)
    for g in a.grids.split(","):
        nx, ny = (int(v) for v in g.split("x"))
        for b in (int(v) for v in a.batches.split(",")):
            best, table = autotune(nx, ny, b, nz=a.nz, save=a.out)
            print("%dx%d batch %d best %s %s" % (nx, ny, b, best,
                  {k: round(v["us_per_step"], 2) for k, v in table.items()}), flush=True)


if __name__ == "__main__":
    main()
