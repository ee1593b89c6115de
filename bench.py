#!/usr/bin/env python3
"""Benchmark: plain Marching Cubes on a 1024^3 value-noise field, one MI355X.

A step is one full isosurface extraction of a field already resident in HBM:
msh_count (classify + count + decoupled look-back scan + the 16-byte size
read-back) followed by msh_extract into preallocated device buffers, i.e.
everything the drop-in `isosurface(A, MarchingCubes(iso=0))` does except
allocating its result arrays.

Workload (default): Float64 1024^3 value noise (32^3 lattice of seeded
Uniform(-1,1) values, trilinearly upsampled, plus 0.25 x a 128^3 octave),
MarchingCubes(iso=0), X=Y=Z=-1:1.  The same field in Float32 is measured in
the same run and reported under "float32".

Printed: one JSON line.  `value` is Gvoxels/s (samples per second / 1e9).
`roofline` is for the dominant kernel (the Marching Cubes emit kernel,
k_emit_mc4), from HIP events on the stream the library launches on; its
`traffic` is the HBM bytes per launch from the rocprofv3 FETCH_SIZE and
WRITE_SIZE summaries kept in profiles/ (null when absent).  `cpu_baseline`
times the CPU oracle (single thread) on an x-slab of the same field.  A
1 GiB device-to-device copy is timed at the end as the practical HBM rate
of the same box (and is the known-size kernel the counter scale factor is
read from).
"""
import argparse
import ctypes
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK = 8.0e12  # bytes/s, MI355X


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1024, help="cube edge")
    ap.add_argument("--method", default="mc", choices=["mc", "mt", "sn"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-planes", type=int, default=512,
                    help="x-planes of the field the CPU baseline meshes")
    return ap.parse_args()


class Runner:
    """Drives the C ABI directly with preallocated outputs."""

    def __init__(self, m, field, method):
        import torch
        from importlib import import_module
        self.torch = torch
        self.capi = import_module(m.__name__ + "._capi")
        self.L = self.capi.lib()
        self.f, _, _ = m._field(field)
        self.field = field
        self.p = m._params(method, (-1, 1), (-1, 1), (-1, 1))
        self.ctx = self.capi.context(field.device.index)
        self.stream = torch.cuda.current_stream()
        self.ctx.check(self.L.msh_set_stream(self.ctx.handle,
                                             ctypes.c_void_p(self.stream.cuda_stream)))
        nv, nf = self.count()
        ok = ctypes.c_int32()
        self.ctx.check(self.L.msh_output_kind(ctypes.byref(self.f), ctypes.byref(self.p),
                                              ctypes.byref(ok)))
        self.out_dtype = torch.float32 if ok.value == self.capi.MSH_FLOAT32 else torch.float64
        self.V = torch.empty((nv, 3), dtype=self.out_dtype, device=field.device)
        self.F = torch.empty((nf, 3), dtype=torch.int64, device=field.device)
        self.nv, self.nf = nv, nf

    def count(self):
        nv, nf = ctypes.c_int64(), ctypes.c_int64()
        self.ctx.check(self.L.msh_count(self.ctx.handle, ctypes.byref(self.f),
                                        ctypes.byref(self.p), ctypes.byref(nv), ctypes.byref(nf)))
        return nv.value, nf.value

    def extract(self):
        self.ctx.check(self.L.msh_extract(self.ctx.handle, ctypes.byref(self.f),
                                          ctypes.byref(self.p),
                                          ctypes.c_void_p(self.V.data_ptr()),
                                          ctypes.c_void_p(self.F.data_ptr()),
                                          self.capi.MSH_DEVICE))

    def step(self):
        nv, nf = self.count()
        assert nv == self.nv and nf == self.nf
        self.extract()


def measure(m, dtype, n, method_name, steps, warmup):
    import torch
    meth = {"mc": m.MarchingCubes(iso=0), "mt": m.MarchingTetrahedra(iso=0),
            "sn": m.NaiveSurfaceNets(iso=0)}[method_name]
    field = m.generate_field("value_noise", (n, n, n), dtype, seed=1)
    torch.cuda.synchronize()
    r = Runner(m, field, meth)
    for _ in range(warmup):
        r.step()
    torch.cuda.synchronize()
    # whole steps
    t0 = time.perf_counter()
    for _ in range(steps):
        r.step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    # per-phase device time with events on the library's stream
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    t_count, t_emit = 0.0, 0.0
    for _ in range(steps):
        ev[0].record()
        r.count()
        ev[1].record()
        r.extract()
        ev[2].record()
        torch.cuda.synchronize()
        t_count += ev[0].elapsed_time(ev[1]) / 1e3
        t_emit += ev[1].elapsed_time(ev[2]) / 1e3
    t_count /= steps
    t_emit /= steps
    esz = 4 if dtype == "float32" else 8
    osz = 4 if r.out_dtype == torch.float32 else 8
    N = n ** 3
    field_bytes = N * esz
    out_bytes = r.nv * 3 * osz + r.nf * 3 * 8
    res = dict(dtype=dtype, samples=N, nverts=r.nv, nfaces=r.nf, step_s=dt, count_s=t_count,
               emit_s=t_emit, field_bytes=field_bytes, out_bytes=out_bytes,
               gvox_s=N / dt / 1e9, tris_s=r.nf / dt)
    del r, field
    torch.cuda.empty_cache()
    return res


def copy_probe():
    """Time a 1 GiB device-to-device tensor copy (2 GiB of HBM traffic)."""
    import torch
    a = torch.empty(256 * 1024 * 1024, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1.0)
    best = None
    for _ in range(4):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        b.copy_(a)
        e1.record()
        torch.cuda.synchronize()
        t = e0.elapsed_time(e1) / 1e3
        best = t if best is None else min(best, t)
    del a, b
    torch.cuda.empty_cache()
    return dict(bytes=2 * 1024 ** 3, best_ms=round(best * 1e3, 4),
                GBps=round(2 * 1024 ** 3 / best / 1e9, 1))


def traffic_of(method, n, dtype):
    """HBM bytes per emit launch from the committed rocprofv3 counter summary."""
    tfile = ROOT / "profiles" / f"traffic_{method}_value_noise_{n}.json"
    if not tfile.exists():
        return None
    return json.loads(tfile.read_text()).get(dtype, {}).get("bytes_per_launch")


def cpu_baseline(m, n, planes):
    """The CPU oracle (single-threaded restatement of the reference loop) on
    the first `planes` x-planes of the same Float64 field."""
    import numpy as np
    import torch
    import __graft_entry__ as entry
    O = entry.load_oracle()
    field = m.generate_field("value_noise", (n, n, n), "float64", seed=1)
    slab = field[:planes].cpu().numpy()  # [x, y, z], x-slab
    del field
    torch.cuda.empty_cache()
    slab = np.asfortranarray(slab)
    t0 = time.perf_counter()
    V, F = O.isosurface(slab, "mc", iso=0)
    dt = time.perf_counter() - t0
    return dict(value=slab.size / dt / 1e9, unit="Gvoxels/s", cores=1, kind="port",
                sample=f"MarchingCubes(iso=0) on x-planes 1..{planes} of the Float64 {n}^3 "
                       f"value-noise field ({slab.size} samples, {len(F)} faces), CPU oracle "
                       f"(single-threaded restatement), {dt:.2f} s")


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not "
                                   "part of it (see DESIGN.md for the x-slab sharding plan)"}))
        sys.exit(2)
    import torch
    import __graft_entry__ as entry
    m = entry.load_package()
    torch.cuda.init()
    f64 = measure(m, "float64", a.n, a.method, a.steps, a.warmup)
    f32 = measure(m, "float32", a.n, a.method, a.steps, a.warmup)
    main_r = f64
    copy = copy_probe()
    cpu = None if a.no_cpu else cpu_baseline(m, a.n, a.cpu_planes)
    kname = {"mc": "k_emit_mc4", "mt": "k_emit (tetrahedra)", "sn": "k_emit (surface nets)"}[a.method]

    def roofline(r):
        b = r["field_bytes"] + r["out_bytes"]
        ach = b / r["emit_s"]
        return {"bound": "hbm", "achieved": round(ach / 1e9, 2), "peak": HBM_PEAK / 1e9,
                "unit": "GB/s", "frac": round(ach / HBM_PEAK, 4),
                "traffic": traffic_of(a.method, a.n, r["dtype"]),
                "kernel": kname, "algorithmic_bytes_per_launch": b,
                "avg_launch_ms": round(r["emit_s"] * 1e3, 4)}

    def summary(r):
        b = r["field_bytes"] + r["out_bytes"]
        return dict(gvoxels_s=round(r["gvox_s"], 3), tris_s=round(r["tris_s"], 1),
                    ms_per_step=round(r["step_s"] * 1e3, 4),
                    count_ms=round(r["count_s"] * 1e3, 4), emit_ms=round(r["emit_s"] * 1e3, 4),
                    nverts=r["nverts"], nfaces=r["nfaces"], algorithmic_bytes=b,
                    step_hbm_frac=round(b / r["step_s"] / HBM_PEAK, 4),
                    roofline=roofline(r))

    out = {
        "metric": "Gvoxels/s + % HBM roofline (1024^3 MC)",
        "value": round(main_r["gvox_s"], 4),
        "unit": "Gvoxels/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(main_r["step_s"] * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (on-device seeded value noise)",
        "config": {"workload": f"{'MarchingCubes' if a.method == 'mc' else a.method}(iso=0) on "
                               f"Float64 {a.n}^3 value noise (32^3 lattice + 0.25 x 128^3 "
                               "octave, seed 1), X=Y=Z=-1:1",
                   "shape": [a.n] * 3, "seed": 1},
        "roofline": roofline(main_r),
        "cpu_baseline": cpu,
        "hbm_copy_probe": copy,
        "float64": summary(f64),
        "float32": summary(f32),
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
