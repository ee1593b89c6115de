#!/usr/bin/env python3
"""Benchmark: plain Marching Cubes on a 1024^3 value-noise field, one MI355X.

A step is one full isosurface extraction of a field already resident in HBM:
msh_count (classify + count + reduce-then-scan + the 16-byte size
read-back) followed by msh_extract into preallocated device buffers, i.e.
everything the drop-in `isosurface(A, MarchingCubes(iso=0))` does except
allocating its result arrays.

Workload (default): Float64 1024^3 value noise (32^3 lattice of seeded
Uniform(-1,1) values, trilinearly upsampled, plus 0.25 x a 128^3 octave),
MarchingCubes(iso=0), X=Y=Z=-1:1.  The same field in Float32 is measured in
the same run and reported under "float32"; "sphere_1024_f32" is the sparse
point (Float32 sphere of radius 0.5, few outputs) and "gyroid_256_f32" the
Marching Tetrahedra / surface nets / Marching Cubes workloads.

Printed: one JSON line.  `value` is Gvoxels/s (samples per second / 1e9).
`roofline` is for the dominant kernel (the Marching Cubes emit kernel,
k_emit_mc5; on the sparse sphere the count kernel, since the emit fetches
only the z chunks with active cells), from HIP events on the stream the
library launches on; its
`traffic` is the HBM bytes per launch from the rocprofv3 FETCH_SIZE and
WRITE_SIZE summaries kept in profiles/ (null when absent).  `cpu_baseline`
times the CPU oracle (single thread) on an x-slab of the same field, and
`cpu_baseline_all_cores` the same sample on up to 16 threads.  "weld" under
"float64"/"float32" times the opt-in vertex welding (weld()) of the same
mesh (algorithmic bytes: vertices read, faces read and written, welded
vertices written).  "vertex_error_float32_vs_float64" compares the mesh of
each Float32 field with the mesh of the same samples as Float64.  A
1 GiB device-to-device copy is timed at the end as the practical HBM rate
of the same box (and is the known-size kernel the counter scale factor is
read from).
"""
import argparse
import ctypes
import json
import os
import re
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
    ap.add_argument("--batch-volumes", type=int, default=512,
                    help="volumes of the batched-vs-plain measurement (0 skips it)")
    ap.add_argument("--serve-requests", type=int, default=200,
                    help="requests of the serving measurement (0 skips it)")
    ap.add_argument("--serve-qps", default="200,400,800,1200,1600,2000,2400,0",
                    help="offered request rates of the serving sweep (0: all at once)")
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


def measure(m, dtype, n, method_name, steps, warmup, kind="value_noise", with_weld=False):
    import torch
    meth = {"mc": m.MarchingCubes(iso=0), "mt": m.MarchingTetrahedra(iso=0),
            "sn": m.NaiveSurfaceNets(iso=0)}[method_name]
    field = m.generate_field(kind, (n, n, n), dtype, seed=1)
    torch.cuda.synchronize()
    family = m.autotune(field, meth)  # the emit kernel family for this (method, dtype)
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
               family=m.KERNEL_FAMILIES[family], family_id=family,
               emit_s=t_emit, field_bytes=field_bytes, out_bytes=out_bytes,
               gvox_s=N / dt / 1e9, tris_s=r.nf / dt, weld=None)
    if with_weld:
        # the opt-in welding step on this mesh: one weld() call (output
        # allocation, the kernels, the welded-count read-back), best of 3
        ts = []
        for _ in range(3):
            ev[0].record()
            Vw, Fw = m.weld(r.V, r.F)
            ev[1].record()
            torch.cuda.synchronize()
            ts.append(ev[0].elapsed_time(ev[1]) / 1e3)
        nu = Vw.shape[0]
        del Vw, Fw
        tw = min(ts)
        wb = r.nv * 3 * osz + 2 * r.nf * 3 * 8 + nu * 3 * osz
        res["weld"] = dict(ms=round(tw * 1e3, 3), nverts=r.nv, nunique=nu,
                           algorithmic_bytes=wb, achieved_GB_s=round(wb / tw / 1e9, 1),
                           frac=round(wb / tw / HBM_PEAK, 4))
    del r, field
    torch.cuda.empty_cache()
    return res


class BatchRunner:
    """The batched form through the C ABI, outputs preallocated per volume."""

    def __init__(self, m, fields, methods):
        import torch
        from importlib import import_module
        self.capi = capi = import_module(m.__name__ + "._capi")
        self.L = L = capi.lib()
        n = self.n = len(fields)
        self.fields = fields
        self.F = (capi.MshField * n)(*[m._field(f)[0] for f in fields])
        self.P = (capi.MshParams * n)(*[m._params(me, (-1, 1), (-1, 1), (-1, 1))
                                         for me in methods])
        self.ctx = capi.context(fields[0].device.index)
        # a stream of its own (default priority): a repeated batch then runs
        # each phase from a captured hipGraph, which the null stream cannot hold
        self.stream = torch.cuda.Stream(device=fields[0].device)
        self.ctx.check(L.msh_set_stream(self.ctx.handle, ctypes.c_void_p(self.stream.cuda_stream)))
        self.nv = (ctypes.c_int64 * n)()
        self.nf = (ctypes.c_int64 * n)()
        self.ctx.check(L.msh_batch_count(self.ctx.handle, n, self.F, self.P, self.nv, self.nf))
        self.vp = (ctypes.c_void_p * n)()
        self.fp = (ctypes.c_void_p * n)()
        self.keep = []
        for v in range(n):
            ok = ctypes.c_int32()
            self.ctx.check(L.msh_output_kind(ctypes.byref(self.F[v]), ctypes.byref(self.P[v]),
                                             ctypes.byref(ok)))
            dt = torch.float32 if ok.value == capi.MSH_FLOAT32 else torch.float64
            V = torch.empty((max(self.nv[v], 1), 3), dtype=dt, device=fields[v].device)
            Fv = torch.empty((max(self.nf[v], 1), 3), dtype=torch.int64, device=fields[v].device)
            self.keep += [V, Fv]
            self.vp[v], self.fp[v] = V.data_ptr(), Fv.data_ptr()

    def step(self):
        L, c = self.L, self.ctx
        c.check(L.msh_set_stream(c.handle, ctypes.c_void_p(self.stream.cuda_stream)))
        c.check(L.msh_batch_count(c.handle, self.n, self.F, self.P, self.nv, self.nf))
        c.check(L.msh_batch_extract(c.handle, self.n, self.F, self.P, self.vp, self.fp,
                                    self.capi.MSH_DEVICE))


def batch_workload(m, nvol, seed):
    """Mixed-size batch: axes iid Uniform{32..256}; field a seeded choice of
    sphere / torus (R=0.6, r=0.2) / gyroid / 16-Gaussian blobs; method a
    seeded choice of MarchingCubes / MarchingTetrahedra / NaiveSurfaceNets,
    iso=0; Float32 fields."""
    import numpy as np
    rng = np.random.default_rng(seed)
    kinds = ["sphere", "torus", "gyroid", "blobs"]
    meths = [m.MarchingCubes(iso=0), m.MarchingTetrahedra(iso=0), m.NaiveSurfaceNets(iso=0)]
    fields, methods = [], []
    for v in range(nvol):
        shape = tuple(int(x) for x in rng.integers(32, 257, size=3))
        fields.append(m.generate_field(kinds[int(rng.integers(4))], shape, "float32",
                                       seed=seed * 100003 + v))
        methods.append(meths[int(rng.integers(3))])
    return fields, methods


def measure_batch(m, nvol, steps, seed=7):
    """Plain (one count + extract per volume, back to back) against batched
    (one msh_batch_count + msh_batch_extract for all volumes)."""
    import torch
    fields, methods = batch_workload(m, nvol, seed)
    torch.cuda.synchronize()
    # kernel families chosen per (method, element type, size class) by
    # timing the batch's own groups as batched launches; the plain loop then
    # uses the same choices
    chosen = m.autotune_batch(fields, methods)
    tuned = {}
    for f, me, fam in zip(fields, methods, chosen):
        cells = (f.shape[0] - 1) * (f.shape[1] - 1) * (f.shape[2] - 1)
        tuned[f"{['mc', 'mt', 'sn'][me.method]}/"
              f"{0 if cells < 2 ** 20 else 1 if cells < 2 ** 23 else 2}"] = m.KERNEL_FAMILIES[fam]
    tuned = dict(sorted(tuned.items()))
    plain = [Runner(m, f, me) for f, me in zip(fields, methods)]
    br = BatchRunner(m, fields, methods)
    for v in range(nvol):
        assert (plain[v].nv, plain[v].nf) == (br.nv[v], br.nf[v])
    samples = sum(f.numel() for f in fields)
    # warm-up
    for r in plain:
        r.step()
    br.step()
    torch.cuda.synchronize()
    t_plain = []
    for _ in range(steps):
        t0 = time.perf_counter()
        for r in plain:
            r.step()
        torch.cuda.synchronize()
        t_plain.append(time.perf_counter() - t0)
    t_batch = []
    for _ in range(steps):
        t0 = time.perf_counter()
        br.step()
        torch.cuda.synchronize()
        t_batch.append(time.perf_counter() - t0)
    tp, tb = min(t_plain), min(t_batch)
    g = (ctypes.c_int64(), ctypes.c_int64())
    br.ctx.check(br.L.msh_batch_graph_launches(br.ctx.handle, ctypes.byref(g[0]),
                                               ctypes.byref(g[1])))
    res = dict(volumes=nvol, samples=samples, seed=seed,
               graph_launches={"count_phase": g[0].value, "emit_phase": g[1].value},
               autotuned_emit_kernels=tuned,
               faces=sum(br.nf[v] for v in range(nvol)),
               plain_ms=round(tp * 1e3, 3), batched_ms=round(tb * 1e3, 3),
               plain_req_s=round(nvol / tp, 1), batched_req_s=round(nvol / tb, 1),
               batched_over_plain=round(tp / tb, 3),
               plain_gvoxels_s=round(samples / tp / 1e9, 3),
               batched_gvoxels_s=round(samples / tb / 1e9, 3))
    del plain, br, fields
    torch.cuda.empty_cache()
    return res


def serve_workload(m, nreq, seed):
    """Serving requests: value noise (as the 1024^3 field, at the request's
    size), cubic size log-uniform in 64^3..512^3, Float32 80% / Float64 20%,
    method a seeded choice of MC / MT / SN with iso=0.  Fields are generated
    and resident before the run."""
    import numpy as np
    rng = np.random.default_rng(seed)
    meths = [m.MarchingCubes(iso=0), m.MarchingTetrahedra(iso=0), m.NaiveSurfaceNets(iso=0)]
    out = []
    for v in range(nreq):
        n = int(round(64 * 8 ** rng.random()))
        dt = "float64" if rng.random() < 0.2 else "float32"
        f = m.generate_field("value_noise", (n, n, n), dt, seed=seed * 100003 + v)
        out.append((f, meths[int(rng.integers(3))]))
    return out


def plain_request(m, field, method, capi, L, ctx):
    """One plain extraction the way a caller without batching runs it:
    count, allocate the outputs, extract."""
    import torch
    f = m._field(field)[0]
    p = m._params(method, (-1, 1), (-1, 1), (-1, 1))
    nv, nf, ok = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int32()
    ctx.check(L.msh_count(ctx.handle, ctypes.byref(f), ctypes.byref(p), ctypes.byref(nv),
                          ctypes.byref(nf)))
    ctx.check(L.msh_output_kind(ctypes.byref(f), ctypes.byref(p), ctypes.byref(ok)))
    dt = torch.float32 if ok.value == capi.MSH_FLOAT32 else torch.float64
    V = torch.empty((nv.value, 3), dtype=dt, device=field.device)
    F = torch.empty((nf.value, 3), dtype=torch.int64, device=field.device)
    ctx.check(L.msh_extract(ctx.handle, ctypes.byref(f), ctypes.byref(p),
                            ctypes.c_void_p(V.data_ptr()), ctypes.c_void_p(F.data_ptr()),
                            capi.MSH_DEVICE))
    return V, F


def measure_serve(m, nreq, qps_list, seed=11):
    """Plain sequential extraction of the request set (count, allocate,
    extract, hand the result to the client, which drops it), then the
    continuous-batching server under Poisson arrivals at each offered rate;
    the client takes each result in a completion callback and drops it."""
    import numpy as np
    import torch
    from importlib import import_module
    capi = import_module(m.__name__ + "._capi")
    L = capi.lib()
    reqs = serve_workload(m, nreq, seed)
    torch.cuda.synchronize()
    samples = sum(f.numel() for f, _ in reqs)
    ctx = capi.context(reqs[0][0].device.index)
    ctx.check(L.msh_set_stream(ctx.handle,
                               ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
    for f, me in reqs:  # warm-up: workspace and the allocator's cache
        plain_request(m, f, me, capi, L, ctx)
    torch.cuda.synchronize()
    lat = []
    t0 = time.perf_counter()
    for f, me in reqs:
        t1 = time.perf_counter()
        out = plain_request(m, f, me, capi, L, ctx)
        torch.cuda.synchronize()
        del out
        lat.append(time.perf_counter() - t1)
    tp = time.perf_counter() - t0
    res = dict(requests=nreq, samples=samples, seed=seed,
               plain=dict(req_s=round(nreq / tp, 1), p50_ms=round(np.percentile(lat, 50) * 1e3, 3),
                          p99_ms=round(np.percentile(lat, 99) * 1e3, 3)),
               served=[])
    # warm-up: one untimed pass with every request queued at once brings
    # the allocator's cache for the server streams to its steady state
    with m.Server(hbm_budget=64 << 30, max_batch=128) as srv:
        for fu in [srv.submit(f, me, on_done=lambda r, l: None) for f, me in reqs]:
            fu.result()
    rng = np.random.default_rng(seed + 1)
    for qps in qps_list:
        gaps = rng.exponential(1.0 / qps, size=nreq) if qps else np.zeros(nreq)
        lats = []
        with m.Server(hbm_budget=64 << 30, max_batch=128) as srv:
            for f, me in reqs[:8]:  # warm the server's contexts and workspaces
                srv.submit(f, me).result()
            base = dict(srv.stats)
            futs = []
            t0 = time.perf_counter()
            due = t0
            for (f, me), gap in zip(reqs, gaps):
                due += gap
                d = due - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                futs.append(srv.submit(f, me, on_done=lambda r, l: lats.append(l)))
            for fu in futs:
                fu.result()
            span = time.perf_counter() - t0
            st = {k: srv.stats[k] - base[k] for k in ("batches", "requests")}
        res["served"].append(dict(offered_qps=qps or "all at t=0", req_s=round(nreq / span, 1),
                                  p50_ms=round(np.percentile(lats, 50) * 1e3, 3),
                                  p99_ms=round(np.percentile(lats, 99) * 1e3, 3),
                                  batches=st["batches"],
                                  mean_batch=round(st["requests"] / max(st["batches"], 1), 2)))
    best = max(x["req_s"] for x in res["served"])
    res["served_over_plain"] = round(best / res["plain"]["req_s"], 3)
    del reqs
    torch.cuda.empty_cache()
    return res


def vertex_error(m, kind, n, method_name):
    """The mesh of a Float32 field against the mesh of the same samples as
    Float64 (the values convert exactly, so classification and faces are the
    same and only the arithmetic type differs): max |difference| of the
    vertex coordinates, that over the cell size and in Float32 ulps of the
    axis end, and the share of coordinates within one Float32 ulp of the
    Float64 coordinate (near zero an ulp is tiny, so a few are not)."""
    import numpy as np
    import torch
    meth = {"mc": m.MarchingCubes(iso=0), "mt": m.MarchingTetrahedra(iso=0),
            "sn": m.NaiveSurfaceNets(iso=0)}[method_name]
    f32 = m.generate_field(kind, (n, n, n), "float32", seed=1)
    f64 = f32.permute(2, 1, 0).double().permute(2, 1, 0)  # x fastest, exact values
    V32, F32 = m.isosurface(f32, meth)
    V64, F64 = m.isosurface(f64, meth)
    del f32, f64
    same = V32.shape == V64.shape and torch.equal(F32, F64)
    res = dict(method=method_name, field=f"{kind} {n}^3", out_dtype=str(V32.dtype).split(".")[-1],
               same_faces=bool(same))
    if same and V32.numel():
        d = (V32.double() - V64).abs()
        e = torch.frexp(V64.abs().clamp_min(2.0 ** -126))[1]
        ulp = torch.ldexp(torch.ones_like(V64), e - 24)
        half = 2 * np.pi if kind == "gyroid" else 1.0  # the axes run over [-half, half]
        cell = 2 * half / (n - 1)
        ulp_range = float(np.spacing(np.float32(half)))
        res.update(max_abs=float(d.max()), max_over_cell=float(d.max()) / cell,
                   max_in_f32_ulps_of_axis_end=float(d.max()) / ulp_range,
                   frac_within_1_f32_ulp_of_coordinate=round(float((d <= ulp).double().mean()), 6))
        del d, e, ulp
    del V32, F32, V64, F64
    torch.cuda.empty_cache()
    return res


def drop_in_host(m, n=1024, dtype="float32"):
    """The drop-in call as the reference is called: a host array in, host
    arrays out (PCIe upload of the field, count, emit, download of the
    mesh), on the same value-noise field; best of 2 after a warm-up call."""
    import numpy as np
    import torch
    f = m.generate_field("value_noise", (n, n, n), dtype, seed=1)
    host = np.asfortranarray(f.cpu().numpy())  # [x, y, z], x fastest
    del f
    torch.cuda.empty_cache()
    meth = m.MarchingCubes(iso=0)
    ts = []
    for _ in range(3):
        t0 = time.perf_counter()
        V, F = m.isosurface(host, meth)
        ts.append(time.perf_counter() - t0)
        nv, nf, vb = len(V), len(F), V.nbytes + F.nbytes
        del V, F
    t = min(ts[1:])
    del host
    return dict(workload=f"MarchingCubes(iso=0), {dtype} {n}^3 value noise, numpy field in, "
                         "numpy mesh out", seconds=round(t, 4), gvoxels_s=round(n ** 3 / t / 1e9, 3),
                field_bytes=n ** 3 * (4 if dtype == "float32" else 8), mesh_bytes=vb,
                nverts=nv, nfaces=nf)


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


TRAFFIC_FILE = ROOT / "profiles" / "traffic_emit_kernels.json"


def traffic_of(workload, family_id):
    """HBM bytes per emit launch of `workload` from the committed rocprofv3
    counter summary (FETCH_SIZE + WRITE_SIZE, scaled by a copy kernel of
    known size), when the summary was taken on the same emit kernel family
    as the one timed here; otherwise None."""
    if not TRAFFIC_FILE.exists():
        return None
    e = json.loads(TRAFFIC_FILE.read_text()).get("workloads", {}).get(workload)
    if not e or e.get("family_id") != family_id:
        return None
    return e["bytes_per_launch"]


# the emit kernel an autotuned family launches, per method: (name, template
# prefix after the method parameter); MC's brick depth per element type and
# family (family 3 halves it)
_EMIT_NAMES = {"mc": {3: "k_emit_mc5", 2: "k_emit_mc5", 1: "k_emit_mc4", 0: "k_emit"},
               "mt": {2: "k_emit_h5", 1: "k_emit_halo", 0: "k_emit"},
               "sn": {3: "k_emit_h5", 2: "k_emit_h5", 1: "k_emit_halo", 0: "k_emit"}}
_MC5_ZC = {("float32", 2): 16, ("float32", 3): 8, ("float64", 2): 8, ("float64", 3): 4}
# brick depth of the tetrahedra / surface-nets emit per family
_H5_ZC = {("mt", 2): 4, ("sn", 2): 8, ("sn", 3): 4}


def _emit_pattern(method, dtype, family_id):
    """A regular expression matching the demangled name of the emit kernel
    that `family_id` launches for this (method, field type)."""
    import re
    name = _EMIT_NAMES[method][family_id]
    t = "float" if dtype == "float32" else "double"
    if name == "k_emit_mc5":
        return re.compile(rf"k_emit_mc5<{t}, [a-z]+, [a-z]+, {_MC5_ZC[(dtype, family_id)]},")
    mid = {"mc": 0, "mt": 1, "sn": 2}[method]
    if name == "k_emit":
        return re.compile(rf"k_emit<{mid}, {t},")
    if name == "k_emit_h5":
        return re.compile(rf"k_emit_h5<{mid}, {t}, [a-z]+, [a-z]+, [a-z]+, "
                          rf"{_H5_ZC[(method, family_id)]},")
    return re.compile(rf"{name}<{mid}, {t},")


def summarize_counters(fetch_dir, write_dir, line_file, out_file=TRAFFIC_FILE, copy_bytes=1 << 30):
    """Turn the counter runs of this benchmark (rocprofv3 --pmc FETCH_SIZE,
    then --pmc WRITE_SIZE, each over `bench.py --steps 1 --warmup 0 --no-cpu
    --batch-volumes 0 --serve-requests 0`) into the traffic summary that the
    roofline reads.  `line_file` holds the JSON line the counter run printed:
    it says which emit family each workload timed, which picks the kernel
    instantiation.  Counter values are KiB per dispatch; each counter is
    scaled by what it reads for the closing device copies of `copy_bytes`
    (the last four __amd_rocclr_copyBuffer dispatches)."""
    import csv
    line = json.loads(Path(line_file).read_text().strip().splitlines()[-1])

    def rows(d, counter):
        path = next(Path(d).rglob("*counter_collection.csv"))
        with open(path, newline="") as fh:
            return [r for r in csv.DictReader(fh) if r["Counter_Name"] == counter]

    fetch, write = rows(fetch_dir, "FETCH_SIZE"), rows(write_dir, "WRITE_SIZE")
    scale = {}
    for counter, rs in (("FETCH_SIZE", fetch), ("WRITE_SIZE", write)):
        cp = [float(r["Counter_Value"]) for r in rs if "copyBuffer" in r["Kernel_Name"]][-4:]
        scale[counter] = (copy_bytes / 1024.0) / (sum(cp) / len(cp))
    # the workloads in the order the bench measures them
    loads = [("mc_value_noise_1024_float64", "mc", "float64", line["float64"], max),
             ("mc_value_noise_1024_float32", "mc", "float32", line["float32"], max)]
    for meth, g in line.get("gyroid_256_f32", {}).items():
        loads.append((f"{meth}_gyroid_256_float32", meth, "float32", g, min))
    for k, g in line.items():
        if k.startswith("sphere_") and k.endswith("_f32") and isinstance(g, dict):
            loads.append((f"mc_{k[:-4]}_float32", "mc", "float32", g, max))
    out = {"note": "HBM bytes per emit launch (FETCH_SIZE and WRITE_SIZE, KiB per dispatch, "
                   "averaged over the dispatches of the instantiation, times the scale a "
                   "1 GiB device copy gives each counter in the same run)",
           "scale": {k: round(v, 4) for k, v in scale.items()}, "workloads": {}}
    per = [{} for _ in loads]
    for counter, rs in (("FETCH_SIZE", fetch), ("WRITE_SIZE", write)):
        # each workload starts by generating its field, so workload i (in
        # bench order) owns the dispatches from the i-th field generation up
        # to the next: other workloads' autotuning launches the same
        # instantiations at the same grids
        gens = sorted(int(x["Dispatch_Id"]) for x in rs if "k_generate" in x["Kernel_Name"])
        for i, (key, meth, dtype, r, pick) in enumerate(loads):
            lo = gens[i] if i < len(gens) else 0
            hi = gens[i + 1] if i + 1 < len(gens) else float("inf")
            t = "float" if dtype == "float32" else "double"
            mid = {"mc": 0, "mt": 1, "sn": 2}[meth]
            for part, pat in (("emit", _emit_pattern(meth, dtype, r["family_id"])),
                              ("count", re.compile(rf"k_count4<{t}, {mid},"))):
                grids = {}
                for x in rs:
                    if lo <= int(x["Dispatch_Id"]) < hi and pat.search(x["Kernel_Name"]):
                        grids.setdefault((x["Kernel_Name"], int(x["Grid_Size"])), []).append(
                            float(x["Counter_Value"]))
                if not grids:
                    continue
                # the 1024^3 launches have the largest grid of their
                # instantiation, the 256^3 ones the smallest
                g = pick(grids, key=lambda k: k[1])
                vals = grids[g]
                per[i].setdefault(part, {})[counter] = dict(
                    kernel=g[0], grid=g[1], dispatches=len(vals), kib=sum(vals) / len(vals))
    for (key, meth, dtype, r, pick), pw in zip(loads, per):
        for part, pc in pw.items():
            if len(pc) < 2:
                continue
            fam = r["family_id"]
            fb = pc["FETCH_SIZE"]["kib"] * 1024 * scale["FETCH_SIZE"]
            wb = pc["WRITE_SIZE"]["kib"] * 1024 * scale["WRITE_SIZE"]
            # emit: field + outputs; count: the field (one read of every sample)
            alg = r.get("algorithmic_bytes") if part == "emit" else r.get("field_bytes")
            e = dict(kernel=pc["FETCH_SIZE"]["kernel"], grid=pc["FETCH_SIZE"]["grid"],
                     family_id=fam, dispatches=pc["FETCH_SIZE"]["dispatches"],
                     fetch_bytes=int(fb), write_bytes=int(wb), bytes_per_launch=int(fb + wb))
            if alg:
                e.update(algorithmic_bytes=alg, traffic_over_algorithmic=round((fb + wb) / alg, 3))
            out["workloads"][key if part == "emit" else key + "_count"] = e
    Path(out_file).write_text(json.dumps(out, indent=1) + "\n")
    return out


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
    nf = len(F)
    del V, F
    one = dict(value=slab.size / dt / 1e9, unit="Gvoxels/s", cores=1, kind="port",
               sample=f"MarchingCubes(iso=0) on x-planes 1..{planes} of the Float64 {n}^3 "
                      f"value-noise field ({slab.size} samples, {nf} faces), CPU oracle "
                      f"(single-threaded restatement), {dt:.2f} s")
    # the same sample over all host cores this process may use (x-slabs on
    # threads; at most 16, the CPU share of one GPU)
    nt = max(1, min(16, len(os.sched_getaffinity(0))))
    t0 = time.perf_counter()
    V, F = O.isosurface(slab, "mc", iso=0, threads=nt)
    dtt = time.perf_counter() - t0
    assert len(F) == nf
    del V, F
    allc = dict(value=slab.size / dtt / 1e9, unit="Gvoxels/s", cores=nt, kind="port",
                sample=f"as cpu_baseline, x-slabs of the CPU oracle on {nt} threads, {dtt:.2f} s")
    return one, allc


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "--summarize-counters":
        # maintenance: bench.py --summarize-counters FETCH_DIR WRITE_DIR LINE_FILE
        print(json.dumps(summarize_counters(*sys.argv[2:5]), indent=1))
        return
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part "
                                   "of it (one volume would shard by x-slabs with one "
                                   "all-gather of two counts per rank)"}))
        sys.exit(2)
    import torch
    import __graft_entry__ as entry
    m = entry.load_package()
    torch.cuda.init()
    def roofline(r, meth, workload, count=False):
        if count:
            # a sparse field: the emit fetches only the z chunks with active
            # cells, and the count pass, which reads every sample once,
            # dominates (its time includes the scan and the size read-back)
            b, t, kernel = r["field_bytes"], r["count_s"], "k_count4"
            traffic = traffic_of(workload + "_count", r["family_id"])
        else:
            b, t = r["field_bytes"] + r["out_bytes"], r["emit_s"]
            kernel = _EMIT_NAMES[meth].get(r["family_id"], "k_emit")
            traffic = traffic_of(workload, r["family_id"])
        ach = b / t
        out = {"bound": "hbm", "achieved": round(ach / 1e9, 2), "peak": HBM_PEAK / 1e9,
               "unit": "GB/s", "frac": round(ach / HBM_PEAK, 4), "traffic": traffic,
               "kernel": kernel, "algorithmic_bytes_per_launch": b,
               "avg_launch_ms": round(t * 1e3, 4)}
        if traffic:
            # the bytes the counters saw the kernel move, over the same launch time
            out["counter_GBps"] = round(traffic / t / 1e9, 1)
            out["counter_frac"] = round(traffic / t / HBM_PEAK, 4)
        return out

    f64 = measure(m, "float64", a.n, a.method, a.steps, a.warmup, with_weld=True)
    f32 = measure(m, "float32", a.n, a.method, a.steps, a.warmup, with_weld=True)
    main_r = f64
    gyroid = {}
    for meth in ("mt", "sn", "mc"):
        r = measure(m, "float32", 256, meth, a.steps, a.warmup, kind="gyroid")
        gyroid[meth] = dict(gvoxels_s=round(r["gvox_s"], 3), tris_s=round(r["tris_s"], 1),
                            autotuned_emit_kernels=r["family"], family_id=r["family_id"],
                            ms_per_step=round(r["step_s"] * 1e3, 4),
                            count_ms=round(r["count_s"] * 1e3, 4),
                            emit_ms=round(r["emit_s"] * 1e3, 4),
                            nverts=r["nverts"], nfaces=r["nfaces"], field_bytes=r["field_bytes"],
                            algorithmic_bytes=r["field_bytes"] + r["out_bytes"],
                            roofline=roofline(r, meth, f"{meth}_gyroid_256_float32"))
    # the sparse 1024^3 point: Float32 sphere of radius 0.5 (few outputs, so
    # the reads dominate)
    sph = measure(m, "float32", a.n, "mc", a.steps, a.warmup, kind="sphere")
    sphere = dict(gvoxels_s=round(sph["gvox_s"], 3), tris_s=round(sph["tris_s"], 1),
                  autotuned_emit_kernels=sph["family"], family_id=sph["family_id"],
                  ms_per_step=round(sph["step_s"] * 1e3, 4),
                  count_ms=round(sph["count_s"] * 1e3, 4), emit_ms=round(sph["emit_s"] * 1e3, 4),
                  nverts=sph["nverts"], nfaces=sph["nfaces"], field_bytes=sph["field_bytes"],
                  algorithmic_bytes=sph["field_bytes"] + sph["out_bytes"],
                  step_hbm_frac=round((sph["field_bytes"] + sph["out_bytes"]) / sph["step_s"]
                                      / HBM_PEAK, 4),
                  roofline=roofline(sph, "mc", f"mc_sphere_{a.n}_float32", count=True))
    batch = measure_batch(m, a.batch_volumes, max(2, min(a.steps, 5))) if a.batch_volumes else None
    serve = (measure_serve(m, a.serve_requests, [float(q) for q in a.serve_qps.split(",")])
             if a.serve_requests else None)
    verr = [vertex_error(m, "value_noise", a.n, a.method)] + \
        [vertex_error(m, "gyroid", 256, meth) for meth in ("mt", "sn", "mc")]
    host = drop_in_host(m, a.n) if a.n >= 256 else None
    copy = copy_probe()
    cpu, cpu_all = (None, None) if a.no_cpu else cpu_baseline(m, a.n, a.cpu_planes)

    def summary(r):
        b = r["field_bytes"] + r["out_bytes"]
        return dict(gvoxels_s=round(r["gvox_s"], 3), tris_s=round(r["tris_s"], 1),
                    ms_per_step=round(r["step_s"] * 1e3, 4),
                    count_ms=round(r["count_s"] * 1e3, 4), emit_ms=round(r["emit_s"] * 1e3, 4),
                    nverts=r["nverts"], nfaces=r["nfaces"], algorithmic_bytes=b,
                    field_bytes=r["field_bytes"],
                    autotuned_emit_kernels=r["family"], family_id=r["family_id"],
                    step_hbm_frac=round(b / r["step_s"] / HBM_PEAK, 4),
                    roofline=roofline(r, a.method, f"{a.method}_value_noise_{a.n}_{r['dtype']}"),
                    weld=r["weld"])

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
        "roofline": roofline(main_r, a.method, f"{a.method}_value_noise_{a.n}_float64"),
        "cpu_baseline": cpu,
        "cpu_baseline_all_cores": cpu_all,
        "hbm_copy_probe": copy,
        "drop_in_host_arrays": host,
        "gyroid_256_f32": gyroid,
        f"sphere_{a.n}_f32": sphere,
        "batched": batch,
        "serving": serve,
        "vertex_error_float32_vs_float64": verr,
        "float64": summary(f64),
        "float32": summary(f32),
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
