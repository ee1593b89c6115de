"""The serving form: every result of a continuously batched request stream
equals the same extraction run alone, bit for bit, whatever batch it landed
in and whatever budget limited admission."""
import random
import threading
import time

import pytest

pytestmark = pytest.mark.gpu


def same(a, b):
    import torch
    (V, F), (Vs, Fs) = a, b
    assert V.dtype == Vs.dtype and V.shape == Vs.shape and F.shape == Fs.shape
    assert torch.equal(F, Fs)
    iv = torch.int32 if V.dtype == torch.float32 else torch.int64
    assert torch.equal(V.view(iv), Vs.view(iv))


def requests(pkg, n, seed):
    rng = random.Random(seed)
    kinds = ["sphere", "torus", "gyroid", "value_noise", "blobs"]
    meths = [pkg.MarchingCubes(iso=0), pkg.MarchingTetrahedra(iso=0), pkg.NaiveSurfaceNets(iso=0),
             pkg.MarchingCubes(iso=0.05)]
    out = []
    for v in range(n):
        shape = tuple(rng.randint(16, 90) for _ in range(3))
        dt = "float64" if rng.random() < 0.2 else "float32"
        spec = pkg.FieldSpec(rng.choice(kinds), shape, dt, seed=seed * 1000 + v)
        out.append((spec, rng.choice(meths)))
    return out


def alone(pkg, spec, method):
    f = pkg.generate_field(spec.kind, spec.shape, spec.dtype, seed=spec.seed)
    return pkg.isosurface(f, method)


@pytest.mark.parametrize("budget", [1 << 34, 24 << 20])
def test_served_results_equal_alone(pkg, budget):
    reqs = requests(pkg, 40, seed=budget & 0xFFFF)
    with pkg.Server(hbm_budget=budget, max_batch=16) as srv:
        futs = []
        rng = random.Random(5)

        def client():
            for spec, m in reqs:
                futs.append(srv.submit(spec, m))
                time.sleep(rng.random() * 0.002)

        t = threading.Thread(target=client)
        t.start()
        t.join()
        results = [f.result(timeout=60) for f in futs]
        stats = dict(srv.stats)
    assert stats["requests"] == len(reqs)
    assert stats["batches"] >= 3  # max_batch 16 alone forces several
    for (spec, m), r in zip(reqs, results):
        same(r, alone(pkg, spec, m))


def test_resident_tensor_requests_and_oversized(pkg):
    f_big = pkg.generate_field("gyroid", (96, 80, 72), "float32", seed=1)
    f_small = pkg.generate_field("sphere", (24, 25, 26), "float64", seed=2)
    m = pkg.MarchingTetrahedra(iso=0)
    # a budget below one request: each runs alone, still correct
    with pkg.Server(hbm_budget=1 << 16) as srv:
        a = srv.submit(f_big, m)
        b = srv.submit(f_small, pkg.MarchingCubes(iso=0), axes=((0, 1), (0, 2), (0, 3)))
        c = srv.submit(pkg.FieldSpec("torus", (50, 40, 30), "float32", seed=3),
                       pkg.NaiveSurfaceNets(iso=0))
        ra, rb, rc = a.result(timeout=60), b.result(timeout=60), c.result(timeout=60)
    same(ra, pkg.isosurface(f_big, m))
    same(rb, pkg.isosurface(f_small, pkg.MarchingCubes(iso=0), (0, 1), (0, 2), (0, 3)))
    same(rc, alone(pkg, pkg.FieldSpec("torus", (50, 40, 30), "float32", seed=3),
                   pkg.NaiveSurfaceNets(iso=0)))


def test_close_drains_queue(pkg):
    reqs = requests(pkg, 12, seed=9)
    srv = pkg.Server(max_batch=4)
    futs = [srv.submit(s, m) for s, m in reqs]
    srv.close()
    assert all(f.done() for f in futs)
    for (spec, m), f in zip(reqs, futs):
        same(f.result(), alone(pkg, spec, m))


def test_on_done_callback_streams_results(pkg):
    reqs = requests(pkg, 10, seed=21)
    got = {}
    with pkg.Server(max_batch=4) as srv:
        futs = [srv.submit(s, m, on_done=lambda r, lat, i=i: got.__setitem__(i, (r, lat)))
                for i, (s, m) in enumerate(reqs)]
        assert all(f.result(timeout=60) is None for f in futs)
    assert sorted(got) == list(range(len(reqs)))
    for i, (spec, m) in enumerate(reqs):
        r, lat = got[i]
        assert lat >= 0
        same(r, alone(pkg, spec, m))


def test_served_value_noise_64_to_512(pkg):
    # the serving workload at its real sizes: value noise (as the 1024^3
    # field, at the request's size), cubic sizes log-uniform in 64^3..512^3
    # (the ends forced in), Float32 80% / Float64 20%, methods mixed, all
    # requests in one stream so they share batches
    rng = random.Random(17)
    meths = [pkg.MarchingCubes(iso=0), pkg.MarchingTetrahedra(iso=0), pkg.NaiveSurfaceNets(iso=0)]
    sizes = [512, 64] + [int(round(64 * 8 ** rng.random())) for _ in range(6)]
    reqs = []
    for v, n in enumerate(sizes):
        dt = "float64" if (v == 1 or rng.random() < 0.2) else "float32"
        reqs.append((pkg.FieldSpec("value_noise", (n, n, n), dt, seed=4242 + v), meths[v % 3]))
    print("serving sizes", [(s.shape[0], s.dtype, type(m).__name__) for s, m in reqs], flush=True)
    with pkg.Server(hbm_budget=64 << 30, max_batch=8) as srv:
        futs = [srv.submit(s, m) for s, m in reqs]
        results = [f.result(timeout=120) for f in futs]
        stats = dict(srv.stats)
    assert stats["requests"] == len(reqs)
    print("batches", stats["batches"], "largest", stats["max_batch_seen"], flush=True)
    for (spec, m), r in zip(reqs, results):
        same(r, alone(pkg, spec, m))


def test_c_order_tensors_in_one_batch(pkg):
    # C-order tensors (not x-fastest): the server makes contiguous copies,
    # which must stay alive until the batch has run (two same-shape
    # requests, so a freed copy would be reused by the second)
    import torch
    g = torch.Generator(device="cuda").manual_seed(3)
    a = torch.randn((40, 33, 27), device="cuda", generator=g, dtype=torch.float32)
    b = torch.randn((40, 33, 27), device="cuda", generator=g, dtype=torch.float32)
    m = pkg.MarchingCubes(iso=0)
    with pkg.Server(max_batch=8) as srv:
        srv.pause()
        fa, fb = srv.submit(a, m), srv.submit(b, pkg.NaiveSurfaceNets(iso=0))
        fc = srv.submit(a.clone(), pkg.MarchingTetrahedra(iso=0))
        srv.resume()
        ra, rb, rc = fa.result(timeout=60), fb.result(timeout=60), fc.result(timeout=60)
        assert srv.stats["max_batch_seen"] == 3
    same(ra, pkg.isosurface(a, m))
    same(rb, pkg.isosurface(b, pkg.NaiveSurfaceNets(iso=0)))
    same(rc, pkg.isosurface(a, pkg.MarchingTetrahedra(iso=0)))


def test_field_written_on_client_stream(pkg):
    # the server reads a resident field only after the work the client's
    # stream had queued at submit()
    import torch
    s = torch.cuda.Stream()
    f = torch.empty((64, 64, 64), device="cuda", dtype=torch.float32)
    src = pkg.generate_field("gyroid", (64, 64, 64), "float32", seed=1)
    torch.cuda.synchronize()
    m = pkg.MarchingCubes(iso=0)
    with pkg.Server() as srv:
        with torch.cuda.stream(s):
            torch.cuda._sleep(20_000_000)  # the copy below starts late
            f.copy_(src)
            fut = srv.submit(f, m)
        r = fut.result(timeout=60)
    same(r, pkg.isosurface(src, m))


def test_budget_is_hard_after_counting(pkg):
    # the output guess is far too small, so counted batches exceed their
    # reservation: the server must wait or split, never exceed the budget
    reqs = [(pkg.FieldSpec("value_noise", (n, n, n), "float32", seed=900 + n),
             pkg.MarchingTetrahedra(iso=0)) for n in (48, 64, 56, 72, 40, 64, 52, 60)]
    alone_sizes = []
    for spec, m in reqs:
        V, F = alone(pkg, spec, m)
        alone_sizes.append(spec.nbytes + V.numel() * V.element_size() + F.numel() * 8
                           + 24 * (spec.shape[0] - 1) * (spec.shape[1] - 1))
    budget = int(max(alone_sizes) * 2.5)
    with pkg.Server(hbm_budget=budget, max_batch=8, out_guess=0.01) as srv:
        srv.pause()
        futs = [srv.submit(s, m) for s, m in reqs]
        srv.resume()
        results = [f.result(timeout=120) for f in futs]
        stats = dict(srv.stats)
    assert stats["peak_in_use"] <= budget, (stats["peak_in_use"], budget)
    assert stats["splits"] >= 1
    for (spec, m), r in zip(reqs, results):
        same(r, alone(pkg, spec, m))


def test_oversize_reject_and_cancel(pkg):
    from concurrent.futures import CancelledError
    small = [(pkg.FieldSpec("sphere", (30 + v, 31, 32), "float32", seed=v),
              pkg.MarchingCubes(iso=0)) for v in range(6)]
    big = (pkg.FieldSpec("gyroid", (128, 128, 128), "float64", seed=1),
           pkg.MarchingTetrahedra(iso=0))  # about 47 MB of outputs
    with pkg.Server(hbm_budget=24 << 20, max_batch=4, oversize="reject") as srv:
        srv.pause()
        f0 = srv.submit(*small[0])
        fb = srv.submit(*big)
        rest = [srv.submit(*x) for x in small[1:]]
        assert rest[1].cancel() and rest[3].cancel()
        srv.resume()
        with pytest.raises(MemoryError, match="budget"):
            fb.result(timeout=60)
        r0 = f0.result(timeout=60)
        got = {}
        for i, f in enumerate(rest):
            if i in (1, 3):
                with pytest.raises(CancelledError):
                    f.result(timeout=60)
            else:
                got[i] = f.result(timeout=60)
        stats = dict(srv.stats)
    assert stats["rejected"] == 1 and stats["cancelled"] == 2
    assert stats["peak_in_use"] <= 24 << 20
    same(r0, alone(pkg, *small[0]))
    for i, r in got.items():
        same(r, alone(pkg, *small[i + 1]))


def test_two_servers_at_once_and_batch_views(pkg):
    # two servers open together share the device's two server streams; each
    # one's results still equal the extraction alone, also a result kept
    # after the others of its batch were dropped and their memory reused
    import torch
    reqs = requests(pkg, 24, seed=31)
    with pkg.Server(max_batch=8) as a, pkg.Server(max_batch=8) as b:
        futs = [(a if v % 2 else b).submit(spec, m) for v, (spec, m) in enumerate(reqs)]
        results = [f.result(timeout=60) for f in futs]
    keep = results[5]
    del results[:5]
    torch.cuda.empty_cache()
    junk = torch.full((1 << 22,), -1.0, device="cuda")  # reuses freed memory
    same(keep, alone(pkg, *reqs[5]))
    for (spec, m), r in zip(reqs[5:], results):
        same(r, alone(pkg, spec, m))
    del junk
