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
