"""Serving form: a request loop over the batched C ABI with continuous
batching and HBM-budget admission.

A request is a field (a CUDA tensor indexed [x, y, z], or a spec that the
server materialises on the device when it admits the request) plus an
algorithm object and axes.  ``Server.submit`` returns a
``concurrent.futures.Future`` whose result is ``(vertices, faces)`` on the
device, equal bit for bit to ``isosurface(field, method, X, Y, Z)`` called
alone.  The results were allocated on a server stream and are complete when
the future resolves; a client that hands them to asynchronous work on
another stream and then drops them should call ``record_stream`` first, as
for any tensor shared across streams.

The scheduler thread forms a batch from every queued request that fits the
HBM budget (fields + workspace + an output estimate, corrected to the exact
output size once the batch is counted), runs ``msh_batch_count`` and
``msh_batch_extract`` (``msh_count`` and ``msh_extract`` for a batch of one)
on one of two HIP streams, and moves on: the next
batch's classify/count runs on the other stream while this batch's emit is
still on the GPU.  A completion thread waits for each batch's event,
resolves its futures and returns its memory to the budget, which lets the
scheduler admit the requests that arrived meanwhile.  Streams keep the
default priority.

There is no counterpart in the reference (Meshing.jl is a library call);
the extraction itself is the reference's ``isosurface`` (src/isosurface.jl:30-32).
"""
import collections
import ctypes
import threading
import time
from concurrent.futures import Future
from dataclasses import dataclass, field as dc_field

from . import _capi
from ._capi import MSH_DEVICE, MSH_FLOAT32, MshField, MshParams

__all__ = ["FieldSpec", "Server"]


@dataclass(frozen=True)
class FieldSpec:
    """A synthetic field the server generates on admission (see
    ``generate_field``): kind in sphere/torus/gyroid/value_noise/blobs."""

    kind: str
    shape: tuple
    dtype: str = "float32"
    seed: int = 0
    p0: float = None
    p1: float = None

    @property
    def nbytes(self):
        nx, ny, nz = self.shape
        return nx * ny * nz * (4 if self.dtype == "float32" else 8)


@dataclass
class _Req:
    field: object
    method: object
    axes: tuple
    future: Future
    t_submit: float
    nbytes: int
    units: int
    reserved: int = 0
    t_done: float = 0.0
    result: object = None
    extra: dict = dc_field(default_factory=dict)


class Server:
    """Continuous-batching isosurface server on one GPU.

    hbm_budget: bytes the server may hold for admitted requests (fields it
        materialises, workspace, outputs until they are handed over).
    max_batch: most requests in one batch.
    out_guess: output bytes assumed per field byte before a batch is counted.
    """

    def __init__(self, device=0, hbm_budget=16 << 30, max_batch=256, out_guess=1.0):
        import torch
        from . import generate_field, _field, _params  # noqa: F401
        self.torch = torch
        self.device = torch.device("cuda", device)
        self.budget = int(hbm_budget)
        self.max_batch = max_batch
        self.out_guess = out_guess
        self.L = _capi.lib()
        self.slots = []
        for _ in range(2):
            ctx = _capi.Context(device)
            stream = torch.cuda.Stream(device=self.device)
            ctx.check(self.L.msh_set_stream(ctx.handle, ctypes.c_void_p(stream.cuda_stream)))
            self.slots.append((ctx, stream))
        self.slot = 0
        self.lock = threading.Condition()
        self.pending = collections.deque()
        self.in_use = 0
        self.running = True
        self.done_q = collections.deque()
        self.done_cv = threading.Condition()
        # host seconds spent per phase (count + size read-back, output
        # allocation, emit launch, waiting for completion)
        self.stats = dict(batches=0, requests=0, max_batch_seen=0, t_count=0.0, t_alloc=0.0,
                          t_extract=0.0, t_wait=0.0)
        self.sched = threading.Thread(target=self._schedule, name="msh-sched", daemon=True)
        self.compl = threading.Thread(target=self._complete, name="msh-complete", daemon=True)
        self.sched.start()
        self.compl.start()

    # -- client side ----------------------------------------------------
    def submit(self, field, method, axes=None, on_done=None):
        """Queue one extraction; returns a Future of (vertices, faces).

        With ``on_done``, the completion thread calls ``on_done(result,
        latency_s)`` instead and the future resolves to None, so the server
        keeps no reference to the output once the callback returns (a
        client that streams results out does not pile them up)."""
        if isinstance(field, FieldSpec):
            nbytes = field.nbytes
            nx, ny, nz = field.shape
        else:
            if not (isinstance(field, self.torch.Tensor) and field.is_cuda):
                raise TypeError("served fields are CUDA tensors or FieldSpec")
            nbytes = 0  # resident already; not the server's memory
            nx, ny, nz = field.shape
        units = max(nx - 1, 0) * max(ny - 1, 0)
        fut = Future()
        r = _Req(field, method, axes or ((-1, 1),) * 3, fut, time.perf_counter(), nbytes, units)
        r.extra["on_done"] = on_done
        with self.lock:
            if not self.running:
                raise RuntimeError("server is closed")
            self.pending.append(r)
            self.lock.notify_all()
        return fut

    def close(self):
        with self.lock:
            self.running = False
            self.lock.notify_all()
        self.sched.join()
        with self.done_cv:
            self.done_q.append(None)
            self.done_cv.notify_all()
        self.compl.join()
        for ctx, stream in self.slots:
            stream.synchronize()
            ctx.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- scheduler ------------------------------------------------------
    def _estimate(self, r):
        return r.nbytes + int(r.nbytes * self.out_guess) + 24 * r.units + (1 << 16)

    def _admit(self):
        """Requests for the next batch, in arrival order, within the budget.
        Called with the lock held; waits while nothing fits.  After close()
        the queue is drained before it returns an empty batch."""
        while True:
            if not self.running and not self.pending:
                return []
            batch, want = [], 0
            while self.pending and len(batch) < self.max_batch:
                est = self._estimate(self.pending[0])
                # a request larger than the whole budget runs alone
                if self.in_use + want + est > self.budget and (batch or self.in_use):
                    break
                r = self.pending.popleft()
                r.reserved = est
                want += est
                batch.append(r)
            if batch:
                self.in_use += want
                return batch
            self.lock.wait()

    def _schedule(self):
        torch = self.torch
        from . import generate_field, _field, _params
        while True:
            with self.lock:
                batch = self._admit()
            if not batch:
                return
            ctx, stream = self.slots[self.slot]
            self.slot ^= 1
            try:
                with torch.cuda.stream(stream):
                    fields = []
                    for r in batch:
                        if isinstance(r.field, FieldSpec):
                            s = r.field
                            f = generate_field(s.kind, s.shape, s.dtype, seed=s.seed, p0=s.p0,
                                               p1=s.p1, device=self.device.index)
                        else:
                            f = r.field
                        r.extra["field"] = f
                        fields.append(f)
                    n = len(batch)
                    F = (MshField * n)(*[_field(f)[0] for f in fields])
                    P = (MshParams * n)(*[_params(r.method, *r.axes) for r in batch])
                    nv = (ctypes.c_int64 * n)()
                    nf = (ctypes.c_int64 * n)()
                    t0 = time.perf_counter()
                    # a lone request takes the single-volume entry points
                    # (same result, fewer launches and uploads)
                    if n == 1:
                        ctx.check(self.L.msh_count(ctx.handle, ctypes.byref(F[0]),
                                                   ctypes.byref(P[0]), nv, nf))
                    else:
                        ctx.check(self.L.msh_batch_count(ctx.handle, n, F, P, nv, nf))
                    t1 = time.perf_counter()
                    vp = (ctypes.c_void_p * n)()
                    fp = (ctypes.c_void_p * n)()
                    extra = 0
                    for v, r in enumerate(batch):
                        ok = ctypes.c_int32()
                        ctx.check(self.L.msh_output_kind(ctypes.byref(F[v]), ctypes.byref(P[v]),
                                                         ctypes.byref(ok)))
                        dt = torch.float32 if ok.value == MSH_FLOAT32 else torch.float64
                        V = torch.empty((nv[v], 3), dtype=dt, device=self.device)
                        Fc = torch.empty((nf[v], 3), dtype=torch.int64, device=self.device)
                        vp[v] = V.data_ptr() if nv[v] else None
                        fp[v] = Fc.data_ptr() if nf[v] else None
                        r.result = (V, Fc)
                        exact = r.nbytes + V.numel() * V.element_size() + Fc.numel() * 8 + \
                            24 * r.units
                        extra += exact - r.reserved
                        r.reserved = exact
                    with self.lock:
                        self.in_use += extra
                    t2 = time.perf_counter()
                    if n == 1:
                        ctx.check(self.L.msh_extract(ctx.handle, ctypes.byref(F[0]),
                                                     ctypes.byref(P[0]), vp[0], fp[0], MSH_DEVICE))
                    else:
                        ctx.check(self.L.msh_batch_extract(ctx.handle, n, F, P, vp, fp,
                                                           MSH_DEVICE))
                    ev = torch.cuda.Event()
                    ev.record(stream)
                    t3 = time.perf_counter()
                    self.stats["t_count"] += t1 - t0
                    self.stats["t_alloc"] += t2 - t1
                    self.stats["t_extract"] += t3 - t2
            except BaseException as e:  # hand the error to every request of the batch
                for r in batch:
                    r.future.set_exception(e)
                with self.lock:
                    self.in_use -= sum(r.reserved for r in batch)
                    self.lock.notify_all()
                continue
            self.stats["batches"] += 1
            self.stats["requests"] += len(batch)
            self.stats["max_batch_seen"] = max(self.stats["max_batch_seen"], len(batch))
            with self.done_cv:
                self.done_q.append((ev, batch))
                self.done_cv.notify_all()

    # -- completion -----------------------------------------------------
    def _complete(self):
        while True:
            with self.done_cv:
                while not self.done_q:
                    self.done_cv.wait()
                item = self.done_q.popleft()
            if item is None:
                return
            ev, batch = item
            tw = time.perf_counter()
            ev.synchronize()
            t = time.perf_counter()
            self.stats["t_wait"] += t - tw
            freed = 0
            for r in batch:
                r.t_done = t
                r.extra.pop("field", None)
                freed += r.reserved
                r.future.latency = t - r.t_submit
                cb = r.extra.pop("on_done", None)
                if cb is None:
                    r.future.set_result(r.result)
                else:
                    try:
                        cb(r.result, r.future.latency)
                        r.future.set_result(None)
                    except BaseException as e:
                        r.future.set_exception(e)
                r.result = None
            batch = None
            with self.lock:
                self.in_use -= freed
                self.lock.notify_all()
