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
for any tensor shared across streams.  The vertices and faces of a result
are views of one allocation (half the allocator calls on the scheduler's
critical path), freed when both are dropped.

The scheduler thread forms a batch from every queued request that fits the
HBM budget (fields + workspace + an output estimate), runs
``msh_batch_count`` (``msh_count`` for a batch of one) on one of two HIP
streams, and then holds the batch to its exact size: when the exact outputs
do not fit beside the memory other batches hold, it waits for them to
finish, and when they still do not fit it keeps the longest prefix of the
batch that does, re-counts it and puts the rest back at the head of the
queue.  So the memory the server accounts for never exceeds the budget,
except for a single request that is larger than the whole budget, which
runs alone (or fails with ``MemoryError`` under ``oversize="reject"``).
It then allocates the outputs, runs ``msh_batch_extract`` (``msh_extract``)
and moves on: the next batch's classify/count runs on the other stream
while this batch's emit is still on the GPU.  A completion thread waits for
each batch's event, resolves its futures and returns its memory to the
budget, which lets the scheduler admit the requests that arrived meanwhile.
The two streams are the device's, shared by every Server the process
opens on it (servers open at the same time queue behind each other's
work), and keep the default priority.

Input contract: a resident field is read on a server stream after the
work that the client's current stream has queued when ``submit`` is called
(the server stream waits for an event recorded there), so the field must be
complete on that stream; the server keeps the tensor, and any contiguous
copy it makes of it, alive until the request completes.  A request whose
future is cancelled before its batch is formed is dropped.

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


_streams = {}
_streams_lock = threading.Lock()


def _server_streams(torch, device):
    """The two streams every Server on ``device`` runs on.  torch caches
    freed memory per stream and serves an allocation only from its own
    stream's cache, so servers that a process opens one after another
    reuse what the earlier ones cached rather than each growing a cache of
    its own until the device runs out (torch then frees every cache at once
    and maps the memory anew: seconds, at the sizes served here)."""
    with _streams_lock:
        if device.index not in _streams:
            _streams[device.index] = [torch.cuda.Stream(device=device) for _ in range(2)]
        return _streams[device.index]


def _align(nbytes):
    return (nbytes + 255) & ~255


def _result_views(torch, buf, fofs, nv, nf, kind):
    """(vertices, faces) of one request: views of its output allocation,
    the faces from byte ``fofs`` on."""
    dt, osz = (torch.float32, 4) if kind == MSH_FLOAT32 else (torch.float64, 8)
    V = buf[:nv * 3 * osz].view(dt).view(nv, 3)
    F = buf[fofs:fofs + nf * 24].view(torch.int64).view(nf, 3)
    return V, F


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
    oversize: what to do with a single request whose exact need exceeds the
        whole budget: "alone" runs it by itself, "reject" fails its future
        with MemoryError.
    """

    def __init__(self, device=0, hbm_budget=16 << 30, max_batch=256, out_guess=1.0,
                 oversize="alone"):
        if oversize not in ("alone", "reject"):
            raise ValueError("oversize must be 'alone' or 'reject'")
        import torch
        from . import generate_field, _field, _params  # noqa: F401
        self.torch = torch
        self.device = torch.device("cuda", device)
        self.budget = int(hbm_budget)
        self.max_batch = max_batch
        self.oversize = oversize
        self.out_guess = out_guess
        self.L = _capi.lib()
        self.slots = []
        for stream in _server_streams(torch, self.device):
            ctx = _capi.Context(device)
            ctx.check(self.L.msh_set_stream(ctx.handle, ctypes.c_void_p(stream.cuda_stream)))
            self.slots.append((ctx, stream))
        self.slot = 0
        self.lock = threading.Condition()
        self.pending = collections.deque()
        self.in_use = 0
        self.running = True
        self.paused = False
        self.done_q = collections.deque()
        self.done_cv = threading.Condition()
        # host seconds spent per phase (count + size read-back, output
        # allocation, emit launch, waiting for completion)
        self.stats = dict(batches=0, requests=0, max_batch_seen=0, t_count=0.0, t_alloc=0.0,
                          t_extract=0.0, t_wait=0.0, peak_in_use=0, splits=0, cancelled=0,
                          rejected=0)
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
        ready = None
        if isinstance(field, FieldSpec):
            nbytes = field.nbytes
            nx, ny, nz = field.shape
        else:
            if not (isinstance(field, self.torch.Tensor) and field.is_cuda):
                raise TypeError("served fields are CUDA tensors or FieldSpec")
            if field.dim() != 3:
                raise ValueError("sdf must be 3-dimensional")
            nbytes = 0  # resident already; not the server's memory
            nx, ny, nz = field.shape
            # the server stream reads the field after what the client's
            # stream has queued so far
            ready = self.torch.cuda.Event()
            ready.record(self.torch.cuda.current_stream(field.device))
        units = max(nx - 1, 0) * max(ny - 1, 0)
        fut = Future()
        r = _Req(field, method, axes or ((-1, 1),) * 3, fut, time.perf_counter(), nbytes, units)
        r.extra["on_done"] = on_done
        r.extra["ready"] = ready
        with self.lock:
            if not self.running:
                raise RuntimeError("server is closed")
            self.pending.append(r)
            self.lock.notify_all()
        return fut

    def pause(self):
        """Stop admitting requests (those already on the GPU finish);
        submissions queue up until resume()."""
        with self.lock:
            self.paused = True

    def resume(self):
        with self.lock:
            self.paused = False
            self.lock.notify_all()

    def close(self):
        """Drain the queue (a paused server resumes), finish every request
        and release the server's contexts."""
        with self.lock:
            self.paused = False
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
        self.slots = []

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- scheduler ------------------------------------------------------
    def _base(self, r):
        """Bytes needed to count a request: its field (when the server
        materialises it) and the count workspace."""
        return r.nbytes + 24 * r.units + (1 << 16)

    def _estimate(self, r):
        return self._base(r) + int(r.nbytes * self.out_guess)

    def _take(self, amount):
        """Account `amount` more bytes (lock held)."""
        self.in_use += amount
        if self.in_use > self.stats["peak_in_use"]:
            self.stats["peak_in_use"] = self.in_use

    def _admit(self):
        """Requests for the next batch, in arrival order, within the budget.
        Called with the lock held; waits while nothing fits.  Cancelled
        requests are dropped here.  After close() the queue is drained
        before it returns an empty batch."""
        while True:
            if not self.running and not self.pending:
                return []
            if self.paused:
                self.lock.wait()
                continue
            batch, want = [], 0
            while self.pending and len(batch) < self.max_batch:
                r = self.pending[0]
                started = r.extra.get("running", False)  # put back after a split
                if not started and r.future.cancelled():
                    self.pending.popleft()
                    self.stats["cancelled"] += 1
                    continue
                est = self._estimate(r)
                if self.in_use + want + est > self.budget:
                    if batch or self.in_use:
                        break
                    # alone and still too large by the estimate: reserve what
                    # counting needs; the count gives the exact size.  A
                    # request that cannot even be counted within the budget
                    # runs alone or is rejected.
                    est = self._base(r)
                    if est > self.budget and self.oversize == "reject":
                        self.pending.popleft()
                        if started or r.future.set_running_or_notify_cancel():
                            self.stats["rejected"] += 1
                            r.future.set_exception(MemoryError(
                                f"request needs at least {est} bytes, more than the server's "
                                f"HBM budget of {self.budget}"))
                        else:
                            self.stats["cancelled"] += 1
                        continue
                self.pending.popleft()
                if not started:
                    if not r.future.set_running_or_notify_cancel():
                        self.stats["cancelled"] += 1
                        continue
                    r.extra["running"] = True
                r.reserved = est
                want += est
                batch.append(r)
            if batch:
                self._take(want)
                return batch
            self.lock.wait()

    def _fit(self, batch, exact):
        """Hold the counted batch to the budget (lock held).  Waits while
        other batches hold the memory it needs; then returns how many
        requests of the batch (a prefix, at least one) fit, and for a lone
        request that can never fit, 0 under oversize="reject"."""
        mine = sum(r.reserved for r in batch)
        need = sum(exact)
        while self.in_use - mine + need > self.budget and self.in_use > mine:
            self.lock.wait()
        others = self.in_use - mine
        if others + need <= self.budget:
            return len(batch)
        k, acc = 1, exact[0]
        while k < len(batch) and others + acc + exact[k] <= self.budget:
            acc += exact[k]
            k += 1
        if k == 1 and others + exact[0] > self.budget and self.oversize == "reject":
            return 0
        return k

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
                    for r in batch:
                        if isinstance(r.field, FieldSpec):
                            s = r.field
                            r.extra["field"] = generate_field(s.kind, s.shape, s.dtype,
                                                              seed=s.seed, p0=s.p0, p1=s.p1,
                                                              device=self.device.index)
                        else:
                            ready = r.extra.get("ready")
                            if ready is not None:
                                stream.wait_event(ready)
                            r.extra["field"] = r.field
                        # the descriptor's tensor (a contiguous copy when the
                        # field is not x-fastest) stays alive with the request
                        desc, view, _ = _field(r.extra["field"])
                        r.extra["desc"] = desc
                        r.extra["view"] = view
                    t0 = time.perf_counter()
                    while True:
                        n = len(batch)
                        F = (MshField * n)(*[r.extra["desc"] for r in batch])
                        P = (MshParams * n)(*[_params(r.method, *r.axes) for r in batch])
                        nv = (ctypes.c_int64 * n)()
                        nf = (ctypes.c_int64 * n)()
                        # a lone request takes the single-volume entry points
                        # (same result, fewer launches and uploads)
                        if n == 1:
                            ctx.check(self.L.msh_count(ctx.handle, ctypes.byref(F[0]),
                                                       ctypes.byref(P[0]), nv, nf))
                        else:
                            ctx.check(self.L.msh_batch_count(ctx.handle, n, F, P, nv, nf))
                        kinds, exact = [], []
                        for v, r in enumerate(batch):
                            ok = ctypes.c_int32()
                            ctx.check(self.L.msh_output_kind(ctypes.byref(F[v]),
                                                             ctypes.byref(P[v]), ctypes.byref(ok)))
                            kinds.append(ok.value)
                            osz = 4 if ok.value == MSH_FLOAT32 else 8
                            exact.append(r.nbytes + nv[v] * 3 * osz + nf[v] * 24 + 24 * r.units)
                        with self.lock:
                            k = self._fit(batch, exact)
                            if k == n:
                                self._take(sum(exact) - sum(r.reserved for r in batch))
                                for r, e in zip(batch, exact):
                                    r.reserved = e
                                break
                            if k == 0:  # the first request alone exceeds the budget
                                r = batch[0]
                                self.in_use -= r.reserved
                                r.reserved = 0
                                for key in ("field", "desc", "view", "ready"):
                                    r.extra.pop(key, None)
                                self.stats["rejected"] += 1
                                r.future.set_exception(MemoryError(
                                    f"request needs {exact[0]} bytes, more than the server's "
                                    f"HBM budget of {self.budget}"))
                            # keep the prefix that fits; the rest goes back to
                            # the head of the queue and is re-admitted later
                            rest = batch[max(k, 1):]
                            batch = batch[:k]
                            for r in reversed(rest):
                                self.in_use -= r.reserved
                                r.reserved = 0
                                for key in ("field", "desc", "view"):
                                    r.extra.pop(key, None)
                                self.pending.appendleft(r)
                            if rest:
                                self.stats["splits"] += 1
                            self.lock.notify_all()
                            if not batch:
                                break
                    if not batch:
                        continue
                    t1 = time.perf_counter()
                    vp = (ctypes.c_void_p * n)()
                    fp = (ctypes.c_void_p * n)()
                    # one allocation per request holds its vertices and
                    # faces; the completion thread makes the two views
                    for v, r in enumerate(batch):
                        osz = 4 if kinds[v] == MSH_FLOAT32 else 8
                        vb = _align(nv[v] * 3 * osz)
                        buf = torch.empty(max(vb + nf[v] * 24, 1), dtype=torch.uint8,
                                          device=self.device)
                        vp[v] = buf.data_ptr() if nv[v] else None
                        fp[v] = buf.data_ptr() + vb if nf[v] else None
                        r.extra["out"] = (buf, vb, nv[v], nf[v], kinds[v])
                    del buf
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
                    r.result = None
                    for key in ("field", "desc", "view", "out"):
                        r.extra.pop(key, None)
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
                for key in ("field", "desc", "view", "ready"):
                    r.extra.pop(key, None)
                freed += r.reserved
                r.future.latency = t - r.t_submit
                out = r.extra.pop("out", None)
                if out is not None:
                    r.result = _result_views(self.torch, *out)
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
