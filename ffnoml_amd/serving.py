"""Continuous batching of predict, VJP and JVP requests on one GPU.

The reference serves one request at a time over a single connection (ffno_fsdp_service.py:189-323). Here requests
from any number of clients go into one queue, and four threads move them through a pipeline:

  scheduler  groups waiting requests that share a bucket (model, operation, tile shape, spacings and column heights,
             so one plan serves them all; the oldest waiting request picks the bucket) and admits a batch while
             fewer than `slots` batches are in the pipeline and its HBM need fits the budget left unclaimed;
  planner    builds the batch's plan from the heights in host memory, with every table the call will read (a new
             tile shape costs up to a few hundred ms of host work), while the device runs the batches before it;
  executor   copies a batch's inputs in (its own copy stream, through pinned buffers) and enqueues the batch's whole
             launch sequence on the one compute stream, then goes on to the next batch without waiting: the device
             runs batches back to back while the host enqueues the next, and nothing in the sequence waits for the
             device;
  finisher   waits for a batch's end, copies its results out (a second copy stream, pinned buffers), checks them and
             answers the requests.

One compute stream, because the caching allocator keeps the blocks freed on a stream for that stream: a batch reuses
the memory of the batch before it in stream order, so the batches in the pipeline need the largest one's working set
plus every batch's inputs and outputs, and that is what admission counts. (Batches on several streams each keep their
own cache; their sum outgrows HBM, and the allocator's retry then frees every cache after synchronising the whole
device, stalling every thread behind the longest batch queued.) GroupNorm statistics are per sample, so a request
batched with others gets what it would get alone.

Inputs and outputs cross PCIe through pinned double buffers (ffnoml_amd.staging): a pageable copy runs at about
1 GB/s, a pinned one at 12-14 GB/s.

A VJP batch keeps as many blocks whole (every intermediate its adjoint reads, so the block is not recomputed) as
fit within `full_share` of the budget and within what is unclaimed now: all of them (about a third less device time
than the checkpointed adjoint), some, or none (checkpoint stride 1: the block inputs kept, each block recomputed
once). A VJP batch is cut to the largest size that keeps as many blocks whole as its first request alone would: a
recomputed block costs a block's forward for every voxel of the batch, more than batching saves once tiles fill
the GPU.

Forward-only batches of small tiles (forward working set at most `graph_bytes`) replay a captured HIP graph of the
whole forward: a batch of n requests runs on the graph of its bucket for the next power of two at or above n (the
other samples of the static input hold earlier requests' inputs; GroupNorm statistics are per sample, so they change
nothing), which the executor captures after running the first such batch eagerly and before that batch's results are
handed out. The capture holds the device alone (HIP refuses another thread's copies while a stream is capturing): the
planner's and the finisher's device work waits for it. A caller in the same process that launches GPU work from a
thread of its own while a capture runs (a request's results arrive only after its batch's capture, so a client waiting
on its own requests never does) can meet torch's capture-state errors, e.g. from its RNG. The memory a graph holds is
counted against the budget; graphs are evicted least recently replayed first, and an evicted graph's memory is
released from the count once its last replay is enqueued.

Admission never stalls: when the oldest request does not fit the unclaimed budget while nothing is in the pipeline,
the graphs and the engines' unused tables are dropped; a request that still does not fit is failed with MemoryError.
"""
import collections
import contextlib
import hashlib
import queue
import threading
import time
from concurrent.futures import Future

import numpy as np
import torch

from .staging import PinnedStage

# float32 values per voxel of a batch's inputs and outputs while it is in the pipeline
IO_FLOATS = {"PREDICT": 6 + 1 + 12, "VJP": 6 + 1 + 12 + 6 + 1, "JVP": 6 + 1 + 6 + 12}


class Request:
    __slots__ = ("op", "model", "features", "z", "dx", "dy", "cotangent", "future", "t_submit", "t_start", "t_done",
                 "key", "voxels")

    def __init__(self, op, model, features, z, dx, dy, cotangent):
        self.op, self.model = op, model
        self.features, self.z, self.dx, self.dy, self.cotangent = features, z, float(dx), float(dy), cotangent
        self.future = Future()
        self.t_submit = time.perf_counter()
        self.t_start = self.t_done = None
        nz, nx, ny = z.shape
        self.voxels = nz * nx * ny
        zh = hashlib.sha1(np.ascontiguousarray(z - z.min(axis=0, keepdims=True)).tobytes()).hexdigest()
        self.key = (model, op, z.shape, self.dx, self.dy, zh)


class DeviceGate:
    """Shared holds for copies, an exclusive hold for a graph capture: a waiting capture blocks new shared holds, so
    the copies running finish and no other thread touches the device while a stream is capturing."""

    def __init__(self):
        self.cv = threading.Condition()
        self.readers = 0
        self.writer = False
        self.waiting = 0

    @contextlib.contextmanager
    def shared(self):
        with self.cv:
            while self.writer or self.waiting:
                self.cv.wait()
            self.readers += 1
        try:
            yield
        finally:
            with self.cv:
                self.readers -= 1
                self.cv.notify_all()

    @contextlib.contextmanager
    def exclusive(self):
        with self.cv:
            self.waiting += 1
            while self.writer or self.readers:
                self.cv.wait()
            self.waiting -= 1
            self.writer = True
        try:
            yield
        finally:
            with self.cv:
                self.writer = False
                self.cv.notify_all()


class _Batch:
    """An admitted batch: its requests, the working set and input/output bytes claimed for it, the VJP's stride, and
    what the executor hands the finisher."""

    __slots__ = ("reqs", "work", "io", "stride", "keep", "plan", "outs", "bad", "done", "capture", "t0", "t_staged",
                 "t_enqueued")

    def __init__(self, reqs, work, io, stride, keep=0):
        self.reqs, self.work, self.io, self.stride, self.keep = reqs, work, io, stride, keep
        self.plan = self.outs = self.bad = self.done = self.capture = None
        self.t0 = self.t_staged = self.t_enqueued = None


class ContinuousBatcher:
    """submit() -> Future of the backend's result: {"populations"} for PREDICT, {"features", "z_scale"} for VJP,
    {"population_tangent"} for JVP."""

    def __init__(self, backends, slots=3, max_batch=8, memory_fraction=0.8, graphs=True, max_graphs=24,
                 graph_bytes=4 << 30, budget=None, full_share=0.8, segment_share=0.85):
        self.backends = dict(backends)
        self.slots = int(slots)
        self.max_batch = int(max_batch)
        self.full_share = float(full_share)
        self.graphs_enabled = bool(graphs)
        self.max_graphs = int(max_graphs)
        self.graph_bytes = int(graph_bytes)
        self.graph_held = 0                             # bytes held by captured graphs' memory pools
        self.gate = DeviceGate()
        self.eager_runs = collections.Counter()         # forward batches run eagerly, per (bucket, batch size)
        dev = next(iter(self.backends.values())).device
        self.device = dev
        self.cuda = torch.device(dev).type == "cuda"
        if budget is None:
            # blocks cached for other streams are no use to the compute stream, and a failed allocation would free
            # them only after synchronising the device: release them now, then budget what is free
            torch.cuda.synchronize(dev)
            torch.cuda.empty_cache()
            free, _ = torch.cuda.mem_get_info(dev)
            budget = int(free * memory_fraction)
        self.budget = int(budget)
        self.claims = {}                                # id(batch) -> (working set, io bytes) of batches in the pipeline
        self.io_claimed = 0
        self.inflight = 0
        self.pending = collections.deque()
        self.cv = threading.Condition()
        self.plans = queue.Queue()
        self.work = queue.Queue()
        self.finish = queue.Queue()
        self.stop = False
        self.graphs = collections.OrderedDict()
        self.no_graph = set()
        self.graph_lock = threading.Lock()
        self.completed = []
        self.peak_inflight = 0
        self.queue_samples = []                          # (time, requests waiting, batches in the pipeline) per admission
        self.batch_log = []                              # (start, op, batch size, voxels, staging, compute, output s)
        self.captures = 0
        self.replays = 0
        self.segment = 0
        if self.cuda:
            self.compute = torch.cuda.Stream(device=dev)
            self.h2d = torch.cuda.Stream(device=dev)
            self.d2h = torch.cuda.Stream(device=dev)
            self._reserve(segment_share)
        self.scheduler = threading.Thread(target=self._schedule, daemon=True, name="ffno-scheduler")
        self.planner, self.executor, self.finisher = (
            threading.Thread(target=self._guard, args=(f,), daemon=True, name=f"ffno{f.__name__.replace('_', '-')}")
            for f in (self._planner, self._executor, self._finisher))
        for t in (self.finisher, self.executor, self.planner, self.scheduler):
            t.start()

    def _reserve(self, share):
        """One cached segment of `share` of the budget for the compute stream: the batches' working sets are then
        carved out of it and coalesce back when they are freed, instead of the caching allocator asking for fresh
        blocks of each new batch's sizes (a hipMalloc of many GB takes 100+ ms per GB, and one that does not fit makes
        the allocator free every cache after synchronising the device)."""
        self.segment = int(self.budget * share) // (2 << 20) * (2 << 20)
        if self.segment > 0:
            with torch.cuda.stream(self.compute):
                torch.empty(self.segment, dtype=torch.uint8, device=self.device)
            self.compute.synchronize()

    # ------------------------------------------------------------------------------------------------ public
    def submit(self, op, model, features, z, dx, dy, cotangent=None):
        if op not in ("PREDICT", "VJP", "JVP"):
            raise ValueError(f"unsupported operation {op}")
        backend = self.backends[model]
        features = np.asarray(features, np.float32)
        z = np.asarray(z, np.float32)
        if features.ndim != 4 or features.shape[0] != 6 or tuple(features.shape[1:]) != tuple(z.shape):
            raise ValueError("features must be (6,nz,nx,ny) and z (nz,nx,ny)")
        if not np.isfinite(features).all() or not np.isfinite(z).all():
            raise ValueError("FFNO request contains NaN or Inf")
        if not (features[0] > 0).all():
            raise ValueError("FFNO request temperature must be positive")
        if op == "VJP":
            cotangent = np.asarray(cotangent, np.float32)
            if tuple(cotangent.shape) != tuple(z.shape) + (len(backend.level_names),):
                raise ValueError("population_cotangent has the wrong shape")
            if not np.isfinite(cotangent).all():
                raise ValueError("FFNO VJP cotangent contains NaN or Inf")
        if op == "JVP":
            cotangent = np.asarray(cotangent, np.float32)          # the feature tangent rides in the same slot
            if tuple(cotangent.shape) != tuple(features.shape):
                raise ValueError("feature_tangent must have shape (6,nz,nx,ny)")
            if not np.isfinite(cotangent).all():
                raise ValueError("FFNO JVP tangent contains NaN or Inf")
        req = Request(op, model, features, z, dx, dy, cotangent)
        need = self._footprint([req]) + self._io([req])
        if need > self.budget:
            req.future.set_exception(MemoryError(f"request needs {need / 2**30:.1f} GiB, more than the "
                                                 f"{self.budget / 2**30:.1f} GiB budget"))
            return req.future
        with self.cv:
            self.pending.append(req)
            self.cv.notify_all()
        return req.future

    def close(self):
        with self.cv:
            self.stop = True
            self.cv.notify_all()
        self.scheduler.join()                           # the scheduler admits what waits, then ends
        self.plans.put(None)                            # each later thread ends after the batches before it
        for t in (self.planner, self.executor, self.finisher):
            t.join()
        # release the device memory the batcher holds: its captured graphs (their pools and static buffers), then the
        # reserved segment and every other cached block
        with self.graph_lock:
            self.graphs.clear()
            self.graph_held = 0
        if self.cuda:
            torch.cuda.synchronize(self.device)
            torch.cuda.empty_cache()

    def latencies_ms(self):
        return np.array([(r.t_done - r.t_submit) * 1e3 for r in self.completed])

    # ------------------------------------------------------------------------------------------------ scheduler
    def _footprint(self, reqs, keep=0):
        """The working set of a batch of these requests on the device (a VJP keeping `keep` blocks whole)."""
        eng = self.backends[reqs[0].model].model.engine()
        nz, nx, ny = reqs[0].z.shape
        span = float(np.max(reqs[0].z.max(axis=0) - reqs[0].z.min(axis=0))) / 1e6       # metres -> Mm
        return eng.footprint_bytes(len(reqs), nz, nx, ny, vjp=reqs[0].op == "VJP", jvp=reqs[0].op == "JVP",
                                   span=span, keep=keep)

    @staticmethod
    def _io(reqs):
        """Device bytes of a batch's inputs and outputs, alive from its copy in to its copy out."""
        return 4 * IO_FLOATS[reqs[0].op] * sum(r.voxels for r in reqs)

    def _free(self):
        """The budget not held by captured graphs or by the engines' cached plans and tables."""
        return self.budget - self.graph_held - sum(b.model.engine().cached_bytes() for b in self.backends.values())

    def _claimed(self, work=0, io=0):
        """HBM the pipeline needs with one more batch of this working set and io (cv held): the batches run one after
        the other on the compute stream, so the largest working set, plus every batch's inputs and outputs."""
        return max([work] + [w for w, _ in self.claims.values()]) + self.io_claimed + io

    def _release_caches(self):
        """With nothing in the pipeline: evict every captured graph and the engines' unused plans and tables."""
        with self.graph_lock:
            while self.graphs:
                self._evict_oldest()
        for b in self.backends.values():
            b.model.engine().drop_unused_tables()

    def _schedule(self):
        while True:
            try:
                if self._schedule_once():
                    return
            except Exception as e:                      # never let the scheduler die silently: fail what waits
                import sys
                import traceback
                print(f"ffno scheduler: {e!r}\n{traceback.format_exc()}", file=sys.stderr, flush=True)
                with self.cv:
                    failed, self.pending = list(self.pending), collections.deque()
                for r in failed:
                    if not r.future.done():
                        r.future.set_exception(e)

    def _schedule_once(self):
        """One scheduling decision; True when the batcher is closed and drained."""
        with self.cv:
            while not self.stop and (not self.pending or self.inflight >= self.slots):
                self.cv.wait()
            if self.stop and not self.pending:
                return True
            if self.inflight >= self.slots:
                self.cv.wait(timeout=0.05)              # closing: the pipeline drains before the rest is admitted
                return False
            head = self.pending[0]
            batch = [r for r in self.pending if r.key == head.key][: self.max_batch]
            free = self._free()
            while len(batch) > 1 and self._claimed(self._footprint(batch), self._io(batch)) > free:
                batch.pop()
            work, io = self._footprint(batch), self._io(batch)
            if self._claimed(work, io) > free:
                if self.inflight == 0:
                    # nothing will release memory by itself: drop the caches, and fail a request that still
                    # does not fit instead of waiting for good
                    self._release_caches()
                    if work + io > self._free():
                        self.pending.remove(head)
                        head.t_done = time.perf_counter()
                        self.completed.append(head)
                        head.future.set_exception(MemoryError(
                            f"request needs {(work + io) / 2**30:.1f} GiB; {max(self._free(), 0) / 2**30:.1f} GiB "
                            f"of the {self.budget / 2**30:.1f} GiB budget are free with nothing in flight"))
                    return False
                self.cv.wait(timeout=0.05)              # wait for memory to be released by a finishing batch
                return False
            stride, keep = None, 0
            if head.op == "VJP":
                layers = self.backends[head.model].model.n_layers

                def kept(reqs):
                    # the most blocks kept whole that fit a share of the budget and what is unclaimed now
                    for k in range(layers, 0, -1):
                        need = self._footprint(reqs, keep=k)
                        if need <= self.full_share * self.budget and self._claimed(need, self._io(reqs)) <= free:
                            return k, need
                    return 0, self._footprint(reqs)

                keep, work = kept(batch)
                if len(batch) > 1 and keep < layers:
                    # every block not kept whole is recomputed, a forward of that block for every voxel of the batch:
                    # a batch keeping fewer blocks than the head alone would costs more per voxel than it saves
                    alone = kept(batch[:1])[0]
                    while len(batch) > 1 and keep < alone:
                        batch.pop()
                        keep, work = kept(batch)
                    io = self._io(batch)
                stride = 0 if keep == layers else 1
            for r in batch:
                self.pending.remove(r)
            b = _Batch(batch, work, io, stride, keep)
            self.claims[id(b)] = (work, io)
            self.io_claimed += io
            self.inflight += 1
            self.peak_inflight = max(self.peak_inflight, self.inflight)
            self.queue_samples.append((time.perf_counter(), len(self.pending), self.inflight))
        self.plans.put(b)
        return False

    def _unclaim(self, b):
        with self.cv:
            if self.claims.pop(id(b), None) is not None:
                self.io_claimed -= b.io
                self.inflight -= 1
            self.cv.notify_all()

    # ------------------------------------------------------------------------------------------------ executor
    def _guard(self, loop):
        """Run a pipeline thread's loop; should it fail outside a batch, fail everything waiting rather than hang."""
        try:
            loop()
        except Exception as e:
            import sys
            import traceback
            print(f"ffno {loop.__name__}: {e!r}\n{traceback.format_exc()}", file=sys.stderr, flush=True)
            with self.cv:
                self.stop = True
                failed, self.pending = list(self.pending), collections.deque()
                self.cv.notify_all()
            for r in failed:
                if not r.future.done():
                    r.future.set_exception(e)

    def _fail(self, b, e):
        done = time.perf_counter()
        with self.cv:
            for r in b.reqs:
                if r.t_done is None:
                    r.t_done = done
                    self.completed.append(r)
        for r in b.reqs:
            if not r.future.done():
                r.future.set_exception(e)
        self._unclaim(b)

    def _planner(self):
        try:
            while True:
                b = self.plans.get()
                if b is None:
                    return
                if self.cuda:
                    r0 = b.reqs[0]
                    backend = self.backends[r0.model]
                    try:
                        # one bucket holds one set of shifted heights: the plan from the first request's
                        with self.gate.shared():
                            zh = torch.from_numpy(r0.z).expand(len(b.reqs), *r0.z.shape)
                            b.plan = backend.plan_for(zh, r0.dx, r0.dy).prepare(adjoint=r0.op == "VJP")
                    except Exception as e:
                        self._fail(b, e)
                        continue
                self.work.put(b)
        finally:
            self.work.put(None)

    def _executor(self):
        stage = PinnedStage() if self.cuda else None
        try:
            while True:
                b = self.work.get()
                if b is None:
                    return
                b.t0 = time.perf_counter()
                for r in b.reqs:
                    r.t_start = b.t0
                try:
                    self._enqueue(b, stage)
                except Exception as e:                  # one failed batch fails its own requests only
                    self._fail(b, e)
                    continue
                capture, b.capture = b.capture, None
                if capture is not None:
                    # before the batch's results go out: a client woken by them may launch GPU work of its own
                    # (torch's RNG refuses to advance while a stream of the process is capturing)
                    self._capture(*capture)
                self.finish.put(b)
        finally:
            self.finish.put(None)

    def _enqueue(self, b, stage):
        """Copy the batch in and enqueue its launch sequence; b.outs, b.bad (a device flag) and b.done (the compute
        stream's event after the batch) for the finisher."""
        reqs = b.reqs
        r0 = reqs[0]
        backend = self.backends[r0.model]
        dev = backend.device
        names = ("features", "z") + (("cotangent",) if r0.op != "PREDICT" else ())
        if self.cuda:
            with torch.cuda.stream(self.h2d):
                ins = [stage.to_device([getattr(r, k) for r in reqs], dev, self.h2d) for k in names]
                staged = torch.cuda.Event()
                staged.record(self.h2d)
            self.compute.wait_event(staged)
            for t in ins:
                t.record_stream(self.compute)           # allocated on the copy stream, read on the compute stream
            plan = b.plan
        else:
            ins = [torch.from_numpy(np.stack([getattr(r, k) for r in reqs])) for k in names]
            plan = None
        b.t_staged = time.perf_counter()
        x, z = ins[0], ins[1]
        with (torch.cuda.stream(self.compute) if self.cuda else contextlib.nullcontext()):
            if r0.op == "PREDICT":
                n = self._predict(b, backend, r0, x, z, plan)
                b.outs = [n.permute(0, 2, 3, 4, 1).contiguous()]
                b.bad = (~torch.isfinite(n)).any() | (n <= 0).any()
            elif r0.op == "JVP":
                _, nd = backend.jvp_tensor(x, z, r0.dx, r0.dy, ins[2], plan=plan)
                b.outs = [nd.permute(0, 2, 3, 4, 1).contiguous()]
                b.bad = (~torch.isfinite(nd)).any()
            else:
                # the stride admission budgeted for: the engine's own choice reads the HBM free now, which the
                # batches queued beside this one will use, and could fall back to recomputing the forward many times
                cot = ins[2].permute(0, 4, 1, 2, 3).contiguous()
                _, fb, zb = backend.vjp_tensor(x, z, r0.dx, r0.dy, cot, stride=1 if b.stride is None else b.stride,
                                               plan=plan, keep=b.keep)
                b.outs = [fb.float().contiguous(), zb.float().contiguous()]
                b.bad = (~torch.isfinite(fb)).any() | (~torch.isfinite(zb)).any()
            if self.cuda:
                b.done = torch.cuda.Event()
                b.done.record(self.compute)
                for t in b.outs + [b.bad]:
                    t.record_stream(self.d2h)           # read by the finisher's copy stream
        b.t_enqueued = time.perf_counter()

    def _predict(self, b, backend, r0, x, z, plan):
        """Linear populations on the device: the bucket's graph replayed when there is one, else the eager forward
        (and a capture request in b.capture once the (bucket, batch size) recurs)."""
        nb = len(b.reqs)
        pad = 1 << (nb - 1).bit_length()               # the graph's batch size
        nz, nx, ny = r0.z.shape
        held = backend.model.engine().footprint_bytes(pad, nz, nx, ny, vjp=False)
        if not self.cuda or not self.graphs_enabled or held > self.graph_bytes:
            return backend.predict_tensor(x, z, r0.dx, r0.dy, plan=plan)
        key = (r0.key, pad)
        with self.graph_lock:
            g = self.graphs.get(key)
            uncapturable = key in self.no_graph
            if g is not None:
                self.graphs.move_to_end(key)            # least recently replayed graphs are evicted first
                g.users += 1
        if g is None:
            n = backend.predict_tensor(x, z, r0.dx, r0.dy, plan=plan)
            with self.graph_lock:
                self.eager_runs[key] += 1
            if not uncapturable:
                b.capture = (backend, key, x, z, r0, pad, held)
            return n
        try:
            # one graph's static buffers serve one replay at a time: the compute stream orders the replays
            g.sx[:nb].copy_(x)
            g.sz[:nb].copy_(z)
            g.graph.replay()
            self.replays += 1
            return g.sn[:nb].clone()
        finally:
            self._release(g)

    def _release(self, g):
        """A replay of g is enqueued: the last user of an evicted graph frees its count of held memory."""
        with self.graph_lock:
            g.users -= 1
            if g.evicted and g.users == 0:
                self.graph_held -= g.held

    def _evict_oldest(self):
        """Evict the least recently replayed graph (graph_lock held). Its held memory stays counted until its last
        replay, being enqueued now, has released it."""
        _, old = self.graphs.popitem(last=False)
        old.evicted = True
        if old.users == 0:
            self.graph_held -= old.held

    def _capture(self, backend, key, x, z, r0, pad, held):
        with self.graph_lock:
            if key in self.graphs or key in self.no_graph:
                return
        with self.cv:
            if self._claimed() + self.graph_held + held > self.budget:
                return                                  # no room to keep another graph's memory now
        plan = backend.plan_for(torch.from_numpy(r0.z).expand(pad, *r0.z.shape), r0.dx, r0.dy).prepare()
        dx, dy = r0.dx, r0.dy
        # the device to this thread alone while it captures: the finisher's copies wait
        with self.gate.exclusive(), torch.cuda.stream(self.compute):
            extra = pad - x.shape[0]                    # padded with the batch's last request
            sx = torch.cat([x, x[-1:].expand(extra, *x.shape[1:])]) if extra else x.clone()
            sz = torch.cat([z, z[-1:].expand(extra, *z.shape[1:])]) if extra else z.clone()
            self.compute.synchronize()
            graph = torch.cuda.CUDAGraph()
            try:
                # the capture calls themselves: torch.cuda.graph would first empty the allocator's caches, the
                # reserved segment with them
                graph.capture_begin(capture_error_mode="thread_local")
                try:
                    sn = backend.predict_tensor(sx, sz, dx, dy, plan=plan)
                finally:
                    graph.capture_end()
            except Exception:
                with self.graph_lock:
                    self.no_graph.add(key)              # this bucket stays eager
                return
            self.compute.synchronize()
        with self.graph_lock:
            self.graphs[key] = _Graph(graph, sx, sz, sn, held)
            self.graph_held += held
            self.captures += 1
            while len(self.graphs) > self.max_graphs:
                self._evict_oldest()

    # ------------------------------------------------------------------------------------------------ finisher
    def _finisher(self):
        stage = PinnedStage() if self.cuda else None
        while True:
            b = self.finish.get()
            if b is None:
                return
            try:
                results, t_computed = self._answer(b, stage)
            except Exception as e:
                self._fail(b, e)
                continue
            done = time.perf_counter()
            with self.cv:                               # recorded before any caller can see its answer
                for r in b.reqs:
                    r.t_done = done
                self.completed.extend(b.reqs)
                self.batch_log.append((b.t0, b.reqs[0].op, len(b.reqs), sum(r.voxels for r in b.reqs),
                                       b.t_staged - b.t0, t_computed - b.t_staged, done - t_computed))
            b.outs = b.bad = None
            self._unclaim(b)
            for r, res in zip(b.reqs, results):
                r.future.set_result(res)

    def _answer(self, b, stage):
        """(results, the time the batch's device work was seen finished)."""
        r0, nb = b.reqs[0], len(b.reqs)
        if self.cuda:
            with self.gate.shared(), torch.cuda.stream(self.d2h):
                b.done.synchronize()
                t_computed = time.perf_counter()
                bad = bool(b.bad.cpu())
                arrays = [] if bad else [stage.to_host(t, self.d2h) for t in b.outs]
        else:
            t_computed = time.perf_counter()
            bad = bool(b.bad)
            arrays = [t.float().numpy() for t in b.outs]
        tile = tuple(r0.z.shape)
        if r0.op == "PREDICT":
            if bad:
                raise RuntimeError(f"FFNO returned non-positive or non-finite populations (batch of {nb}, tile {tile})")
            return [{"populations": arrays[0][i]} for i in range(nb)], t_computed
        if r0.op == "JVP":
            if bad:
                raise RuntimeError(f"FFNO JVP returned NaN or Inf (batch of {nb}, tile {tile})")
            return [{"population_tangent": arrays[0][i]} for i in range(nb)], t_computed
        if bad:
            raise RuntimeError(f"FFNO VJP returned NaN or Inf (batch of {nb}, tile {tile})")
        return [{"features": arrays[0][i], "z_scale": arrays[1][i]} for i in range(nb)], t_computed


class _Graph:
    """A captured forward of one (bucket, batch size): its static buffers, the memory it holds, the replays using it
    now, and whether it has been evicted."""

    __slots__ = ("graph", "sx", "sz", "sn", "held", "users", "evicted")

    def __init__(self, graph, sx, sz, sn, held):
        self.graph, self.sx, self.sz, self.sn, self.held = graph, sx, sz, sn, held
        self.users = 0
        self.evicted = False
