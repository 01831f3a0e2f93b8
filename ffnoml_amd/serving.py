"""Continuous batching of predict and VJP requests on one GPU.

The reference serves one request at a time over a single connection (ffno_fsdp_service.py:189-323). Here requests
from any number of clients go into one queue; a scheduler thread groups waiting requests that share a bucket (model,
operation, tile shape, spacings and column heights, so one plan serves them all), admits a batch as soon as an
execution slot is free and its HBM footprint fits the budget still unclaimed by batches in flight, and hands it to a
worker with its own HIP stream, so batches of different buckets run on the device at the same time. GroupNorm
statistics are per sample, so a request batched with others gets what it would get alone.

Forward-only batches of small tiles (forward footprint at most `graph_bytes`), once their bucket has been seen,
replay a captured HIP graph of the whole forward (static input/output buffers per (bucket, batch size)); there the
launch overhead of the ~15 launches per layer is a visible share of the batch time. When a (bucket, batch size)
recurs, the worker that ran it eagerly for the second time captures it right after answering, holding the device
alone for the capture (HIP refuses another thread's copies while a stream is capturing): new batches wait for it and
the ones in flight drain first, a stall of about one batch per captured shape. The memory a graph holds is counted
against the HBM budget;
graphs are evicted least recently replayed first, and an evicted graph's memory is released from the count only
when its last replay has finished. Larger tiles and VJP batches run eagerly.

Admission never stalls: when the oldest request does not fit the unclaimed budget while nothing is in flight, the
graphs and the engines' unused tables are dropped; a request that still does not fit is failed with MemoryError.
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
    """Shared holds for batches, an exclusive hold for a graph capture: a waiting capture blocks new shared holds, so
    the batches in flight drain and no other thread touches the device while a stream is capturing."""

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


class ContinuousBatcher:
    """submit() -> Future of the backend's result ({"populations"} or {"populations", "features", "z_scale"})."""

    def __init__(self, backends, slots=3, max_batch=8, memory_fraction=0.85, graphs=True, max_graphs=24,
                 graph_bytes=4 << 30, budget=None):
        self.backends = dict(backends)
        for backend in self.backends.values():
            # admission budgets the adjoint at checkpoint stride 1; keeping every intermediate would outgrow it
            backend.model.engine().full_save = False
        self.slots = int(slots)
        self.max_batch = int(max_batch)
        self.graphs_enabled = bool(graphs)
        self.max_graphs = int(max_graphs)
        self.graph_bytes = int(graph_bytes)
        self.graph_held = 0                             # bytes held by captured graphs' memory pools
        self.gate = DeviceGate()
        self.eager_runs = collections.Counter()         # forward batches run eagerly, per (bucket, batch size)
        dev = next(iter(self.backends.values())).device
        self.device = dev
        if budget is None:
            free, _ = torch.cuda.mem_get_info(dev)
            budget = int(free * memory_fraction)
        self.budget = int(budget)
        self.claimed = 0
        self.inflight = 0
        self.pending = collections.deque()
        self.cv = threading.Condition()
        self.work = queue.Queue()
        self.stop = False
        self.graphs = collections.OrderedDict()
        self.no_graph = set()
        self.graph_lock = threading.Lock()
        self.completed = []
        self.peak_inflight = 0
        self.queue_samples = []                          # (time, requests waiting, batches in flight) per admission
        self.captures = 0
        self.replays = 0
        self.workers = [threading.Thread(target=self._worker, daemon=True, name=f"ffno-worker-{i}")
                        for i in range(self.slots)]
        for w in self.workers:
            w.start()
        self.scheduler = threading.Thread(target=self._schedule, daemon=True, name="ffno-scheduler")
        self.scheduler.start()

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
        need = self._footprint([req])
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
        self.scheduler.join()
        for _ in self.workers:
            self.work.put(None)
        for w in self.workers:
            w.join()

    def latencies_ms(self):
        return np.array([(r.t_done - r.t_submit) * 1e3 for r in self.completed])

    # ------------------------------------------------------------------------------------------------ scheduler
    def _footprint(self, reqs):
        eng = self.backends[reqs[0].model].model.engine()
        nz, nx, ny = reqs[0].z.shape
        span = float(np.max(reqs[0].z.max(axis=0) - reqs[0].z.min(axis=0))) / 1e6       # metres -> Mm
        return eng.footprint_bytes(len(reqs), nz, nx, ny, vjp=reqs[0].op == "VJP", jvp=reqs[0].op == "JVP",
                                   span=span)

    def _free(self):
        """The budget not held by captured graphs or by the engines' cached plans and tables."""
        return self.budget - self.graph_held - sum(b.model.engine().cached_bytes() for b in self.backends.values())

    def _release_caches(self):
        """With nothing in flight: evict every captured graph and the engines' unused plans and tables."""
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
            # the oldest waiting request picks the bucket; same-bucket requests join it, oldest first
            head = self.pending[0]
            batch = [r for r in self.pending if r.key == head.key][: self.max_batch]
            free = self._free()
            while len(batch) > 1 and self.claimed + self._footprint(batch) > free:
                batch.pop()
            need = self._footprint(batch)
            if self.claimed + need > free:
                if self.inflight == 0:
                    # nothing will release memory by itself: drop the caches, and fail a request that still
                    # does not fit instead of waiting for good
                    self._release_caches()
                    if need > self._free():
                        self.pending.remove(head)
                        head.t_done = time.perf_counter()
                        self.completed.append(head)
                        head.future.set_exception(MemoryError(
                            f"request needs {need / 2**30:.1f} GiB; {max(self._free(), 0) / 2**30:.1f} GiB of the "
                            f"{self.budget / 2**30:.1f} GiB budget are free with nothing in flight"))
                    return False
                self.cv.wait(timeout=0.05)              # wait for memory to be released by a finishing batch
                return False
            for r in batch:
                self.pending.remove(r)
            self.claimed += need
            self.inflight += 1
            self.peak_inflight = max(self.peak_inflight, self.inflight)
            self.queue_samples.append((time.perf_counter(), len(self.pending), self.inflight))
        self.work.put((batch, need))
        return False

    # ------------------------------------------------------------------------------------------------ workers
    def _worker(self):
        cuda = torch.device(self.device).type == "cuda"
        stream = torch.cuda.Stream(device=self.device) if cuda else None
        while True:
            item = self.work.get()
            if item is None:
                return
            batch, need = item
            t0 = time.perf_counter()
            for r in batch:
                r.t_start = t0
            capture = None
            try:
                with self.gate.shared():
                    if cuda:
                        with torch.cuda.stream(stream):
                            results, capture = self._run(batch)
                        stream.synchronize()
                    else:
                        results, capture = self._run(batch)
                done = time.perf_counter()
                for r in batch:
                    r.t_done = done
                with self.cv:                           # recorded before any caller can see its answer
                    self.completed.extend(batch)
                for r, res in zip(batch, results):
                    r.future.set_result(res)
            except Exception as e:                      # one failed batch fails its own requests only
                done = time.perf_counter()
                with self.cv:
                    for r in batch:
                        if r.t_done is None:
                            r.t_done = done
                            self.completed.append(r)
                for r in batch:
                    if not r.future.done():
                        r.future.set_exception(e)
            try:
                if capture is not None:                 # after the answers went out: capture costs no latency
                    self._capture(stream, *capture)
            finally:
                with self.cv:
                    self.claimed -= need
                    self.inflight -= 1
                    self.cv.notify_all()

    def _stack(self, backend, batch):
        x = torch.from_numpy(np.stack([r.features for r in batch])).to(backend.device, non_blocking=False)
        z = torch.from_numpy(np.stack([r.z for r in batch])).to(backend.device, non_blocking=False)
        return x, z

    def _run(self, batch):
        """(results, capture request or None)."""
        backend = self.backends[batch[0].model]
        r0 = batch[0]
        x, z = self._stack(backend, batch)
        if r0.op == "PREDICT":
            n, capture = self._predict(backend, r0, x, z, len(batch))
            if not torch.isfinite(n).all() or torch.any(n <= 0):
                raise RuntimeError(f"FFNO returned non-positive or non-finite populations (batch of {len(batch)}, "
                                   f"tile {tuple(r0.z.shape)})")
            out = n.permute(0, 2, 3, 4, 1).float().cpu().numpy()
            return [{"populations": out[i]} for i in range(len(batch))], capture
        if r0.op == "JVP":
            td = torch.from_numpy(np.stack([r.cotangent for r in batch])).to(backend.device)
            _, nd = backend.jvp_tensor(x, z, r0.dx, r0.dy, td)
            if not torch.isfinite(nd).all():
                raise RuntimeError(f"FFNO JVP returned NaN or Inf (batch of {len(batch)}, tile {tuple(r0.z.shape)})")
            out = nd.permute(0, 2, 3, 4, 1).float().cpu().numpy()
            return [{"population_tangent": out[i]} for i in range(len(batch))], None
        cot = torch.from_numpy(np.stack([r.cotangent for r in batch])).to(backend.device).permute(0, 4, 1, 2, 3)
        n, fb, zb = backend.vjp_tensor(x, z, r0.dx, r0.dy, cot.contiguous())
        if not torch.isfinite(fb).all() or not torch.isfinite(zb).all():
            raise RuntimeError(f"FFNO VJP returned NaN or Inf (batch of {len(batch)}, tile {tuple(r0.z.shape)})")
        fb, zb = fb.float().cpu().numpy(), zb.float().cpu().numpy()
        return [{"features": fb[i], "z_scale": zb[i]} for i in range(len(batch))], None

    def _predict(self, backend, r0, x, z, nb):
        """(populations, capture request or None): replays the bucket's graph when there is one."""
        nz, nx, ny = r0.z.shape
        held = backend.model.engine().footprint_bytes(nb, nz, nx, ny, vjp=False)
        if not self.graphs_enabled or held > self.graph_bytes:
            return backend.predict_tensor(x, z, r0.dx, r0.dy), None
        key = (r0.key, nb)
        with self.graph_lock:
            g = self.graphs.get(key)
            uncapturable = key in self.no_graph
            if g is not None:
                self.graphs.move_to_end(key)            # least recently replayed graphs are evicted first
                g.users += 1
        if g is None:
            # batches of this (bucket, batch size) run eagerly until it recurs (the first builds the plan and its tables
            # and warms the allocator); at the second the worker captures, afterwards
            n = backend.predict_tensor(x, z, r0.dx, r0.dy)
            with self.graph_lock:
                self.eager_runs[key] += 1
                recurring = self.eager_runs[key] >= 2
            return n, (None if uncapturable or not recurring else (backend, key, x, z, r0.dx, r0.dy, held))
        try:
            # one graph's static buffers serve one batch at a time: a replay on this worker's stream waits until the
            # previous replay (on whichever stream) has copied its output out
            with g.lock:
                stream = torch.cuda.current_stream()
                if g.done is not None:
                    stream.wait_event(g.done)
                g.sx.copy_(x)
                g.sz.copy_(z)
                g.graph.replay()
                self.replays += 1
                out = g.sn.clone()
                g.done = torch.cuda.Event()
                g.done.record(stream)
        finally:
            self._release(g)
        return out, None

    def _release(self, g):
        """A replay of g is over (enqueued): the last user of an evicted graph frees its count of held memory."""
        with self.graph_lock:
            g.users -= 1
            if g.evicted and g.users == 0:
                self.graph_held -= g.held

    def _evict_oldest(self):
        """Evict the least recently replayed graph (graph_lock held). Its held memory stays counted until its last
        replay, running now or enqueued, has released it."""
        _, old = self.graphs.popitem(last=False)
        old.evicted = True
        if old.users == 0:
            self.graph_held -= old.held

    def _capture(self, stream, backend, key, x, z, dx, dy, held):
        with self.graph_lock:
            if key in self.graphs or key in self.no_graph:
                return
            if self.claimed + self.graph_held + held > self.budget:
                return                                  # no room to keep another graph's memory now
        # the device to this thread alone while it captures (HIP refuses other threads' copies to a capturing
        # device): new batches wait, the ones in flight drain
        with self.gate.exclusive(), torch.cuda.stream(stream):
            plan = backend.plan_for(z, dx, dy)          # cached by the eager batches: no host synchronisation here
            sx, sz = x.clone(), z.clone()
            stream.synchronize()
            graph = torch.cuda.CUDAGraph()
            try:
                with torch.cuda.graph(graph, stream=stream, capture_error_mode="thread_local"):
                    sn = backend.predict_tensor(sx, sz, dx, dy, plan=plan)
            except Exception:
                with self.graph_lock:
                    self.no_graph.add(key)              # this bucket stays eager
                return
            stream.synchronize()
        with self.graph_lock:
            self.graphs[key] = _Graph(graph, sx, sz, sn, held)
            self.graph_held += held
            self.captures += 1
            while len(self.graphs) > self.max_graphs:
                self._evict_oldest()


class _Graph:
    """A captured forward of one (bucket, batch size): its static buffers, the memory it holds, the event of its
    last replay, the replays using it now, and whether it has been evicted."""

    __slots__ = ("graph", "sx", "sz", "sn", "held", "lock", "done", "users", "evicted")

    def __init__(self, graph, sx, sz, sn, held):
        self.graph, self.sx, self.sz, self.sn, self.held = graph, sx, sz, sn, held
        self.lock = threading.Lock()
        self.done = None
        self.users = 0
        self.evicted = False
