"""Continuous batching of predict and VJP requests on one GPU.

The reference serves one request at a time over a single connection (ffno_fsdp_service.py:189-323). Here requests
from any number of clients go into one queue; a scheduler thread groups waiting requests that share a bucket (model,
operation, tile shape, spacings and column heights, so one plan serves them all), admits a batch as soon as an
execution slot is free and its HBM footprint fits the budget still unclaimed by batches in flight, and hands it to a
worker with its own HIP stream, so batches of different buckets run on the device at the same time. GroupNorm
statistics are per sample, so a request batched with others gets what it would get alone.

Forward-only batches of small tiles (forward footprint at most `graph_bytes`), once their bucket has been seen,
replay a captured HIP graph of the whole forward (static input/output buffers per (bucket, batch size)); there the
launch overhead of the ~15 launches per layer is a visible share of the batch time. The memory a graph holds is
counted against the HBM budget. A capture takes the device alone: batches run under shared holds of a gate, and a
capture happens only when the gate is idle (no batch in flight), so no other stream's host-synchronising work can
invalidate it and no batch ever waits for one; a busy device defers the capture to a later batch of the bucket.
Larger tiles and VJP batches run eagerly.
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
    """Shared holds for batches, exclusive holds for graph captures; a waiting capture blocks new shared holds."""

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
    def try_exclusive(self):
        """An exclusive hold only if the device is idle right now (yields False, holding nothing, otherwise)."""
        with self.cv:
            ok = not self.writer and not self.readers and not self.waiting
            if ok:
                self.writer = True
        try:
            yield ok
        finally:
            if ok:
                with self.cv:
                    self.writer = False
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

    def __init__(self, backends, slots=3, max_batch=8, memory_fraction=0.85, graphs=True, max_graphs=32,
                 graph_bytes=1 << 30):
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
        dev = next(iter(self.backends.values())).device
        self.device = dev
        free, total = torch.cuda.mem_get_info(dev)
        self.budget = int(free * memory_fraction)
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
        need = backend.model.engine().footprint_bytes(1, *z.shape, vjp=op == "VJP", jvp=op == "JVP")
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
        return eng.footprint_bytes(len(reqs), nz, nx, ny, vjp=reqs[0].op == "VJP", jvp=reqs[0].op == "JVP")

    def _schedule(self):
        while True:
            with self.cv:
                while not self.stop and (not self.pending or self.inflight >= self.slots):
                    self.cv.wait()
                if self.stop and not self.pending:
                    return
                # the oldest waiting request picks the bucket; same-bucket requests join it, oldest first
                head = self.pending[0]
                batch = [r for r in self.pending if r.key == head.key][: self.max_batch]
                # per-mode tables the engines cache are device memory outside any batch's footprint
                free = self.budget - self.graph_held - sum(b.model.engine().cached_bytes()
                                                           for b in self.backends.values())
                while len(batch) > 1 and self.claimed + self._footprint(batch) > free:
                    batch.pop()
                need = self._footprint(batch)
                if self.claimed + need > free:
                    self.cv.wait(timeout=0.05)          # wait for memory to be released by a finishing batch
                    continue
                for r in batch:
                    self.pending.remove(r)
                self.claimed += need
                self.inflight += 1
                self.peak_inflight = max(self.peak_inflight, self.inflight)
            self.work.put((batch, need))

    # ------------------------------------------------------------------------------------------------ workers
    def _worker(self):
        stream = torch.cuda.Stream(device=self.device)
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
                    with torch.cuda.stream(stream):
                        results, capture = self._run(batch)
                    stream.synchronize()
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
        if g is None:
            # first batch of this shape runs eagerly (building the plan and warming the allocator); the worker
            # captures afterwards, alone on the device
            n = backend.predict_tensor(x, z, r0.dx, r0.dy)
            return n, (None if uncapturable else (backend, key, x, z, r0.dx, r0.dy, held))
        graph, sx, sz, sn, _, lock, _ = g
        # one graph's static buffers serve one batch at a time: a replay on this worker's stream waits until the
        # previous replay (on whichever stream) has copied its output out
        with lock:
            stream = torch.cuda.current_stream()
            if g[6] is not None:
                stream.wait_event(g[6])
            sx.copy_(x)
            sz.copy_(z)
            graph.replay()
            out = sn.clone()
            done = torch.cuda.Event()
            done.record(stream)
            g[6] = done
        return out, None

    def _capture(self, stream, backend, key, x, z, dx, dy, held):
        with self.graph_lock:
            if key in self.graphs or key in self.no_graph:
                return
            if self.claimed + self.graph_held + held > self.budget:
                return                                  # no room to keep another graph's memory now
        with self.gate.try_exclusive() as idle, torch.cuda.stream(stream):
            if not idle:
                return                                  # other batches are running: capture on a later batch
            plan = backend.plan_for(z, dx, dy)          # the host synchronisation of planning happens here
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
            self.graphs[key] = [graph, sx, sz, sn, held, threading.Lock(), None]
            self.graph_held += held
            while len(self.graphs) > self.max_graphs:
                _, old = self.graphs.popitem(last=False)
                with old[5]:                            # its last replay has finished before its pool is freed
                    if old[6] is not None:
                        old[6].synchronize()
                self.graph_held -= old[4]
