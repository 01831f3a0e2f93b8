"""The batcher's admission on the CPU, with stand-in backends (no GPU): a request that fits the budget but not the
memory the engines' caches leave free, with nothing in flight, is admitted once the caches are dropped -- or failed
with MemoryError when they cannot be -- instead of waiting for good; a scheduler error fails the waiting requests
instead of ending the scheduler silently; captured graphs are evicted least recently replayed first, and an evicted
graph's memory stays counted until its last replay releases it; a VJP batch is no larger than keeps as many blocks
whole as its first request alone."""
import numpy as np
import pytest
import torch

from ffnoml_amd.serving import ContinuousBatcher, _Graph


class _Engine:
    def __init__(self, need, cached, droppable):
        self.need, self.cached, self.droppable = need, cached, droppable
        self.full_save = None
        self.dropped = 0

    def footprint_bytes(self, B, nz, nx, ny, vjp=False, jvp=False, span=0.0, keep=0):
        return self.need * B * (1 + keep)

    def cached_bytes(self):
        return self.cached

    def drop_unused_tables(self):
        self.dropped += 1
        if self.droppable:
            self.cached = 0


class _Model:
    def __init__(self, engine):
        self._e = engine

    def engine(self):
        return self._e


class _Backend:
    device = torch.device("cpu")
    level_names = ["L0", "L1"]

    def __init__(self, engine):
        self.model = _Model(engine)

    def predict_tensor(self, x, z, dx, dy, plan=None):
        return x[:, :2].double() + 1.0                   # [B, 2, nz, nx, ny], positive and finite


def _request():
    f = np.ones((6, 3, 2, 2), np.float32)
    z = np.broadcast_to(np.arange(3, dtype=np.float32)[:, None, None] * 1e6, (3, 2, 2)).copy()
    return f, z


@pytest.mark.parametrize("droppable", [True, False])
def test_a_lone_request_is_not_left_waiting(droppable):
    engine = _Engine(need=80_000, cached=50_000, droppable=droppable)
    batcher = ContinuousBatcher({"m": _Backend(engine)}, slots=2, graphs=False, budget=100_000)
    try:
        f, z = _request()
        fut = batcher.submit("PREDICT", "m", f, z, 48e3, 48e3)
        if droppable:
            out = fut.result(timeout=10)["populations"]
            assert out.shape == (3, 2, 2, 2) and np.allclose(out, 2.0)
        else:
            with pytest.raises(MemoryError):
                fut.result(timeout=10)
        assert engine.dropped >= 1
        # a request larger than the whole budget is refused at once
        engine.need = 200_000
        with pytest.raises(MemoryError):
            batcher.submit("PREDICT", "m", f, z, 48e3, 48e3).result(timeout=10)
    finally:
        batcher.close()


def test_scheduler_error_fails_waiting_requests():
    engine = _Engine(need=10_000, cached=0, droppable=True)
    batcher = ContinuousBatcher({"m": _Backend(engine)}, slots=1, graphs=False, budget=100_000)
    try:
        def broken():
            raise RuntimeError("cache accounting failed")
        engine.cached_bytes = broken
        f, z = _request()
        with pytest.raises(RuntimeError, match="cache accounting"):
            batcher.submit("PREDICT", "m", f, z, 48e3, 48e3).result(timeout=10)
        engine.cached_bytes = lambda: 0                  # the scheduler is still alive
        out = batcher.submit("PREDICT", "m", f, z, 48e3, 48e3).result(timeout=10)
        assert np.allclose(out["populations"], 2.0)
    finally:
        batcher.close()


def test_graph_eviction_is_lru_and_counts_until_released():
    engine = _Engine(need=10, cached=0, droppable=True)
    batcher = ContinuousBatcher({"m": _Backend(engine)}, slots=1, graphs=False, budget=1000, max_graphs=2)
    try:
        with batcher.graph_lock:
            for key in ("a", "b"):
                batcher.graphs[key] = _Graph(None, None, None, None, held=100)
                batcher.graph_held += 100
            batcher.graphs.move_to_end("a")              # "a" replayed last: "b" is the least recently used
            a_graph = batcher.graphs["a"]
            a_graph.users = 1                            # a replay of "a" in progress
            batcher._evict_oldest()
            assert list(batcher.graphs) == ["a"] and batcher.graph_held == 100
            batcher._evict_oldest()                      # "a" evicted while in use: still counted
            assert batcher.graph_held == 100
        batcher._release(a_graph)                        # its replay ends: now released
        assert batcher.graph_held == 0
    finally:
        batcher.close()


def test_vjp_batch_keeps_as_many_blocks_whole_as_one_request():
    """Three same-bucket VJP requests: together they would keep one of four blocks whole (the other three recomputed
    for every voxel of the batch), one alone keeps all four, two keep three: the scheduler admits the first alone,
    with every intermediate kept (stride 0). Forward requests of the bucket still batch together."""
    import queue
    engine = _Engine(need=10_000, cached=0, droppable=True)
    backend = _Backend(engine)
    backend.model.n_layers = 4
    batcher = ContinuousBatcher({"m": backend}, slots=2, graphs=False, budget=100_000)
    planner_queue = batcher.plans
    try:
        admitted = queue.Queue()
        with batcher.cv:
            batcher.plans = admitted                    # the planner stays on the old queue: batches stop here
            batcher.inflight = batcher.slots            # hold the scheduler until all requests wait
        f, z = _request()
        cot = np.zeros(z.shape + (2,), np.float32)
        for _ in range(3):
            batcher.submit("VJP", "m", f, z, 48e3, 48e3, cot)
        with batcher.cv:
            batcher.inflight = 0
            batcher.cv.notify_all()
        b = admitted.get(timeout=10)
        assert len(b.reqs) == 1 and b.keep == 4 and b.stride == 0
        with batcher.cv:
            batcher.inflight = batcher.slots
        for _ in range(3):
            batcher.submit("PREDICT", "m", f, z, 48e3, 48e3)
        with batcher.cv:
            batcher.inflight = 1                        # one slot free: the next VJP is admitted alone, then no more
            batcher.cv.notify_all()
        b = admitted.get(timeout=10)
        assert len(b.reqs) == 1 and b.reqs[0].op == "VJP"
    finally:
        with batcher.cv:                                # what was admitted here never runs: forget it, then close
            batcher.inflight = 0
            batcher.pending.clear()
            batcher.plans = planner_queue
        batcher.close()
