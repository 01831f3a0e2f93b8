"""Host <-> device copies of served arrays through a small pinned double buffer.

A copy between pageable host memory and the GPU runs at about 1 GB/s on this platform (the runtime stages it
through its own small pinned buffer, one chunk at a time); from pinned memory the DMA engines move 12-14 GB/s. A
served batch's arrays are hundreds of MB to a few GB, so each worker copies them in pinned chunks: the device-side
copy of one chunk runs while the host copies the previous one between the pinned buffer and the caller's array.
Two chunks of pinned memory per direction per worker, whatever the batch size.
"""
import numpy as np
import torch

CHUNK_BYTES = 128 << 20


class PinnedStage:
    """One worker's pinned buffers (not shared between threads). Every copy is ordered on the given stream and is
    complete when the call returns."""

    def __init__(self, chunk_bytes=CHUNK_BYTES):
        self.chunk = int(chunk_bytes)
        self._bufs = None
        self._events = [None, None]

    def _buffers(self):
        if self._bufs is None:
            self._bufs = [torch.empty(self.chunk, dtype=torch.uint8, pin_memory=True) for _ in range(2)]
        return self._bufs

    def _wait(self, i):
        if self._events[i] is not None:
            self._events[i].synchronize()
            self._events[i] = None

    def to_host(self, t, stream):
        """A float32 numpy array with t's shape and values (t on the device; made contiguous there first)."""
        t = t.float().contiguous()
        out = np.empty(tuple(t.shape), dtype=np.float32)
        src = t.reshape(-1).view(torch.uint8)
        dst = out.reshape(-1).view(np.uint8)
        bufs = self._buffers()
        n, off, i, pending = src.numel(), 0, 0, None
        while off < n or pending is not None:
            issued = None
            if off < n:
                k = min(self.chunk, n - off)
                b = i % 2
                self._wait(b)
                bufs[b][:k].copy_(src[off:off + k], non_blocking=True)
                ev = torch.cuda.Event()
                ev.record(stream)
                self._events[b] = ev
                issued = (b, off, k)
                off += k
                i += 1
            if pending is not None:
                pb, poff, pk = pending
                self._wait(pb)
                dst[poff:poff + pk] = bufs[pb][:pk].numpy()
            pending = issued
        return out

    def to_device(self, arrays, device, stream):
        """torch float32 tensor [len(arrays), *shape] on the device from same-shaped host arrays (stacked)."""
        first = np.asarray(arrays[0])
        out = torch.empty((len(arrays),) + first.shape, dtype=torch.float32, device=device)
        flat = out.view(len(arrays), -1).view(torch.uint8)
        bufs = self._buffers()
        i = 0
        for j, a in enumerate(arrays):
            src = np.ascontiguousarray(a, dtype=np.float32).reshape(-1).view(np.uint8)
            n, off = src.size, 0
            while off < n:
                k = min(self.chunk, n - off)
                b = i % 2
                self._wait(b)                           # the buffer's previous copy to the device has finished
                bufs[b][:k].numpy()[:] = src[off:off + k]
                flat[j, off:off + k].copy_(bufs[b][:k], non_blocking=True)
                ev = torch.cuda.Event()
                ev.record(stream)
                self._events[b] = ev
                off += k
                i += 1
        for b in (0, 1):
            self._wait(b)
        return out
