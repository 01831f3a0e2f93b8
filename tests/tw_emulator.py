"""A float64 CPU model of the twiddle-GEMM and mix entry points over host tensors, for checking the engine's
launch sequences (index maps, factorised DFT stages, twiddles, layouts) with numbers on a machine without a GPU.
ffno_tw_gemm follows the header's contract (Y[b][m][n] = sum_k A[m][k] X[b][k][n] with the two-level index maps,
split-bf16 planes of A summed exactly, the optional twiddle of a first DFT stage); ffno_mix is replaced by a known
per-mode complex factor (its conjugate for the adjoint's two-plane form), so the spectral pipeline around it is
fixed by numpy.fft."""
import ctypes

import numpy as np
import torch


class Resolver:
    """Maps a raw host pointer to (numpy view of the owning root tensor, element offset)."""

    def __init__(self, tensors, more=None):
        self.roots = []
        self.more = more            # called for further tensors when a pointer is in none registered so far
        self.add(tensors)

    def add(self, tensors):
        for t in tensors:
            root = t._base if t._base is not None else t
            if not any(r is root for r in self.roots):
                self.roots.append(root)

    def __call__(self, ptr, dtype, retry=True):
        ptr = ptr.value if isinstance(ptr, ctypes.c_void_p) else ptr
        for r in self.roots:
            base, nbytes = r.data_ptr(), r.numel() * r.element_size()
            if base <= ptr < base + nbytes:
                arr = (r.view(torch.int16) if dtype == "bf16" else r).numpy().reshape(-1)
                return arr, (ptr - base) // arr.itemsize
        if retry and self.more is not None:
            self.add(self.more())
            return self(ptr, dtype, retry=False)
        raise AssertionError(f"pointer {ptr:#x} is in no registered tensor")


def _imap(i, split, hi, lo):
    return (i // split) * hi + (i % split) * lo


def _bf16_planes(res, g, mp, kp):
    a = np.zeros((mp, kp))
    for ptr in (g.a_hi, g.a_lo, g.a_l3):
        if not ptr:
            continue
        arr, off = res(ptr, "bf16")
        bits = arr[off:off + mp * kp].astype(np.int32) << 16
        a += bits.astype(np.uint32).view(np.float32).astype(np.float64).reshape(mp, kp)
    return a


def tw_gemm(res, g):
    mp, kp = g.m_pad, g.k_pad
    a = _bf16_planes(res, g, mp, kp)[:g.m, :g.k]
    x, xo = res(g.x, "f32")
    y, yo = res(g.y, "f32")
    b = np.arange(g.batch)
    n = np.arange(g.n)
    xb = _imap(b, g.b_split, g.x_b_hi, g.x_b_lo)
    yb = _imap(b, g.b_split, g.y_b_hi, g.y_b_lo)
    koff = _imap(np.arange(g.k), g.k_split, g.x_k_hi, g.x_k_lo)
    moff = _imap(np.arange(g.m), g.m_split, g.y_m_hi, g.y_m_lo)
    xv = x[xo + xb[:, None, None] + koff[None, :, None] + n[None, None, :]].astype(np.float64)
    out = np.einsum("mk,bkn->bmn", a, xv)
    rows = np.arange(g.m)
    if g.twiddle:
        tw, to = res(g.twiddle, "f32")
        t = (n // g.tw_cols)[None, :] if g.tw_cols else (b % g.tw_bsplit)[:, None] + 0 * n[None, :]
        for k1 in range(g.tw_n1):
            idx = to + 2 * (k1 * g.tw_n2 + t)
            w = tw[idx].astype(np.float64) + 1j * tw[idx + 1].astype(np.float64)
            z = (out[:, k1] + 1j * out[:, g.m_split + k1]) * w
            out[:, k1], out[:, g.m_split + k1] = z.real, z.imag
        rows = rows[rows % g.m_split < g.tw_n1]
    dst = yo + yb[:, None, None] + moff[rows][None, :, None] + n[None, None, :]
    val = out[:, rows]
    if g.accumulate:
        val = val + y[dst]
    y[dst] = val


class Emulator:
    """Stands in for the library's call(): ffno_tw_gemm as above, ffno_mix as the per-mode factor `factor`
    ([modes] complex, mode = ky * Wf + kx)."""

    def __init__(self, tensors, factor):
        from ffnoml_amd import engine as E

        def blocks():              # column blocks the engine cut from large matrices
            return [t for _, cache in E._K_BLOCKS.values() for blks in cache.values() for _, sub in blks
                    for t in sub if isinstance(t, torch.Tensor)]
        self.res = Resolver(tensors, blocks)
        self.factor = factor
        self.calls = []

    def __call__(self, name, *args):
        self.calls.append(name)
        if name == "ffno_tw_gemm":
            tw_gemm(self.res, args[0]._obj)
        elif name == "ffno_mix":
            m = args[0]._obj
            re, ro = self.res(m.re, "f32")
            im, io = self.res(m.im, "f32")
            f = self.factor if m.parts == 3 else np.conj(self.factor)
            for r in range(m.rows):
                off = _imap(r, m.r_split, m.r_hi, m.r_lo)
                idx = off + np.arange(m.modes)[:, None] * m.mode_stride + np.arange(m.C)[None, :]
                z = (re[ro + idx].astype(np.float64) + 1j * im[io + idx]) * f[:, None]
                re[ro + idx], im[io + idx] = z.real, z.imag
        else:
            raise AssertionError(f"{name} is not emulated")
        return 0
