"""CPU bounds check of every HIP launch the engine issues: the engine's own forward / adjoint schedule runs on meta
tensors (shapes, no data, no GPU), every device pointer it passes gets a synthetic address inside the tensor that
owns it, and each entry point's call is checked against a model of the bytes its kernel reads and writes -- the
largest offset from the kernel's index arithmetic (row-major activations, the two-level GEMM index maps, per-mode
tables, partial-statistics slots) against the end of the owning tensor. This is how a launch at a size no GPU test
covers (a whole 504 x 504 x 256 volume) is shown to stay in bounds before it runs on a GPU."""
import ctypes
import types

import numpy as np
import torch

from ffnoml_amd import _lib, engine as E
from ffnoml_amd.transforms import round_up

F32, BF = 4, 2


class Registry:
    """Synthetic device addresses: each root tensor gets its own 1 TiB-aligned range."""

    def __init__(self):
        self.roots = {}
        self.keep = []
        self.next = 1 << 40

    def address(self, t):
        if t is None:
            return 0
        root = t._base if t._base is not None else t
        key = id(root)
        if key not in self.roots:
            self.roots[key] = (self.next, root.numel() * root.element_size())
            self.keep.append(root)
            self.next += 1 << 40
        base, _ = self.roots[key]
        return base + t.storage_offset() * t.element_size()

    def owner(self, addr):
        for base, nbytes in self.roots.values():
            if base <= addr < base + max(nbytes, 1):
                return base, nbytes
        raise AssertionError(f"pointer {addr:#x} is not inside any tensor")


class Checker:
    def __init__(self, reg):
        self.reg = reg
        self.calls = []

    # ---- helpers
    def span(self, name, what, ptr, nbytes, optional=False):
        """ptr .. ptr + nbytes must lie inside one tensor."""
        ptr = ptr.value if isinstance(ptr, ctypes.c_void_p) else ptr
        if not ptr:
            assert optional, f"{name}: {what} is NULL"
            return
        assert nbytes > 0, f"{name}: {what} has an empty access"
        base, size = self.reg.owner(ptr)
        assert ptr + nbytes <= base + size, (f"{name}: {what} reads/writes {ptr - base + nbytes} bytes of a "
                                             f"{size}-byte tensor")

    def act(self, name, what, ptr, rows, optional=False):
        self.span(name, what, ptr, rows.B * rows.voxels * rows.C * F32, optional)

    def asm(self, name, a, rows):
        self.act(name, "assembly base", a.base, rows, optional=True)
        self.act(name, "assembly t", a.t, rows, optional=True)
        if a.t:
            self.span(name, "assembly aff_a", a.aff_a, rows.B * rows.C * F32)
            self.span(name, "assembly aff_b", a.aff_b, rows.B * rows.C * F32)
        self.span(name, "assembly scale", a.scale, F32, optional=True)

    def stats(self, name, s, rows):
        self.span(name, "depth_w", s.depth_w, rows.B * rows.voxels * F32)
        self.span(name, "partials", s.partials, rows.B * rows.nwg * 8 * 3 * F32, optional=True)

    def gnback(self, name, g, rows):
        self.act(name, "gn t", g.t, rows)
        self.span(name, "gn stats", g.stats, rows.B * 8 * 3 * F32)
        self.span(name, "gn gamma", g.gamma, rows.C * F32)
        self.span(name, "gn aff_a", g.aff_a, rows.B * rows.C * F32)
        self.span(name, "gn aff_b", g.aff_b, rows.B * rows.C * F32)
        self.span(name, "gn scale", g.scale, F32, optional=True)
        self.span(name, "gn sums", g.sums, rows.B * 8 * 2 * F32, optional=True)
        self.span(name, "gn depth_w", g.depth_w, rows.B * rows.voxels * F32)
        self.span(name, "gn wbar", g.wbar, rows.B * rows.voxels * F32, optional=True)

    def weight(self, name, what, ptr, n_rows, k, parts=2):
        """A packed split-bf16 matrix [parts][n_rows][round_up(k, 32) + WPAD] as the kernels stage it (rows padded by WPAD
        bf16; parts=3: the forward's three-way pointwise layers)."""
        self.span(name, what, ptr, parts * n_rows * (round_up(k, 32) + E.WPAD) * BF)

    @staticmethod
    def imap_max(count, split, hi, lo):
        assert split > 0 and hi >= 0 and lo >= 0 and count > 0
        q = (count - 1) // split
        best = q * hi + ((count - 1) % split) * lo
        if q >= 1:
            best = max(best, (q - 1) * hi + (split - 1) * lo)
        return best

    # ---- entry points
    def __call__(self, name, *a):
        self.calls.append(name)
        getattr(self, name)(name, *a)
        return 0

    def ffno_lift(self, n, rows, x_in, w, out, st, stream):
        self.span(n, "x_in", x_in, rows.B * 6 * rows.voxels * F32)
        self.span(n, "w_lift", w, rows.C * 6 * F32)
        self.act(n, "out", out, rows)
        self.stats(n, st, rows)

    def ffno_gn_finalize(self, n, B, nwg, C, part, gamma, beta, stats, a, b, stream):
        self.span(n, "partials", part, B * nwg * 8 * 3 * F32)
        self.span(n, "gamma", gamma, C * F32)
        self.span(n, "beta", beta, C * F32)
        self.span(n, "stats", stats, B * 8 * 3 * F32)
        self.span(n, "aff_a", a, B * C * F32)
        self.span(n, "aff_b", b, B * C * F32)

    def ffno_block_in(self, n, rows, a, x_out, g1, g2, v1, v2, z_proj, z, xs, xv, stream):
        C = rows.C
        self.asm(n, a, rows)
        self.act(n, "x_out", x_out, rows, optional=True)
        for what, w in (("g1", g1), ("g2", g2), ("v1", v1), ("v2", v2)):
            self.weight(n, what, w, C, C, parts=3)
        self.span(n, "z_proj", z_proj, C * F32)
        self.span(n, "z", z, rows.B * rows.voxels * F32)
        self.act(n, "xs", xs, rows, optional=True)
        self.act(n, "xv", xv, rows)

    def ffno_post(self, n, rows, u, w1, w2, y, st, stream):
        C = rows.C
        self.act(n, "u", u, rows)
        self.weight(n, "w1", w1, 2 * C, C, parts=3)
        self.weight(n, "w2", w2, C, 2 * C, parts=3)
        self.act(n, "y", y, rows)
        self.stats(n, st, rows)

    def ffno_post_gn_noalloc(self, n, rows, u, w1, w2, y, depth_w, gamma, beta, stats, aff_a, aff_b, stream):
        C = rows.C
        self.act(n, "u", u, rows)
        self.weight(n, "w1", w1, 2 * C, C, parts=3)
        self.weight(n, "w2", w2, C, 2 * C, parts=3)
        self.act(n, "y", y, rows)
        self.span(n, "depth_w", depth_w, rows.B * rows.voxels * F32)
        self.span(n, "gamma", gamma, C * F32)
        self.span(n, "beta", beta, C * F32)
        self.span(n, "stats", stats, rows.B * 8 * 3 * F32, optional=True)
        self.span(n, "aff_a", aff_a, rows.B * C * F32)
        self.span(n, "aff_b", aff_b, rows.B * C * F32)
        assert rows.B * rows.nwg <= 4096, "ffno_post_gn_noalloc: more partial slots than its static scratch holds"

    def ffno_fuse(self, n, rows, x, sp, vp, sa, sb, va, vb, bg1, bg2, bias, f1, f2, fp, st, gates, stream):
        C = rows.C
        for what, p in (("x", x), ("sp", sp), ("vp", vp), ("fp", fp)):
            self.act(n, what, p, rows)
        for what, p in (("s_a", sa), ("s_b", sb), ("v_a", va), ("v_b", vb)):
            self.span(n, what, p, rows.B * C * F32)
        self.span(n, "bg1", bg1, 2 * C * (3 * round_up(C, 32) + E.WPAD) * BF)
        self.weight(n, "bg2", bg2, 2 * C, C)
        self.span(n, "bg2_bias", bias, 2 * C * F32)
        self.span(n, "f1", f1, 2 * 2 * C * (2 * round_up(C, 32) + E.WPAD) * BF)
        self.weight(n, "f2", f2, C, 2 * C)
        self.stats(n, st, rows)
        self.span(n, "gates_out", gates, rows.B * rows.voxels * 2 * C * F32, optional=True)

    def ffno_resmlp(self, n, rows, a, x_out, w1, w2, y, st, stream):
        C = rows.C
        self.asm(n, a, rows)
        self.act(n, "x_out", x_out, rows, optional=True)
        self.weight(n, "w1", w1, 2 * C, C, parts=3)
        self.weight(n, "w2", w2, C, 2 * C, parts=3)
        self.act(n, "y", y, rows)
        self.stats(n, st, rows)

    def ffno_head(self, n, rows, a, x_out, p1, p2, bias, x_in, desc, y, stream):
        d = desc._obj
        n_head = sum(d.levels[s] + 1 for s in range(d.n_species))
        n_out = sum(d.levels[s] for s in range(d.n_species))
        nth = (n_head + 15) // 16
        C = rows.C
        self.asm(n, a, rows)
        self.act(n, "x_out", x_out, rows, optional=True)
        self.weight(n, "p1", p1, 2 * C, C, parts=3)
        self.span(n, "p2", p2, 3 * nth * 16 * (2 * C + E.WPAD) * BF)
        self.span(n, "p2_bias", bias, nth * 16 * F32)
        self.span(n, "x_in", x_in, rows.B * 6 * rows.voxels * F32)
        self.span(n, "y", y, rows.B * n_out * rows.voxels * F32, optional=True)
        assert x_out or y, "ffno_head: nothing to write"

    def ffno_tw_gemm(self, n, gref, stream):
        g = gref._obj
        assert g.n % 4 == 0 and g.m_pad % 16 == 0 and g.k_pad % 32 == 0 and g.m <= g.m_pad and g.k <= g.k_pad
        self.span(n, "a_hi", g.a_hi, g.m_pad * g.k_pad * BF)
        self.span(n, "a_lo", g.a_lo, g.m_pad * g.k_pad * BF)
        xb = self.imap_max(g.batch, g.b_split, g.x_b_hi, g.x_b_lo)
        xk = self.imap_max(g.k, g.k_split, g.x_k_hi, g.x_k_lo)
        self.span(n, "x", g.x, (xb + xk + g.n) * F32)
        yb = self.imap_max(g.batch, g.b_split, g.y_b_hi, g.y_b_lo)
        if g.twiddle:
            # the first stage of a factorised DFT: rows k1 < tw_n1 of each part are written, and the twiddle table
            # [tw_n1][tw_n2] complex is read at the column groups present
            assert g.m_pad <= 128 and g.k_pad <= 128 and g.m_pad % 32 == 0 and g.m_split == g.m_pad // 2
            assert g.m == g.m_split + g.tw_n1 and 0 < g.tw_n1 <= g.m_split and not g.accumulate
            assert g.tw_cols >= 0 and g.tw_cols % 4 == 0
            if g.tw_cols:
                tmax = -(-g.n // g.tw_cols) - 1                    # column group of the last column
            else:
                assert 0 < g.tw_bsplit <= g.tw_n2
                tmax = min(g.batch, g.tw_bsplit) - 1               # batch item's place in its cycle
            assert tmax < g.tw_n2
            ym = max((m // g.m_split) * g.y_m_hi + (m % g.m_split) * g.y_m_lo
                     for m in range(g.m) if m % g.m_split < g.tw_n1)
            self.span(n, "twiddle", g.twiddle, ((g.tw_n1 - 1) * g.tw_n2 + tmax + 1) * 2 * F32)
        else:
            ym = self.imap_max(g.m, g.m_split, g.y_m_hi, g.y_m_lo)
        self.span(n, "y", g.y, (yb + ym + g.n) * F32)

    def ffno_col_gemm(self, n, gref, stream):
        g = gref._obj
        assert g.C % 4 == 0 and g.m_pad % 16 == 0 and g.k_pad % 32 == 0 and g.m <= g.m_pad and g.k <= g.k_pad
        n_a = g.B * g.hw * g.m_pad * g.k_pad * BF
        self.span(n, "a_hi", g.a_hi, n_a)
        self.span(n, "a_lo", g.a_lo, n_a)
        self.span(n, "a_l3", g.a_l3, n_a, optional=True)
        self.span(n, "x", g.x, g.B * g.k * g.hw * g.C * F32)
        self.span(n, "y", g.y, g.B * g.m * g.hw * g.C * F32)

    def ffno_mix(self, n, mref, stream):
        m = mref._obj
        C = m.C
        KP = round_up(C, 32)
        top = self.imap_max(m.rows, m.r_split, m.r_hi, m.r_lo) + (m.modes - 1) * m.mode_stride + C
        self.span(n, "re", m.re, top * F32)
        self.span(n, "im", m.im, top * F32)
        self.span(n, "wc", m.wc, m.parts * 2 * C * (2 * KP + E.WPAD) * BF)
        self.span(n, "gate", m.gate, m.modes * 2 * F32)
        if m.rank:
            self.span(n, "v", m.v_p, m.modes * m.parts * 16 * (KP + E.WPAD) * BF)
            self.span(n, "u", m.u_p, m.modes * m.parts * C * (32 + E.WPAD) * BF)
        self.span(n, "factor", m.factor, m.modes * F32, optional=True)

    def ffno_gn_grad_stats(self, n, rows, gb, ubar, part, stream):
        self.gnback(n, gb, rows)
        self.act(n, "ubar", ubar, rows)
        self.span(n, "partials", part, rows.B * rows.nwg * 8 * 2 * F32)

    def ffno_gn_grad_finalize(self, n, B, nwg, part, sums, stream):
        self.span(n, "partials", part, B * nwg * 8 * 2 * F32)
        self.span(n, "sums", sums, B * 8 * 2 * F32)

    def next_sums(self, n, ref, part, rows):
        """An optional fused GroupNorm-gradient-sums target: the next GroupNorm's descriptor and its partials."""
        if part is None or not getattr(part, "value", part):
            return
        assert ref is not None, f"{n}: partials without a GroupNorm descriptor"
        self.gnback(n, ref._obj, rows)
        self.span(n, "next partials", part, rows.B * rows.nwg * 8 * 2 * F32)

    def ffno_mlp_bwd(self, n, rows, ubar, gb, x, w1, w2t, w1t, residual, xbar, nxt, npart, stream):
        C = rows.C
        self.act(n, "ubar", ubar, rows)
        self.gnback(n, gb, rows)
        assert gb.sums, "ffno_mlp_bwd: GroupNorm gradient sums missing"
        self.asm(n, x, rows)
        self.weight(n, "w1", w1, 2 * C, C)
        self.weight(n, "w2t", w2t, 2 * C, C)
        self.weight(n, "w1t", w1t, C, 2 * C)
        self.act(n, "xbar", xbar, rows)
        self.next_sums(n, nxt, npart, rows)

    def ffno_fuse_bwd(self, n, rows, x1bar, gb, h, sp, vp, sa, sb, va, vb, bg1, bg2, bias, f1, f2t, f1t, bg2t, bg1t,
                      hbar, sbar, vbar, stream):
        C = rows.C
        KP = round_up(C, 32)
        for what, p in (("x1bar", x1bar), ("h", h), ("sp", sp), ("vp", vp), ("hbar", hbar), ("sbar", sbar),
                        ("vbar", vbar)):
            self.act(n, what, p, rows)
        self.gnback(n, gb, rows)
        for what, p in (("s_a", sa), ("s_b", sb), ("v_a", va), ("v_b", vb)):
            self.span(n, what, p, rows.B * C * F32)
        self.span(n, "bg1", bg1, 2 * C * (3 * KP + E.WPAD) * BF)
        self.weight(n, "bg2", bg2, 2 * C, C)
        self.span(n, "bg2_bias", bias, 2 * C * F32)
        self.span(n, "f1", f1, 2 * 2 * C * (2 * KP + E.WPAD) * BF)
        self.weight(n, "f2t", f2t, 2 * C, C)
        self.span(n, "f1t", f1t, 2 * 2 * C * (2 * C + E.WPAD) * BF)
        self.weight(n, "bg2t", bg2t, C, 2 * C)
        self.span(n, "bg1t", bg1t, 2 * 3 * C * (KP + E.WPAD) * BF)

    def ffno_fuse_mlp_bwd(self, n, rows, x1bar, gb, sp, vp, sa, sb, va, vb, gates, f1, f2t, f1t, smbar, vmbar, stream):
        C = rows.C
        KP = round_up(C, 32)
        for what, p in (("x1bar", x1bar), ("sp", sp), ("vp", vp), ("smbar", smbar), ("vmbar", vmbar)):
            self.act(n, what, p, rows)
        self.gnback(n, gb, rows)
        assert gb.sums, "ffno_fuse_mlp_bwd: GroupNorm gradient sums missing"
        for what, p in (("s_a", sa), ("s_b", sb), ("v_a", va), ("v_b", vb)):
            self.span(n, what, p, rows.B * C * F32)
        self.span(n, "gates", gates, rows.B * rows.voxels * 2 * C * F32)
        self.span(n, "f1", f1, 2 * 2 * C * (2 * KP + E.WPAD) * BF)
        self.weight(n, "f2t", f2t, 2 * C, C)
        self.span(n, "f1t", f1t, 2 * 2 * C * (2 * C + E.WPAD) * BF)

    def ffno_fuse_gates_bwd(self, n, rows, xh, h, sp, vp, sa, sb, va, vb, gates, bg1, bg2t, bg1t, svs, svv, gs, ps,
                            gv, pv, stream):
        C = rows.C
        KP = round_up(C, 32)
        for what, p in (("x1bar/hbar", xh), ("h", h), ("sp", sp), ("vp", vp), ("smbar/sbar", svs),
                        ("vmbar/vbar", svv)):
            self.act(n, what, p, rows)
        for what, p in (("s_a", sa), ("s_b", sb), ("v_a", va), ("v_b", vb)):
            self.span(n, what, p, rows.B * C * F32)
        self.span(n, "gates", gates, rows.B * rows.voxels * 2 * C * F32)
        self.span(n, "bg1", bg1, 2 * C * (3 * KP + E.WPAD) * BF)
        self.weight(n, "bg2t", bg2t, C, 2 * C)
        self.span(n, "bg1t", bg1t, 2 * 3 * C * (KP + E.WPAD) * BF)
        assert bool(getattr(ps, "value", ps)) == bool(getattr(pv, "value", pv)), f"{n}: one of two sums targets"
        self.next_sums(n, gs, ps, rows)
        self.next_sums(n, gv, pv, rows)

    def ffno_lowrank_tables(self, n, modes, C, r, hidden, hid, wu, bu, wv, bv, parts, adjoint, v, u, stream):
        kc = round_up(C, 32)
        self.span(n, "hid", hid, modes * hidden * F32)
        for what, p in (("wu_t", wu), ("wv_t", wv)):
            self.span(n, what, p, hidden * C * r * F32)
        for what, p in (("bu", bu), ("bv", bv)):
            self.span(n, what, p, C * r * F32)
        self.span(n, "v_out", v, modes * parts * 16 * (kc + E.WPAD) * BF)
        self.span(n, "u_out", u, modes * parts * C * (32 + E.WPAD) * BF)

    def ffno_gates_bwd(self, n, rows, h, xsb, xvb, hin, g1, g2, g2t, g1t, v1, v2, v2t, v1t, z_proj, hout, zbar,
                       stream):
        C = rows.C
        for what, p in (("h", h), ("xsbar", xsb), ("xvbar", xvb), ("hbar_in", hin), ("hbar_out", hout)):
            self.act(n, what, p, rows)
        for what, w in (("g1", g1), ("g2", g2), ("g2t", g2t), ("g1t", g1t), ("v1", v1), ("v2", v2), ("v2t", v2t),
                        ("v1t", v1t)):
            self.weight(n, what, w, C, C)
        self.span(n, "z_proj", z_proj, C * F32)
        self.span(n, "zbar", zbar, rows.B * rows.voxels * F32, optional=True)

    def ffno_head_bwd(self, n, rows, h, p1, p2, p2t, p1t, bias, x_in, desc, ybar, hbar, xbar_in, stream):
        d = desc._obj
        n_head = sum(d.levels[s] + 1 for s in range(d.n_species))
        n_out = sum(d.levels[s] for s in range(d.n_species))
        nth = (n_head + 15) // 16
        C = rows.C
        self.act(n, "h", h, rows)
        self.weight(n, "p1", p1, 2 * C, C)
        self.span(n, "p2", p2, 2 * nth * 16 * (2 * C + E.WPAD) * BF)
        self.weight(n, "p2t", p2t, 2 * C, n_head)
        self.weight(n, "p1t", p1t, C, 2 * C)
        self.span(n, "p2_bias", bias, nth * 16 * F32)
        self.span(n, "x_in", x_in, rows.B * 6 * rows.voxels * F32)
        self.span(n, "ybar", ybar, rows.B * n_out * rows.voxels * F32)
        self.act(n, "hbar", hbar, rows)
        self.span(n, "xbar_in", xbar_in, rows.B * 6 * rows.voxels * F32)

    def ffno_dot_reduce(self, n, rows, D, a, delta, r, zbar, stream):
        assert rows.voxels % D == 0 and delta in (-1, 0, 1)
        self.act(n, "a", a, rows)          # a is read at row + delta * hw only when that level exists
        self.act(n, "r", r, rows)
        self.span(n, "zbar", zbar, rows.B * rows.voxels * F32)

    def ffno_zgrad_columns(self, n, g, stream):
        g = g._obj
        assert g.d_pad % 16 == 0 and g.r_pad % 16 == 0 and g.d_pad >= g.D and g.r_pad >= g.R and g.C % 16 == 0
        KP = round_up(g.C, 32)
        # per column the kernel reads channels up to the last 32-block's end only where they exist (c0 < C)
        lev = g.B * g.D * g.hw * g.C * F32
        spec = g.B * g.R * g.hw * g.C * F32
        assert KP - 32 < g.C
        for what, p, nb in (("xv", g.xv, lev), ("xbar_spec", g.xbar_spec, spec), ("yv_bar", g.yv_bar, lev),
                            ("y_spec", g.y_spec, spec)):
            self.span(n, what, p, nb)
        self.span(n, "tables", g.tables, 4 * g.r_pad * g.d_pad * F32)
        self.span(n, "zbar", g.zbar, g.B * g.D * g.hw * F32)

    def ffno_lift_bwd(self, n, rows, ubar, gb, w, xbar_in, stream):
        self.act(n, "ubar", ubar, rows)
        self.gnback(n, gb, rows)
        self.span(n, "w_lift", w, rows.C * 6 * F32)
        self.span(n, "xbar_in", xbar_in, rows.B * 6 * rows.voxels * F32)


    # ---- tangent (JVP) entry points
    def ffno_gn_tangent_stats(self, n, rows, y, yd, stats, depth_w, part, stream):
        self.act(n, "y", y, rows)
        self.act(n, "yd", yd, rows)
        self.span(n, "stats", stats, rows.B * 8 * 3 * F32)
        self.span(n, "depth_w", depth_w, rows.B * rows.voxels * F32)
        self.span(n, "partials", part, rows.B * rows.nwg * 8 * 2 * F32)

    def ffno_gn_tangent_finalize(self, n, B, nwg, C, part, stats, gamma, tb, tc, stream):
        self.span(n, "partials", part, B * nwg * 8 * 2 * F32)
        self.span(n, "stats", stats, B * 8 * 3 * F32)
        self.span(n, "gamma", gamma, C * F32)
        self.span(n, "tan_b", tb, B * C * F32)
        self.span(n, "tan_c", tc, B * C * F32)

    def ffno_assemble_t(self, n, rows, a, out, stream):
        assert rows.C % 4 == 0
        self.act(n, "base_dot", a.base_dot, rows, optional=True)
        self.act(n, "t", a.t, rows)
        self.act(n, "t_dot", a.t_dot, rows)
        for what, p in (("aff_a", a.aff_a), ("aff_b", a.aff_b), ("tan_b", a.tan_b), ("tan_c", a.tan_c)):
            self.span(n, what, p, rows.B * rows.C * F32)
        self.span(n, "scale", a.scale, F32, optional=True)
        self.act(n, "out", out, rows)

    def ffno_lift_t(self, n, rows, xd_in, w, yd, stream):
        self.span(n, "xd_in", xd_in, rows.B * 6 * rows.voxels * F32)
        self.span(n, "w_lift", w, rows.C * 6 * F32)
        self.act(n, "yd", yd, rows)

    def ffno_gates_t(self, n, rows, h, hd, g1, g2, v1, v2, xsd, xvd, stream):
        C = rows.C
        for what, p in (("h", h), ("hd", hd), ("xsd", xsd), ("xvd", xvd)):
            self.act(n, what, p, rows)
        for what, w in (("g1", g1), ("g2", g2), ("v1", v1), ("v2", v2)):
            self.weight(n, what, w, C, C)

    def tan_sums(self, n, ts, rows):
        if not ts.partials:
            return
        self.act(n, "ts.y", ts.y, rows)
        self.span(n, "ts.stats", ts.stats, rows.B * 8 * 3 * F32)
        self.span(n, "ts.depth_w", ts.depth_w, rows.B * rows.voxels * F32)
        self.span(n, "ts.partials", ts.partials, rows.B * rows.nwg * 8 * 2 * F32)

    def ffno_mlp_t(self, n, rows, a, xd, w1, w2, yd, ts, stream):
        C = rows.C
        self.asm(n, a, rows)
        self.act(n, "xd", xd, rows)
        self.weight(n, "w1", w1, 2 * C, C)
        self.weight(n, "w2", w2, C, 2 * C)
        self.act(n, "yd", yd, rows)
        self.tan_sums(n, ts, rows)

    def ffno_fuse_t(self, n, rows, h, hd, st, vt, gates, bg1, bg2, f1, f2, fpd, ts, stream):
        C = rows.C
        KP = round_up(C, 32)
        for what, p in (("h", h), ("hd", hd), ("fpd", fpd)):
            self.act(n, what, p, rows)
        for tag, e in (("s", st), ("v", vt)):
            assert e.act_gelu == 1 and not e.base_dot and not e.scale, f"{n}: {tag} assembly"
            self.act(n, f"{tag}.t", e.t, rows)
            self.act(n, f"{tag}.t_dot", e.t_dot, rows)
            for what, p in (("aff_a", e.aff_a), ("aff_b", e.aff_b), ("tan_b", e.tan_b), ("tan_c", e.tan_c)):
                self.span(n, f"{tag}.{what}", p, rows.B * C * F32)
        self.span(n, "gates", gates, rows.B * rows.voxels * 2 * C * F32)
        self.span(n, "bg1", bg1, 2 * C * (3 * KP + E.WPAD) * BF)
        self.weight(n, "bg2", bg2, 2 * C, C)
        self.span(n, "f1", f1, 2 * 2 * C * (2 * KP + E.WPAD) * BF)
        self.weight(n, "f2", f2, C, 2 * C)
        self.tan_sums(n, ts, rows)

    def ffno_head_t(self, n, rows, a, xd, p1, p2, bias, x_in, xd_in, desc, yd, stream):
        d = desc._obj
        n_head = sum(d.levels[s] + 1 for s in range(d.n_species))
        n_out = sum(d.levels[s] for s in range(d.n_species))
        nth = (n_head + 15) // 16
        C = rows.C
        self.asm(n, a, rows)
        self.act(n, "xd", xd, rows)
        self.weight(n, "p1", p1, 2 * C, C)
        self.span(n, "p2", p2, 2 * nth * 16 * (2 * C + E.WPAD) * BF)
        self.span(n, "p2_bias", bias, nth * 16 * F32)
        self.span(n, "x_in", x_in, rows.B * 6 * rows.voxels * F32)
        self.span(n, "xd_in", xd_in, rows.B * 6 * rows.voxels * F32)
        self.span(n, "yd", yd, rows.B * n_out * rows.voxels * F32)

    def _head_desc_counts(self, desc):
        d = desc._obj
        return sum(d.levels[s] + 1 for s in range(d.n_species)), sum(d.levels[s] for s in range(d.n_species))

    def ffno_head_output(self, n, B, voxels, head, x_in, desc, y, stream):
        n_head, n_out = self._head_desc_counts(desc)
        self.span(n, "head", head, B * n_head * voxels * F32)
        self.span(n, "x_in", x_in, B * 6 * voxels * F32)
        self.span(n, "y", y, B * n_out * voxels * F32)

    def ffno_head_output_bwd(self, n, B, voxels, head, x_in, desc, ybar, head_bar, x_bar, stream):
        n_head, n_out = self._head_desc_counts(desc)
        self.span(n, "head", head, B * n_head * voxels * F32)
        self.span(n, "x_in", x_in, B * 6 * voxels * F32)
        self.span(n, "ybar", ybar, B * n_out * voxels * F32)
        self.span(n, "head_bar", head_bar, B * n_head * voxels * F32)
        self.span(n, "x_bar", x_bar, B * 6 * voxels * F32)


def dry_engine(model):
    """An Engine over `model` (kept on the host) whose tensors live on the meta device and whose launches go to a
    Checker instead of the library. Returns (engine, checker, restore)."""
    reg = Registry()
    chk = Checker(reg)
    eng = E.Engine(model)
    meta = torch.device("meta")
    eng.device = meta
    eng.weights = E.PackedWeights(model, meta)
    eng.weights.adjoint()
    eng.sync = lambda dev: None

    def prepare(self, x, z, dx, dy, plan):
        B = x.shape[0]
        return meta, plan, x, z.reshape(B, -1)

    saved = (E._p, _lib.call, E.Engine._stream, E.Engine._prepare)
    E._p = lambda t: ctypes.c_void_p(reg.address(t))
    _lib.call = chk
    E.Engine._stream = staticmethod(lambda: ctypes.c_void_p(0))
    E.Engine._prepare = prepare

    def restore():
        E._p, _lib.call, E.Engine._stream, E.Engine._prepare = saved
    return eng, chk, restore


def check_shape(model, B, D, H, W, stride=1, want_z=True, column=None, two_pass=True, percol=False, fuse_chunk=None):
    """Run forward, the saving forward, the adjoint and the tangent (JVP) pass of `model` at [B, D, H, W] through the
    checker (with the two-pass fusion adjoint where the stride allows it); returns the number of launches checked."""
    eng, chk, restore = dry_engine(model)
    eng.two_pass_fuse = two_pass
    eng.fuse_chunk = fuse_chunk
    try:
        col = np.linspace(14.0, -1.0, D) if column is None else np.asarray(column, np.float64)
        zs = col - col.min()
        span = float(col.max() - col.min())
        meta = torch.device("meta")
        if percol:              # every column its own heights (host float64, as the engine builds them)
            n = B * H * W
            jitter = 0.01 * torch.rand(n, D, generator=torch.Generator().manual_seed(D), dtype=torch.float64)
            cols = torch.from_numpy(zs)[None] + jitter
            cols = cols - cols.amin(dim=1, keepdim=True)
            plan = E.Plan(eng, B, D, H, W, None, span, 48e3, 48e3, columns=cols)
        else:
            plan = E.Plan(eng, B, D, H, W, zs, span, 48e3, 48e3)
        x = torch.empty(B, 6, D, H, W, device=meta)
        z = torch.empty(B, D, H, W, device=meta)
        eng.forward(x, z, 48e3, 48e3, plan=plan)
        y, saved = eng.forward_saved(x, z, 48e3, 48e3, plan=plan, stride=stride)
        ybar = torch.empty_like(y)
        eng.vjp(x, z, 48e3, 48e3, ybar, saved, want_z=want_z, plan=plan)
        eng.jvp(x, z, 48e3, 48e3, torch.empty_like(x), plan=plan)
        head = torch.empty(B, model.out_channels + len(model.species_levels), D, H, W, device=meta)
        eng.head_output(head, x)
        eng.head_output_bwd(head, x, ybar)
        # the allocation-free post + GroupNorm entry point, on the first block's post-network weights
        rows, C = plan.rows, plan.C
        blk = eng.weights.blocks[0]
        u, out = torch.empty(B * plan.voxels, C, device=meta), torch.empty(B * plan.voxels, C, device=meta)
        gn = blk["norm_spec"]
        E._lib.call("ffno_post_gn_noalloc", rows, E._p(u), E._p(blk["spec.p1"]), E._p(blk["spec.p2"]), E._p(out),
                    E._p(plan.depth_w), E._p(gn[0]), E._p(gn[1]), E._p(torch.empty(B, 8, 3, device=meta)),
                    E._p(torch.empty(B, C, device=meta)), E._p(torch.empty(B, C, device=meta)), E.Engine._stream())
    finally:
        restore()
    return len(chk.calls)
