"""The HIP engine of FFNO3D: packs a module's weights for the kernels, builds per-request plans (transform matrices,
per-mode gates and low-rank tables, depth weights) and runs the forward schedule through libffno_hip.so.

Activations live channel-last ([B, D, H, W, C] "voxel rows") in float32 HBM; every pointwise layer runs as bf16 MFMA
with float32 accumulation, every transform as a split-bf16 GEMM. The reference forward this replaces is
models/ffno_model.py:800-832 (eval mode, no dropout).
"""
import collections
import ctypes
import hashlib
import math
import threading
import weakref

import numpy as np
import torch
import torch.nn.functional as F

from . import _lib
from . import transforms as T
from ._lib import FfnoAssemble, FfnoHeadDesc, FfnoMix, FfnoRows, FfnoStatsOut, FfnoTwGemm

SUPPORTED_WIDTHS = (16, 32, 64, 96, 128)
GROUPS = 8
ROW_TILES_PER_WG = 24         # 16-voxel tiles per workgroup of the voxel-row kernels, at least (small requests)
ROW_WGS_PER_CU = 3            # workgroups per CU and batch item, at most (large requests: three rounds of staging)
WPAD = 16                    # zero bf16 after every packed weight row (the kernels' row stride is K_pad + WPAD)
TABLE_CACHE_BYTES = 12 << 30  # a branch's per-mode tables of all blocks stay cached when at most this large (a
#                               256 x 256 tile's horizontal set, 8 blocks at width 64, is 9.4 GiB)
MODE_CACHE_BYTES = 24 << 30  # all cached per-mode tables of an engine (least recently used sets are dropped)
H_CHUNK_BYTES = 2 << 30      # horizontal half-spectrum scratch per buffer: (b, d) planes are transformed in chunks
DENSE_DFT_MAX = 128          # horizontal DFT matrices up to this many rows and columns run as one LDS-staged GEMM;
#                              larger ones run factorised (two stages) when the length splits into factors <= 64
WORKING_UNITS = 9            # activation-sized buffers of the adjoint besides the saved block inputs
GATE_UNITS = 2               # the stored branch gates of the two-pass fusion adjoint
FUSE_CHUNK_BYTES = 2 << 30   # gates of one chunk of voxels when a whole block's do not fit (one sample)

# MULTI3D constants in cgs (models/ffno_model.py:111-113)
_EE, _HH, _EM, _BK, _PI = 1.602189e-12, 6.626176e-27, 9.109534e-28, 1.380662e-16, 3.14159265359
_LN_CCON = math.log(0.5 * (_HH / math.sqrt(2.0 * _PI * _EM) / math.sqrt(_BK)) ** 3)


def _settle(device):
    """Wait for the current stream's work: tables and weights built once are then read from other workers' streams
    (a HIP stream does not wait for another's writes)."""
    if torch.device(device).type == "cuda":
        torch.cuda.current_stream(device).synchronize()


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


_K_BLOCKS = {}        # id of a split matrix's first plane -> (weak reference to it, {row block: column blocks})


def _k_cache(hi):
    c = _K_BLOCKS.get(id(hi))
    if c is None or c[0]() is not hi:
        c = (weakref.ref(hi), {})
        _K_BLOCKS[id(hi)] = c
        weakref.finalize(hi, _K_BLOCKS.pop, id(hi), None)
    return c[1]


def _k_blocks(mat, m0):
    """[(k0, split matrix)] of rows m0 .. m0 + DENSE_DFT_MAX of a split matrix in column blocks of at most
    DENSE_DFT_MAX (row blocks are views of the planes; column blocks are copies, made once per matrix)."""
    hi, lo, m, k, _, kp = mat[:6]
    planes = (hi, lo) + tuple(mat[6:])
    rows = min(DENSE_DFT_MAX, m - m0)
    rp = T.round_up(rows, 16)
    if kp <= DENSE_DFT_MAX:
        sub = tuple(pl[m0:m0 + rp] for pl in planes)
        return [(0, (sub[0], sub[1], rows, k, rp, kp) + sub[2:])]
    cache = _k_cache(hi)
    if m0 not in cache:
        out = []
        for k0 in range(0, k, DENSE_DFT_MAX):
            kb = min(DENSE_DFT_MAX, k - k0)
            kbp = T.round_up(kb, 32)
            sub = tuple(torch.nn.functional.pad(pl[m0:m0 + rp, k0:k0 + kb], (0, kbp - kb)).contiguous()
                        for pl in planes)
            out.append((k0, (sub[0], sub[1], rows, kb, rp, kbp) + sub[2:]))
        cache[m0] = out
    return cache[m0]


def horizontal_factors(H, W):
    """(factors of H or None, factors of W or None): the horizontal DFTs that run factorised."""
    wf = W // 2 + 1
    fh = T.dft_factors(H) if 2 * H > DENSE_DFT_MAX else None
    fw = T.dft_factors(W) if max(T.round_up(2 * wf, 16), T.round_up(W, 32)) > DENSE_DFT_MAX else None
    return fh, fw


def zgrad_tables(bands, v_dz):
    """The tables of ffno_zgrad_columns from one column's Filon bands {delta: G_delta [D, 2M]} and synthesis
    derivative V_dz [D, 2M] (host float64): [4][R_pad][D_pad] with [delta + 1][r][d'] = G_delta[d' - delta, r] (zero
    where d' - delta leaves the column) and [3][r][d] = V_dz[d, r]; R_pad, D_pad = 2M, D rounded up to 16."""
    D, R = v_dz.shape
    rp, dp = T.round_up(R, 16), T.round_up(D, 16)
    tab = np.zeros((4, rp, dp))
    for k, delta in enumerate((-1, 0, 1)):
        lo, hi = max(0, delta), min(D, D + delta)          # levels d' whose d' - delta lies in the column
        tab[k, :R, lo:hi] = bands[delta][lo - delta:hi - delta, :].T
    tab[3, :R, :D] = v_dz.T
    return tab


def horizontal_planes(H, W, C):
    """(floats of one (b, d) plane's half spectrum, floats per plane of the second horizontal scratch buffer: the
    half spectrum, or with W factorised the larger of that and its full-length complex intermediate)."""
    plane = 2 * H * (W // 2 + 1) * C
    return plane, (max(plane, 2 * H * W * C) if horizontal_factors(H, W)[1] is not None else plane)


def pack_mlp_weight(w, segments=None, n_pad=None, parts=2):
    """Split-bf16 [.., parts, N_pad, K_pad + 8] for the MFMA kernels from a float weight [.., N, K] (PyTorch
    [out, in]): each input segment padded to a multiple of 32, k permuted within every 32-block as the kernels' lane
    layout wants (k = 32s + 16h + 4q + j is stored at 32s + 8q + 4h + j), N padded to a multiple of 16, and every
    row followed by 8 zeros (the kernels' row stride K_pad + 8 keeps the 16 rows of an LDS fragment read on distinct
    banks). Planes: hi = bf16(w),
    lo = bf16(w - hi), and with parts=3 a third, bf16(w - hi - lo), for the forward's three-way products (the first
    two planes are then still the two-way split the adjoint kernels read). A float32 weight is split in float32:
    each remainder (w - hi, w - hi - mid) is exact in float32, so the planes equal those of a float64 split."""
    w = w.double() if w.dtype == torch.float64 and w.device.type == "cpu" else w.float()
    lead = w.shape[:-2]
    N, K = w.shape[-2:]
    segs = segments or [K]
    assert sum(segs) == K
    pieces, off = [], 0
    for s in segs:
        piece = w[..., off:off + s]
        off += s
        pieces.append(F.pad(piece, (0, T.round_up(s, 32) - s)))
    wp = torch.cat(pieces, -1)
    npad = n_pad or T.round_up(N, 16)
    wp = F.pad(wp, (0, 0, 0, npad - N))
    kp = wp.shape[-1]
    wp = wp.reshape(*lead, npad, kp // 32, 2, 4, 4).transpose(-3, -2).reshape(*lead, npad, kp)
    wp = F.pad(wp, (0, WPAD))
    planes = []
    rest = wp
    for _ in range(parts):
        plane = rest.to(torch.bfloat16)
        planes.append(plane)
        rest = rest - plane.to(rest.dtype)
    return torch.stack(planes, -3).contiguous().view(torch.int16)


def _conv_w(p):
    return p.detach().reshape(p.shape[0], -1)


def _gelu(x):
    return 0.5 * x * (1.0 + torch.erf(x * (1.0 / math.sqrt(2.0))))


def _freq_mlp(prefix, sd, feats):
    h = _gelu(F.linear(feats, sd[prefix + "freq_mlp.0.weight"], sd[prefix + "freq_mlp.0.bias"]))
    h = _gelu(F.linear(h, sd[prefix + "freq_mlp.2.weight"], sd[prefix + "freq_mlp.2.bias"]))
    return F.linear(h, sd[prefix + "freq_mlp.4.weight"], sd[prefix + "freq_mlp.4.bias"])


def _lowrank_hidden(prefix, sd, feats):
    h = _gelu(F.linear(feats, sd[prefix + "freq_lowrank.0.weight"], sd[prefix + "freq_lowrank.0.bias"]))
    return _gelu(F.linear(h, sd[prefix + "freq_lowrank.2.weight"], sd[prefix + "freq_lowrank.2.bias"]))


def _freq_lowrank(prefix, sd, feats, C, r):
    h = _lowrank_hidden(prefix, sd, feats)
    u = F.linear(h, sd[prefix + "freq_u.weight"], sd[prefix + "freq_u.bias"]).reshape(-1, C, r)
    v = F.linear(h, sd[prefix + "freq_v.weight"], sd[prefix + "freq_v.bias"]).reshape(-1, r, C)
    return u, v


def lowrank_tables_torch(prefix, sd, feats, C, r, adjoint):
    """The packed low-rank tables (v, u) of every mode by torch ops (host-side plans; on the GPU the check of
    ffno_lowrank_tables): u = freq_u(h) [C x r], v = freq_v(h) [r x C] through pack_mlp_weight."""
    u, v = _freq_lowrank(prefix, sd, feats, C, r)
    if adjoint:
        return pack_mlp_weight(u.transpose(-1, -2), n_pad=16), pack_mlp_weight(v.transpose(-1, -2))
    return pack_mlp_weight(v, n_pad=16, parts=3), pack_mlp_weight(u, parts=3)


def _lowrank_tables(prefix, sd, feats, C, r, adjoint):
    """The packed low-rank tables (v, u) of every mode: the hidden features h(k) of the frequency low-rank network
    (models/ffno_model.py:404-429) by torch, then u = freq_u(h), v = freq_v(h) and their split-bf16 packing by
    ffno_lowrank_tables (forward: v [r x C] and u [C x r], three planes; adjoint: u^T and v^T, two planes), the same
    layouts pack_mlp_weight writes. Plans built in host memory (tests) take the torch form."""
    if feats.device.type == "cpu":
        return lowrank_tables_torch(prefix, sd, feats, C, r, adjoint)
    h = _lowrank_hidden(prefix, sd, feats).float().contiguous()
    modes, hidden = h.shape
    parts = 2 if adjoint else 3
    kc = T.round_up(C, 32)
    v = torch.zeros(modes, parts, 16, kc + WPAD, dtype=torch.int16, device=h.device)
    u = torch.zeros(modes, parts, C, 32 + WPAD, dtype=torch.int16, device=h.device)
    wu = sd[prefix + "freq_u.weight"].float().t().contiguous()          # [hidden, C r]
    wv = sd[prefix + "freq_v.weight"].float().t().contiguous()          # [hidden, r C]
    bu, bv = sd[prefix + "freq_u.bias"].float().contiguous(), sd[prefix + "freq_v.bias"].float().contiguous()
    _lib.call("ffno_lowrank_tables", modes, C, r, hidden, _p(h), _p(wu), _p(bu), _p(wv), _p(bv), parts,
              1 if adjoint else 0, _p(v), _p(u), Engine._stream())
    return v, u


class PackedWeights:
    """The module's pointwise weights in kernel form, rebuilt whenever a parameter changes."""

    def __init__(self, module, device):
        C = module.width
        # float32 on the device whatever the module's dtype: no float64 arithmetic runs on the GPU
        sd = {k: (v.detach().float() if v.is_floating_point() else v.detach()).to(device)
              for k, v in module.state_dict().items()}
        self.sd = sd
        self.C = C
        self._adj = None
        self.wc_real = []
        self.lift = _conv_w(sd["lift.0.weight"]).float().contiguous()          # [C, 6]
        self.lift_gn = (sd["lift.1.weight"].float().contiguous(), sd["lift.1.bias"].float().contiguous())
        self.blocks = []
        for i in range(module.n_layers):
            p = f"blocks.{i}."
            b = {}
            for br in ("spec", "vertical"):
                b[br + ".g1"] = pack_mlp_weight(_conv_w(sd[p + br + ".input_gate.0.weight"]), parts=3)
                b[br + ".g2"] = pack_mlp_weight(_conv_w(sd[p + br + ".input_gate.2.weight"]), parts=3)
                b[br + ".p1"] = pack_mlp_weight(_conv_w(sd[p + br + ".post.0.weight"]), parts=3)
                b[br + ".p2"] = pack_mlp_weight(_conv_w(sd[p + br + ".post.2.weight"]), parts=3)
                wr, wi = sd[p + br + ".weight_real"].float(), sd[p + br + ".weight_imag"].float()   # [in, out]
                # [Ar; Ai] = [[Wr^T, -Wi^T], [Wi^T, Wr^T]] [Yr; Yi]: the complex product Y (Wr + i Wi) over channels
                wc = torch.cat([torch.cat([wr.t(), -wi.t()], 1), torch.cat([wi.t(), wr.t()], 1)], 0)
                b[br + ".wc"] = pack_mlp_weight(wc, segments=[C, C], parts=3)
                if br == "spec":
                    self.wc_real.append({})
                self.wc_real[-1][br] = wc
            b["z_proj"] = sd[p + "vertical.z_proj.weight"].reshape(C).float().contiguous()
            b["bg1"] = pack_mlp_weight(_conv_w(sd[p + "branch_gate.0.weight"]), segments=[C, C, C])
            b["bg2"] = pack_mlp_weight(_conv_w(sd[p + "branch_gate.2.weight"]))
            b["bg2_bias"] = sd[p + "branch_gate.2.bias"].float().contiguous()
            b["f1"] = pack_mlp_weight(_conv_w(sd[p + "fuse.0.weight"]), segments=[C, C])
            b["f2"] = pack_mlp_weight(_conv_w(sd[p + "fuse.2.weight"]))
            b["pw1"] = pack_mlp_weight(_conv_w(sd[p + "pw.0.weight"]), parts=3)
            b["pw2"] = pack_mlp_weight(_conv_w(sd[p + "pw.2.weight"]), parts=3)
            b["m1"] = pack_mlp_weight(_conv_w(sd[p + "mlp.net.0.weight"]), parts=3)
            b["m2"] = pack_mlp_weight(_conv_w(sd[p + "mlp.net.3.weight"]), parts=3)
            for n in ("norm_spec", "norm_vert", "norm_fuse", "norm_pw", "norm_mlp"):
                b[n] = (sd[p + n + ".weight"].float().contiguous(), sd[p + n + ".bias"].float().contiguous())
            for n in ("res_fused", "res_pw", "res_mlp"):
                b[n] = sd[p + n].float().contiguous()
            self.blocks.append(b)
        self.p1 = pack_mlp_weight(_conv_w(sd["proj1.weight"]), parts=3)
        self.p2 = pack_mlp_weight(_conv_w(sd["proj2.weight"]), parts=3)
        nh = sd["proj2.weight"].shape[0]
        self.p2_bias = F.pad(sd["proj2.bias"].float(), (0, T.round_up(nh, 16) - nh)).contiguous()
        self.head = self._head_desc(module, module.state_dict())     # host scalars, read where the module lives

    def adjoint(self):
        """Transposed weights for the adjoint kernels, packed on first use."""
        if getattr(self, "_adj", None) is None:
            sd = self.sd
            C = self.C
            blocks = []
            for i, b in enumerate(self.blocks):
                p = f"blocks.{i}."
                t = {}
                for br in ("spec", "vertical"):
                    t[br + ".g2t"] = pack_mlp_weight(_conv_w(sd[p + br + ".input_gate.2.weight"]).t())
                    t[br + ".g1t"] = pack_mlp_weight(_conv_w(sd[p + br + ".input_gate.0.weight"]).t())
                    t[br + ".p2t"] = pack_mlp_weight(_conv_w(sd[p + br + ".post.2.weight"]).t())
                    t[br + ".p1t"] = pack_mlp_weight(_conv_w(sd[p + br + ".post.0.weight"]).t())
                    t[br + ".wc"] = pack_mlp_weight(self.wc_real[i][br].t(), segments=[C, C])
                t["bg2t"] = pack_mlp_weight(_conv_w(sd[p + "branch_gate.2.weight"]).t())
                t["bg1t"] = pack_mlp_weight(_conv_w(sd[p + "branch_gate.0.weight"]).t())
                t["f2t"] = pack_mlp_weight(_conv_w(sd[p + "fuse.2.weight"]).t())
                t["f1t"] = pack_mlp_weight(_conv_w(sd[p + "fuse.0.weight"]).t())
                t["pw2t"] = pack_mlp_weight(_conv_w(sd[p + "pw.2.weight"]).t())
                t["pw1t"] = pack_mlp_weight(_conv_w(sd[p + "pw.0.weight"]).t())
                t["m2t"] = pack_mlp_weight(_conv_w(sd[p + "mlp.net.3.weight"]).t())
                t["m1t"] = pack_mlp_weight(_conv_w(sd[p + "mlp.net.0.weight"]).t())
                blocks.append(t)
            self._adj = dict(blocks=blocks, p2t=pack_mlp_weight(_conv_w(sd["proj2.weight"]).t()),
                             p1t=pack_mlp_weight(_conv_w(sd["proj1.weight"]).t()))
        return self._adj

    @staticmethod
    def _head_desc(module, sd):
        d = FfnoHeadDesc()
        levels = module.species_levels
        d.n_species = len(levels)
        for i, n in enumerate(levels):
            d.levels[i] = n
            d.log10_mass[i] = float(sd["log10_particle_mass"][i])
        log_g, ev, ndxi, dion = module.lte_tables
        for i in range(len(log_g)):
            d.log_g[i], d.ev[i], d.ndxi[i], d.dion[i] = log_g[i], ev[i], ndxi[i], dion[i]
        for c in range(6):
            d.mean_x[c] = float(sd["mean_X"][c])
            d.std_x[c] = float(sd["std_X"][c])
        for c in range(sum(levels)):
            d.mean_y[c] = float(sd["mean_Y"][c])
            d.std_y[c] = float(sd["std_Y"][c])
        d.t_channel, d.ne_channel, d.rho_channel = module.temperature_channel, module.ne_channel, module.rho_channel
        d.total_bound = module.total_bound
        d.ln_ccon = _LN_CCON
        d.ln_k = math.log(_BK)
        d.e_over_k = _EE / _BK
        d.dxi_coef = 4.98e-4 * 5040.0
        return d


class ModeTables:
    """Per-mode tables of one branch for every block (models/ffno_model.py:394-429, :285-296): the gate g(k) and the
    rank-r factors V(k), U(k) packed for the mix kernel (three planes: the forward's three-way products), and for the
    adjoint, built on first use, conj(g) and the low-rank roles swapped (V' = U^T, U' = V^T, two planes). They depend
    on the mode features only -- (H, W, dx, dy) for the horizontal branch, the wavenumbers for the vertical one -- so
    every plan with the same modes shares one set. A set whose blocks would exceed TABLE_CACHE_BYTES is rebuilt per
    block on use instead of kept."""

    def __init__(self, sd, branch, feats, C, rank, n_layers):
        self.sd, self.branch, self.feats, self.C, self.rank, self.n_layers = sd, branch, feats, C, rank, n_layers
        self._lock = threading.Lock()
        self._fwd, self._adj = {}, {}
        first = self._build(0, False)
        per_mode = sum(t.numel() * t.element_size() for t in first.values() if t is not None)
        per_mode_adj = per_mode * 2 // 3 if rank else per_mode       # two planes instead of three
        self.block_bytes = per_mode + per_mode_adj
        self.cache = self.block_bytes * n_layers <= TABLE_CACHE_BYTES
        _settle(feats.device)                   # feats and block 0's tables are read from other streams later
        if self.cache:
            self._fwd[0] = first

    def cached_bytes(self):
        with self._lock:            # worker threads insert into the stores under this lock
            return sum(t.numel() * t.element_size() for d in (self._fwd, self._adj) for e in d.values()
                       for t in e.values() if t is not None)

    def _build(self, i, adjoint):
        prefix = f"blocks.{i}.{self.branch}."
        gate = _freq_mlp(prefix, self.sd, self.feats).float()
        if adjoint:
            gate[:, 1] = -gate[:, 1]
        e = dict(gate=gate.contiguous(), v=None, u=None)
        if self.rank:
            # [modes, parts, 16, KS*32 + 8] and [modes, parts, C, 32 + 8]
            e["v"], e["u"] = _lowrank_tables(prefix, self.sd, self.feats, self.C, self.rank, adjoint)
        return e

    def get(self, i, adjoint=False):
        store = self._adj if adjoint else self._fwd
        with self._lock:
            e = store.get(i)
            if e is None:
                e = self._build(i, adjoint)
                if self.cache:
                    _settle(self.feats.device)          # kept tables are read by other workers' streams
                    store[i] = e
        return e


class Plan:
    """Everything that depends on the request's shape, heights and spacings but not on its data: transform
    matrices (split-bf16), per-mode gates and low-rank tables of every block (cached when small, else rebuilt per
    block on use), depth weights, the chunking of the horizontal transforms."""

    def __init__(self, engine, B, D, H, W, column, span, dx, dy, columns=None):
        m = engine.module
        dev = engine.device
        C, r = m.width, m.freq_mix_rank
        self.B, self.D, self.H, self.W, self.C, self.rank = B, D, H, W, C, r
        self.n_layers = m.n_layers
        self.wf = W // 2 + 1
        self.voxels = D * H * W
        # columns: [B*H*W, D] float64 shifted heights of every column, on the host (the reference's per-column
        # path); else the one shared column. Every matrix below is built on the host in float64 and uploaded split.
        self.percol = columns is not None
        self.length, k, factor = T.vertical_box(span, m.z_reference_length, m.z_min_wavelength,
                                                int(math.floor(m.z_reference_length / m.z_min_wavelength)) + 1)
        self.modes = len(k)
        hm = T.horizontal_matrices(H, W)
        named = [("w_fwd", hm["w_fwd"]), ("h_fwd", hm["h_fwd"]), ("h_inv", hm["h_inv"]), ("w_inv", hm["w_inv"])]
        if self.percol:
            # depth weights (models/ffno_model.py:28-34, :67-74) of every column, in float32 as the reference has
            # them
            zc = columns.float()
            if D == 1:
                tw = torch.ones_like(zc)
            else:
                half = 0.5 * (zc[:, 1:] - zc[:, :-1]).abs()
                tw = F.pad(half, (0, 1)) + F.pad(half, (1, 0))
            w = tw / tw.mean(dim=1, keepdim=True).clamp_min(1e-12)                 # [B*HW, D]
            self.depth_w = w.reshape(B, H * W, D).transpose(1, 2).reshape(B, -1).contiguous().to(dev)
            v_fwd, v_inv = T.vertical_matrices_t(columns, k, factor, self.length)
            self.colmats = dict(v_fwd=T.split_bf16_columns(v_fwd, dev, parts=3),
                                v_inv=T.split_bf16_columns(v_inv, dev, parts=3))
            del v_fwd, v_inv
            self._columns = columns
            self._col_adj = None
            column = None
        else:
            column = np.asarray(column, np.float64)
            # depth weights of the shared column, in float32 as the reference has them, broadcast to every voxel
            zc = torch.from_numpy(column).float()
            if D == 1:
                tw = torch.ones(1)
            else:
                half = 0.5 * (zc[1:] - zc[:-1]).abs()
                tw = F.pad(half, (0, 1)) + F.pad(half, (1, 0))
            w = tw / tw.mean().clamp_min(1e-12)
            self.depth_w = w.to(dev)[None, :, None].expand(B, D, H * W).reshape(B, -1).contiguous()
            v_fwd, v_inv = T.vertical_matrices(column, k, factor, self.length)
            named += [("v_fwd", v_fwd), ("v_inv", v_inv)]
        self.tw = {name: T.split_bf16(a, dev, parts=3) for name, a in named}
        # along H beyond 64 rows (a [2H x 2H] matrix no longer fits the LDS-staged kernel, and its cost grows as
        # H^2 per column) the complex DFTs and their transposes run factorised: two small GEMMs, the first with
        # the twiddle factors applied to its output (forward ones three-way, adjoint ones two-way)
        self.hfac = self.wfac = None
        f, fw = horizontal_factors(H, W)
        if fw is not None:
            # along W: real rows -> half spectrum (forward; with the irfft weights and 1/W the adjoint of the
            # inverse) and back (inverse with the irfft weights and 1/W; unweighted the adjoint of the forward)
            self.wfac = dict(n=fw)
            for name, real_in, scale, weighted, parts in (("w_fwd", True, 1.0, False, 3), ("w_inv", False, 1.0 / W, True, 3),
                                                          ("w_inv_T", True, 1.0 / W, True, 2),
                                                          ("w_fwd_T", False, 1.0, False, 2)):
                sp = lambda a: T.split_bf16(a, dev, parts=parts)      # noqa: E731
                if real_in:
                    a, tw, runs = T.rdft_stages(W, fw[0], fw[1], scale, weighted)
                    self.wfac[name] = (True, sp(a), torch.from_numpy(tw).to(dev),
                                       [(st, cnt, sp(b)) for st, cnt, b in runs])
                else:
                    runs, tw, b = T.irdft_stages(W, fw[0], fw[1], scale, weighted)
                    self.wfac[name] = (False, [(st, cnt, sp(a)) for st, cnt, a in runs],
                                       torch.from_numpy(tw).to(dev), sp(b))
        if f is not None:
            self.hfac = dict(n=f)
            for name, sign, scale, parts in (("h_fwd", -1, 1.0, 3), ("h_inv", 1, 1.0 / H, 3),
                                             ("h_inv_T", -1, 1.0 / H, 2), ("h_fwd_T", 1, 1.0, 2)):
                a, tw, b = T.dft_stages(H, f[0], f[1], sign, scale)
                self.hfac[name] = (T.split_bf16(a, dev, parts=parts), torch.from_numpy(tw).to(dev),
                                   T.split_bf16(b, dev, parts=parts))
        # the adjoint chain runs the same GEMMs with the transposed matrices in reverse order
        self._mats = dict(named)
        self._tw_adj = None
        self._column, self._k, self._factor = column, k, factor
        self._height = None
        self._lock = threading.Lock()
        # horizontal transforms run over chunks of (b, d) planes so that their half-spectrum scratch stays bounded
        self.plane, self.plane2 = horizontal_planes(H, W, C)
        self.h_chunk = max(1, min(B * D, H_CHUNK_BYTES // (4 * self.plane2)))
        # per-mode tables, shared with every plan of the same modes
        self._mode_tables = dict(spec=engine.mode_tables("spec", H, W, float(dx), float(dy)),
                                 vertical=engine.mode_tables("vertical", self.length, tuple(k.tolist())))
        nt = (self.voxels + 15) // 16
        # workgroups per batch item of the voxel-row kernels: enough tiles per workgroup that its waves all have work
        # (small requests), at most ROW_WGS_PER_CU per CU across the batch (large ones: each extra round restages the
        # workgroup's weights into LDS)
        cus = torch.cuda.get_device_properties(dev).multi_processor_count if dev.type == "cuda" else 256
        self.nwg = max(1, min(-(-nt // ROW_TILES_PER_WG), max(1, ROW_WGS_PER_CU * cus // B)))
        self.rows = FfnoRows(B, self.voxels, C, self.nwg)
        self.device = dev
        _settle(dev)                            # the plan's matrices are read from every worker's stream

    def device_bytes(self):
        """Device memory the plan's own matrices and tables hold (per-mode table sets not included)."""
        seen, total = set(), 0
        stack = [v for k, v in vars(self).items() if k != "_mode_tables"]
        while stack:
            v = stack.pop()
            if torch.is_tensor(v):
                if v.device.type != "cpu" and id(v) not in seen:
                    seen.add(id(v))
                    total += v.numel() * v.element_size()
            elif isinstance(v, dict):
                stack.extend(v.values())
            elif isinstance(v, (list, tuple)):
                stack.extend(v)
        return total

    def tables(self, i, adjoint=False):
        """{"spec": ..., "vertical": ...} per-mode tables of block i (the adjoint's with adjoint=True)."""
        return {br: t.get(i, adjoint) for br, t in self._mode_tables.items()}

    def uncached_table_bytes(self):
        """Per-mode tables that are rebuilt per block on use (not kept): one block's worth is alive at a time."""
        return sum(t.block_bytes for t in self._mode_tables.values() if not t.cache)

    def _column_height_tables(self):
        """height_tables for per-column grids: every column's Filon bands, synthesis derivative and taper terms, the
        median interval and sign per column and the mask of each column's lowest level(s)."""
        with self._lock:
            if self._height is None:
                cols, dev = self._columns, self.device
                bands = T.filon_height_bands_t(cols, self._k)
                v_dz, v0, dfac, i_med, sgn = T.vertical_height_terms_t(cols, self._k, self._factor, self.length)
                lowest = (cols == cols.min(dim=1, keepdim=True).values).float()        # [n, D]
                any_taper = bool((dfac != 0).any())
                self._height = dict(
                    bands={d: T.split_bf16_columns(bands[d], dev) for d in (-1, 0, 1)},
                    v_dz=T.split_bf16_columns(v_dz, dev),
                    v0=T.split_bf16_columns(v0.transpose(1, 2), dev) if any_taper else None,
                    dfac=dfac.float().to(dev), i_med=i_med.to(dev), sign=sgn.float().to(dev),
                    lowest=(lowest / lowest.sum(1, keepdim=True)).to(dev))
                del bands, v_dz, v0
                _settle(self.device)
            return self._height

    def column_adjoint(self):
        """The transposed per-column matrices (two planes): V_inv^T [2M, D] and V_fwd^T [D, 2M] of every column."""
        with self._lock:
            if self._col_adj is None:
                v_fwd, v_inv = T.vertical_matrices_t(self._columns, self._k, self._factor, self.length)
                self._col_adj = dict(v_inv_T=T.split_bf16_columns(v_inv.transpose(1, 2), self.device),
                                     v_fwd_T=T.split_bf16_columns(v_fwd.transpose(1, 2), self.device))
                del v_fwd, v_inv
                _settle(self.device)
            return self._col_adj

    def height_tables(self):
        """Height derivatives of the vertical transforms for the z-gradient: the three bands of the Filon weights'
        Jacobian, the synthesis' derivative, and for the taper (non-zero only for modes between 2 pi / (3 dz) and
        pi / dz) the transposed unscaled synthesis and the per-mode factor * d taper / d dz."""
        if self.percol:
            return self._column_height_tables()
        with self._lock:
            if self._height is None:
                bands = T.filon_height_bands(self._column, self._k)
                v_dz, v0, dfac, i_med, sgn = T.vertical_height_terms(self._column, self._k, self._factor, self.length)
                col = np.asarray(self._column)
                ties = np.flatnonzero(col == col.min())
                self._height = dict(
                    bands={d: T.split_bf16(bands[d], self.device) for d in (-1, 0, 1)},
                    v_dz=T.split_bf16(v_dz, self.device),
                    columns=self._zgrad_tables(bands, v_dz),
                    v0=T.split_bf16(np.ascontiguousarray(v0.T), self.device) if np.any(dfac != 0) else None,
                    dfac=torch.from_numpy(dfac).float().to(self.device), i_med=i_med, sign=sgn, argmin=ties)
                _settle(self.device)
            return self._height

    def _zgrad_tables(self, bands, v_dz):
        """zgrad_tables on the device in float32: (tables, R_pad, D_pad)."""
        tab = zgrad_tables(bands, v_dz)
        return torch.from_numpy(tab).float().to(self.device), tab.shape[1], tab.shape[2]

    def prepare(self, adjoint=False):
        """Build now what a call would otherwise build on first use (the kept per-mode tables of every block, and for
        the adjoint its transposed matrices and height tables), so that a caller can make a plan ready on one thread
        while another launches work on the device."""
        for t in self._mode_tables.values():
            if t.cache:
                for i in range(self.n_layers):
                    t.get(i, False)
                    if adjoint:
                        t.get(i, True)
        if adjoint:
            self.adjoint_tables()
            self.height_tables()
            if self.percol:
                self.column_adjoint()
        return self

    def adjoint_tables(self):
        """Transposed transform matrices, so that the same GEMM kernel applies the transpose of every forward
        transform."""
        with self._lock:
            if self._tw_adj is None:
                self._tw_adj = {name + "_T": T.split_bf16(np.ascontiguousarray(a.T), self.device)
                                for name, a in self._mats.items()}
                _settle(self.device)
            return self._tw_adj

    # ------------------------------------------------------------------------------------------------ memory
    def unit_bytes(self):
        """One float32 activation tensor [B, D, H, W, C]."""
        return self.B * self.voxels * self.C * 4

    def scratch_bytes(self):
        """Half-spectrum chunks, both vertical spectra, GroupNorm partials, inputs and outputs."""
        v1 = self.B * 2 * self.modes * self.H * self.W * self.C * 4
        return self.h_chunk * (self.plane + self.plane2) * 4 + 2 * v1 + self.B * self.voxels * 4 * 32

    def full_bytes(self, want_z=True):
        """The adjoint keeping every block's intermediates: 11 units per block (input, both branch outputs, five
        GroupNorm inputs, x1, the gates) and the vertical spectrum with want_z, plus the head input and two working
        units."""
        v1 = self.B * 2 * self.modes * self.H * self.W * self.C * 4
        return (self.unit_bytes() * (11 * self.n_layers + 3) + (self.n_layers if want_z else 0) * v1 +
                self.scratch_bytes() + 2 * self.uncached_table_bytes())

    def keep_bytes(self, keep, want_z=True):
        """The stride-1 two-pass adjoint with the last `keep` blocks kept whole: the larger of its own count and the
        L - keep saved inputs, 11 units per kept block (and its vertical spectrum with want_z) and four more."""
        if keep >= self.n_layers:
            return self.full_bytes(want_z)
        v1 = self.B * 2 * self.modes * self.H * self.W * self.C * 4
        units = max(self.n_layers + 1 + WORKING_UNITS + GATE_UNITS, self.n_layers - keep + 11 * keep + 4)
        return (self.unit_bytes() * units + (keep if want_z else 0) * v1 + self.scratch_bytes() +
                2 * self.uncached_table_bytes())

    def vjp_bytes(self, stride, two_pass=False):
        saved = -(-self.n_layers // stride) + (1 if stride == 1 else 0) + (stride - 1)
        units = saved + WORKING_UNITS + (GATE_UNITS if two_pass else 0)
        chunk = 0 if two_pass else min(FUSE_CHUNK_BYTES, GATE_UNITS * self.unit_bytes())
        return self.unit_bytes() * units + chunk + self.scratch_bytes() + 2 * self.uncached_table_bytes()


class Saved:
    """What the forward keeps for the adjoint: the inputs of blocks 0, s, 2s, ... (s = stride) and, with s = 1, the
    head's input. The adjoint recomputes the inputs of the other blocks from the nearest kept one. With full, the
    forward keeps every intermediate the adjoint reads instead (about 11 activation-sized tensors per block), so that
    nothing is recomputed -- the choice when HBM has room for it. In between, keep = k (stride 1) keeps the last k
    blocks whole and the inputs of the others: the adjoint recomputes L - k blocks, in the memory of the stride-1
    adjoint plus about 10 activation-sized tensors per kept block."""

    def __init__(self, stride=1, two_pass=False, full=False, keep=0, n_layers=0):
        self.stride = int(stride)
        self.two_pass = bool(two_pass) or bool(full)   # fusion adjoint in two LDS-staged passes over stored gates
        self.full = bool(full)             # every block's intermediates kept: the adjoint recomputes nothing
        # blocks from full_from on are kept whole (None: none)
        self.full_from = 0 if full else (n_layers - int(keep) if keep and self.stride == 1 else None)
        self.inputs = {}
        self.blocks = {}                   # block -> the buffers and GroupNorm records of its forward, if kept whole
        self.head = None


def _asm(base, t=None, a=None, b=None, scale=None, act=0):
    return FfnoAssemble(_p(base), _p(t), _p(a), _p(b), _p(scale), act)


class Engine:
    """Runs FFNO3D forward and input-VJP on the GPU for one module (ffnoml_amd.model.FFNO3D)."""

    def __init__(self, module):
        if module.width not in SUPPORTED_WIDTHS:
            raise ValueError(f"width {module.width} is not supported (one of {SUPPORTED_WIDTHS})")
        if module.freq_mix_rank > 16:
            raise ValueError("freq_mix_rank above 16 is not supported")
        self.module = module
        self.device = None
        self.weights = None
        self._version = None
        self._plans = {}
        self._building = {}                 # plan key -> event set when the thread building that plan is done
        self._lock = threading.RLock()      # the serving threads share one engine
        self.checkpoint_stride = None       # None: the smallest stride whose adjoint fits the free HBM
        self.two_pass_fuse = None           # None: whenever its gate buffer fits at stride 1
        self.full_save = None               # None: keep every intermediate when that fits; False: never
        self.fuse_chunk = None              # voxels per chunk of the one-sample fusion adjoint without a block's gate
        #                                     buffer (None: FUSE_CHUNK_BYTES of gates; False: the one-pass kernel)
        self.last_stride = None
        self.last_keep = None
        self._modes = collections.OrderedDict()     # per-mode table sets shared by plans, least recently used first

    def vertical_modes(self, span=0.0):
        """Vertical modes of a request whose tallest column spans `span` (Mm): the reference's box rule
        (models/ffno_model.py:248-259) gives more modes when twice the span exceeds the reference length."""
        m = self.module
        return len(T.vertical_box(float(span), m.z_reference_length, m.z_min_wavelength,
                                  int(math.floor(m.z_reference_length / m.z_min_wavelength)) + 1)[1])

    def footprint_bytes(self, B, D, H, W, vjp=False, jvp=False, span=0.0, full=False, keep=0):
        """HBM one request of this shape needs (for admission; the tables of a cached plan not included):
        forward 3 activation-sized buffers, adjoint (L+1) saved inputs and 9 working buffers (full: 11 per block
        and its vertical spectrum, as Plan.full_bytes counts them; keep = k blocks kept whole: the larger of the
        stride-1 count and L - k inputs, 11 per kept block and 4, and k spectra), JVP 16 (one block's intermediates,
        its gates and four tangents), plus the scratch; span (Mm): the tallest column's height range, which sets the
        vertical mode count."""
        C, L, r = self.module.width, self.module.n_layers, self.module.freq_mix_rank
        unit = B * D * H * W * C * 4
        plane, plane2 = horizontal_planes(H, W, C)
        chunk = max(1, min(B * D, H_CHUNK_BYTES // (4 * plane2))) * (plane + plane2) * 4
        modes = self.vertical_modes(span)
        v1 = B * 2 * modes * H * W * C * 4
        if vjp and (full or keep >= L):
            units, spectra = 11 * L + 3, L + 2
        elif vjp and keep:
            units = max((L + 1) + WORKING_UNITS + GATE_UNITS, (L - keep) + 11 * keep + 4)
            spectra = keep + 2
        elif vjp:
            units, spectra = (L + 1) + WORKING_UNITS + GATE_UNITS, 2
        else:
            units, spectra = (16 if jvp else 3), 1
        return unit * units + chunk + spectra * v1 + B * D * H * W * 4 * 32 + (64 << 20)

    def _param_version(self):
        return tuple((p.data_ptr(), p._version) for p in self.module.parameters()) + \
            tuple((b.data_ptr(), b._version) for b in self.module.buffers())

    def sync(self, device):
        with self._lock:
            ver = self._param_version()
            if self.weights is None or ver != self._version or device != self.device:
                self.device = device
                self.weights = PackedWeights(self.module, device)
                self.weights.adjoint()
                self._version = ver
                _settle(device)
                self._plans.clear()
                self._modes.clear()

    def mode_tables(self, branch, *key):
        """The shared ModeTables of a branch for these modes: key (H, W, dx, dy) for "spec", (L, k) for "vertical".
        Sets are dropped least recently used first while all cached ones exceed MODE_CACHE_BYTES (a plan holding a
        dropped set keeps it alive until the plan itself goes)."""
        with self._lock:
            t = self._modes.get((branch,) + key)
            if t is not None:
                self._modes.move_to_end((branch,) + key)
                return t
            m, dev = self.module, self.device
            if branch == "spec":
                H, W, dx, dy = key
                ky = torch.fft.fftfreq(H, d=float(dy), device=dev)
                kx = torch.fft.rfftfreq(W, d=float(dx), device=dev)
                ky, kx = torch.meshgrid(ky, kx, indexing="ij")
                ky, kx = ky * 1e5, kx * 1e5      # models/ffno_model.py:375-392, :443-447 (float32)
                feats = torch.stack([kx, ky, torch.sqrt(kx ** 2 + ky ** 2 + 1e-12), torch.sign(ky)], -1)
                feats = feats.reshape(-1, 4)
            else:
                kt = torch.tensor(key[1], dtype=torch.float32, device=dev)
                qf = kt * m.z_feature_wavelength / (2 * math.pi)       # models/ffno_model.py:231-237
                tf = qf / (1 + qf)
                feats = torch.stack([tf, tf * tf], -1)
            t = ModeTables(self.weights.sd, branch, feats, m.width, m.freq_mix_rank, m.n_layers)
            self._modes[(branch,) + key] = t
            while len(self._modes) > 1 and sum(v.cached_bytes() for v in self._modes.values()) > MODE_CACHE_BYTES:
                self._modes.popitem(last=False)
            return t

    def cached_bytes(self):
        """Device memory held by the engine's caches (the serving admission counts it): the cached plans' matrices
        and tables, and every per-mode table set alive in the engine's LRU or in a cached plan."""
        with self._lock:
            plans = list(self._plans.values())
            sets = {id(v): v for v in self._modes.values()}
        for p in plans:
            for t in p._mode_tables.values():
                sets.setdefault(id(t), t)
        return sum(p.device_bytes() for p in plans) + sum(v.cached_bytes() for v in sets.values())

    def clear_caches(self):
        """Drop every cached plan and per-mode table set (the memory returns to torch's caching allocator once no
        running call holds them)."""
        with self._lock:
            self._plans.clear()
            self._modes.clear()

    def drop_unused_tables(self):
        """Drop the per-mode table sets no cached plan uses, and the plans beyond the most recent few: what the
        serving admission frees when a request waits for memory with nothing in flight."""
        with self._lock:
            while len(self._plans) > 4:
                self._plans.pop(next(iter(self._plans)))
            used = {id(t) for p in self._plans.values() for t in p._mode_tables.values()}
            for key in [k for k, v in self._modes.items() if id(v) not in used]:
                del self._modes[key]

    def plan(self, B, D, H, W, z, dx, dy):
        """The plan for these heights, cached by content: the reference's shared-grid decision
        (models/ffno_model.py:267-269) and its box span (:252) are read on the host here, once per call, and every
        later decision uses them (no host synchronisation inside the launch sequence)."""
        zc = z.reshape(B, D, H * W).float()
        zs = zc - zc.amin(dim=1, keepdim=True)
        flags = torch.stack([(zs == zs[:, :, :1]).all() & (zs == zs[:1]).all(),
                             (zc.amax(dim=1) - zc.amin(dim=1)).max()])
        shared, span = flags.tolist()
        columns = None
        if shared:
            # the shared column, shifted in float64 on the host (the spacings then equal the float64 reference's bit
            # for bit, so a tie between spacings at the taper's median resolves to the same interval)
            z0 = zc[0, :, 0].cpu().double()
            column = (z0 - z0.min()).numpy()
            key = (B, D, H, W, float(dx), float(dy), float(span), column.tobytes())
        else:
            # the reference's per-column path (models/ffno_model.py:277-283): every column's own shifted heights, in
            # float64 on the host
            zd = zc.cpu().double()
            columns = (zd - zd.amin(dim=1, keepdim=True)).transpose(1, 2).reshape(B * H * W, D).contiguous()
            column = None
            key = (B, D, H, W, float(dx), float(dy), float(span), hashlib.sha1(columns.numpy().tobytes()).hexdigest())
        # built outside the engine's lock (a plan takes up to a few hundred ms of host work, while other threads
        # launch calls that take the lock); a second caller of the same key waits for the first one's plan
        while True:
            with self._lock:
                p = self._plans.get(key)
                if p is not None:
                    return p
                building = self._building.get(key)
                if building is None:
                    building = self._building[key] = threading.Event()
                    break
            building.wait()
        try:
            p = Plan(self, B, D, H, W, column, span, dx, dy, columns=columns)
            with self._lock:
                if len(self._plans) > 64:
                    self._plans.pop(next(iter(self._plans)))
                self._plans[key] = p
        finally:
            with self._lock:
                del self._building[key]
            building.set()
        return p

    def choose_checkpoints(self, plan, stride=None, want_z=True):
        """(stride, two_pass, full, keep) of the adjoint. stride 0 asks for every intermediate kept (full).
        Otherwise the configured (or given) stride, else: full when it fits the HBM that is free now (the caching
        allocator's unused reserve included), then stride 1 with the two-pass fusion adjoint and as many of the last
        blocks kept whole as fit, then the smallest stride that fits."""
        stride = self.checkpoint_stride if stride is None else stride
        if stride == 0:
            return 1, True, True, 0
        if stride and self.two_pass_fuse is not None:
            return int(stride), bool(self.two_pass_fuse) and int(stride) == 1, False, 0
        free, _ = torch.cuda.mem_get_info(plan.device)
        free += torch.cuda.memory_reserved(plan.device) - torch.cuda.memory_allocated(plan.device)
        budget = 0.97 * free
        if stride:
            return int(stride), int(stride) == 1 and plan.vjp_bytes(1, True) <= budget, False, 0
        if self.full_save is not False and self.two_pass_fuse is not False and plan.full_bytes(want_z) <= budget:
            return 1, True, True, 0
        if self.two_pass_fuse is not False and plan.vjp_bytes(1, True) <= budget:
            keep = 0
            if self.full_save is not False:
                keep = next((k for k in range(plan.n_layers - 1, 0, -1) if plan.keep_bytes(k, want_z) <= budget), 0)
            return 1, True, False, keep
        for s in range(1, plan.n_layers + 1):
            if plan.vjp_bytes(s) <= budget:
                return s, False, False, 0
        raise MemoryError(f"the adjoint of a [{plan.B}, {plan.D}, {plan.H}, {plan.W}] request needs "
                          f"{plan.vjp_bytes(plan.n_layers) / 2**30:.1f} GiB even with one saved input")

    # ------------------------------------------------------------------------------------------------ launches
    @staticmethod
    def _stream():
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    def _gn(self, plan, partials, gn, dev):
        a = torch.empty(plan.B, plan.C, device=dev)
        b = torch.empty(plan.B, plan.C, device=dev)
        stats = torch.empty(plan.B, GROUPS, 3, device=dev)
        _lib.call("ffno_gn_finalize", plan.B, plan.nwg, plan.C, _p(partials), _p(gn[0]), _p(gn[1]), _p(stats),
                  _p(a), _p(b), self._stream())
        return a, b, stats

    def _tw(self, mat, x, y, batch, n, bmap_x, bmap_y, kmap, mmap, accumulate=0, twiddle=None, block=True):
        hi, lo, m, k, mp, kp = mat[:6]
        if (block and twiddle is None and (mp > DENSE_DFT_MAX or kp > DENSE_DFT_MAX) and
                (mp <= DENSE_DFT_MAX or mmap[0] >= m) and (kp <= DENSE_DFT_MAX or kmap[0] >= k)):
            # the large dimensions have single-level maps: blocks of the LDS-staged size, the k blocks accumulated
            xf, yf = x.view(-1), y.view(-1)
            for m0 in range(0, m, DENSE_DFT_MAX):
                for j, (k0, blk) in enumerate(_k_blocks(mat, m0)):
                    self._tw(blk, xf[k0 * kmap[2]:], yf[m0 * mmap[2]:], batch, n, bmap_x, bmap_y, kmap, mmap,
                             accumulate if j == 0 else 1, block=False)
            return
        g = FfnoTwGemm()
        g.a_hi, g.a_lo = _p(hi), _p(lo)
        g.a_l3 = _p(mat[6] if len(mat) > 6 else None)
        g.m, g.k, g.n, g.m_pad, g.k_pad, g.batch = m, k, n, mp, kp, batch
        g.x, g.y = _p(x), _p(y)
        g.b_split, g.x_b_hi, g.x_b_lo = bmap_x
        g.y_b_hi, g.y_b_lo = bmap_y
        g.k_split, g.x_k_hi, g.x_k_lo = kmap
        g.m_split, g.y_m_hi, g.y_m_lo = mmap
        g.accumulate = accumulate
        if twiddle is not None:
            g.twiddle = _p(twiddle[0])
            g.tw_n1, g.tw_n2, g.tw_cols, g.tw_bsplit = twiddle[1:]
        _lib.call("ffno_tw_gemm", ctypes.byref(g), self._stream())

    @staticmethod
    def _w_side(plan):
        """(part stride, h stride) of the half spectrum where the W-direction transforms meet the H-direction ones:
        [part][H][Wf][C] per plane, or with W factorised row-major [H][part][Wf][C] (each voxel row's spectrum in
        one piece, so that the W stages address it with two-level maps)."""
        row = plan.wf * plan.C
        return (row, 2 * row) if plan.wfac is not None else (plan.H * row, row)

    def _hdft(self, plan, stages, s, tmp, pc, into_spectrum):
        """The complex DFT along H of pc (b, d) planes of s in place, factorised (transforms.dft_stages): stage A
        into tmp (layout [part][k1][j2][Wf][C]), then stage B back into s. into_spectrum: from the W-side layout
        into [part][H][Wf][C] (the layout the mix reads); else the reverse."""
        H, row, pl = plan.H, plan.wf * plan.C, plan.plane
        ps, hs = self._w_side(plan)
        n1, n2 = plan.hfac["n"]
        a, tw, b = stages
        big = 1 << 40
        r = a[4] // 2
        in_ps, in_hs = (ps, hs) if into_spectrum else (H * row, row)
        out_ps, out_hs = (H * row, row) if into_spectrum else (ps, hs)
        if in_hs == row:
            # rows j2 of a j1 are contiguous: columns (j2, kx, c), the twiddle by column group
            self._tw(a, s, tmp, pc, n2 * row, (big, 0, pl), (0, pl), (n1, in_ps, n2 * row), (r, H * row, n2 * row),
                     twiddle=(tw, n1, n2, row, 0))
        else:
            # batch items (plane, j2), columns (kx, c), the twiddle by batch item
            self._tw(a, s, tmp, pc * n2, row, (n2, pl, in_hs), (pl, row), (n1, in_ps, n2 * in_hs),
                     (r, H * row, n2 * row), twiddle=(tw, n1, n2, 0, n2))
        self._tw(b, tmp, s, pc * n1, row, (n1, pl, n2 * row), (pl, out_hs), (n2, H * row, row),
                 (n2, out_ps, n1 * out_hs))

    def _wdft_r(self, plan, st, x, s, tmp, rows):
        """Real voxel rows (W x C, row-major) -> their half spectra in the W-side layout, factorised
        (transforms.rdft_stages): stage A into tmp (per row [part][k1][j2][C]), stage B by runs of k1 into s."""
        W, C, wf = plan.W, plan.C, plan.wf
        n1, n2 = plan.wfac["n"]
        _, a, tw, runs = st
        big = 1 << 40
        wc = W * C
        self._tw(a, x, tmp, rows, n2 * C, (big, 0, wc), (0, 2 * wc), (big, 0, n2 * C), (a[4] // 2, wc, n2 * C),
                 twiddle=(tw, n1, n2, C, 0))
        for k0, cnt, b in runs:
            K2 = b[2] // 2
            self._tw(b, tmp[k0 * n2 * C:], s[k0 * C:], rows * cnt, C, (cnt, 2 * wc, n2 * C), (2 * wf * C, C),
                     (n2, wc, C), (K2, wf * C, n1 * C))

    def _wdft_i(self, plan, st, s, out, tmp, rows):
        """Half spectra in the W-side layout -> real voxel rows (W x C), factorised (transforms.irdft_stages): stage A
        by runs of j2 into tmp (per row [part][k1][j2][C]), stage B into out."""
        W, C, wf = plan.W, plan.C, plan.wf
        n1, n2 = plan.wfac["n"]
        _, runs, tw, b = st
        big = 1 << 40
        wc = W * C
        twf = tw.view(-1)
        for j0, cnt, a in runs:
            J1 = a[3] // 2
            self._tw(a, s[j0 * C:], tmp[j0 * C:], rows * cnt, C, (cnt, 2 * wf * C, C), (2 * wc, C),
                     (J1, wf * C, n2 * C), (a[4] // 2, wc, n2 * C), twiddle=(twf[2 * j0:], n1, n2, 0, cnt))
        self._tw(b, tmp, out, rows * n1, C, (n1, 2 * wc, n2 * C), (wc, C), (n2, wc, C), (big, 0, n1 * C))

    def _mix(self, plan, re, im, rows, modes, mode_stride, rmap, wc, e, adjoint=False):
        mx = FfnoMix()
        mx.re, mx.im = _p(re), _p(im)
        mx.C, mx.rows, mx.modes, mx.rank = plan.C, rows, modes, plan.rank
        mx.mode_stride = mode_stride
        mx.r_split, mx.r_hi, mx.r_lo = rmap
        mx.wc, mx.gate = _p(wc), _p(e["gate"])
        mx.v_p, mx.u_p = _p(e["v"]), _p(e["u"])
        mx.factor = ctypes.c_void_p(0)
        mx.parts = 2 if adjoint else 3
        _lib.call("ffno_mix", ctypes.byref(mx), self._stream())

    def horizontal(self, plan, xs, out, i, scr, tables, adjoint=False):
        """rfft2 -> mix -> irfft2 of xs [B,D,H,W,C] into out (models/ffno_model.py:441-448); with adjoint, the
        transpose of that linear map (W_fwd^T H_fwd^T Mix^T H_inv^T W_inv^T), tables then being the adjoint's.
        (b, d) planes are independent: they go through in chunks of plan.h_chunk, each chunk's five launches back to
        back. out may be xs."""
        B, D, H, W, C, wf = plan.B, plan.D, plan.H, plan.W, plan.C, plan.wf
        plane = plan.plane
        hwc = H * W * C
        big = 1 << 40
        if adjoint:
            ta = plan.adjoint_tables()
            m1, m2, m3, m4 = ta["w_inv_T"], ta["h_inv_T"], ta["h_fwd_T"], ta["w_fwd_T"]
            wc = self.weights.adjoint()["blocks"][i]["spec.wc"]
        else:
            m1, m2, m3, m4 = plan.tw["w_fwd"], plan.tw["h_fwd"], plan.tw["h_inv"], plan.tw["w_inv"]
            wc = self.weights.blocks[i]["spec.wc"]
        s1, s2 = scr["s1"], scr["s2"]
        hf, wv = plan.hfac, plan.wfac
        if hf is not None:
            f2, f3 = (hf["h_inv_T"], hf["h_fwd_T"]) if adjoint else (hf["h_fwd"], hf["h_inv"])
        if wv is not None:
            w1, w4 = (wv["w_inv_T"], wv["w_fwd_T"]) if adjoint else (wv["w_fwd"], wv["w_inv"])
        ps, hs = self._w_side(plan)
        row = wf * C
        spec = s1 if hf is not None else s2        # where the spectrum is mixed
        xf, of = xs.view(-1), out.view(-1)
        for p0 in range(0, B * D, plan.h_chunk):
            pc = min(plan.h_chunk, B * D - p0)
            xc, oc = xf[p0 * hwc:], of[p0 * hwc:]
            # along W, real -> half spectrum: one GEMM per voxel row (b, d, h), or its factorised form through s2
            if wv is not None:
                self._wdft_r(plan, w1, xc, s1, s2, pc * H)
            else:
                self._tw(m1, xc, s1, pc * H, C, (H, hwc, W * C), (plane, hs), (big, 0, C), (wf, ps, C))
            # along H, complex: one GEMM per plane (b, d) over (part, h) rows, columns (kx, c), or its factorised
            # form in place in s1
            if hf is not None:
                self._hdft(plan, f2, s1, s2, pc, True)
            else:
                self._tw(m2, s1, s2, pc, row, (big, 0, plane), (0, plane), (H, ps, hs), (H, H * row, row))
            self._mix(plan, spec, spec[H * row:], pc, H * wf, C, (big, 0, plane), wc, tables["spec"], adjoint)
            if hf is not None:
                self._hdft(plan, f3, s1, s2, pc, False)
            else:
                self._tw(m3, s2, s1, pc, row, (big, 0, plane), (0, plane), (H, H * row, row), (H, ps, hs))
            # half spectrum -> real along W
            if wv is not None:
                self._wdft_i(plan, w4, s1, oc, s2, pc * H)
            else:
                self._tw(m4, s1, oc, pc * H, C, (H, plane, hs), (hwc, W * C), (wf, ps, C), (big, 0, C))

    def _colgemm(self, plan, mat, x, y):
        hi, lo, m, k, mp, kp = mat[:6]
        g = _lib.FfnoColGemm()
        g.a_hi, g.a_lo = _p(hi), _p(lo)
        g.a_l3 = _p(mat[6] if len(mat) > 6 else None)
        g.m, g.k, g.m_pad, g.k_pad = m, k, mp, kp
        g.B, g.hw, g.C = plan.B, plan.H * plan.W, plan.C
        g.x, g.y = _p(x), _p(y)
        _lib.call("ffno_col_gemm", ctypes.byref(g), self._stream())

    def _vgemm(self, plan, mat, x, y, to_spectrum):
        """One vertical GEMM of the height gradient: levels [B, D, HW, C] -> spectrum [B, 2M, HW, C]
        (to_spectrum) or back, with the shared matrix or every column's own."""
        if plan.percol:
            self._colgemm(plan, mat, x, y)
            return
        B, D, M = plan.B, plan.D, plan.modes
        hwc = plan.H * plan.W * plan.C
        big = 1 << 40
        if to_spectrum:
            self._tw(mat, x, y, B, hwc, (big, 0, D * hwc), (0, 2 * M * hwc), (big, 0, hwc), (big, 0, hwc))
        else:
            self._tw(mat, x, y, B, hwc, (big, 0, 2 * M * hwc), (0, D * hwc), (big, 0, hwc), (big, 0, hwc))

    def vertical(self, plan, xv, v1, out, i, tables, adjoint=False):
        """Filon transform along D -> mix -> synthesis at the levels, /L folded in (models/ffno_model.py:270-305);
        with adjoint, its transpose (tables then being the adjoint's). The spectrum after the mix stays in v1. out may
        be xv."""
        B, D, C, M = plan.B, plan.D, plan.C, plan.modes
        hwc = plan.H * plan.W * C
        big = 1 << 40
        if plan.percol:
            # every column its own analysis and synthesis matrices (the reference's per-column path)
            if adjoint:
                ca = plan.column_adjoint()
                m1, m2 = ca["v_inv_T"], ca["v_fwd_T"]
                wc = self.weights.adjoint()["blocks"][i]["vertical.wc"]
            else:
                m1, m2 = plan.colmats["v_fwd"], plan.colmats["v_inv"]
                wc = self.weights.blocks[i]["vertical.wc"]
            self._colgemm(plan, m1, xv, v1)
            hw = plan.H * plan.W
            self._mix(plan, v1, v1[M * hwc:], B * hw, M, hwc, (hw, 2 * M * hwc, C), wc, tables["vertical"], adjoint)
            self._colgemm(plan, m2, v1, out)
            return
        if adjoint:
            ta = plan.adjoint_tables()
            m1, m2 = ta["v_inv_T"], ta["v_fwd_T"]
            wc = self.weights.adjoint()["blocks"][i]["vertical.wc"]
        else:
            m1, m2 = plan.tw["v_fwd"], plan.tw["v_inv"]
            wc = self.weights.blocks[i]["vertical.wc"]
        # spectra [B, 2, M, H, W, C]: row i of the [2M, D] analysis (real parts, then imaginary) sits at i * hwc, a
        # single-level map, so that 2M > 128 rows run as LDS-staged 128-row blocks
        self._tw(m1, xv, v1, B, hwc, (big, 0, D * hwc), (0, 2 * M * hwc), (big, 0, hwc), (big, 0, hwc))
        hw = plan.H * plan.W
        self._mix(plan, v1, v1[M * hwc:], B * hw, M, hwc, (hw, 2 * M * hwc, C), wc, tables["vertical"], adjoint)
        self._tw(m2, v1, out, B, hwc, (big, 0, 2 * M * hwc), (0, D * hwc), (big, 0, hwc), (big, 0, hwc))

    def _scratch(self, plan, dev, adjoint=False):
        f32 = dict(device=dev, dtype=torch.float32)
        scr = dict(s1=torch.empty(plan.h_chunk * plan.plane, **f32), s2=torch.empty(plan.h_chunk * plan.plane2, **f32),
                   v1=torch.empty(plan.B * 2 * plan.modes * plan.H * plan.W * plan.C, **f32),
                   part=torch.empty(plan.B, plan.nwg, GROUPS, 3, **f32))
        if adjoint:
            scr["v1y"] = torch.empty_like(scr["v1"])
            scr["part2"] = torch.empty(plan.B, plan.nwg, GROUPS, 2, **f32)
            scr["part3"] = torch.empty(plan.B, plan.nwg, GROUPS, 2, **f32)   # a second GroupNorm's in flight
        return scr

    def _block(self, plan, i, prev, x_out, h, hb, s, v, scr, bz, spectra="v1"):
        """Forward of block i on three activation buffers. prev: the assembled block input, stored to x_out (may
        be None when h already holds it); h: the tensor holding the block input. Schedule, with every write in place
        of a buffer whose last reader is the writing kernel itself:
            block_in: prev -> x_out, xs -> s, xv -> v        horizontal: s -> s     post: s -> s (sp)
            vertical: v -> v                                 post: v -> v (vp)      fuse: h, s, v -> s (fp)
            resmlp: h + res gelu(GN(fp)) -> hb (x1), MLP -> v (pp)
            resmlp: hb + res GN(pp) -> hb (x2), MLP -> s (mp)
        Returns the next block's input as an assembly (x2 + res GN(mp)) and the five GroupNorms."""
        dev = s.device
        blk = self.weights.blocks[i]
        tables = plan.tables(i)
        st = self._stream()
        rows = plan.rows
        stats = FfnoStatsOut(_p(plan.depth_w), _p(scr["part"]))
        _lib.call("ffno_block_in", rows, prev, _p(x_out), _p(blk["spec.g1"]), _p(blk["spec.g2"]),
                  _p(blk["vertical.g1"]), _p(blk["vertical.g2"]), _p(blk["z_proj"]), _p(bz), _p(s), _p(v), st)
        gn = {}
        self.horizontal(plan, s, s, i, scr, tables)
        _lib.call("ffno_post", rows, _p(s), _p(blk["spec.p1"]), _p(blk["spec.p2"]), _p(s), stats, st)
        gn["s"] = self._gn(plan, scr["part"], blk["norm_spec"], dev)
        self.vertical(plan, v, scr[spectra], v, i, tables)
        _lib.call("ffno_post", rows, _p(v), _p(blk["vertical.p1"]), _p(blk["vertical.p2"]), _p(v), stats, st)
        gn["v"] = self._gn(plan, scr["part"], blk["norm_vert"], dev)
        self._tail(plan, i, h, s, v, s, hb, v, hb, s, scr, gn)
        ma, mb, _ = gn["m"]
        return _asm(hb, s, ma, mb, blk["res_mlp"]), gn

    def _block_keep(self, plan, i, prev, x_out, bk, scr, bz, x2=None):
        """Forward of block i keeping what its adjoint reads: the input (assembled from prev into x_out, or already
        in bk["h"]), the branch transforms' outputs in bk["s"], bk["v"] (the vertical spectrum after the mix in
        bk["v1y"] or the scratch), the five GroupNorm inputs, x1 and the gates; x2 to x2 when given. Returns the
        GroupNorm records."""
        dev = bk["s"].device
        blk = self.weights.blocks[i]
        tables = plan.tables(i)
        st = self._stream()
        rows = plan.rows
        s, v = bk["s"], bk["v"]
        _lib.call("ffno_block_in", rows, prev, _p(x_out), _p(blk["spec.g1"]), _p(blk["spec.g2"]),
                  _p(blk["vertical.g1"]), _p(blk["vertical.g2"]), _p(blk["z_proj"]), _p(bz), _p(s), _p(v), st)
        stats = FfnoStatsOut(_p(plan.depth_w), _p(scr["part"]))
        gn = {}
        self.horizontal(plan, s, s, i, scr, tables)
        _lib.call("ffno_post", rows, _p(s), _p(blk["spec.p1"]), _p(blk["spec.p2"]), _p(bk["sp"]), stats, st)
        gn["s"] = self._gn(plan, scr["part"], blk["norm_spec"], dev)
        v1y = bk.get("v1y")
        self.vertical(plan, v, v1y if v1y is not None else scr["v1y" if "v1y" in scr else "v1"], v, i, tables)
        _lib.call("ffno_post", rows, _p(v), _p(blk["vertical.p1"]), _p(blk["vertical.p2"]), _p(bk["vp"]), stats, st)
        gn["v"] = self._gn(plan, scr["part"], blk["norm_vert"], dev)
        self._tail(plan, i, bk["h"], bk["sp"], bk["vp"], bk["fp"], bk["x1"], bk["pp"], x2, bk["mp"], scr, gn,
                   bk.get("gates"))
        return gn

    def _tail(self, plan, i, h, sp, vp, fp, x1, pp, x2, mp, scr, gn, gates=None):
        """fuse -> pw -> mlp of block i (models/ffno_model.py:524-551), buffers as named (x2 may be None; gates, if
        given, receives the branch gates)."""
        dev = sp.device
        blk = self.weights.blocks[i]
        st = self._stream()
        rows = plan.rows
        stats = FfnoStatsOut(_p(plan.depth_w), _p(scr["part"]))
        sa, sb, _ = gn["s"]
        va, vb, _ = gn["v"]
        _lib.call("ffno_fuse", rows, _p(h), _p(sp), _p(vp), _p(sa), _p(sb), _p(va), _p(vb), _p(blk["bg1"]),
                  _p(blk["bg2"]), _p(blk["bg2_bias"]), _p(blk["f1"]), _p(blk["f2"]), _p(fp), stats, _p(gates), st)
        gn["f"] = self._gn(plan, scr["part"], blk["norm_fuse"], dev)
        fa, fb, _ = gn["f"]
        _lib.call("ffno_resmlp", rows, _asm(h, fp, fa, fb, blk["res_fused"], 1), _p(x1), _p(blk["pw1"]),
                  _p(blk["pw2"]), _p(pp), stats, st)
        gn["p"] = self._gn(plan, scr["part"], blk["norm_pw"], dev)
        pa, pb, _ = gn["p"]
        _lib.call("ffno_resmlp", rows, _asm(x1, pp, pa, pb, blk["res_pw"]), _p(x2), _p(blk["m1"]), _p(blk["m2"]),
                  _p(mp), stats, st)
        gn["m"] = self._gn(plan, scr["part"], blk["norm_mlp"], dev)

    def _prepare(self, x, z, dx, dy, plan):
        if not x.is_cuda:
            raise RuntimeError("the FFNO3D HIP engine runs on the GPU only; move the inputs to a cuda device")
        dev = x.device
        self.sync(dev)
        B, cin, D, H, W = x.shape
        if cin != 6:
            raise ValueError("FFNO3D expects 6 input channels")
        z = z.reshape(B, D, H, W)
        if plan is None:
            plan = self.plan(B, D, H, W, z, dx, dy)
        return dev, plan, x.contiguous().float(), z.reshape(B, -1).contiguous().float()

    def forward(self, x, z, dx, dy, plan=None, saved=None):
        """Normalised log10 populations [B, Cout, D, H, W] from x [B, 6, D, H, W] and z_scale [B, D, H, W] (Mm).
        saved: a Saved that receives the block inputs the adjoint needs (its stride says which)."""
        dev, plan, x, bz = self._prepare(x, z, dx, dy, plan)
        Wt = self.weights
        B, D, H, W, C = plan.B, plan.D, plan.H, plan.W, plan.C
        N = B * plan.voxels
        st = self._stream()
        f32 = dict(device=dev, dtype=torch.float32)
        hb, s, v = (torch.empty(N, C, **f32) for _ in range(3))
        scr = self._scratch(plan, dev)
        rows = plan.rows
        _lib.call("ffno_lift", rows, _p(x), _p(Wt.lift), _p(s), FfnoStatsOut(_p(plan.depth_w), _p(scr["part"])), st)
        a, b, _ = self._gn(plan, scr["part"], Wt.lift_gn, dev)
        prev = _asm(None, s, a, b, None, 1)
        keep_from = saved.full_from if saved is not None and saved.full_from is not None else len(Wt.blocks)
        for i in range(len(Wt.blocks)):
            if i >= keep_from:
                # every intermediate to a buffer of its own; x2 to a running buffer (hb), the next input from it
                bk = {k: torch.empty(N, C, **f32) for k in ("h", "s", "v", "sp", "vp", "fp", "pp", "mp", "x1")}
                bk["gates"] = torch.empty(N, 2 * C, **f32)
                bk["v1y"] = torch.empty_like(scr["v1"]) if getattr(saved, "want_z", True) else None
                bk["gn"] = self._block_keep(plan, i, prev, bk["h"], bk, scr, bz, hb)
                saved.blocks[i] = bk
                ma, mb, _ = bk["gn"]["m"]
                prev = _asm(hb, bk["mp"], ma, mb, self.weights.blocks[i]["res_mlp"])
                continue
            h = hb
            if saved is not None and i % saved.stride == 0:
                h = torch.empty(N, C, **f32)
                saved.inputs[i] = h
            prev, _ = self._block(plan, i, prev, h, h, hb, s, v, scr, bz)
        y = torch.empty(B, self.module.out_channels, D, H, W, **f32)
        h_last = None
        if saved is not None and saved.stride == 1:
            h_last = saved.head = torch.empty(N, C, **f32)
        _lib.call("ffno_head", rows, prev, _p(h_last), _p(Wt.p1), _p(Wt.p2), _p(Wt.p2_bias), _p(x),
                  ctypes.byref(Wt.head), _p(y), st)
        return y

    # ------------------------------------------------------------------------------------------------ head only
    def _head_args(self, head, x_in):
        if head.device.type == "cpu" or x_in.device.type == "cpu":
            raise RuntimeError("the FFNO3D HIP engine runs on the GPU only; move head and x_in to a cuda device")
        self.sync(head.device)
        m = self.module
        n_head = m.out_channels + len(m.species_levels)
        if head.dim() < 3 or head.shape[1] != n_head:
            raise ValueError(f"head must have shape [B, {n_head}, ...]; got {tuple(head.shape)}")
        if tuple(x_in.shape) != (head.shape[0], 6) + tuple(head.shape[2:]):
            raise ValueError(f"x_in must have shape {(head.shape[0], 6) + tuple(head.shape[2:])}; got "
                             f"{tuple(x_in.shape)}")
        voxels = int(math.prod(head.shape[2:]))
        return head.shape[0], voxels, head.contiguous().float(), x_in.contiguous().float()

    def head_output(self, head, x_in):
        """FFNO3D.head_output (models/ffno_model.py:730-749): normalised log10 n [B, Cout, ...] from the head channels
        [B, Cout + S, ...] and the normalised inputs [B, 6, ...] (ffno_head_output)."""
        B, voxels, head, x_in = self._head_args(head, x_in)
        y = torch.empty((B, self.module.out_channels) + tuple(head.shape[2:]), device=head.device)
        _lib.call("ffno_head_output", B, voxels, _p(head), _p(x_in), ctypes.byref(self.weights.head), _p(y),
                  self._stream())
        return y

    def head_output_bwd(self, head, x_in, ybar):
        """(head_bar, x_in_bar) of <ybar, head_output(head, x_in)> (ffno_head_output_bwd)."""
        B, voxels, head, x_in = self._head_args(head, x_in)
        if tuple(ybar.shape) != (B, self.module.out_channels) + tuple(head.shape[2:]):
            raise ValueError(f"ybar has shape {tuple(ybar.shape)}; expected {(B, self.module.out_channels) + tuple(head.shape[2:])}")
        ybar = ybar.contiguous().float()
        head_bar = torch.empty_like(head)
        x_bar = torch.empty_like(x_in)
        _lib.call("ffno_head_output_bwd", B, voxels, _p(head), _p(x_in), ctypes.byref(self.weights.head), _p(ybar),
                  _p(head_bar), _p(x_bar), self._stream())
        return head_bar, x_bar

    # ------------------------------------------------------------------------------------------------ tangent
    def _gn_tangent(self, plan, y, yd, rec, gamma, part2):
        """(tan_b, tan_c) [B, C] of a GroupNorm's tangent n_dot = a yd + tan_b y + tan_c (a the forward affine)."""
        _, _, stats = rec
        _lib.call("ffno_gn_tangent_stats", plan.rows, _p(y), _p(yd), _p(stats), _p(plan.depth_w), _p(part2),
                  self._stream())
        return self._gn_tangent_finalize(plan, rec, gamma, part2)

    @staticmethod
    def _tan_sums(plan, y, rec, part2):
        """The producer of a GroupNorm input's tangent accumulates that GroupNorm's tangent sums (y its primal
        input) into part2."""
        return _lib.FfnoTanStatsOut(_p(y), _p(rec[2]), _p(plan.depth_w), _p(part2))

    def _gn_tangent_finalize(self, plan, rec, gamma, part2):
        tb = torch.empty(plan.B, plan.C, device=part2.device)
        tc = torch.empty(plan.B, plan.C, device=part2.device)
        _lib.call("ffno_gn_tangent_finalize", plan.B, plan.nwg, plan.C, _p(part2), _p(rec[2]), _p(gamma), _p(tb),
                  _p(tc), self._stream())
        return tb, tc

    def _assemble_t(self, plan, out, t, td, rec, tan, act=0, scale=None, base_dot=None):
        """out = base_dot + scale * act'(a t + b) * (a td + tan_b t + tan_c): the tangent of a GroupNorm output
        (GELU after it with act) scaled and added to a residual's tangent; elementwise, out may alias td or
        base_dot."""
        a, b, _ = rec
        tb, tc = tan
        at = _lib.FfnoAssembleT(_p(base_dot), _p(t), _p(td), _p(a), _p(b), _p(tb), _p(tc), _p(scale), act)
        _lib.call("ffno_assemble_t", plan.rows, at, _p(out), self._stream())

    def jvp(self, x, z, dx, dy, xdot, plan=None):
        """(y, y_dot): the forward and its directional derivative along xdot [B, 6, D, H, W] (normalised input
        features; the heights carry no tangent), as the reference's forward-AD JVP (ffno_fsdp_runtime.py:350-372)
        computes it. Block by block: the block's forward keeping its intermediates, then its tangent (gate, MLP,
        fusion and head tangent kernels; the spectral transforms and mixes are linear and run on the tangent as they
        are; every GroupNorm's tangent from its two tangent sums). Memory: about 16 activation-sized buffers whatever
        the depth."""
        dev, plan, x, bz = self._prepare(x, z, dx, dy, plan)
        xdot = xdot.contiguous().float()
        if tuple(xdot.shape) != tuple(x.shape):
            raise ValueError(f"tangent shape {tuple(xdot.shape)} differs from the input's {tuple(x.shape)}")
        Wt = self.weights
        B, D, H, W, C = plan.B, plan.D, plan.H, plan.W, plan.C
        N = B * plan.voxels
        st = self._stream()
        rows = plan.rows
        f32 = dict(device=dev, dtype=torch.float32)
        bk = {k: torch.empty(N, C, **f32) for k in ("h", "s", "v", "sp", "vp", "fp", "pp", "mp", "x1")}
        bk["gates"] = torch.empty(N, 2 * C, **f32)
        hb = torch.empty(N, C, **f32)
        hd, t1, t2, t3 = (torch.empty(N, C, **f32) for _ in range(4))
        scr = self._scratch(plan, dev)
        part2 = torch.empty(B, plan.nwg, GROUPS, 2, **f32)
        # lift: h0 = gelu(GN(W x)); its tangent from W xdot
        lift = bk["s"]
        _lib.call("ffno_lift", rows, _p(x), _p(Wt.lift), _p(lift), FfnoStatsOut(_p(plan.depth_w), _p(scr["part"])),
                  st)
        rec = self._gn(plan, scr["part"], Wt.lift_gn, dev)
        _lib.call("ffno_lift_t", rows, _p(xdot), _p(Wt.lift), _p(t1), st)
        self._assemble_t(plan, hd, lift, t1, rec, self._gn_tangent(plan, lift, t1, rec, Wt.lift_gn[0], part2), act=1)
        prev = _asm(None, lift, rec[0], rec[1], None, 1)
        for i in range(len(Wt.blocks)):
            blk = Wt.blocks[i]
            tables = plan.tables(i)
            gn = self._block_keep(plan, i, prev, bk["h"], bk, scr, bz, hb)
            # gates and the two branches' transforms and post networks; each post network's tangent kernel also
            # accumulates its GroupNorm's tangent sums
            _lib.call("ffno_gates_t", rows, _p(bk["h"]), _p(hd), _p(blk["spec.g1"]), _p(blk["spec.g2"]),
                      _p(blk["vertical.g1"]), _p(blk["vertical.g2"]), _p(t1), _p(t2), st)
            self.horizontal(plan, t1, t1, i, scr, tables)
            _lib.call("ffno_mlp_t", rows, _asm(bk["s"]), _p(t1), _p(blk["spec.p1"]), _p(blk["spec.p2"]), _p(t1),
                      self._tan_sums(plan, bk["sp"], gn["s"], part2), st)
            tan_s = self._gn_tangent_finalize(plan, gn["s"], blk["norm_spec"][0], part2)
            self.vertical(plan, t2, scr["v1"], t2, i, tables)
            _lib.call("ffno_mlp_t", rows, _asm(bk["v"]), _p(t2), _p(blk["vertical.p1"]), _p(blk["vertical.p2"]),
                      _p(t2), self._tan_sums(plan, bk["vp"], gn["v"], part2), st)
            tan_v = self._gn_tangent_finalize(plan, gn["v"], blk["norm_vert"][0], part2)
            # the fusion forms s = gelu(GN_s(sp)), v = gelu(GN_v(vp)) and their tangents itself
            sa, sb_, _ = gn["s"]
            va, vb, _ = gn["v"]
            st_s = _lib.FfnoAssembleT(None, _p(bk["sp"]), _p(t1), _p(sa), _p(sb_), _p(tan_s[0]), _p(tan_s[1]), None, 1)
            st_v = _lib.FfnoAssembleT(None, _p(bk["vp"]), _p(t2), _p(va), _p(vb), _p(tan_v[0]), _p(tan_v[1]), None, 1)
            _lib.call("ffno_fuse_t", rows, _p(bk["h"]), _p(hd), st_s, st_v, _p(bk["gates"]), _p(blk["bg1"]),
                      _p(blk["bg2"]), _p(blk["f1"]), _p(blk["f2"]), _p(t3),
                      self._tan_sums(plan, bk["fp"], gn["f"], part2), st)
            tan_f = self._gn_tangent_finalize(plan, gn["f"], blk["norm_fuse"][0], part2)
            # x1 = h + res_fused gelu(GN_f(fp)) -> t1; pp tangent -> t2
            self._assemble_t(plan, t1, bk["fp"], t3, gn["f"], tan_f, act=1, scale=blk["res_fused"], base_dot=hd)
            _lib.call("ffno_mlp_t", rows, _asm(bk["x1"]), _p(t1), _p(blk["pw1"]), _p(blk["pw2"]), _p(t2),
                      self._tan_sums(plan, bk["pp"], gn["p"], part2), st)
            tan_p = self._gn_tangent_finalize(plan, gn["p"], blk["norm_pw"][0], part2)
            # x2 = x1 + res_pw GN_p(pp) -> t1; mp tangent -> t3
            self._assemble_t(plan, t1, bk["pp"], t2, gn["p"], tan_p, scale=blk["res_pw"], base_dot=t1)
            pa, pb, _ = gn["p"]
            _lib.call("ffno_mlp_t", rows, _asm(bk["x1"], bk["pp"], pa, pb, blk["res_pw"]), _p(t1), _p(blk["m1"]),
                      _p(blk["m2"]), _p(t3), self._tan_sums(plan, bk["mp"], gn["m"], part2), st)
            tan_m = self._gn_tangent_finalize(plan, gn["m"], blk["norm_mlp"][0], part2)
            # the next block's input: x2 + res_mlp GN_m(mp)
            self._assemble_t(plan, hd, bk["mp"], t3, gn["m"], tan_m, scale=blk["res_mlp"], base_dot=t1)
            ma, mb, _ = gn["m"]
            prev = _asm(hb, bk["mp"], ma, mb, blk["res_mlp"])
        y = torch.empty(B, self.module.out_channels, D, H, W, **f32)
        yd = torch.empty_like(y)
        _lib.call("ffno_head", rows, prev, None, _p(Wt.p1), _p(Wt.p2), _p(Wt.p2_bias), _p(x), ctypes.byref(Wt.head),
                  _p(y), st)
        _lib.call("ffno_head_t", rows, prev, _p(hd), _p(Wt.p1), _p(Wt.p2), _p(Wt.p2_bias), _p(x), _p(xdot),
                  ctypes.byref(Wt.head), _p(yd), st)
        return y, yd

    def forward_saved(self, x, z, dx, dy, plan=None, stride=None, want_z=True, keep=0):
        """(y, saved): the forward that keeps what the adjoint needs: every intermediate (stride 0, or by default
        when it fits the free HBM), else the block inputs at the given checkpoint stride or the one that fits.
        want_z: the adjoint will be asked for the height gradient too (full keeps the vertical spectra then).
        keep (with stride 1): the last `keep` blocks kept whole, the others recomputed (two-pass fusion adjoint)."""
        dev, plan, _, _ = self._prepare(x, z, dx, dy, plan)
        if keep and stride == 1:
            keep = min(int(keep), plan.n_layers)
            saved = Saved(1, True, keep == plan.n_layers, keep=keep, n_layers=plan.n_layers)
        else:
            stride, two_pass, full, keep = self.choose_checkpoints(plan, stride, want_z)
            saved = Saved(stride, two_pass, full, keep=keep, n_layers=plan.n_layers)
        saved.want_z = want_z
        self.last_stride = 0 if saved.full else saved.stride     # informational (the latest saving forward's)
        self.last_keep = 0 if saved.full_from is None else plan.n_layers - saved.full_from   # blocks kept whole
        return self.forward(x, z, dx, dy, plan=plan, saved=saved), saved

    # ------------------------------------------------------------------------------------------------ adjoint
    def _gn_sums(self, plan, gb, ubar, part2):
        """A GroupNorm backward's S1, S2 sums from u_bar in a pass of their own, into gb.sums; returns the sums
        tensor, which the caller keeps alive until gb's consumer is enqueued."""
        _lib.call("ffno_gn_grad_stats", plan.rows, gb, _p(ubar), _p(part2), self._stream())
        return self._gn_finalize_sums(plan, gb, part2)

    def _gn_finalize_sums(self, plan, gb, part2):
        """gb.sums from the per-workgroup partials a kernel accumulated; returns the sums tensor (kept alive by the
        caller until gb's consumer is enqueued)."""
        sums = torch.empty(plan.B, GROUPS, 2, device=part2.device)
        _lib.call("ffno_gn_grad_finalize", plan.B, plan.nwg, _p(part2), _p(sums), self._stream())
        gb.sums = _p(sums)
        return sums
        return sums

    def vjp(self, x, z, dx, dy, ybar, saved, want_z=False, plan=None):
        """(x_bar [B, 6, D, H, W], z_bar [B, D, H, W] or None) of <ybar, forward(x, z)>: the hand-written adjoint,
        block by block in reverse, each block recomputed from its input (the reference's non-reentrant
        checkpointing, models/ffno_model.py:751-780); inputs that were not saved are recomputed from the nearest
        saved one. With want_z the height gradient is computed too, per column: through the depth weights of every
        GroupNorm, z_proj, and the vertical transform's Filon weights, phases, taper and per-column shift
        (models/ffno_model.py:28-74, :239-309).

        Buffers: nine activation-sized ones besides the saved inputs (g: the gradient flowing down; s, v: the
        branch transforms' outputs; sp, vp, fp, pp, mp: the five GroupNorms' inputs; x1), plus t for a recomputed
        block input when the stride is above 1. The backward writes in place wherever the writer is the last reader:
        g carries out_bar -> x2_bar -> x1_bar -> h_bar; sp, vp, fp, pp, mp, x1 are reused once dead."""
        dev, plan, x, bz = self._prepare(x, z, dx, dy, plan)
        Wt = self.weights
        A = Wt.adjoint()
        plan.adjoint_tables()
        B, D, H, W, C, M = plan.B, plan.D, plan.H, plan.W, plan.C, plan.modes
        N = B * plan.voxels
        L = len(Wt.blocks)
        stride = saved.stride
        ybar = ybar.contiguous().float()
        st = self._stream()
        rows = plan.rows
        f32 = dict(device=dev, dtype=torch.float32)
        g = torch.empty(N, C, **f32)
        # the working buffers of the recomputed blocks, allocated at the first one (the blocks kept whole, whose
        # adjoints run first, have been released by then)
        work = {}

        def working():
            if not work:
                work.update(zip(("s", "v", "sp", "vp", "fp", "pp", "mp", "x1"),
                                (torch.empty(N, C, **f32) for _ in range(WORKING_UNITS - 1))))
                work["tbuf"] = torch.empty(N, C, **f32) if stride > 1 else None
                work["gates"] = torch.empty(N, 2 * C, **f32) if saved.two_pass else None
                work["chunk"] = self._fuse_chunk(plan) if (work["gates"] is None and B == 1) else 0
                work["gates_c"] = (torch.empty(min(work["chunk"], plan.voxels), 2 * C, **f32) if work["chunk"]
                                   else None)
            return work

        s = chunk = gates_c = None
        scr = self._scratch(plan, dev, adjoint=True)
        part2 = scr["part2"]
        ones = torch.ones(B, C, **f32)
        zeros = torch.zeros(B, C, **f32)
        xbar = torch.empty(B, 6, D, H, W, **f32)
        if want_z:
            ht = plan.height_tables()
            zbar = torch.zeros(B, plan.voxels, **f32)        # z_proj terms
            zbar_s = torch.zeros(B, plan.voxels, **f32)      # vertical transform, with respect to the shifted heights
            wbar = torch.zeros(B, plan.voxels, **f32)        # depth weights
            dfac_col = torch.zeros(B, H * W, **f32) if ht["v0"] is not None else None
        else:
            zbar = zbar_s = wbar = dfac_col = None

        def gnback(t, rec, norm, scale, act):
            a, b, stats = rec
            gb = _lib.FfnoGnBack()
            gb.t, gb.stats, gb.gamma, gb.aff_a, gb.aff_b = _p(t), _p(stats), _p(norm[0]), _p(a), _p(b)
            gb.scale = _p(scale)
            gb.act_gelu = act
            gb.sums = ctypes.c_void_p(0)
            gb.depth_w = _p(plan.depth_w)
            gb.wbar = _p(wbar)
            return gb

        def head_bwd(h):
            _lib.call("ffno_head_bwd", rows, _p(h), _p(Wt.p1), _p(Wt.p2), _p(A["p2t"]), _p(A["p1t"]),
                      _p(Wt.p2_bias), _p(x), ctypes.byref(Wt.head), _p(ybar), _p(g), _p(xbar), st)

        if saved.head is not None:
            head_bwd(saved.head)
        for i in reversed(range(L)):
            blk, adj = Wt.blocks[i], A["blocks"][i]
            last_unsaved = i == L - 1 and saved.head is None
            bk = saved.blocks.pop(i, None)                  # kept whole: released when this block's adjoint is done
            if bk is not None:
                h, s, v, sp, vp, fp, pp, mp, x1 = (bk[k] for k in ("h", "s", "v", "sp", "vp", "fp", "pp", "mp", "x1"))
                gates, gn = bk["gates"], bk["gn"]
                v1y = bk["v1y"] if bk["v1y"] is not None else scr["v1y"]
                chunk, gates_c = 0, None
            else:
                wk = working()
                s, v, sp, vp, fp, pp, mp, x1 = (wk[k] for k in ("s", "v", "sp", "vp", "fp", "pp", "mp", "x1"))
                tbuf, gates, chunk, gates_c = wk["tbuf"], wk["gates"], wk["chunk"], wk["gates_c"]
                base = (i // stride) * stride
                prev = _asm(None, saved.inputs[base], ones, zeros)
                if i == base:
                    h, x_out = saved.inputs[base], None
                else:
                    # forward from the nearest saved input up to block i's input (x1 as the running buffer)
                    for j in range(base, i):
                        hj = saved.inputs[base] if j == base else x1
                        prev, _ = self._block(plan, j, prev, None if j == base else x1, hj, x1, s, v, scr, bz)
                    h, x_out = tbuf, tbuf
                # recompute block i: branch outputs stay in s and v, the vertical spectrum after the mix in v1y
                bk = dict(h=h, s=s, v=v, sp=sp, vp=vp, fp=fp, pp=pp, mp=mp, x1=x1, gates=gates, v1y=scr["v1y"])
                gn = self._block_keep(plan, i, prev, x_out, bk, scr, bz, g if last_unsaved else None)
                v1y = scr["v1y"]
            pa, pb, _ = gn["p"]
            if last_unsaved:
                # the head input x2 + res_mlp GN(mp), assembled in g by the forward head's own code (bit for bit what
                # the forward fed the head), then the head's adjoint in place
                ma, mb, _ = gn["m"]
                _lib.call("ffno_head", rows, _asm(g, mp, ma, mb, blk["res_mlp"]), _p(g), _p(Wt.p1), _p(Wt.p2),
                          _p(Wt.p2_bias), _p(x), ctypes.byref(Wt.head), None, st)
                head_bwd(g)
            # The pw GroupNorm's gradient sums are accumulated by the mlp adjoint, which writes its u_bar (x2_bar);
            # the others (behind a GELU, or fed from the block above) take a pass of their own. keep holds the sums
            # tensors until their consumers are enqueued.
            keep = []
            gb_m = gnback(mp, gn["m"], blk["norm_mlp"], blk["res_mlp"], 0)
            gb_p = gnback(pp, gn["p"], blk["norm_pw"], blk["res_pw"], 0)
            gb_f = gnback(fp, gn["f"], blk["norm_fuse"], blk["res_fused"], 1)
            gb_s = gnback(sp, gn["s"], blk["norm_spec"], None, 1)
            gb_v = gnback(vp, gn["v"], blk["norm_vert"], None, 1)
            # mlp: out = x2 + res_mlp GN(mp), x2 = x1 + res_pw GN(pp)
            keep.append(self._gn_sums(plan, gb_m, g, part2))
            _lib.call("ffno_mlp_bwd", rows, _p(g), gb_m, _asm(x1, pp, pa, pb, blk["res_pw"]), _p(blk["m1"]),
                      _p(adj["m2t"]), _p(adj["m1t"]), 1, _p(g), ctypes.byref(gb_p), _p(part2), st)
            keep.append(self._gn_finalize_sums(plan, gb_p, part2))
            # pw: x2 = x1 + res_pw GN(pp)
            _lib.call("ffno_mlp_bwd", rows, _p(g), gb_p, _asm(x1), _p(blk["pw1"]), _p(adj["pw2t"]), _p(adj["pw1t"]), 1,
                      _p(g), None, None, st)
            keep.append(self._gn_sums(plan, gb_f, g, part2))
            # fusion: x1 = h + res_fused gelu(GN(fp)); s_bar -> mp, v_bar -> pp
            sa, sb_, _ = gn["s"]
            va, vb, _ = gn["v"]
            if gates is not None:
                # two LDS-staged passes over the stored gates: [sm_bar, vm_bar] -> mp, pp; then in place, the second
                # also accumulating the branch GroupNorms' gradient sums (their u_bar are the s_bar, v_bar it writes)
                _lib.call("ffno_fuse_mlp_bwd", rows, _p(g), gb_f, _p(sp), _p(vp), _p(sa), _p(sb_), _p(va), _p(vb),
                          _p(gates), _p(blk["f1"]), _p(adj["f2t"]), _p(adj["f1t"]), _p(mp), _p(pp), st)
                _lib.call("ffno_fuse_gates_bwd", rows, _p(g), _p(h), _p(sp), _p(vp), _p(sa), _p(sb_), _p(va), _p(vb),
                          _p(gates), _p(blk["bg1"]), _p(adj["bg2t"]), _p(adj["bg1t"]), _p(mp), _p(pp),
                          ctypes.byref(gb_s), _p(part2), ctypes.byref(gb_v), _p(scr["part3"]), st)
                keep.append(self._gn_finalize_sums(plan, gb_s, part2))
                keep.append(self._gn_finalize_sums(plan, gb_v, scr["part3"]))
            else:
                if gates_c is not None:
                    self._fuse_bwd_chunked(plan, chunk, gates_c, blk, adj, gb_f, g, h, sp, vp, fp, mp, pp, x1,
                                           (sa, sb_, va, vb), wbar, scr, st)
                else:
                    _lib.call("ffno_fuse_bwd", rows, _p(g), gb_f, _p(h), _p(sp), _p(vp), _p(sa), _p(sb_), _p(va),
                              _p(vb), _p(blk["bg1"]), _p(blk["bg2"]), _p(blk["bg2_bias"]), _p(blk["f1"]),
                              _p(adj["f2t"]), _p(adj["f1t"]), _p(adj["bg2t"]), _p(adj["bg1t"]), _p(g), _p(mp), _p(pp),
                              st)
                keep.append(self._gn_sums(plan, gb_s, mp, part2))
                keep.append(self._gn_sums(plan, gb_v, pp, scr["part3"]))
            tables = plan.tables(i, adjoint=True)         # both branches' adjoint tables, built once per block
            # horizontal branch: spec = gelu(GN_s(post(Spec(xs)))); ys_bar -> s, xs_bar -> mp
            _lib.call("ffno_mlp_bwd", rows, _p(mp), gb_s, _asm(s), _p(blk["spec.p1"]), _p(adj["spec.p2t"]),
                      _p(adj["spec.p1t"]), 0, _p(s), None, None, st)
            self.horizontal(plan, s, mp, i, scr, tables, adjoint=True)
            # vertical branch: yv_bar -> v, the spectrum gradient in v1, xv_bar -> pp
            _lib.call("ffno_mlp_bwd", rows, _p(pp), gb_v, _asm(v), _p(blk["vertical.p1"]), _p(adj["vertical.p2t"]),
                      _p(adj["vertical.p1t"]), 0, _p(v), None, None, st)
            v1 = scr["v1"]
            if want_z and ht["v0"] is not None:
                # taper: dL/df_m per column = sum_c,part (V0^T yv_bar) * Y, before the mix adjoint overwrites v1
                self._vgemm(plan, ht["v0"], v, v1, True)
                gsum = (v1.view(B, 2, M, H * W, C) * v1y.view(B, 2, M, H * W, C)).sum((1, 4))   # [B, M, HW]
                if plan.percol:
                    dfac_col += (gsum.transpose(1, 2) * ht["dfac"].view(B, H * W, M)).sum(2)
                else:
                    dfac_col += (gsum * ht["dfac"][None, :, None]).sum(1)
            self.vertical(plan, v, v1, pp, i, tables, adjoint=True)
            if want_z:
                # the vertical branch's input again, into fp (dead here)
                _lib.call("ffno_block_in", rows, _asm(None, h, ones, zeros), None, _p(blk["spec.g1"]),
                          _p(blk["spec.g2"]), _p(blk["vertical.g1"]), _p(blk["vertical.g2"]), _p(blk["z_proj"]),
                          _p(bz), None, _p(fp), st)
                if "columns" in ht:
                    # Filon weights and synthesis phases per column: zbar_s[d] += sum_delta sum_r G_delta[d, r]
                    # (xv X_bar^T)[d + delta, r] + sum_r V_dz[d, r] (yv_bar Y^T)[d, r]
                    tab, rp, dp = ht["columns"]
                    zg = _lib.FfnoZGrad(B, H * W, D, 2 * M, C, rp, dp, _p(fp), _p(v1), _p(v), _p(v1y), _p(tab),
                                        _p(zbar_s))
                    _lib.call("ffno_zgrad_columns", ctypes.byref(zg), st)
                else:
                    # Filon weights: zbar_s[e] += sum_delta sum_c xv[e + delta] (G_delta X_bar)[e]   (x1 as scratch)
                    for delta in (-1, 0, 1):
                        self._vgemm(plan, ht["bands"][delta], v1, x1, False)
                        _lib.call("ffno_dot_reduce", rows, D, _p(fp), delta, _p(x1), _p(zbar_s), st)
                    # synthesis phases: zbar_s[d] += sum_c yv_bar[d] (V_dz Y)[d]
                    self._vgemm(plan, ht["v_dz"], v1y, x1, False)
                    _lib.call("ffno_dot_reduce", rows, D, _p(v), 0, _p(x1), _p(zbar_s), st)
            # input gates (and z_proj): h_bar in g
            _lib.call("ffno_gates_bwd", rows, _p(h), _p(mp), _p(pp), _p(g), _p(blk["spec.g1"]), _p(blk["spec.g2"]),
                      _p(adj["spec.g2t"]), _p(adj["spec.g1t"]), _p(blk["vertical.g1"]), _p(blk["vertical.g2"]),
                      _p(adj["vertical.g2t"]), _p(adj["vertical.g1t"]), _p(blk["z_proj"]), _p(g), _p(zbar), st)
        # lift: h0 = gelu(GN(W_lift x))
        _lib.call("ffno_lift", rows, _p(x), _p(Wt.lift), _p(s), FfnoStatsOut(_p(plan.depth_w), _p(scr["part"])), st)
        g0 = self._gn(plan, scr["part"], Wt.lift_gn, dev)
        gb = gnback(s, g0, Wt.lift_gn, None, 1)
        sums = self._gn_sums(plan, gb, g, part2)
        _lib.call("ffno_lift_bwd", rows, _p(g), gb, _p(Wt.lift), _p(xbar), st)
        del sums
        if not want_z:
            return xbar, None
        return xbar, self._height_gradient(plan, z, zbar, zbar_s, wbar, dfac_col)

    def _fuse_chunk(self, plan):
        if self.fuse_chunk is False:
            return 0
        if self.fuse_chunk:
            return T.round_up(int(self.fuse_chunk), 16)
        return max(16, FUSE_CHUNK_BYTES // (8 * plan.C) // 16 * 16)

    def _fuse_bwd_chunked(self, plan, chunk, gates_c, blk, adj, gb_f, g, h, sp, vp, fp, mp, pp, x1, aff, wbar, scr,
                          st):
        """The two-pass branch-fusion adjoint of one sample, chunk by chunk over its voxels, for when a whole block's
        gate buffer does not fit: each chunk's gates are recomputed by the fusion forward (its output goes to x1 and
        its statistics to the forward's partials buffer, neither read here; every voxel's result is its own, so the
        gates equal the whole block's), then both adjoint passes run on the chunk. Row tensors are addressed by
        offset; the GroupNorm's per-sample sums and affines are shared by every chunk."""
        C = plan.C
        sa, sb, va, vb = aff
        row = lambda t, v0: t.view(-1)[v0 * C:]          # noqa: E731
        for v0 in range(0, plan.voxels, chunk):
            cv = min(chunk, plan.voxels - v0)
            rc = FfnoRows(1, cv, C, max(1, min(plan.nwg, -(-((cv + 15) // 16) // ROW_TILES_PER_WG))))
            dw = plan.depth_w.view(-1)[v0:]
            _lib.call("ffno_fuse", rc, _p(row(h, v0)), _p(row(sp, v0)), _p(row(vp, v0)), _p(sa), _p(sb), _p(va),
                      _p(vb), _p(blk["bg1"]), _p(blk["bg2"]), _p(blk["bg2_bias"]), _p(blk["f1"]), _p(blk["f2"]),
                      _p(row(x1, v0)), FfnoStatsOut(_p(dw), _p(scr["part"])), _p(gates_c), st)
            gc = _lib.FfnoGnBack.from_buffer_copy(gb_f)
            gc.t, gc.depth_w = _p(row(fp, v0)), _p(dw)
            gc.wbar = _p(wbar.view(-1)[v0:]) if wbar is not None else None
            _lib.call("ffno_fuse_mlp_bwd", rc, _p(row(g, v0)), gc, _p(row(sp, v0)), _p(row(vp, v0)), _p(sa), _p(sb),
                      _p(va), _p(vb), _p(gates_c), _p(blk["f1"]), _p(adj["f2t"]), _p(adj["f1t"]), _p(row(mp, v0)),
                      _p(row(pp, v0)), st)
            _lib.call("ffno_fuse_gates_bwd", rc, _p(row(g, v0)), _p(row(h, v0)), _p(row(sp, v0)), _p(row(vp, v0)),
                      _p(sa), _p(sb), _p(va), _p(vb), _p(gates_c), _p(blk["bg1"]), _p(adj["bg2t"]), _p(adj["bg1t"]),
                      _p(row(mp, v0)), _p(row(pp, v0)), None, None, None, None, st)

    @staticmethod
    def _height_gradient(plan, z, zbar, zbar_s, wbar, dfac_col):
        """Assemble dL/dz [B, D, H, W] from the per-voxel pieces: zbar (z_proj), zbar_s (the vertical transform,
        taken with respect to the shifted heights z - min z of each column), wbar (the depth weights, through
        w = t / mean_d t with t the trapezoid |dz| of models/ffno_model.py:28-34), and the taper term."""
        B, D, H, W = plan.B, plan.D, plan.H, plan.W
        zs = zbar_s.view(B, D, H * W)
        ht = plan.height_tables()
        if dfac_col is not None:
            i, sgn = ht["i_med"], ht["sign"]
            if plan.percol:
                i = i.view(B, 1, H * W)
                delta = (sgn.view(B, H * W) * dfac_col)[:, None, :]
                zs.scatter_add_(1, i + 1, delta)
                zs.scatter_add_(1, i, -delta)
            else:
                zs[:, i + 1] += sgn * dfac_col
                zs[:, i] -= sgn * dfac_col
        # the per-column shift z - amin(z): the minimum's level(s) take minus the column total, shared evenly
        total = zs.sum(1)
        out = zbar.view(B, D, H * W) + zs
        if plan.percol:
            out -= ht["lowest"].view(B, H * W, D).transpose(1, 2) * total[:, None, :]
        else:
            for j in ht["argmin"]:
                out[:, int(j)] -= total / len(ht["argmin"])
        # depth weights: w_d = t_d / mean(t); t_d = |dz| halves of the intervals either side of level d
        zc = z.reshape(B, D, H * W).float()
        if D > 1:
            dzv = zc[:, 1:] - zc[:, :-1]
            t = torch.zeros_like(zc)
            t[:, :-1] += 0.5 * dzv.abs()
            t[:, 1:] += 0.5 * dzv.abs()
            mt = t.mean(1, keepdim=True).clamp_min(1e-12)
            wb = wbar.view(B, D, H * W)
            tb = wb / mt - (wb * t).sum(1, keepdim=True) / (D * mt * mt)
            g = 0.5 * torch.sign(dzv) * (tb[:, :-1] + tb[:, 1:])      # d/d(z_{d+1} - z_d) of both halves
            out[:, 1:] += g
            out[:, :-1] -= g
        return out.view(B, D, H, W)
