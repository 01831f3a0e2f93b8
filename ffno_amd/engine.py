"""The HIP engine of FFNO3D: packs a module's weights for the kernels, builds per-request plans (transform matrices,
per-mode gates and low-rank tables, depth weights) and runs the forward schedule through libffno_hip.so.

Activations live channel-last ([B, D, H, W, C] "voxel rows") in float32 HBM; every pointwise layer runs as bf16 MFMA
with float32 accumulation, every transform as a split-bf16 GEMM. The reference forward this replaces is
models/ffno_model.py:800-832 (eval mode, no dropout).
"""
import ctypes
import math

import numpy as np
import torch
import torch.nn.functional as F

from . import _lib
from . import transforms as T
from ._lib import FfnoAssemble, FfnoHeadDesc, FfnoMix, FfnoRows, FfnoStatsOut, FfnoTwGemm

SUPPORTED_WIDTHS = (16, 32, 64, 96, 128)
GROUPS = 8
NW = 4                       # waves per workgroup of the voxel-row kernels

# MULTI3D constants in cgs (models/ffno_model.py:111-113)
_EE, _HH, _EM, _BK, _PI = 1.602189e-12, 6.626176e-27, 9.109534e-28, 1.380662e-16, 3.14159265359
_LN_CCON = math.log(0.5 * (_HH / math.sqrt(2.0 * _PI * _EM) / math.sqrt(_BK)) ** 3)


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def pack_mlp_weight(w, segments=None, n_pad=None):
    """bf16 [.., N_pad, K_pad] for the MFMA kernels from a float weight [.., N, K] (PyTorch [out, in]): each input
    segment padded to a multiple of 32, k permuted within every 32-block as the kernels' lane layout wants
    (k = 32s + 16h + 4q + j is stored at 32s + 8q + 4h + j), N padded to a multiple of 16."""
    w = w.float()
    lead = w.shape[:-2]
    N, K = w.shape[-2:]
    segs = segments or [K]
    assert sum(segs) == K
    parts, off = [], 0
    for s in segs:
        part = w[..., off:off + s]
        off += s
        parts.append(F.pad(part, (0, T.round_up(s, 32) - s)))
    wp = torch.cat(parts, -1)
    npad = n_pad or T.round_up(N, 16)
    wp = F.pad(wp, (0, 0, 0, npad - N))
    kp = wp.shape[-1]
    wp = wp.reshape(*lead, npad, kp // 32, 2, 4, 4).transpose(-3, -2).reshape(*lead, npad, kp)
    return wp.to(torch.bfloat16).contiguous().view(torch.int16)


def _conv_w(p):
    return p.detach().reshape(p.shape[0], -1)


def _gelu(x):
    return 0.5 * x * (1.0 + torch.erf(x * (1.0 / math.sqrt(2.0))))


def _freq_mlp(prefix, sd, feats):
    h = _gelu(F.linear(feats, sd[prefix + "freq_mlp.0.weight"], sd[prefix + "freq_mlp.0.bias"]))
    h = _gelu(F.linear(h, sd[prefix + "freq_mlp.2.weight"], sd[prefix + "freq_mlp.2.bias"]))
    return F.linear(h, sd[prefix + "freq_mlp.4.weight"], sd[prefix + "freq_mlp.4.bias"])


def _freq_lowrank(prefix, sd, feats, C, r):
    h = _gelu(F.linear(feats, sd[prefix + "freq_lowrank.0.weight"], sd[prefix + "freq_lowrank.0.bias"]))
    h = _gelu(F.linear(h, sd[prefix + "freq_lowrank.2.weight"], sd[prefix + "freq_lowrank.2.bias"]))
    u = F.linear(h, sd[prefix + "freq_u.weight"], sd[prefix + "freq_u.bias"]).reshape(-1, C, r)
    v = F.linear(h, sd[prefix + "freq_v.weight"], sd[prefix + "freq_v.bias"]).reshape(-1, r, C)
    return u, v


class PackedWeights:
    """The module's pointwise weights in kernel form, rebuilt whenever a parameter changes."""

    def __init__(self, module, device):
        C = module.width
        sd = {k: v.detach().to(device) for k, v in module.state_dict().items()}
        self.sd = sd
        self.lift = _conv_w(sd["lift.0.weight"]).float().contiguous()          # [C, 6]
        self.lift_gn = (sd["lift.1.weight"].float().contiguous(), sd["lift.1.bias"].float().contiguous())
        self.blocks = []
        for i in range(module.n_layers):
            p = f"blocks.{i}."
            b = {}
            for br in ("spec", "vertical"):
                b[br + ".g1"] = pack_mlp_weight(_conv_w(sd[p + br + ".input_gate.0.weight"]))
                b[br + ".g2"] = pack_mlp_weight(_conv_w(sd[p + br + ".input_gate.2.weight"]))
                b[br + ".p1"] = pack_mlp_weight(_conv_w(sd[p + br + ".post.0.weight"]))
                b[br + ".p2"] = pack_mlp_weight(_conv_w(sd[p + br + ".post.2.weight"]))
                wr, wi = sd[p + br + ".weight_real"].float(), sd[p + br + ".weight_imag"].float()   # [in, out]
                wc = torch.cat([torch.cat([wr.t(), -wi.t()], 1), torch.cat([wi.t(), wr.t()], 1)], 0)
                b[br + ".wc"] = pack_mlp_weight(wc, segments=[C, C])
            b["z_proj"] = sd[p + "vertical.z_proj.weight"].reshape(C).float().contiguous()
            b["bg1"] = pack_mlp_weight(_conv_w(sd[p + "branch_gate.0.weight"]), segments=[C, C, C])
            b["bg2"] = pack_mlp_weight(_conv_w(sd[p + "branch_gate.2.weight"]))
            b["bg2_bias"] = sd[p + "branch_gate.2.bias"].float().contiguous()
            b["f1"] = pack_mlp_weight(_conv_w(sd[p + "fuse.0.weight"]), segments=[C, C])
            b["f2"] = pack_mlp_weight(_conv_w(sd[p + "fuse.2.weight"]))
            b["pw1"] = pack_mlp_weight(_conv_w(sd[p + "pw.0.weight"]))
            b["pw2"] = pack_mlp_weight(_conv_w(sd[p + "pw.2.weight"]))
            b["m1"] = pack_mlp_weight(_conv_w(sd[p + "mlp.net.0.weight"]))
            b["m2"] = pack_mlp_weight(_conv_w(sd[p + "mlp.net.3.weight"]))
            for n in ("norm_spec", "norm_vert", "norm_fuse", "norm_pw", "norm_mlp"):
                b[n] = (sd[p + n + ".weight"].float().contiguous(), sd[p + n + ".bias"].float().contiguous())
            for n in ("res_fused", "res_pw", "res_mlp"):
                b[n] = sd[p + n].float().contiguous()
            self.blocks.append(b)
        self.p1 = pack_mlp_weight(_conv_w(sd["proj1.weight"]))
        self.p2 = pack_mlp_weight(_conv_w(sd["proj2.weight"]))
        nh = sd["proj2.weight"].shape[0]
        self.p2_bias = F.pad(sd["proj2.bias"].float(), (0, T.round_up(nh, 16) - nh)).contiguous()
        self.head = self._head_desc(module, sd)

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


class Plan:
    """Everything that depends on the request's shape, heights and spacings but not on its data: transform
    matrices (split-bf16), per-mode gates and low-rank tables for every block, depth weights."""

    def __init__(self, engine, B, D, H, W, z, dx, dy):
        m = engine.module
        dev = engine.device
        C, r = m.width, m.freq_mix_rank
        self.B, self.D, self.H, self.W, self.C, self.rank = B, D, H, W, C, r
        self.wf = W // 2 + 1
        self.voxels = D * H * W
        # depth weights (models/ffno_model.py:28-34, :67-74), float32 as the reference computes them
        zc = z.to(dev, torch.float32).reshape(B, 1, D, H, W)
        if D == 1:
            tw = torch.ones_like(zc)
        else:
            half = 0.5 * (zc[:, :, 1:] - zc[:, :, :-1]).abs()
            tw = F.pad(half, (0, 0, 0, 0, 0, 1)) + F.pad(half, (0, 0, 0, 0, 1, 0))
        self.depth_w = (tw / tw.mean(dim=2, keepdim=True).clamp_min(1e-12)).reshape(B, -1).contiguous()
        self.z = zc.reshape(B, -1).contiguous()
        # vertical grid: the reference's shared-grid decision (models/ffno_model.py:267-269), made on the host once
        zs = zc - zc.amin(dim=2, keepdim=True)
        shared = bool((zs == zs[..., :1, :1]).all()) and bool((zs == zs[:1]).all())
        if not shared:
            raise NotImplementedError("per-column vertical grids are not implemented in the HIP engine; "
                                      "every column must share its heights (after the per-column shift)")
        column = zs[0, 0, :, 0, 0].double().cpu().numpy()
        span = float((zc.amax(dim=2) - zc.amin(dim=2)).max())
        self.length, k, factor = T.vertical_box(span, m.z_reference_length, m.z_min_wavelength,
                                                int(math.floor(m.z_reference_length / m.z_min_wavelength)) + 1)
        self.modes = len(k)
        v_fwd, v_inv = T.vertical_matrices(column, k, factor, self.length)
        hm = T.horizontal_matrices(H, W)
        self.tw = {name: T.split_bf16(a, dev) for name, a in
                   (("w_fwd", hm["w_fwd"]), ("h_fwd", hm["h_fwd"]), ("h_inv", hm["h_inv"]), ("w_inv", hm["w_inv"]),
                    ("v_fwd", v_fwd), ("v_inv", v_inv))}
        # per-mode gates and low-rank tables (models/ffno_model.py:375-392, :443-447, :231-237, :285-296), float32
        sd = engine.weights.sd
        ky = torch.fft.fftfreq(H, d=float(dy), device=dev)
        kx = torch.fft.rfftfreq(W, d=float(dx), device=dev)
        ky, kx = torch.meshgrid(ky, kx, indexing="ij")
        ky, kx = ky * 1e5, kx * 1e5
        hfeat = torch.stack([kx, ky, torch.sqrt(kx ** 2 + ky ** 2 + 1e-12), torch.sign(ky)], -1).reshape(-1, 4)
        kt = torch.from_numpy(k).float().to(dev)
        qf = kt * m.z_feature_wavelength / (2 * math.pi)
        tf = qf / (1 + qf)
        vfeat = torch.stack([tf, tf * tf], -1)
        self.mix = []
        for i in range(m.n_layers):
            entry = {}
            for br, feats in (("spec", hfeat), ("vertical", vfeat)):
                prefix = f"blocks.{i}.{br}."
                gate = _freq_mlp(prefix, sd, feats).float().contiguous()
                e = dict(gate=gate, v=None, u=None)
                if r:
                    u, v = _freq_lowrank(prefix, sd, feats, C, r)
                    e["v"] = pack_mlp_weight(v, n_pad=16)                      # [modes, 16, KS*32]
                    e["u"] = pack_mlp_weight(u)                                # [modes, C, 32]
                entry[br] = e
            self.mix.append(entry)
        nt = (self.voxels + 15) // 16
        self.nwg = max(1, min((nt + NW - 1) // NW, max(1, 2048 // B)))
        self.rows = FfnoRows(B, self.voxels, C, self.nwg)


class Engine:
    """Runs FFNO3D forward on the GPU for one module (ffno_amd.model.FFNO3D)."""

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

    def _param_version(self):
        return tuple((p.data_ptr(), p._version) for p in self.module.parameters()) + \
            tuple((b.data_ptr(), b._version) for b in self.module.buffers())

    def sync(self, device):
        ver = self._param_version()
        if self.weights is None or ver != self._version or device != self.device:
            self.device = device
            self.weights = PackedWeights(self.module, device)
            self._version = ver
            self._plans.clear()

    def plan(self, B, D, H, W, z, dx, dy):
        key = (B, D, H, W, float(dx), float(dy), z.data_ptr(), z._version, tuple(z.shape))
        p = self._plans.get(key)
        if p is None:
            if len(self._plans) > 16:
                self._plans.clear()
            p = Plan(self, B, D, H, W, z, dx, dy)
            self._plans[key] = p
        return p

    # ------------------------------------------------------------------------------------------------ launches
    @staticmethod
    def _stream():
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    def _gn(self, plan, partials, gn, dev):
        a = torch.empty(plan.B, plan.C, device=dev)
        b = torch.empty(plan.B, plan.C, device=dev)
        stats = torch.empty(plan.B, GROUPS, 2, device=dev)
        _lib.call("ffno_gn_finalize", plan.B, plan.nwg, plan.C, _p(partials), _p(gn[0]), _p(gn[1]), _p(stats),
                  _p(a), _p(b), self._stream())
        return a, b, stats

    def _tw(self, mat, x, y, batch, n, bmap_x, bmap_y, kmap, mmap, accumulate=0):
        hi, lo, m, k, mp, kp = mat
        g = FfnoTwGemm()
        g.a_hi, g.a_lo = _p(hi), _p(lo)
        g.m, g.k, g.n, g.m_pad, g.k_pad, g.batch = m, k, n, mp, kp, batch
        g.x, g.y = _p(x), _p(y)
        g.b_split, g.x_b_hi, g.x_b_lo = bmap_x
        g.y_b_hi, g.y_b_lo = bmap_y
        g.k_split, g.x_k_hi, g.x_k_lo = kmap
        g.m_split, g.y_m_hi, g.y_m_lo = mmap
        g.accumulate = accumulate
        _lib.call("ffno_tw_gemm", ctypes.byref(g), self._stream())

    def _mix(self, plan, re, im, rows, modes, mode_stride, rmap, wc, e):
        mx = FfnoMix()
        mx.re, mx.im = _p(re), _p(im)
        mx.C, mx.rows, mx.modes, mx.rank = plan.C, rows, modes, plan.rank
        mx.mode_stride = mode_stride
        mx.r_split, mx.r_hi, mx.r_lo = rmap
        mx.wc, mx.gate = _p(wc), _p(e["gate"])
        mx.v_p, mx.u_p = _p(e["v"]), _p(e["u"])
        mx.factor = ctypes.c_void_p(0)
        _lib.call("ffno_mix", ctypes.byref(mx), self._stream())

    def horizontal(self, plan, xs, s1, s2, out, i):
        """rfft2 -> mix -> irfft2 of xs [B,D,H,W,C] into out (models/ffno_model.py:441-448)."""
        B, D, H, W, C, wf = plan.B, plan.D, plan.H, plan.W, plan.C, plan.wf
        plane = 2 * H * wf * C
        big = 1 << 40
        # rfft along W: one GEMM per voxel row (b, d, h)
        self._tw(plan.tw["w_fwd"], xs, s1, B * D * H, C, (H, H * W * C, W * C), (plane, wf * C),
                 (big, 0, C), (wf, H * wf * C, C))
        # DFT along H: one GEMM per plane (b, d) over (part, h) rows, columns (kx, c)
        self._tw(plan.tw["h_fwd"], s1, s2, B * D, wf * C, (big, 0, plane), (0, plane),
                 (H, H * wf * C, wf * C), (H, H * wf * C, wf * C))
        e = plan.mix[i]["spec"]
        self._mix(plan, s2, s2[H * wf * C:], B * D, H * wf, C, (big, 0, plane), self.weights.blocks[i]["spec.wc"], e)
        self._tw(plan.tw["h_inv"], s2, s1, B * D, wf * C, (big, 0, plane), (0, plane),
                 (H, H * wf * C, wf * C), (H, H * wf * C, wf * C))
        # c2r along W
        self._tw(plan.tw["w_inv"], s1, out, B * D * H, C, (H, plane, wf * C), (H * W * C, W * C),
                 (wf, H * wf * C, C), (big, 0, C))

    def vertical(self, plan, xv, v1, out, i):
        """Filon transform along D -> mix -> synthesis at the levels, /L folded in (models/ffno_model.py:270-305)."""
        B, D, C, M = plan.B, plan.D, plan.C, plan.modes
        hwc = plan.H * plan.W * C
        big = 1 << 40
        self._tw(plan.tw["v_fwd"], xv, v1, B, hwc, (big, 0, D * hwc), (0, 2 * M * hwc),
                 (big, 0, hwc), (M, M * hwc, hwc))
        e = plan.mix[i]["vertical"]
        hw = plan.H * plan.W
        self._mix(plan, v1, v1[M * hwc:], B * hw, M, hwc, (hw, 2 * M * hwc, C), self.weights.blocks[i]["vertical.wc"],
                  e)
        self._tw(plan.tw["v_inv"], v1, out, B, hwc, (big, 0, 2 * M * hwc), (0, D * hwc),
                 (M, M * hwc, hwc), (big, 0, hwc))

    def forward(self, x, z, dx, dy, keep=None):
        """Normalised log10 populations [B, Cout, D, H, W] from x [B, 6, D, H, W] and z_scale [B, D, H, W] (Mm).
        keep: a list that receives every block input (channel-last) for the adjoint."""
        if not x.is_cuda:
            raise RuntimeError("the FFNO3D HIP engine runs on the GPU only; move the inputs to a cuda device")
        dev = x.device
        self.sync(dev)
        Wt = self.weights
        B, cin, D, H, W = x.shape
        if cin != 6:
            raise ValueError("FFNO3D expects 6 input channels")
        z = z.reshape(B, D, H, W)
        plan = self.plan(B, D, H, W, z, dx, dy)
        C = plan.C
        N = B * plan.voxels
        x = x.contiguous().float()
        st = self._stream()
        f32 = dict(device=dev, dtype=torch.float32)
        part = torch.empty(B, plan.nwg, GROUPS, 3, **f32)
        stats = FfnoStatsOut(_p(plan.depth_w), _p(part))
        rows = plan.rows
        t0 = torch.empty(N, C, **f32)
        _lib.call("ffno_lift", rows, _p(x), _p(Wt.lift), _p(t0), stats, st)
        a, b, _ = self._gn(plan, part, Wt.lift_gn, dev)
        prev = FfnoAssemble(None, _p(t0), _p(a), _p(b), None, 1)
        wf, M = plan.wf, plan.modes
        xs = torch.empty(N, C, **f32)
        xv = torch.empty(N, C, **f32)
        s1 = torch.empty(B * D * 2 * H * wf * C, **f32)
        s2 = torch.empty_like(s1)
        v1 = torch.empty(B * 2 * M * H * W * C, **f32)
        sp = torch.empty(N, C, **f32)
        vp = torch.empty(N, C, **f32)
        fp = torch.empty(N, C, **f32)
        pp = torch.empty(N, C, **f32)
        x1 = torch.empty(N, C, **f32)
        x2 = torch.empty(N, C, **f32)
        xin_buf = torch.empty(N, C, **f32) if keep is None else None
        for i, blk in enumerate(Wt.blocks):
            # the block input never aliases x2, which the assembly reads (block i-1's x2 and mlp output)
            xin = torch.empty(N, C, **f32) if keep is not None else xin_buf
            _lib.call("ffno_block_in", rows, prev, _p(xin), _p(blk["spec.g1"]), _p(blk["spec.g2"]),
                      _p(blk["vertical.g1"]), _p(blk["vertical.g2"]), _p(blk["z_proj"]), _p(plan.z), _p(xs), _p(xv),
                      st)
            if keep is not None:
                keep.append(xin)
            self.horizontal(plan, xs, s1, s2, xs, i)
            _lib.call("ffno_post", rows, _p(xs), _p(blk["spec.p1"]), _p(blk["spec.p2"]), _p(sp), stats, st)
            sa, sb, _ = self._gn(plan, part, blk["norm_spec"], dev)
            self.vertical(plan, xv, v1, xv, i)
            _lib.call("ffno_post", rows, _p(xv), _p(blk["vertical.p1"]), _p(blk["vertical.p2"]), _p(vp), stats, st)
            va, vb, _ = self._gn(plan, part, blk["norm_vert"], dev)
            _lib.call("ffno_fuse", rows, _p(xin), _p(sp), _p(vp), _p(sa), _p(sb), _p(va), _p(vb), _p(blk["bg1"]),
                      _p(blk["bg2"]), _p(blk["bg2_bias"]), _p(blk["f1"]), _p(blk["f2"]), _p(fp), stats, st)
            fa, fb, _ = self._gn(plan, part, blk["norm_fuse"], dev)
            _lib.call("ffno_resmlp", rows, FfnoAssemble(_p(xin), _p(fp), _p(fa), _p(fb), _p(blk["res_fused"]), 1),
                      _p(x1), _p(blk["pw1"]), _p(blk["pw2"]), _p(pp), stats, st)
            pa, pb, _ = self._gn(plan, part, blk["norm_pw"], dev)
            mp = fp
            _lib.call("ffno_resmlp", rows, FfnoAssemble(_p(x1), _p(pp), _p(pa), _p(pb), _p(blk["res_pw"]), 0),
                      _p(x2), _p(blk["m1"]), _p(blk["m2"]), _p(mp), stats, st)
            ma, mb, _ = self._gn(plan, part, blk["norm_mlp"], dev)
            prev = FfnoAssemble(_p(x2), _p(mp), _p(ma), _p(mb), _p(blk["res_mlp"]), 0)
        out_ch = self.module.out_channels
        y = torch.empty(B, out_ch, D, H, W, **f32)
        _lib.call("ffno_head", rows, prev, _p(Wt.p1), _p(Wt.p2), _p(Wt.p2_bias), _p(x), ctypes.byref(Wt.head), _p(y),
                  st)
        return y
