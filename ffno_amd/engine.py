"""The HIP engine of FFNO3D: packs a module's weights for the kernels, builds per-request plans (transform matrices,
per-mode gates and low-rank tables, depth weights) and runs the forward schedule through libffno_hip.so.

Activations live channel-last ([B, D, H, W, C] "voxel rows") in float32 HBM; every pointwise layer runs as bf16 MFMA
with float32 accumulation, every transform as a split-bf16 GEMM. The reference forward this replaces is
models/ffno_model.py:800-832 (eval mode, no dropout).
"""
import ctypes
import math
import threading

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
    """Split-bf16 [.., 2, N_pad, K_pad] (hi, lo) for the MFMA kernels from a float weight [.., N, K] (PyTorch
    [out, in]): each input segment padded to a multiple of 32, k permuted within every 32-block as the kernels' lane
    layout wants (k = 32s + 16h + 4q + j is stored at 32s + 8q + 4h + j), N padded to a multiple of 16;
    hi = bf16(w), lo = bf16(w - hi)."""
    w = w.double()
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
    hi = wp.to(torch.bfloat16)
    lo = (wp - hi.double()).to(torch.bfloat16)
    return torch.stack([hi, lo], -3).contiguous().view(torch.int16)


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
                b[br + ".g1"] = pack_mlp_weight(_conv_w(sd[p + br + ".input_gate.0.weight"]))
                b[br + ".g2"] = pack_mlp_weight(_conv_w(sd[p + br + ".input_gate.2.weight"]))
                b[br + ".p1"] = pack_mlp_weight(_conv_w(sd[p + br + ".post.0.weight"]))
                b[br + ".p2"] = pack_mlp_weight(_conv_w(sd[p + br + ".post.2.weight"]))
                wr, wi = sd[p + br + ".weight_real"].float(), sd[p + br + ".weight_imag"].float()   # [in, out]
                # [Ar; Ai] = [[Wr^T, -Wi^T], [Wi^T, Wr^T]] [Yr; Yi]: the complex product Y (Wr + i Wi) over channels
                wc = torch.cat([torch.cat([wr.t(), -wi.t()], 1), torch.cat([wi.t(), wr.t()], 1)], 0)
                b[br + ".wc"] = pack_mlp_weight(wc, segments=[C, C])
                if br == "spec":
                    self.wc_real.append({})
                self.wc_real[-1][br] = wc
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


class Plan:
    """Everything that depends on the request's shape, heights and spacings but not on its data: transform
    matrices (split-bf16), per-mode gates and low-rank tables for every block, depth weights."""

    def __init__(self, engine, B, D, H, W, column, span, dx, dy):
        m = engine.module
        dev = engine.device
        C, r = m.width, m.freq_mix_rank
        self.B, self.D, self.H, self.W, self.C, self.rank = B, D, H, W, C, r
        self.wf = W // 2 + 1
        self.voxels = D * H * W
        column = np.asarray(column, np.float64)
        # depth weights (models/ffno_model.py:28-34, :67-74) of the shared column, in float32 as the reference has
        # them, broadcast to every voxel
        zc = torch.from_numpy(column).float()
        if D == 1:
            tw = torch.ones(1)
        else:
            half = 0.5 * (zc[1:] - zc[:-1]).abs()
            tw = F.pad(half, (0, 1)) + F.pad(half, (1, 0))
        w = tw / tw.mean().clamp_min(1e-12)
        self.depth_w = w.to(dev)[None, :, None].expand(B, D, H * W).reshape(B, -1).contiguous()
        self.length, k, factor = T.vertical_box(span, m.z_reference_length, m.z_min_wavelength,
                                                int(math.floor(m.z_reference_length / m.z_min_wavelength)) + 1)
        self.modes = len(k)
        v_fwd, v_inv = T.vertical_matrices(column, k, factor, self.length)
        hm = T.horizontal_matrices(H, W)
        self.tw = {name: T.split_bf16(a, dev) for name, a in
                   (("w_fwd", hm["w_fwd"]), ("h_fwd", hm["h_fwd"]), ("h_inv", hm["h_inv"]), ("w_inv", hm["w_inv"]),
                    ("v_fwd", v_fwd), ("v_inv", v_inv))}
        # the adjoint chain runs the same GEMMs with the transposed matrices in reverse order
        self._mats = dict(w_fwd=hm["w_fwd"], h_fwd=hm["h_fwd"], h_inv=hm["h_inv"], w_inv=hm["w_inv"], v_fwd=v_fwd,
                          v_inv=v_inv)
        self._tw_adj = None
        self._column, self._k, self._factor = column, k, factor
        self._height = None
        self._lock = threading.Lock()
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
                e = dict(gate=gate, v=None, u=None, u_raw=None, v_raw=None)
                if r:
                    u, v = _freq_lowrank(prefix, sd, feats, C, r)
                    e["v"] = pack_mlp_weight(v, n_pad=16)                      # [modes, 2, 16, KS*32]
                    e["u"] = pack_mlp_weight(u)                                # [modes, 2, C, 32]
                    e["u_raw"], e["v_raw"] = u, v
                entry[br] = e
            self.mix.append(entry)
        nt = (self.voxels + 15) // 16
        self.nwg = max(1, min((nt + NW - 1) // NW, max(1, 2048 // B)))
        self.rows = FfnoRows(B, self.voxels, C, self.nwg)
        self.device = dev

    def height_tables(self):
        """Height derivatives of the vertical transforms for the z-gradient: the three bands of the Filon weights'
        Jacobian, the synthesis' derivative, and for the taper (non-zero only for modes between 2 pi / (3 dz) and
        pi / dz) the transposed unscaled synthesis and the per-mode factor * d taper / d dz."""
        with self._lock:
            return self._height_tables()

    def _height_tables(self):
        if self._height is None:
            bands = T.filon_height_bands(self._column, self._k)
            v_dz, v0, dfac, i_med, sgn = T.vertical_height_terms(self._column, self._k, self._factor, self.length)
            col = np.asarray(self._column)
            ties = np.flatnonzero(col == col.min())
            self._height = dict(
                bands={d: T.split_bf16(bands[d], self.device) for d in (-1, 0, 1)},
                v_dz=T.split_bf16(v_dz, self.device),
                v0=T.split_bf16(np.ascontiguousarray(v0.T), self.device) if np.any(dfac != 0) else None,
                dfac=torch.from_numpy(dfac).float().to(self.device), i_med=i_med, sign=sgn, argmin=ties)
        return self._height

    def adjoint_tables(self):
        """Transposed transform matrices and adjoint mix tables: conj(g), and the low-rank roles swapped
        (V' = U^T, U' = V^T), so that the same kernels apply the transpose of every forward stage."""
        with self._lock:
            return self._adjoint_tables()

    def _adjoint_tables(self):
        if self._tw_adj is None:
            m = self._mats
            tables = {name + "_T": T.split_bf16(np.ascontiguousarray(a.T), self.device) for name, a in m.items()}
            for entry in self.mix:
                for e in entry.values():
                    g = e["gate"].clone()
                    g[:, 1] = -g[:, 1]
                    e["gate_adj"] = g.contiguous()
                    if e["u_raw"] is not None:
                        e["v_adj"] = pack_mlp_weight(e["u_raw"].transpose(-1, -2), n_pad=16)
                        e["u_adj"] = pack_mlp_weight(e["v_raw"].transpose(-1, -2))
                    else:
                        e["v_adj"] = e["u_adj"] = None
            self._tw_adj = tables                      # last: another thread may test it before the loop ends
        return self._tw_adj


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
        self._lock = threading.RLock()      # the serving workers share one engine

    def footprint_bytes(self, B, D, H, W, vjp=False):
        """HBM the engine needs for one request of this shape: float32 channel-last tensors of C channels counted
        in units of B*D*H*W*C*4 bytes (forward: 12 activation buffers and 2 half-spectrum planes; adjoint: the
        saved block inputs plus 24 recompute and gradient buffers), plus 10% for spectra padding and tables."""
        C, L = self.module.width, self.module.n_layers
        unit = B * D * H * W * C * 4
        units = (L + 1) + 24 if vjp else 14
        return int(unit * units * 1.1) + (64 << 20)

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
                self._plans.clear()

    def plan(self, B, D, H, W, z, dx, dy):
        """The plan for these heights, cached by content: the reference's shared-grid decision
        (models/ffno_model.py:267-269) and its box span (:252) are read on the host here, once per call, and every
        later decision uses them (no host synchronisation inside the launch sequence)."""
        zc = z.reshape(B, D, H * W).float()
        zs = zc - zc.amin(dim=1, keepdim=True)
        flags = torch.stack([(zs == zs[:, :, :1]).all() & (zs == zs[:1]).all(),
                             (zc.amax(dim=1) - zc.amin(dim=1)).max()])
        shared, span = flags.tolist()
        if not shared:
            raise NotImplementedError("per-column vertical grids are not implemented in the HIP engine; "
                                      "every column must share its heights (after the per-column shift)")
        column = zs[0, :, 0].double().cpu().numpy()
        key = (B, D, H, W, float(dx), float(dy), float(span), column.tobytes())
        with self._lock:
            p = self._plans.get(key)
            if p is None:
                if len(self._plans) > 64:
                    self._plans.pop(next(iter(self._plans)))
                p = Plan(self, B, D, H, W, column, span, dx, dy)
                self._plans[key] = p
        return p

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

    def _mix(self, plan, re, im, rows, modes, mode_stride, rmap, wc, e, adjoint=False):
        mx = FfnoMix()
        mx.re, mx.im = _p(re), _p(im)
        mx.C, mx.rows, mx.modes, mx.rank = plan.C, rows, modes, plan.rank
        mx.mode_stride = mode_stride
        mx.r_split, mx.r_hi, mx.r_lo = rmap
        sfx = "_adj" if adjoint else ""
        mx.wc, mx.gate = _p(wc), _p(e["gate" + sfx])
        mx.v_p, mx.u_p = _p(e["v" + sfx]), _p(e["u" + sfx])
        mx.factor = ctypes.c_void_p(0)
        _lib.call("ffno_mix", ctypes.byref(mx), self._stream())

    def horizontal(self, plan, xs, s1, s2, out, i, adjoint=False):
        """rfft2 -> mix -> irfft2 of xs [B,D,H,W,C] into out (models/ffno_model.py:441-448); with adjoint, the
        transpose of that linear map (W_fwd^T H_fwd^T Mix^T H_inv^T W_inv^T)."""
        B, D, H, W, C, wf = plan.B, plan.D, plan.H, plan.W, plan.C, plan.wf
        plane = 2 * H * wf * C
        big = 1 << 40
        if adjoint:
            ta = plan.adjoint_tables()
            m1, m2, m3, m4 = ta["w_inv_T"], ta["h_inv_T"], ta["h_fwd_T"], ta["w_fwd_T"]
            wc = self.weights.adjoint()["blocks"][i]["spec.wc"]
        else:
            m1, m2, m3, m4 = plan.tw["w_fwd"], plan.tw["h_fwd"], plan.tw["h_inv"], plan.tw["w_inv"]
            wc = self.weights.blocks[i]["spec.wc"]
        # along W, real -> half spectrum: one GEMM per voxel row (b, d, h)
        self._tw(m1, xs, s1, B * D * H, C, (H, H * W * C, W * C), (plane, wf * C), (big, 0, C), (wf, H * wf * C, C))
        # along H, complex: one GEMM per plane (b, d) over (part, h) rows, columns (kx, c)
        self._tw(m2, s1, s2, B * D, wf * C, (big, 0, plane), (0, plane), (H, H * wf * C, wf * C),
                 (H, H * wf * C, wf * C))
        self._mix(plan, s2, s2[H * wf * C:], B * D, H * wf, C, (big, 0, plane), wc, plan.mix[i]["spec"], adjoint)
        self._tw(m3, s2, s1, B * D, wf * C, (big, 0, plane), (0, plane), (H, H * wf * C, wf * C),
                 (H, H * wf * C, wf * C))
        # half spectrum -> real along W
        self._tw(m4, s1, out, B * D * H, C, (H, plane, wf * C), (H * W * C, W * C), (wf, H * wf * C, C), (big, 0, C))

    def vertical(self, plan, xv, v1, out, i, adjoint=False):
        """Filon transform along D -> mix -> synthesis at the levels, /L folded in (models/ffno_model.py:270-305);
        with adjoint, its transpose."""
        B, D, C, M = plan.B, plan.D, plan.C, plan.modes
        hwc = plan.H * plan.W * C
        big = 1 << 40
        if adjoint:
            ta = plan.adjoint_tables()
            m1, m2 = ta["v_inv_T"], ta["v_fwd_T"]
            wc = self.weights.adjoint()["blocks"][i]["vertical.wc"]
        else:
            m1, m2 = plan.tw["v_fwd"], plan.tw["v_inv"]
            wc = self.weights.blocks[i]["vertical.wc"]
        self._tw(m1, xv, v1, B, hwc, (big, 0, D * hwc), (0, 2 * M * hwc), (big, 0, hwc), (M, M * hwc, hwc))
        hw = plan.H * plan.W
        self._mix(plan, v1, v1[M * hwc:], B * hw, M, hwc, (hw, 2 * M * hwc, C), wc, plan.mix[i]["vertical"], adjoint)
        self._tw(m2, v1, out, B, hwc, (big, 0, 2 * M * hwc), (0, D * hwc), (M, M * hwc, hwc), (big, 0, hwc))

    def _buffers(self, plan, dev, extra=()):
        B, D, H, W, C = plan.B, plan.D, plan.H, plan.W, plan.C
        N = B * plan.voxels
        f32 = dict(device=dev, dtype=torch.float32)
        names = ("xs", "xv", "sp", "vp", "fp", "pp", "x1", "x2") + tuple(extra)
        bufs = {n: torch.empty(N, C, **f32) for n in names}
        bufs["s1"] = torch.empty(B * D * 2 * H * plan.wf * C, **f32)
        bufs["s2"] = torch.empty_like(bufs["s1"])
        bufs["v1"] = torch.empty(B * 2 * plan.modes * H * W * C, **f32)
        bufs["part"] = torch.empty(B, plan.nwg, GROUPS, 3, **f32)
        return bufs

    def _block(self, plan, i, prev, xin, bf, ys=None, yv=None):
        """One block's forward from the assembled input `prev`; writes the block input to xin (unless None) and
        every intermediate into bf. ys/yv: where the branch transforms' outputs go (default: over xs/xv).
        Returns the GroupNorm affines and statistics of the five norms."""
        dev = bf["xs"].device
        blk = self.weights.blocks[i]
        st = self._stream()
        rows = plan.rows
        stats = FfnoStatsOut(_p(plan.depth_w), _p(bf["part"]))
        xs, xv = bf["xs"], bf["xv"]
        ys = xs if ys is None else ys
        yv = xv if yv is None else yv
        _lib.call("ffno_block_in", rows, prev, _p(xin), _p(blk["spec.g1"]), _p(blk["spec.g2"]),
                  _p(blk["vertical.g1"]), _p(blk["vertical.g2"]), _p(blk["z_proj"]), _p(bf["z"]), _p(xs), _p(xv), st)
        gn = {}
        self.horizontal(plan, xs, bf["s1"], bf["s2"], ys, i)
        _lib.call("ffno_post", rows, _p(ys), _p(blk["spec.p1"]), _p(blk["spec.p2"]), _p(bf["sp"]), stats, st)
        gn["s"] = self._gn(plan, bf["part"], blk["norm_spec"], dev)
        self.vertical(plan, xv, bf["v1"], yv, i)
        _lib.call("ffno_post", rows, _p(yv), _p(blk["vertical.p1"]), _p(blk["vertical.p2"]), _p(bf["vp"]), stats, st)
        gn["v"] = self._gn(plan, bf["part"], blk["norm_vert"], dev)
        h = xin if xin is not None else None
        return gn, h

    def _block_tail(self, plan, i, h, bf, gn):
        dev = bf["xs"].device
        blk = self.weights.blocks[i]
        st = self._stream()
        rows = plan.rows
        stats = FfnoStatsOut(_p(plan.depth_w), _p(bf["part"]))
        sa, sb, _ = gn["s"]
        va, vb, _ = gn["v"]
        _lib.call("ffno_fuse", rows, _p(h), _p(bf["sp"]), _p(bf["vp"]), _p(sa), _p(sb), _p(va), _p(vb), _p(blk["bg1"]),
                  _p(blk["bg2"]), _p(blk["bg2_bias"]), _p(blk["f1"]), _p(blk["f2"]), _p(bf["fp"]), stats, st)
        gn["f"] = self._gn(plan, bf["part"], blk["norm_fuse"], dev)
        fa, fb, _ = gn["f"]
        _lib.call("ffno_resmlp", rows, FfnoAssemble(_p(h), _p(bf["fp"]), _p(fa), _p(fb), _p(blk["res_fused"]), 1),
                  _p(bf["x1"]), _p(blk["pw1"]), _p(blk["pw2"]), _p(bf["pp"]), stats, st)
        gn["p"] = self._gn(plan, bf["part"], blk["norm_pw"], dev)
        pa, pb, _ = gn["p"]
        _lib.call("ffno_resmlp", rows, FfnoAssemble(_p(bf["x1"]), _p(bf["pp"]), _p(pa), _p(pb), _p(blk["res_pw"]), 0),
                  _p(bf["x2"]), _p(blk["m1"]), _p(blk["m2"]), _p(bf["mp"]), stats, st)
        gn["m"] = self._gn(plan, bf["part"], blk["norm_mlp"], dev)
        ma, mb, _ = gn["m"]
        return FfnoAssemble(_p(bf["x2"]), _p(bf["mp"]), _p(ma), _p(mb), _p(blk["res_mlp"]), 0)

    def forward(self, x, z, dx, dy, keep=None, plan=None):
        """Normalised log10 populations [B, Cout, D, H, W] from x [B, 6, D, H, W] and z_scale [B, D, H, W] (Mm).
        keep: a list that receives every block input and the head input (channel-last) for the adjoint."""
        if not x.is_cuda:
            raise RuntimeError("the FFNO3D HIP engine runs on the GPU only; move the inputs to a cuda device")
        dev = x.device
        self.sync(dev)
        Wt = self.weights
        B, cin, D, H, W = x.shape
        if cin != 6:
            raise ValueError("FFNO3D expects 6 input channels")
        z = z.reshape(B, D, H, W)
        if plan is None:
            plan = self.plan(B, D, H, W, z, dx, dy)
        C = plan.C
        N = B * plan.voxels
        x = x.contiguous().float()
        st = self._stream()
        f32 = dict(device=dev, dtype=torch.float32)
        bf = self._buffers(plan, dev, extra=("mp",))
        bf["z"] = z.reshape(B, -1).contiguous().float()
        stats = FfnoStatsOut(_p(plan.depth_w), _p(bf["part"]))
        rows = plan.rows
        t0 = torch.empty(N, C, **f32)
        _lib.call("ffno_lift", rows, _p(x), _p(Wt.lift), _p(t0), stats, st)
        a, b, _ = self._gn(plan, bf["part"], Wt.lift_gn, dev)
        prev = FfnoAssemble(None, _p(t0), _p(a), _p(b), None, 1)
        xin_buf = torch.empty(N, C, **f32) if keep is None else None
        for i in range(len(Wt.blocks)):
            # the block input never aliases x2 or mp, which the assembly reads
            xin = torch.empty(N, C, **f32) if keep is not None else xin_buf
            gn, _ = self._block(plan, i, prev, xin, bf)
            if keep is not None:
                keep.append(xin)
            prev = self._block_tail(plan, i, xin, bf, gn)
        out_ch = self.module.out_channels
        y = torch.empty(B, out_ch, D, H, W, **f32)
        h_last = torch.empty(N, C, **f32) if keep is not None else None
        _lib.call("ffno_head", rows, prev, _p(h_last), _p(Wt.p1), _p(Wt.p2), _p(Wt.p2_bias), _p(x),
                  ctypes.byref(Wt.head), _p(y), st)
        if keep is not None:
            keep.append(h_last)
        return y

    # ------------------------------------------------------------------------------------------------ adjoint
    def _gn_sums(self, plan, gb, ubar, part2):
        sums = torch.empty(plan.B, GROUPS, 2, device=ubar.device)
        _lib.call("ffno_gn_grad_stats", plan.rows, gb, _p(ubar), _p(part2), self._stream())
        _lib.call("ffno_gn_grad_finalize", plan.B, plan.nwg, _p(part2), _p(sums), self._stream())
        gb.sums = _p(sums)
        return sums

    def _gnback(self, plan, t, gnrec, norm, scale, act):
        a, b, stats = gnrec
        gb = _lib.FfnoGnBack()
        gb.t, gb.stats, gb.gamma, gb.aff_a, gb.aff_b = _p(t), _p(stats), _p(norm[0]), _p(a), _p(b)
        gb.scale = _p(scale)
        gb.act_gelu = act
        gb.sums = ctypes.c_void_p(0)
        gb.depth_w = _p(plan.depth_w)
        return gb

    def vjp(self, x, z, dx, dy, ybar, keep, want_z=False, plan=None):
        """(x_bar [B, 6, D, H, W], z_bar [B, D, H, W] or None) of <ybar, forward(x, z)>: the hand-written adjoint,
        block by block in reverse, each block recomputed from its saved input (the reference's non-reentrant
        checkpointing, models/ffno_model.py:751-780). With want_z the height gradient is computed too, per column:
        through the depth weights of every GroupNorm, z_proj, and the vertical transform's Filon weights, phases,
        taper and per-column shift (models/ffno_model.py:28-74, :239-309)."""
        dev = x.device
        self.sync(dev)
        Wt = self.weights
        A = Wt.adjoint()
        B, _, D, H, W = x.shape
        z = z.reshape(B, D, H, W)
        if plan is None:
            plan = self.plan(B, D, H, W, z, dx, dy)
        plan.adjoint_tables()
        C = plan.C
        N = B * plan.voxels
        M = plan.modes
        hwc = H * W * C
        big = 1 << 40
        x = x.contiguous().float()
        ybar = ybar.contiguous().float()
        st = self._stream()
        rows = plan.rows
        f32 = dict(device=dev, dtype=torch.float32)
        bf = self._buffers(plan, dev, extra=("mp", "ys", "yv", "gx2", "gx1", "gs", "gv", "gh"))
        bf["v1y"] = torch.empty_like(bf["v1"])
        bf["z"] = z.reshape(B, -1).contiguous().float()
        part2 = torch.empty(B, plan.nwg, GROUPS, 2, **f32)
        ones = torch.ones(B, C, **f32)
        zeros = torch.zeros(B, C, **f32)
        xbar = torch.empty(B, 6, D, H, W, **f32)
        hbar = torch.empty(N, C, **f32)
        if want_z:
            ht = plan.height_tables()
            zbar = torch.zeros(B, plan.voxels, **f32)        # z_proj terms
            zbar_s = torch.zeros(B, plan.voxels, **f32)      # vertical transform, with respect to the shifted heights
            wbar = torch.zeros(B, plan.voxels, **f32)        # depth weights
            rbuf = torch.empty(N, C, **f32)
            dfac_col = torch.zeros(B, H * W, **f32) if ht["v0"] is not None else None
        else:
            zbar = zbar_s = wbar = None

        def gnback(t, rec, norm, scale, act):
            gb = self._gnback(plan, t, rec, norm, scale, act)
            gb.wbar = _p(wbar)
            return gb

        _lib.call("ffno_head_bwd", rows, _p(keep[-1]), _p(Wt.p1), _p(Wt.p2), _p(A["p2t"]), _p(A["p1t"]), _p(Wt.p2_bias),
                  _p(x), ctypes.byref(Wt.head), _p(ybar), _p(hbar), _p(xbar), st)
        for i in reversed(range(len(Wt.blocks))):
            h = keep[i]
            blk, adj = Wt.blocks[i], A["blocks"][i]
            # recompute the block; the branch transforms' outputs go to ys/yv (the post networks' adjoints read them)
            # and the vertical spectrum after the mix stays in v1y (the height gradient reads it)
            bfr = dict(bf)
            bfr["v1"] = bf["v1y"]
            gn, _ = self._block(plan, i, FfnoAssemble(None, _p(h), _p(ones), _p(zeros), None, 0), None, bfr,
                                ys=bf["ys"], yv=bf["yv"])
            self._block_tail(plan, i, h, bf, gn)
            # mlp: h' = x2 + res_mlp GN(mp)
            gb = gnback(bf["mp"], gn["m"], blk["norm_mlp"], blk["res_mlp"], 0)
            self._gn_sums(plan, gb, hbar, part2)
            _lib.call("ffno_mlp_bwd", rows, _p(hbar), gb, _p(bf["x2"]), _p(blk["m1"]), _p(adj["m2t"]), _p(adj["m1t"]),
                      1, _p(bf["gx2"]), st)
            # pw: x2 = x1 + res_pw GN(pp)
            gb = gnback(bf["pp"], gn["p"], blk["norm_pw"], blk["res_pw"], 0)
            self._gn_sums(plan, gb, bf["gx2"], part2)
            _lib.call("ffno_mlp_bwd", rows, _p(bf["gx2"]), gb, _p(bf["x1"]), _p(blk["pw1"]), _p(adj["pw2t"]),
                      _p(adj["pw1t"]), 1, _p(bf["gx1"]), st)
            # fusion: x1 = h + res_fused gelu(GN(fp))
            gb = gnback(bf["fp"], gn["f"], blk["norm_fuse"], blk["res_fused"], 1)
            self._gn_sums(plan, gb, bf["gx1"], part2)
            sa, sb_, _ = gn["s"]
            va, vb, _ = gn["v"]
            _lib.call("ffno_fuse_bwd", rows, _p(bf["gx1"]), gb, _p(h), _p(bf["sp"]), _p(bf["vp"]), _p(sa), _p(sb_),
                      _p(va), _p(vb), _p(blk["bg1"]), _p(blk["bg2"]), _p(blk["bg2_bias"]), _p(blk["f1"]),
                      _p(adj["f2t"]), _p(adj["f1t"]), _p(adj["bg2t"]), _p(adj["bg1t"]), _p(bf["gh"]), _p(bf["gs"]),
                      _p(bf["gv"]), st)
            # horizontal branch: spec = gelu(GN_s(post(Spec(xs)))); x_s bar lands in gs
            gb = gnback(bf["sp"], gn["s"], blk["norm_spec"], None, 1)
            self._gn_sums(plan, gb, bf["gs"], part2)
            _lib.call("ffno_mlp_bwd", rows, _p(bf["gs"]), gb, _p(bf["ys"]), _p(blk["spec.p1"]), _p(adj["spec.p2t"]),
                      _p(adj["spec.p1t"]), 0, _p(bf["gx2"]), st)
            self.horizontal(plan, bf["gx2"], bf["s1"], bf["s2"], bf["gs"], i, adjoint=True)
            # vertical branch; x_v bar lands in gv, the spectrum gradient stays in v1
            gb = gnback(bf["vp"], gn["v"], blk["norm_vert"], None, 1)
            self._gn_sums(plan, gb, bf["gv"], part2)
            _lib.call("ffno_mlp_bwd", rows, _p(bf["gv"]), gb, _p(bf["yv"]), _p(blk["vertical.p1"]),
                      _p(adj["vertical.p2t"]), _p(adj["vertical.p1t"]), 0, _p(bf["gx1"]), st)
            if want_z and ht["v0"] is not None:
                # taper: dL/df_m per column = sum_c,part (V0^T yv_bar) * Y, before the mix adjoint overwrites v1
                self._tw(ht["v0"], bf["gx1"], bf["v1"], B, hwc, (big, 0, D * hwc), (0, 2 * M * hwc), (big, 0, hwc),
                         (M, M * hwc, hwc))
                g = (bf["v1"].view(B, 2, M, H * W, C) * bf["v1y"].view(B, 2, M, H * W, C)).sum((1, 4))   # [B, M, HW]
                dfac_col += (g * ht["dfac"][None, :, None]).sum(1)
            self.vertical(plan, bf["gx1"], bf["v1"], bf["gv"], i, adjoint=True)
            if want_z:
                # Filon weights: zbar_s[e] += sum_delta sum_c xv[e + delta] (G_delta X_bar)[e]
                for delta in (-1, 0, 1):
                    self._tw(ht["bands"][delta], bf["v1"], rbuf, B, hwc, (big, 0, 2 * M * hwc), (0, D * hwc),
                             (M, M * hwc, hwc), (big, 0, hwc))
                    _lib.call("ffno_dot_reduce", rows, D, _p(bf["xv"]), delta, _p(rbuf), _p(zbar_s), st)
                # synthesis phases: zbar_s[d] += sum_c yv_bar[d] (V_dz Y)[d]
                self._tw(ht["v_dz"], bf["v1y"], rbuf, B, hwc, (big, 0, 2 * M * hwc), (0, D * hwc), (M, M * hwc, hwc),
                         (big, 0, hwc))
                _lib.call("ffno_dot_reduce", rows, D, _p(bf["gx1"]), 0, _p(rbuf), _p(zbar_s), st)
            # input gates (and z_proj)
            _lib.call("ffno_gates_bwd", rows, _p(h), _p(bf["gs"]), _p(bf["gv"]), _p(bf["gh"]), _p(blk["spec.g1"]),
                      _p(blk["spec.g2"]), _p(adj["spec.g2t"]), _p(adj["spec.g1t"]), _p(blk["vertical.g1"]),
                      _p(blk["vertical.g2"]), _p(adj["vertical.g2t"]), _p(adj["vertical.g1t"]), _p(blk["z_proj"]),
                      _p(hbar), _p(zbar), st)
        # lift: h0 = gelu(GN(W_lift x))
        t0 = torch.empty(N, C, **f32)
        _lib.call("ffno_lift", rows, _p(x), _p(Wt.lift), _p(t0), FfnoStatsOut(_p(plan.depth_w), _p(bf["part"])), st)
        g0 = self._gn(plan, bf["part"], Wt.lift_gn, dev)
        gb = gnback(t0, g0, Wt.lift_gn, None, 1)
        self._gn_sums(plan, gb, hbar, part2)
        _lib.call("ffno_lift_bwd", rows, _p(hbar), gb, _p(Wt.lift), _p(xbar), st)
        if not want_z:
            return xbar, None
        return xbar, self._height_gradient(plan, z, zbar, zbar_s, wbar, dfac_col)

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
            zs[:, i + 1] += sgn * dfac_col
            zs[:, i] -= sgn * dfac_col
        # the per-column shift z - amin(z): the minimum's level(s) take minus the column total, shared evenly
        total = zs.sum(1)
        out = zbar.view(B, D, H * W) + zs
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
