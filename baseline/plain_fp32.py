"""The plain FP32 GPU path: FFNO3D evaluated with stock PyTorch ops as the reference evaluates it (1x1 Conv3d,
torch.fft.rfft2/irfft2, einsum, autograd through per-block non-reentrant checkpoints; models/ffno_model.py:261-309,
:431-450, :524-553, :730-832). It is the comparison baseline of bench.py ("plain FP32 GPU path"), kept outside the product package: nothing in
ffnoml_amd imports it (ffnoml_amd.FFNO3D's forward and adjoint are the HIP kernels only).

Runs on a state_dict with the reference's keys, channel-first tensors, shared column grids.
"""
import math

import torch
import torch.nn.functional as F
from torch.utils.checkpoint import checkpoint

_EE, _HH, _EM, _BK, _PI = 1.602189e-12, 6.626176e-27, 9.109534e-28, 1.380662e-16, 3.14159265359
_LN_CCON = math.log(0.5 * (_HH / math.sqrt(2.0 * _PI * _EM) / math.sqrt(_BK)) ** 3)


def _c(x, w, b=None):
    return F.conv3d(x, w if w.dim() == 5 else w[..., None, None, None], b)


def _wgn(x, w, weight, bias, groups=8, eps=1e-5):
    B, C = x.shape[:2]
    g = x.reshape(B, groups, C // groups, *x.shape[2:])
    wg = w.unsqueeze(1)
    dims = tuple(range(2, g.dim()))
    total = (wg * torch.ones_like(g)).sum(dim=dims, keepdim=True)
    mean = (wg * g).sum(dim=dims, keepdim=True) / total
    var = (wg * (g - mean) ** 2).sum(dim=dims, keepdim=True) / total
    y = ((g - mean) * torch.rsqrt(var + eps)).reshape_as(x)
    return y * weight.view(1, C, 1, 1, 1) + bias.view(1, C, 1, 1, 1)


def _mlp(sd, p, feats, n):
    h = F.gelu(F.linear(feats, sd[p + f"{n}.0.weight"], sd[p + f"{n}.0.bias"]))
    return F.gelu(F.linear(h, sd[p + f"{n}.2.weight"], sd[p + f"{n}.2.bias"]))


class PlainFFNO3D:
    """FFNO3D with stock torch ops in float32; forward(x, z, dx, dy) like the reference's."""

    def __init__(self, module):
        self.sd = {k: v.detach() for k, v in module.state_dict().items()}
        self.m = module

    def to(self, device):
        self.sd = {k: v.to(device) for k, v in self.sd.items()}
        return self

    def _spec(self, p, x, dx, dy):
        sd, m = self.sd, self.m
        B, C, D, H, W = x.shape
        x = x * torch.sigmoid(_c(F.gelu(_c(x, sd[p + "input_gate.0.weight"])), sd[p + "input_gate.2.weight"]))
        xf = torch.fft.rfft2(x, dim=(-2, -1))
        ky = torch.fft.fftfreq(H, d=dy, device=x.device)
        kx = torch.fft.rfftfreq(W, d=dx, device=x.device)
        ky, kx = torch.meshgrid(ky, kx, indexing="ij")
        ky, kx = ky * 1e5, kx * 1e5
        feats = torch.stack([kx, ky, torch.sqrt(kx ** 2 + ky ** 2 + 1e-12), torch.sign(ky)], -1)
        gate = F.linear(_mlp(sd, p, feats, "freq_mlp"), sd[p + "freq_mlp.4.weight"], sd[p + "freq_mlp.4.bias"])
        wr, wi = sd[p + "weight_real"][:, :, None, None], sd[p + "weight_imag"][:, :, None, None]
        gr, gi = gate[..., 0][None, None], gate[..., 1][None, None]
        w_re, w_im = wr * gr - wi * gi, wr * gi + wi * gr
        xr, xi = xf.real, xf.imag
        out = torch.complex(torch.einsum("bidyx,ioyx->bodyx", xr, w_re) - torch.einsum("bidyx,ioyx->bodyx", xi, w_im),
                            torch.einsum("bidyx,ioyx->bodyx", xr, w_im) + torch.einsum("bidyx,ioyx->bodyx", xi, w_re))
        r = m.freq_mix_rank
        if r:
            h = _mlp(sd, p, feats, "freq_lowrank")
            u = F.linear(h, sd[p + "freq_u.weight"], sd[p + "freq_u.bias"]).reshape(H, W // 2 + 1, C, r).to(xf.dtype)
            v = F.linear(h, sd[p + "freq_v.weight"], sd[p + "freq_v.bias"]).reshape(H, W // 2 + 1, r, C).to(xf.dtype)
            out = out + torch.einsum("brdyx,yxor->bodyx", torch.einsum("bidyx,yxri->brdyx", xf, v), u)
        y = torch.fft.irfft2(out, s=(H, W), dim=(-2, -1))
        return _c(F.gelu(_c(y, sd[p + "post.0.weight"])), sd[p + "post.2.weight"])

    def _vert(self, p, x, zs):
        sd, m = self.sd, self.m
        C = x.shape[1]
        span = float((zs.amax(dim=-3) - zs.amin(dim=-3)).max())
        length = m.z_reference_length
        modes = int(math.floor(length / m.z_min_wavelength)) + 1
        if 2 * span > length:
            length = 2 * span
            modes = int(math.floor(length / m.z_min_wavelength)) + 1
        k = 2 * math.pi * torch.arange(modes, dtype=torch.float32, device=x.device) / length
        factor = torch.cat([k.new_ones(1), 2 * k.new_ones(modes - 1)])
        x = x * torch.sigmoid(_c(F.gelu(_c(x, sd[p + "input_gate.0.weight"])), sd[p + "input_gate.2.weight"])) + \
            _c(zs, sd[p + "z_proj.weight"])
        z = zs - zs.amin(dim=2, keepdim=True)
        col = z[0, 0, :, 0, 0]
        z0, z1 = col[:-1, None], col[1:, None]
        low, hh = torch.minimum(z0, z1), (z1 - z0).abs()
        th = hh * k
        small = th.abs() < 1e-2
        tt = torch.where(small, torch.ones_like(th), th)
        c, s = torch.cos(tt), torch.sin(tt)
        r_re = torch.where(small, 0.5 - th ** 2 / 8, (c + tt * s - 1) / tt ** 2)
        r_im = torch.where(small, -th / 3 + th ** 3 / 30, (tt * c - s) / tt ** 2)
        l_re = torch.where(small, 0.5 - th ** 2 / 24, s / tt - r_re)
        l_im = torch.where(small, -th / 6 + th ** 3 / 120, (c - 1) / tt - r_im)
        e_re, e_im = hh * torch.cos(low * k), -hh * torch.sin(low * k)
        lo_re, lo_im = e_re * l_re - e_im * l_im, e_re * l_im + e_im * l_re
        hi_re, hi_im = e_re * r_re - e_im * r_im, e_re * r_im + e_im * r_re
        asc = z0 <= z1
        q_re = F.pad(torch.where(asc, lo_re, hi_re), (0, 0, 0, 1)) + F.pad(torch.where(asc, hi_re, lo_re), (0, 0, 1, 0))
        q_im = F.pad(torch.where(asc, lo_im, hi_im), (0, 0, 0, 1)) + F.pad(torch.where(asc, hi_im, lo_im), (0, 0, 1, 0))
        xr = torch.einsum("bcdhw,dm->bcmhw", x, q_re)
        xi = torch.einsum("bcdhw,dm->bcmhw", x, q_im)
        dz = (col[1:] - col[:-1]).abs().median()
        kn = math.pi / dz
        taper = (0.5 * (1 + torch.cos(math.pi * ((k - 2 * kn / 3) / (kn / 3)).clamp(0, 1))))[:, None, None]
        q = k * m.z_feature_wavelength / (2 * math.pi)
        t = q / (1 + q)
        feats = torch.stack([t, t * t], -1)
        gate = F.linear(_mlp(sd, p, feats, "freq_mlp"), sd[p + "freq_mlp.4.weight"], sd[p + "freq_mlp.4.bias"])
        gr, gi = gate[:, 0, None, None], gate[:, 1, None, None]
        wr, wi = sd[p + "weight_real"], sd[p + "weight_imag"]
        ar = torch.einsum("bimhw,io->bomhw", xr, wr) - torch.einsum("bimhw,io->bomhw", xi, wi)
        ai = torch.einsum("bimhw,io->bomhw", xi, wr) + torch.einsum("bimhw,io->bomhw", xr, wi)
        yr, yi = ar * gr - ai * gi, ar * gi + ai * gr
        r = m.freq_mix_rank
        if r:
            h = _mlp(sd, p, feats, "freq_lowrank")
            u = F.linear(h, sd[p + "freq_u.weight"], sd[p + "freq_u.bias"]).reshape(modes, C, r)
            v = F.linear(h, sd[p + "freq_v.weight"], sd[p + "freq_v.bias"]).reshape(modes, r, C)
            yr = yr + torch.einsum("brmhw,mor->bomhw", torch.einsum("bimhw,mri->brmhw", xr, v), u)
            yi = yi + torch.einsum("brmhw,mor->bomhw", torch.einsum("bimhw,mri->brmhw", xi, v), u)
        f = factor[:, None, None] * taper
        ph = col[:, None] * k
        y = torch.einsum("bomhw,dm->bodhw", yr * f, torch.cos(ph)) - torch.einsum("bomhw,dm->bodhw", yi * f,
                                                                                torch.sin(ph))
        return _c(F.gelu(_c(y / length, sd[p + "post.0.weight"])), sd[p + "post.2.weight"])

    def _block(self, i, x, zs, w, dx, dy):
        sd, p = self.sd, f"blocks.{i}."
        n = lambda name, t: _wgn(t, w, sd[p + name + ".weight"], sd[p + name + ".bias"])
        spec = F.gelu(n("norm_spec", self._spec(p + "spec.", x, dx, dy)))
        vert = F.gelu(n("norm_vert", self._vert(p + "vertical.", x, zs)))
        gates = torch.sigmoid(_c(F.gelu(_c(torch.cat([x, spec, vert], 1), sd[p + "branch_gate.0.weight"])),
                                 sd[p + "branch_gate.2.weight"], sd[p + "branch_gate.2.bias"]))
        sg, vg = gates.chunk(2, dim=1)
        sm, vm = spec * sg, vert * vg
        fused = _c(F.gelu(_c(torch.cat([sm, vm], 1), sd[p + "fuse.0.weight"])), sd[p + "fuse.2.weight"])
        x1 = x + sd[p + "res_fused"] * F.gelu(n("norm_fuse", fused + 0.5 * (sm + vm)))
        x2 = x1 + sd[p + "res_pw"] * n("norm_pw", _c(F.gelu(_c(x1, sd[p + "pw.0.weight"])), sd[p + "pw.2.weight"]))
        return x2 + sd[p + "res_mlp"] * n("norm_mlp", _c(F.gelu(_c(x2, sd[p + "mlp.net.0.weight"])),
                                                         sd[p + "mlp.net.3.weight"]))

    def _head(self, head, x_in):
        sd, m = self.sd, self.m
        mx, sx = sd["mean_X"], sd["std_X"]
        t, e, r = m.temperature_channel, m.ne_channel, m.rho_channel
        lt = (x_in[:, t].double() * sx[t].double() + mx[t].double())
        le = (x_in[:, e].double() * sx[e].double() + mx[e].double())
        tt = torch.pow(10.0, lt)[:, None]
        ne = torch.pow(10.0, le - 6.0)[:, None]
        tab = [torch.tensor(v, dtype=torch.float64, device=x_in.device).view(1, -1, 1, 1, 1) for v in m.lte_tables]
        conl = _LN_CCON + torch.log(ne) - 1.5 * torch.log(tt)
        dxi = 4.98e-4 * 5040.0 / tt * torch.sqrt(ne * _BK * tt)
        lte = tab[0] - (_EE / _BK) / tt * (tab[1] - dxi * tab[2]) - tab[3] * conl
        out, c, k = [], 0, 0
        log10_rho = x_in[:, r] * sx[r] + mx[r]
        for i, levels in enumerate(m.species_levels):
            frac = (torch.log_softmax(head[:, c:c + levels].double() + lte[:, k:k + levels], 1) / math.log(10.0))
            s = m.total_bound * torch.tanh(head[:, c + levels] / m.total_bound)
            out.append((log10_rho - sd["log10_particle_mass"][i] + s)[:, None] + frac.to(head.dtype))
            c += levels + 1
            k += levels
        y = torch.cat(out, 1)
        return (y - sd["mean_Y"].view(1, -1, 1, 1, 1)) / sd["std_Y"].view(1, -1, 1, 1, 1)

    MAX_ELEMENTS = 2 ** 31 - 1

    def __call__(self, x, z_scale, dx, dy):
        # stock library kernels (convolutions, FFTs) are only run here on activations below 2^31 elements, the sizes
        # this baseline is timed at; a whole 504 x 504 x 256 volume at width 64 (4.2e9 elements) is refused
        B, _, D, H, W = x.shape
        if B * self.m.width * D * H * (2 * (W // 2 + 1)) > self.MAX_ELEMENTS:
            raise ValueError("the plain FP32 path is not run on activations of 2^31 elements or more")
        sd, dx, dy = self.sd, float(dx), float(dy)
        zs = z_scale.unsqueeze(1)
        tw = F.pad(0.5 * (zs[:, :, 1:] - zs[:, :, :-1]).abs(), (0, 0, 0, 0, 0, 1)) + \
            F.pad(0.5 * (zs[:, :, 1:] - zs[:, :, :-1]).abs(), (0, 0, 0, 0, 1, 0))
        w = tw / tw.mean(dim=2, keepdim=True).clamp_min(1e-12)
        grad = torch.is_grad_enabled()
        lift = lambda t: F.gelu(_wgn(_c(t, sd["lift.0.weight"]), w, sd["lift.1.weight"], sd["lift.1.bias"]))
        h = checkpoint(lift, x, use_reentrant=False) if grad else lift(x)
        for i in range(self.m.n_layers):
            fn = lambda t, i=i: self._block(i, t, zs, w, dx, dy)
            h = checkpoint(fn, h, use_reentrant=False) if grad else fn(h)
        head_fn = lambda t: _c(F.gelu(_c(t, sd["proj1.weight"])), sd["proj2.weight"], sd["proj2.bias"])
        head = checkpoint(head_fn, h, use_reentrant=False) if grad else head_fn(h)
        return self._head(head, x)
