"""CPU oracle for the FFNO3D forward pass and its input adjoint.

TEST INFRASTRUCTURE ONLY. Nothing in the product (the ffnoml_amd package, its HIP library, the service) imports this
module; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, and only as the checker or as
the timed CPU baseline.

A functional restatement, in plain PyTorch on the CPU, of the network in the reference's models/ffno_model.py,
operating on a state_dict with the reference's keys. It runs in whatever dtype its inputs have: float64 is the
accuracy oracle, float32 the CPU baseline. Gradients come from torch autograd over this restatement.

Every function cites the reference lines it follows (paths are relative to the reference tree).

Pinning (how each part is checked, in tests/test_oracle.py):
  - forward, x-gradient and z-gradient of the whole network: golden vectors written by
    tests/golden/make_golden.py from the reference itself (tests/golden/*.npz);
  - height_fourier_weights: exact integrals of linear profiles, trapezoid weights at k=0, flip symmetry;
  - weighted_group_norm: equals torch.nn.GroupNorm when every weight is 1;
  - box / cutoff: the reference tests' closed-form values;
  - irfft2-as-matrices conventions: numpy.fft on non-Hermitian spectra;
  - head LTE weights: an independent numpy transcription of MULTI3D calc_lte_populations.
"""
import math

import torch
import torch.nn.functional as F

SQRT_HALF = 1.0 / math.sqrt(2.0)


# ----------------------------------------------------------------------------------------------------------------
# configuration
# ----------------------------------------------------------------------------------------------------------------
def model_config(width=96, n_layers=4, hidden_spectral=128, hidden_vertical=128, freq_mix_rank=8,
                 z_reference_length=40.0, z_min_wavelength=0.04, z_feature_wavelength=0.3, species=None,
                 in_channels=6, out_channels=None, rho_channel=5, total_bound=6.0, temperature_channel=0,
                 ne_channel=4):
    """The constructor arguments of FFNO3D (models/ffno_model.py:628-651) that change the arithmetic."""
    if not species:
        raise ValueError("the level-fraction head needs species")        # models/ffno_model.py:654-655
    levels = [int(s["levels"]) for s in species]
    if out_channels is None:
        out_channels = sum(levels)
    if sum(levels) != out_channels:                                      # models/ffno_model.py:657-658
        raise ValueError("species levels do not add up to out_channels")
    modes = int(math.floor(float(z_reference_length) / float(z_min_wavelength))) + 1   # models/ffno_model.py:200
    return dict(width=int(width), n_layers=int(n_layers), hidden_spectral=int(hidden_spectral),
                hidden_vertical=int(hidden_vertical), rank=int(freq_mix_rank),
                z_reference_length=float(z_reference_length), z_min_wavelength=float(z_min_wavelength),
                z_feature_wavelength=float(z_feature_wavelength), species=list(species), levels=levels,
                in_channels=int(in_channels), out_channels=int(out_channels), rho_channel=int(rho_channel),
                total_bound=float(total_bound), temperature_channel=int(temperature_channel),
                ne_channel=int(ne_channel), modes=modes, groups=group_count(int(width)))


def group_count(channels, max_groups=8):
    """models/ffno_model.py:21-25: the largest g <= 8 dividing the channel count."""
    g = min(max_groups, channels)
    while g > 1 and channels % g != 0:
        g -= 1
    return g


# ----------------------------------------------------------------------------------------------------------------
# elementwise pieces
# ----------------------------------------------------------------------------------------------------------------
def gelu(x):
    """Exact (erf) GELU, torch.nn.GELU() default as used throughout models/ffno_model.py: 0.5 x (1 + erf(x / sqrt 2)),
    evaluated by torch's own kernel (one pass)."""
    return F.gelu(x)


def conv1x1(x, weight, bias=None):
    """A 1x1x1 Conv3d on channel-first x [B, Cin, ...] with weight [Cout, Cin(,1,1,1)]."""
    w = weight.reshape(weight.shape[0], -1)
    y = (w @ x.reshape(x.shape[0], x.shape[1], -1)).reshape((x.shape[0], w.shape[0]) + x.shape[2:])
    if bias is not None:
        y = y + bias.reshape((1, -1) + (1,) * (x.dim() - 2))
    return y


def linear(x, weight, bias):
    return x @ weight.t() + bias


# ----------------------------------------------------------------------------------------------------------------
# heights
# ----------------------------------------------------------------------------------------------------------------
def trapezoid_weights(z):
    """|dz| trapezoid weights of each column along dim 2 of z [B, 1, D, H, W] (models/ffno_model.py:28-34)."""
    if z.shape[2] == 1:
        return torch.ones_like(z)
    half = 0.5 * (z[:, :, 1:] - z[:, :, :-1]).abs()
    w = torch.zeros_like(z)
    w[:, :, :-1] = w[:, :, :-1] + half
    w[:, :, 1:] = w[:, :, 1:] + half
    return w


def depth_weights(z_scale):
    """Trapezoid weight over its column mean, [B, 1, D, H, W] (models/ffno_model.py:67-74)."""
    if z_scale.dim() == 4:
        z_scale = z_scale.unsqueeze(1)
    w = trapezoid_weights(z_scale)
    return w / w.mean(dim=2, keepdim=True).clamp_min(1e-12)


def height_fourier_weights(z, k):
    """Filon weights (q_re, q_im) [..., D, M] with sum_j x_j q_jm = int x(z) exp(-i k_m z) dz for x linear between
    the levels z [..., D] (any order); the small-angle branch below |theta| < 1e-2 uses the same truncated series as
    models/ffno_model.py:37-64."""
    z0, z1 = z[..., :-1, None], z[..., 1:, None]
    low = torch.minimum(z0, z1)
    h = (z1 - z0).abs()
    th = h * k
    small = th.abs() < 1e-2
    tt = torch.where(small, torch.ones_like(th), th)
    c, s = torch.cos(tt), torch.sin(tt)
    # share of the upper end of an interval: int_0^1 u e^{-i th u} du; of the lower end: int_0^1 (1-u) e^{-i th u} du
    up_re = torch.where(small, 0.5 - th ** 2 / 8, (c + tt * s - 1) / tt ** 2)
    up_im = torch.where(small, -th / 3 + th ** 3 / 30, (tt * c - s) / tt ** 2)
    whole_re, whole_im = s / tt, (c - 1) / tt
    lo_re = torch.where(small, 0.5 - th ** 2 / 24, whole_re - up_re)
    lo_im = torch.where(small, -th / 6 + th ** 3 / 120, whole_im - up_im)
    ph = low * k
    e_re, e_im = h * torch.cos(ph), -h * torch.sin(ph)
    a_re, a_im = e_re * lo_re - e_im * lo_im, e_re * lo_im + e_im * lo_re      # lower level of the interval
    b_re, b_im = e_re * up_re - e_im * up_im, e_re * up_im + e_im * up_re      # upper level
    ascending = z0 <= z1
    first_re, first_im = torch.where(ascending, a_re, b_re), torch.where(ascending, a_im, b_im)
    second_re, second_im = torch.where(ascending, b_re, a_re), torch.where(ascending, b_im, a_im)
    shape = list(first_re.shape)
    shape[-2] += 1
    q_re = first_re.new_zeros(shape)
    q_im = first_re.new_zeros(shape)
    q_re[..., :-1, :] += first_re
    q_re[..., 1:, :] += second_re
    q_im[..., :-1, :] += first_im
    q_im[..., 1:, :] += second_im
    return q_re, q_im


def vertical_box(z_scale, cfg, dtype, span=None):
    """(L, k [M], mode factors [M]) for these columns (models/ffno_model.py:248-259). The span is read on the host
    (span: that of a whole volume, when z_scale holds only some of its columns)."""
    if span is None:
        span = float((z_scale.detach().amax(dim=-3) - z_scale.detach().amin(dim=-3)).max())
    length = cfg["z_reference_length"]
    modes = cfg["modes"]
    if 2 * span > length:
        # a longer box computes its k in the module's dtype (models/ffno_model.py:255-257)
        length = 2 * span
        modes = int(math.floor(length / cfg["z_min_wavelength"])) + 1
        k = 2 * math.pi * torch.arange(modes, dtype=dtype) / length
    else:
        # the module's own k buffer is built in float32, then cast with the module (models/ffno_model.py:226)
        k = (2 * math.pi * torch.arange(modes, dtype=torch.float32) / length).to(dtype)
    factor = torch.cat([torch.ones(1, dtype=dtype), 2 * torch.ones(modes - 1, dtype=dtype)])
    return length, k, factor


def vertical_k_features(k, feature_wavelength):
    """[M, 2] = (t, t^2), t = q / (1 + q), q = k * feature_wavelength / 2 pi (models/ffno_model.py:231-237)."""
    q = k * feature_wavelength / (2 * math.pi)
    t = q / (1 + q)
    return torch.stack([t, t * t], dim=-1)


def column_taper(columns, k):
    """Half-cosine taper [..., M] from 2 pi / (3 dz) to pi / dz, dz the columns' lower median spacing
    (models/ffno_model.py:239-246)."""
    dz = (columns[..., 1:] - columns[..., :-1]).abs().median(dim=-1).values[..., None]
    k_nyq = math.pi / dz
    u = ((k - 2 * k_nyq / 3) / (k_nyq / 3)).clamp(0, 1)
    return 0.5 * (1 + torch.cos(math.pi * u))


# ----------------------------------------------------------------------------------------------------------------
# GroupNorm
# ----------------------------------------------------------------------------------------------------------------
def weighted_group_norm(x, w, weight, bias, groups, eps=1e-5):
    """Depth-weighted GroupNorm (models/ffno_model.py:77-95): statistics weighted by w [B, 1, D, H, W], two-pass
    variance, then the affine."""
    B, C = x.shape[:2]
    g = x.reshape(B, groups, C // groups, *x.shape[2:])
    wg = w.unsqueeze(1).expand_as(g)
    dims = tuple(range(2, g.dim()))
    total = wg.sum(dim=dims, keepdim=True)
    mean = (wg * g).sum(dim=dims, keepdim=True) / total
    var = (wg * (g - mean) ** 2).sum(dim=dims, keepdim=True) / total
    y = ((g - mean) * torch.rsqrt(var + eps)).reshape_as(x)
    shape = (1, C) + (1,) * (x.dim() - 2)
    return y * weight.view(shape) + bias.view(shape)


# ----------------------------------------------------------------------------------------------------------------
# frequency MLPs
# ----------------------------------------------------------------------------------------------------------------
def freq_gate(sd, prefix, feats):
    """freq_mlp: Linear-GELU-Linear-GELU-Linear -> [..., 2] (models/ffno_model.py:212-213, :346-352)."""
    h = gelu(linear(feats, sd[prefix + "freq_mlp.0.weight"], sd[prefix + "freq_mlp.0.bias"]))
    h = gelu(linear(h, sd[prefix + "freq_mlp.2.weight"], sd[prefix + "freq_mlp.2.bias"]))
    return linear(h, sd[prefix + "freq_mlp.4.weight"], sd[prefix + "freq_mlp.4.bias"])


def freq_lowrank(sd, prefix, feats, channels, rank):
    """(U [..., C, r], V [..., r, C]) from freq_lowrank, freq_u, freq_v (models/ffno_model.py:215-217, :294-296,
    :423-426)."""
    h = gelu(linear(feats, sd[prefix + "freq_lowrank.0.weight"], sd[prefix + "freq_lowrank.0.bias"]))
    h = gelu(linear(h, sd[prefix + "freq_lowrank.2.weight"], sd[prefix + "freq_lowrank.2.bias"]))
    u = linear(h, sd[prefix + "freq_u.weight"], sd[prefix + "freq_u.bias"])
    v = linear(h, sd[prefix + "freq_v.weight"], sd[prefix + "freq_v.bias"])
    return (u.reshape(*feats.shape[:-1], channels, rank), v.reshape(*feats.shape[:-1], rank, channels))


def horizontal_k_features(H, W, dx, dy, dtype):
    """[H, Wf, 4] = (kx, ky, |k|, sign ky) * 1e5 on fftfreq(H, dy) x rfftfreq(W, dx) (models/ffno_model.py:375-392,
    :443-445). The reference builds the grids in torch's default dtype, so it runs in float64 only with the default
    set to float64; the oracle builds them in the activations' dtype, which is the same thing."""
    ky = torch.fft.fftfreq(H, d=dy, dtype=dtype)
    kx = torch.fft.rfftfreq(W, d=dx, dtype=dtype)
    ky, kx = torch.meshgrid(ky, kx, indexing="ij")
    ky = ky * 1e5
    kx = kx * 1e5
    return torch.stack([kx, ky, torch.sqrt(kx ** 2 + ky ** 2 + 1e-12), torch.sign(ky)], dim=-1)


# ----------------------------------------------------------------------------------------------------------------
# the two spectral branches
# ----------------------------------------------------------------------------------------------------------------
def input_gate(sd, prefix, x):
    """Conv-GELU-Conv-Sigmoid gate (models/ffno_model.py:205-210, :339-344)."""
    return torch.sigmoid(conv1x1(gelu(conv1x1(x, sd[prefix + "input_gate.0.weight"])), sd[prefix + "input_gate.2.weight"]))


def post_mlp(sd, prefix, x):
    """Conv C->2C, GELU, Conv 2C->C (models/ffno_model.py:221-225, :369-373)."""
    return conv1x1(gelu(conv1x1(x, sd[prefix + "post.0.weight"])), sd[prefix + "post.2.weight"])


def horizontal_branch(sd, prefix, x, dx, dy, cfg):
    """SpectralConv2dFull.forward (models/ffno_model.py:431-450) with mix_spectrum (:394-429)."""
    B, C, D, H, W = x.shape
    x = x * input_gate(sd, prefix, x)
    xf = torch.fft.rfft2(x, dim=(-2, -1))
    feats = horizontal_k_features(H, W, dx, dy, x.dtype)
    gate = freq_gate(sd, prefix, feats)                                  # [H, Wf, 2]
    gr, gi = gate[..., 0], gate[..., 1]
    wr, wi = sd[prefix + "weight_real"], sd[prefix + "weight_imag"]      # [Cin, Cout]
    xr, xi = xf.real, xf.imag
    ar = torch.einsum("bidyx,io->bodyx", xr, wr) - torch.einsum("bidyx,io->bodyx", xi, wi)
    ai = torch.einsum("bidyx,io->bodyx", xr, wi) + torch.einsum("bidyx,io->bodyx", xi, wr)
    out_r = ar * gr - ai * gi
    out_i = ar * gi + ai * gr
    if cfg["rank"]:
        u, v = freq_lowrank(sd, prefix, feats, C, cfg["rank"])          # [H, Wf, C, r], [H, Wf, r, C]
        lr = torch.einsum("bidyx,yxri->brdyx", xr, v)
        li = torch.einsum("bidyx,yxri->brdyx", xi, v)
        out_r = out_r + torch.einsum("brdyx,yxor->bodyx", lr, u)
        out_i = out_i + torch.einsum("brdyx,yxor->bodyx", li, u)
    y = torch.fft.irfft2(torch.complex(out_r, out_i), s=(H, W), dim=(-2, -1))
    return post_mlp(sd, prefix, y)


def grid_is_shared(z):
    """True when every column of every sample has the same heights after the per-column shift
    (models/ffno_model.py:267-269). z [B, 1, D, H, W] already shifted."""
    return bool((z == z[..., :1, :1]).all()) and bool((z == z[:1]).all())


def vertical_branch(sd, prefix, x, z_scale, cfg, force_per_column=False, span=None):
    """HeightSpectralConv.forward (models/ffno_model.py:261-309). span: the whole volume's, when x holds some of its
    columns (every column is transformed on its own; only the box length is shared)."""
    if z_scale.dim() == 4:
        z_scale = z_scale.unsqueeze(1)
    C = x.shape[1]
    length, k, factor = vertical_box(z_scale, cfg, x.dtype, span)
    x = x * input_gate(sd, prefix, x) + conv1x1(z_scale, sd[prefix + "z_proj.weight"])
    z = z_scale - z_scale.amin(dim=2, keepdim=True)
    shared = grid_is_shared(z) and not force_per_column
    if shared:
        col = z[0, 0, :, 0, 0]
        q_re, q_im = height_fourier_weights(col, k)                     # [D, M]
        xr = torch.einsum("bcdhw,dm->bcmhw", x, q_re)
        xi = torch.einsum("bcdhw,dm->bcmhw", x, q_im)
        phase = col[:, None] * k
        taper = column_taper(col, k)[:, None, None]                     # [M, 1, 1]
    else:
        cols = z[:, 0].permute(0, 2, 3, 1)                              # [B, H, W, D]
        q_re, q_im = height_fourier_weights(cols, k)                    # [B, H, W, D, M]
        xr = torch.einsum("bcdhw,bhwdm->bcmhw", x, q_re)
        xi = torch.einsum("bcdhw,bhwdm->bcmhw", x, q_im)
        phase = z[:, 0, ..., None] * k                                  # [B, D, H, W, M]
        taper = column_taper(cols, k).permute(0, 3, 1, 2)[:, None]      # [B, 1, M, H, W]
    feats = vertical_k_features(k, cfg["z_feature_wavelength"])
    gate = freq_gate(sd, prefix, feats)                                  # [M, 2]
    gr, gi = gate[:, 0, None, None], gate[:, 1, None, None]
    wr, wi = sd[prefix + "weight_real"], sd[prefix + "weight_imag"]
    ar = torch.einsum("bimhw,io->bomhw", xr, wr) - torch.einsum("bimhw,io->bomhw", xi, wi)
    ai = torch.einsum("bimhw,io->bomhw", xi, wr) + torch.einsum("bimhw,io->bomhw", xr, wi)
    yr, yi = ar * gr - ai * gi, ar * gi + ai * gr
    if cfg["rank"]:
        u, v = freq_lowrank(sd, prefix, feats, C, cfg["rank"])          # [M, C, r], [M, r, C]
        lr = torch.einsum("bimhw,mri->brmhw", xr, v)
        li = torch.einsum("bimhw,mri->brmhw", xi, v)
        yr = yr + torch.einsum("brmhw,mor->bomhw", lr, u)
        yi = yi + torch.einsum("brmhw,mor->bomhw", li, u)
    f = factor[:, None, None] * taper
    yr, yi = yr * f, yi * f
    cos, sin = torch.cos(phase), torch.sin(phase)
    if shared:
        y = torch.einsum("bomhw,dm->bodhw", yr, cos) - torch.einsum("bomhw,dm->bodhw", yi, sin)
    else:
        y = torch.einsum("bomhw,bdhwm->bodhw", yr, cos) - torch.einsum("bomhw,bdhwm->bodhw", yi, sin)
    return post_mlp(sd, prefix, y / length)


# ----------------------------------------------------------------------------------------------------------------
# block, lift, head
# ----------------------------------------------------------------------------------------------------------------
def fuse_branches(sd, p, x, spec, vert):
    """The branch gates and the fusion MLP before its GroupNorm: fuse([s g_s, v g_v]) + (s g_s + v g_v) / 2
    (models/ffno_model.py:524-532)."""
    gates = torch.sigmoid(conv1x1(gelu(conv1x1(torch.cat([x, spec, vert], 1), sd[p + "branch_gate.0.weight"])),
                                  sd[p + "branch_gate.2.weight"], sd[p + "branch_gate.2.bias"]))
    sg, vg = gates.chunk(2, dim=1)
    sm, vm = spec * sg, vert * vg
    fused = conv1x1(gelu(conv1x1(torch.cat([sm, vm], 1), sd[p + "fuse.0.weight"])), sd[p + "fuse.2.weight"])
    return fused + 0.5 * (sm + vm)


def pw_mlp(sd, p, x1):
    """pw: Conv C->2C, GELU, Conv 2C->C (models/ffno_model.py:489-495, :545)."""
    return conv1x1(gelu(conv1x1(x1, sd[p + "pw.0.weight"])), sd[p + "pw.2.weight"])


def point_mlp(sd, p, x2):
    """mlp: PointwiseMLP C->2C->C (models/ffno_model.py:147-160, :548)."""
    return conv1x1(gelu(conv1x1(x2, sd[p + "mlp.net.0.weight"])), sd[p + "mlp.net.3.weight"])


def block(sd, i, x, z_scale, dx, dy, cfg, force_per_column=False):
    """FFNOBlock3dBalanced.forward in eval mode (models/ffno_model.py:536-553, :524-534)."""
    p = f"blocks.{i}."
    G = cfg["groups"]
    w = depth_weights(z_scale)

    def norm(name, t):
        return weighted_group_norm(t, w, sd[p + name + ".weight"], sd[p + name + ".bias"], G)

    spec = gelu(norm("norm_spec", horizontal_branch(sd, p + "spec.", x, dx, dy, cfg)))
    vert = gelu(norm("norm_vert", vertical_branch(sd, p + "vertical.", x, z_scale, cfg, force_per_column)))
    fused = gelu(norm("norm_fuse", fuse_branches(sd, p, x, spec, vert)))
    x1 = x + sd[p + "res_fused"] * fused
    x2 = x1 + sd[p + "res_pw"] * norm("norm_pw", pw_mlp(sd, p, x1))
    return x2 + sd[p + "res_mlp"] * norm("norm_mlp", point_mlp(sd, p, x2))


# MULTI3D constants in cgs (models/ffno_model.py:111-113)
M3D_EE, M3D_HH, M3D_EM, M3D_BK, M3D_PI = 1.602189e-12, 6.626176e-27, 9.109534e-28, 1.380662e-16, 3.14159265359
M3D_LN_CCON = math.log(0.5 * (M3D_HH / math.sqrt(2.0 * M3D_PI * M3D_EM) / math.sqrt(M3D_BK)) ** 3)


def lte_tables(species):
    """Per level, over all species: ln(g_i/g_1), ev_i, Debye multiple ndxi(ion_i), ion_i - ion_1
    (models/ffno_model.py:116-128, :676-677)."""
    out = ([], [], [], [])
    for s in species:
        atom, levels = s["multi3d_lte"], int(s["levels"])
        g, ev, ion = [float(v) for v in atom["g"]], [float(v) for v in atom["ev"]], [int(v) for v in atom["ion"]]
        if not len(g) == len(ev) == len(ion) == levels:
            raise ValueError("multi3d_lte atom data does not match the level count")
        nd = [0] * 21
        for i in range(2, 21):
            nd[i] = nd[i - 1] + (i - 1)
        for i in range(1, 21):              # MULTI3D zeroes ndxi(ion_1) on the way; later stages keep theirs
            nd[i] -= nd[ion[0]]
        out[0].extend(math.log(gi / g[0]) for gi in g)
        out[1].extend(ev)
        out[2].extend(float(nd[i]) for i in ion)
        out[3].extend(float(max(i - ion[0], 0)) for i in ion)
    return out


def lte_log_weights(log10_t, log10_ne, tables):
    """ln of MULTI3D's unnormalised LTE weights [B, levels, ...] in float64 (models/ffno_model.py:131-141)."""
    t = torch.pow(10.0, log10_t.double())[:, None]
    ne = torch.pow(10.0, log10_ne.double() - 6.0)[:, None]
    shape = (1, -1) + (1,) * (t.dim() - 2)
    tab = [torch.tensor(v, dtype=torch.float64).view(shape) for v in tables]
    conl = M3D_LN_CCON + torch.log(ne) - 1.5 * torch.log(t)
    dxi = 4.98e-4 * 5040.0 / t * torch.sqrt(ne * M3D_BK * t)
    return tab[0] - (M3D_EE / M3D_BK) / t * (tab[1] - dxi * tab[2]) - tab[3] * conl


def head_output(sd, head, x_in, cfg):
    """Level-fraction head (models/ffno_model.py:730-749): per species
    log10 n = log10 rho - log10 m + 6 tanh(u/6) + log_softmax(head + ln f*)/ln 10, normalised by mean_Y, std_Y."""
    mx, sx = sd["mean_X"], sd["std_X"]
    r, t, e = cfg["rho_channel"], cfg["temperature_channel"], cfg["ne_channel"]
    log10_rho = x_in[:, r] * sx[r] + mx[r]
    lte = lte_log_weights(x_in[:, t].double() * sx[t].double() + mx[t].double(),
                          x_in[:, e].double() * sx[e].double() + mx[e].double(), lte_tables(cfg["species"]))
    out = []
    c = k = 0
    bound = cfg["total_bound"]
    for i, levels in enumerate(cfg["levels"]):
        logits = head[:, c:c + levels].double() + lte[:, k:k + levels]
        frac = (torch.log_softmax(logits, dim=1) / math.log(10.0)).to(head.dtype)
        s = bound * torch.tanh(head[:, c + levels] / bound)
        out.append((log10_rho - sd["log10_particle_mass"][i] + s)[:, None] + frac)
        c += levels + 1
        k += levels
    y = torch.cat(out, dim=1)
    return (y - sd["mean_Y"][None, :, None, None, None]) / sd["std_Y"][None, :, None, None, None]


def forward(sd, cfg, x, z_scale, dx, dy, force_per_column=False):
    """FFNO3D.forward in eval mode (models/ffno_model.py:800-832): lift (:782-788), blocks, head (:790-798).
    x [B, 6, D, H, W] normalised, z_scale [B, D, H, W] in Mm, dx, dy in m. Returns normalised log10 n."""
    dx, dy = float(dx), float(dy)
    w = depth_weights(z_scale)
    h = conv1x1(x, sd["lift.0.weight"])
    h = gelu(weighted_group_norm(h, w, sd["lift.1.weight"], sd["lift.1.bias"], cfg["groups"]))
    for i in range(cfg["n_layers"]):
        h = block(sd, i, h, z_scale, dx, dy, cfg, force_per_column)
    head = conv1x1(gelu(conv1x1(h, sd["proj1.weight"])), sd["proj2.weight"], sd["proj2.bias"])
    return head_output(sd, head, x, cfg)


def to_dtype(sd, dtype):
    return {k: (v.to(dtype) if torch.is_floating_point(v) else v) for k, v in sd.items()}


def vjp(sd, cfg, x, z_scale, dx, dy, cotangent, force_per_column=False):
    """(y, x_bar, z_bar) by torch autograd over this restatement."""
    x = x.detach().clone().requires_grad_(True)
    z = z_scale.detach().clone().requires_grad_(True)
    y = forward(sd, cfg, x, z, dx, dy, force_per_column)
    xb, zb = torch.autograd.grad(y, (x, z), grad_outputs=cotangent)
    return y.detach(), xb, zb


# ----------------------------------------------------------------------------------------------------------------
# service-level units (ffno_fsdp_runtime.py)
# ----------------------------------------------------------------------------------------------------------------
def model_inputs_si(features, z_m):
    """(T, vx, vy, vz, log10 ne, log10 rho) [B,6,...] and z [m] -> training features and z in Mm
    (ffno_fsdp_runtime.py:48-58)."""
    f = torch.cat((torch.log10(features[:, 0:1]), features[:, 1:4] / 1.0e3, features[:, 4:6]), dim=1)
    return f, z_m / 1.0e6


def predict_si(sd, cfg, features, z_m, dx, dy, force_per_column=False):
    """Linear populations [B, L, D, H, W] from SI features (ffno_fsdp_runtime.py:265-281, :76-78). force_per_column:
    the vertical branch's per-column path even on a shared grid (the same values; the height gradient then goes to
    every column's own heights)."""
    f, z = model_inputs_si(features, z_m)
    mx = sd["mean_X"][None, :, None, None, None]
    sx = sd["std_X"][None, :, None, None, None]
    y = forward(sd, cfg, (f - mx) / sx, z, dx, dy, force_per_column)
    log = y * sd["std_Y"][None, :, None, None, None] + sd["mean_Y"][None, :, None, None, None]
    return torch.pow(10.0, log.clamp(-37.0, 37.0))
