"""FFNO3D as a torch.nn.Module whose parameters and buffers carry the reference's state_dict keys and shapes
(models/ffno_model.py:147-832), so that checkpoints load with strict=True in both directions, and whose forward
runs the HIP engine (ffnoml_amd.engine) on the GPU.

There is no CPU path: forward on a CPU tensor raises. Differentiation goes through ffnoml_amd.autograd.FFNOFunction,
whose backward is the engine's hand-written adjoint.
"""
import math

import torch
import torch.nn as nn


def group_count(channels, max_groups=8):
    """GroupNorm group count of the reference (models/ffno_model.py:21-25)."""
    g = min(max_groups, channels)
    while g > 1 and channels % g != 0:
        g -= 1
    return g


class DepthWeightedGroupNorm(nn.GroupNorm):
    """Parameter holder of the reference's DepthWeightedGroupNorm (models/ffno_model.py:77-95)."""


def _wgn(channels):
    return DepthWeightedGroupNorm(group_count(channels), channels)


def _gate(channels):
    return nn.Sequential(nn.Conv3d(channels, channels, 1, bias=False), nn.GELU(),
                         nn.Conv3d(channels, channels, 1, bias=False), nn.Sigmoid())


def _post(channels):
    return nn.Sequential(nn.Conv3d(channels, 2 * channels, 1, bias=False), nn.GELU(),
                         nn.Conv3d(2 * channels, channels, 1, bias=False))


def _freq_mlp(n_in, hidden):
    return nn.Sequential(nn.Linear(n_in, hidden), nn.GELU(), nn.Linear(hidden, hidden), nn.GELU(),
                         nn.Linear(hidden, 2))


def _freq_lowrank(n_in, hidden):
    return nn.Sequential(nn.Linear(n_in, hidden), nn.GELU(), nn.Linear(hidden, hidden), nn.GELU())


class SpectralConv2dFull(nn.Module):
    """Parameters of the horizontal branch (models/ffno_model.py:315-373)."""

    def __init__(self, channels, hidden=128, rank=0):
        super().__init__()
        self.rank = int(rank)
        scale = 1.0 / math.sqrt(channels)
        self.weight_real = nn.Parameter(scale * torch.randn(channels, channels))
        self.weight_imag = nn.Parameter(scale * torch.randn(channels, channels))
        self.input_gate = _gate(channels)
        self.freq_mlp = _freq_mlp(4, hidden)
        if self.rank:
            self.freq_lowrank = _freq_lowrank(4, hidden)
            self.freq_u = nn.Linear(hidden, channels * self.rank)
            self.freq_v = nn.Linear(hidden, self.rank * channels)
            with torch.no_grad():
                self.freq_u.weight.zero_()
                self.freq_u.bias.zero_()
        self.post = _post(channels)


class HeightSpectralConv(nn.Module):
    """Parameters and wavenumber constants of the vertical branch (models/ffno_model.py:166-229)."""

    def __init__(self, channels, reference_length, min_wavelength, feature_wavelength, hidden=128, rank=0):
        super().__init__()
        if reference_length <= 0 or min_wavelength <= 0 or feature_wavelength <= 0:
            raise ValueError("the height operator needs z_reference_length, z_min_wavelength and "
                             "z_feature_wavelength > 0")
        self.reference_length = float(reference_length)
        self.min_wavelength = float(min_wavelength)
        self.feature_wavelength = float(feature_wavelength)
        self.modes = int(math.floor(self.reference_length / self.min_wavelength)) + 1
        self.rank = int(rank)
        scale = 1.0 / math.sqrt(channels)
        self.weight_real = nn.Parameter(scale * torch.randn(channels, channels))
        self.weight_imag = nn.Parameter(scale * torch.randn(channels, channels))
        self.input_gate = _gate(channels)
        self.z_proj = nn.Conv3d(1, channels, 1, bias=False)
        self.freq_mlp = _freq_mlp(2, hidden)
        if self.rank:
            self.freq_lowrank = _freq_lowrank(2, hidden)
            self.freq_u = nn.Linear(hidden, channels * self.rank)
            self.freq_v = nn.Linear(hidden, self.rank * channels)
            with torch.no_grad():
                self.freq_u.weight.zero_()
                self.freq_u.bias.zero_()
        self.post = _post(channels)


class PointwiseMLP(nn.Module):
    """models/ffno_model.py:147-160 (the Identity keeps the Dropout slot, so the second conv is net.3)."""

    def __init__(self, width, expansion=2):
        super().__init__()
        hidden = width * expansion
        self.net = nn.Sequential(nn.Conv3d(width, hidden, 1, bias=False), nn.GELU(), nn.Identity(),
                                 nn.Conv3d(hidden, width, 1, bias=False))


class FFNOBlock3dBalanced(nn.Module):
    """Parameters of one block (models/ffno_model.py:456-522)."""

    def __init__(self, width, hidden_spectral, hidden_vertical, freq_mix_rank, z_reference_length, z_min_wavelength,
                 z_feature_wavelength):
        super().__init__()
        self.spec = SpectralConv2dFull(width, hidden=hidden_spectral, rank=freq_mix_rank)
        self.vertical = HeightSpectralConv(width, z_reference_length, z_min_wavelength, z_feature_wavelength,
                                           hidden=hidden_vertical, rank=freq_mix_rank)
        self.pw = _post(width)
        self.mlp = PointwiseMLP(width, expansion=2)
        self.norm_spec = _wgn(width)
        self.norm_vert = _wgn(width)
        self.branch_gate = nn.Sequential(nn.Conv3d(3 * width, width, 1, bias=False), nn.GELU(),
                                         nn.Conv3d(width, 2 * width, 1, bias=True), nn.Sigmoid())
        self.fuse = nn.Sequential(nn.Conv3d(2 * width, 2 * width, 1, bias=False), nn.GELU(),
                                  nn.Conv3d(2 * width, width, 1, bias=False))
        self.norm_fuse = _wgn(width)
        self.norm_pw = _wgn(width)
        self.norm_mlp = _wgn(width)
        self.res_fused = nn.Parameter(torch.tensor([1.0]))
        self.res_pw = nn.Parameter(torch.tensor([0.5]))
        self.res_mlp = nn.Parameter(torch.tensor([0.5]))


def lte_tables(species):
    """Per level over all species: ln(g_i/g_1), ev_i, the Debye-lowering multiple and ion_i - ion_1, as MULTI3D's
    calc_lte_populations uses them (models/ffno_model.py:116-128)."""
    out = ([], [], [], [])
    for s in species:
        atom, levels = s["multi3d_lte"], int(s["levels"])
        g, ev, ion = [float(v) for v in atom["g"]], [float(v) for v in atom["ev"]], [int(v) for v in atom["ion"]]
        if not len(g) == len(ev) == len(ion) == levels:
            raise ValueError(f"multi3d_lte atom data has {len(g)}, {len(ev)}, {len(ion)} entries for {levels} levels")
        stage = [0] * 21                  # Fortran ndxi(1:20) by ionisation stage: 0, 0, 1, 3, 6, ...
        for i in range(2, 21):
            stage[i] = stage[i - 1] + (i - 1)
        # MULTI3D subtracts ndxi(ion_1) in a loop that zeroes ndxi(ion_1) on its way: stages below ion_1 are
        # shifted, ion_1 becomes 0 and the stages above keep their values (Ca III: 3, not 2)
        for i in range(1, 21):
            stage[i] -= stage[ion[0]]
        out[0].extend(math.log(gi / g[0]) for gi in g)
        out[1].extend(ev)
        out[2].extend(float(stage[i]) for i in ion)
        out[3].extend(float(max(i - ion[0], 0)) for i in ion)
    return out


class FFNO3D(nn.Module):
    """Drop-in for models/ffno_model.py:614-832: same constructor, same state_dict, forward(x, z_scale, dx, dy)
    -> normalised log10 populations [B, out_channels, D, H, W], computed by the HIP engine."""

    def __init__(self, in_channels=6, out_channels=6, width=96, n_layers=4, dropout=0.0, spec_dropout=0.0,
                 vertical_dropout=0.0, spec_dropout_layers=None, vertical_dropout_layers=None, checkpoint_blocks=True,
                 hidden_spectral=128, hidden_vertical=128, freq_mix_rank=8, z_reference_length=40.0,
                 z_min_wavelength=0.04, z_feature_wavelength=0.3, species=None, rho_channel=5, total_bound=6.0,
                 temperature_channel=0, ne_channel=4):
        super().__init__()
        if not species:
            raise ValueError("FFNO3D needs species (levels, log10_particle_mass, multi3d_lte)")
        self.species_levels = [int(s["levels"]) for s in species]
        if sum(self.species_levels) != out_channels:
            raise ValueError(f"species levels {self.species_levels} do not add up to out_channels={out_channels}")
        missing = [i for i, s in enumerate(species) if "multi3d_lte" not in s]
        if missing:
            raise ValueError(f"the level-fraction head needs each species' multi3d_lte atom data (g, ev, ion); "
                             f"missing for {missing}")
        if in_channels != 6:
            raise ValueError("the level-fraction head reads T, ne and rho from the six input channels")
        head_channels = out_channels + len(self.species_levels)
        self.in_channels, self.out_channels, self.width, self.n_layers = in_channels, out_channels, width, n_layers
        self.hidden_spectral, self.hidden_vertical, self.freq_mix_rank = hidden_spectral, hidden_vertical, freq_mix_rank
        self.z_reference_length, self.z_min_wavelength = float(z_reference_length), float(z_min_wavelength)
        self.z_feature_wavelength = float(z_feature_wavelength)
        self.total_bound = float(total_bound)
        self.rho_channel, self.temperature_channel, self.ne_channel = int(rho_channel), int(temperature_channel), \
            int(ne_channel)
        self.species = [dict(s) for s in species]
        self.register_buffer("log10_particle_mass",
                             torch.tensor([float(s["log10_particle_mass"]) for s in species], dtype=torch.float32))
        self.register_buffer("mean_X", torch.zeros(in_channels))
        self.register_buffer("std_X", torch.ones(in_channels))
        self.register_buffer("mean_Y", torch.zeros(out_channels))
        self.register_buffer("std_Y", torch.ones(out_channels))
        self.lte_tables = lte_tables(species)
        self.checkpoint_blocks = checkpoint_blocks
        self.external_block_checkpointing = False
        self.lift = nn.Sequential(nn.Conv3d(in_channels, width, 1, bias=False), _wgn(width), nn.GELU())
        self.blocks = nn.ModuleList([
            FFNOBlock3dBalanced(width, hidden_spectral, hidden_vertical, freq_mix_rank, z_reference_length,
                                z_min_wavelength, z_feature_wavelength)
            for _ in range(n_layers)])
        self.proj1 = nn.Conv3d(width, width * 2, 1, bias=False)
        self.proj2 = nn.Conv3d(width * 2, head_channels, 1)
        with torch.no_grad():                   # the total corrections start at s = 0 (models/ffno_model.py:714-720)
            c = 0
            for levels in self.species_levels:
                self.proj2.weight[c + levels].zero_()
                self.proj2.bias[c + levels].zero_()
                c += levels + 1
        self._engine = None

    def set_normalization(self, mean_X, std_X, mean_Y, std_Y):
        """Copy the dataset's normalisation into the buffers (models/ffno_model.py:724-728)."""
        for name, value in (("mean_X", mean_X), ("std_X", std_X), ("mean_Y", mean_Y), ("std_Y", std_Y)):
            buffer = getattr(self, name)
            buffer.copy_(torch.as_tensor(value, dtype=buffer.dtype).broadcast_to(buffer.shape))

    def engine(self):
        """The HIP engine bound to this module's parameters (built on first use, repacked when they change)."""
        from .engine import Engine
        if self._engine is None:
            self._engine = Engine(self)
        return self._engine

    def head_output(self, head, x_in):
        """Normalised log10 populations [B, Cout, D, H, W] from the head's channels head [B, Cout + S, D, H, W] and
        the normalised inputs x_in [B, 6, D, H, W] (models/ffno_model.py:730-749), by the HIP library in float32;
        differentiable in both arguments."""
        from .autograd import head_output_apply
        return head_output_apply(self, head, x_in)

    def forward(self, x, z_scale, dx, dy, collect_stats=False, branch_mask=None):
        """models/ffno_model.py:800-832. collect_stats and branch_mask, the reference's training-statistics and
        branch-ablation switches, are accepted at their defaults only."""
        if collect_stats or branch_mask is not None:
            raise NotImplementedError("collect_stats and branch_mask (per-block statistics and branch ablation) are "
                                      "not part of the HIP path; call forward with their defaults")
        from .autograd import ffno_apply
        return ffno_apply(self, x, z_scale, dx, dy)
