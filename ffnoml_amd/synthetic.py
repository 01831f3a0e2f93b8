"""Synthetic stratified atmospheres and seeded weights for tests and benchmarks.

Atmosphere recipe (one per sample, seed-determined):
  heights  z = linspace(14, -1, D) Mm in every column (index 0 at the top), dx = dy = 48 km;
  log10 T  = 3.7 -> 6.0 along a tanh in height (centre 2 Mm, width 0.5 Mm) + a 0.05-dex Gaussian random field;
  vx,vy,vz = Gaussian random fields with sigma 3 km/s (model units: km/s);
  log10 ne = 20 -> 16 linear in height + 0.05-dex noise (m^-3);
  log10 rho= -3.5 -> -11 linear in height + 0.05-dex noise;
  every channel z-scored with the generated data's own per-channel mean and standard deviation.
The random fields are white noise smoothed by a Gaussian of sigma = 4 cells (circular, by FFT) and rescaled to unit
variance.
"""
import math

import numpy as np
import torch

DX = DY = 48.0e3
CORRELATION_CELLS = 4.0


def _smooth_field(rng, shape, sigma_cells):
    """Unit-variance Gaussian random field [*shape] (float64), white noise smoothed by a Gaussian of sigma_cells."""
    noise = rng.standard_normal(shape)
    spec = np.fft.rfftn(noise)
    freqs = [np.fft.fftfreq(n) for n in shape[:-1]] + [np.fft.rfftfreq(shape[-1])]
    grids = np.meshgrid(*freqs, indexing="ij")
    k2 = sum(g * g for g in grids)
    spec *= np.exp(-2.0 * (math.pi * sigma_cells) ** 2 * k2)
    field = np.fft.irfftn(spec, s=shape, axes=tuple(range(len(shape))))
    return field / max(field.std(), 1e-30)


def heights(D):
    return np.linspace(14.0, -1.0, D)


def atmosphere_features(B, D, H, W, seed=0):
    """Un-normalised model features [B, 6, D, H, W] float64 (log10 T, v [km/s], log10 ne, log10 rho) and z [D] Mm."""
    rng = np.random.default_rng(seed)
    z = heights(D)
    zc = z[None, :, None, None]
    up = (z - z.min()) / max(z.max() - z.min(), 1e-30)                  # 0 at the bottom, 1 at the top
    upc = up[None, :, None, None]
    feats = np.empty((B, 6, D, H, W))
    shape = (B, D, H, W)
    feats[:, 0] = 3.7 + 2.3 * 0.5 * (1.0 + np.tanh((zc - 2.0) / 0.5)) \
        + 0.05 * _smooth_field(rng, shape, CORRELATION_CELLS)
    for c in (1, 2, 3):
        feats[:, c] = 3.0 * _smooth_field(rng, shape, CORRELATION_CELLS)
    feats[:, 4] = 20.0 + (16.0 - 20.0) * upc + 0.05 * rng.standard_normal(shape)
    feats[:, 5] = -3.5 + (-11.0 + 3.5) * upc + 0.05 * rng.standard_normal(shape)
    return feats, z


def atmosphere(B, D, H, W, seed=0):
    """(x [B,6,D,H,W] float32 z-scored, z_scale [B,D,H,W] float32 in Mm, mean_X [6], std_X [6])."""
    feats, z = atmosphere_features(B, D, H, W, seed)
    mean = feats.mean(axis=(0, 2, 3, 4))
    std = feats.std(axis=(0, 2, 3, 4))
    x = (feats - mean[None, :, None, None, None]) / std[None, :, None, None, None]
    zs = np.broadcast_to(z[None, :, None, None], (B, D, H, W)).copy()
    return (torch.from_numpy(x.astype(np.float32)), torch.from_numpy(zs.astype(np.float32)),
            mean.astype(np.float32), std.astype(np.float32))


def perturb_state(model, seed=0):
    """Make the parts that start at a fixed value live, as the reference tests do (tests/test_batched_backend.py:29-32,
    tests/test_level_fraction_head.py:116): freq_u ~ N(0, 0.1), the head's total rows ~ N(0, 0.3); and give the
    GroupNorm affines and residual scalars non-trivial values so that a transposed or dropped term shows."""
    g = torch.Generator().manual_seed(int(seed) + 12345)
    with torch.no_grad():
        for name, p in model.named_parameters():
            if ".freq_u." in name:
                p.copy_(0.1 * torch.randn(p.shape, generator=g))
            elif name.endswith(("norm_spec.weight", "norm_vert.weight", "norm_fuse.weight", "norm_pw.weight",
                                "norm_mlp.weight", "lift.1.weight")):
                p.copy_(1.0 + 0.1 * torch.randn(p.shape, generator=g))
            elif name.endswith(("norm_spec.bias", "norm_vert.bias", "norm_fuse.bias", "norm_pw.bias",
                                "norm_mlp.bias", "lift.1.bias")):
                p.copy_(0.1 * torch.randn(p.shape, generator=g))
            elif name.endswith(("res_fused", "res_pw", "res_mlp")):
                p.copy_(p + 0.1 * torch.randn(p.shape, generator=g))
        c = 0
        for levels in model.species_levels:
            model.proj2.weight[c + levels].copy_(0.3 * torch.randn(model.proj2.weight[c + levels].shape, generator=g))
            model.proj2.bias[c + levels].copy_(0.3 * torch.randn((), generator=g))
            c += levels + 1
    return model


def random_model(seed=0, **config):
    """An FFNO3D (ffnoml_amd.model) with torch's default initialisation under manual_seed(seed), then perturb_state."""
    from .model import FFNO3D
    torch.manual_seed(int(seed))
    return perturb_state(FFNO3D(**config).eval(), seed)


def si_request(D, H, W, seed=None, levels=12, cotangent_scale=1e-15):
    """One service request in the SI units the client sends: features (6, nz, nx, ny) float32 = (T [K], vx, vy, vz
    [m/s], log10 ne, log10 rho) of a synthetic atmosphere, heights (nz, nx, ny) [m] and a population cotangent
    (nz, nx, ny, levels) of N(0, 1) * cotangent_scale. seed None: D + H + W."""
    seed = D + H + W if seed is None else int(seed)
    feats, zc = atmosphere_features(1, D, H, W, seed=seed)
    si = feats[0].copy()
    si[0] = 10.0 ** si[0]
    si[1:4] *= 1e3
    z = np.broadcast_to(zc[:, None, None] * 1e6, si.shape[1:]).astype(np.float32).copy()
    cot = np.random.default_rng(seed).standard_normal((D, H, W, levels), dtype=np.float32) * np.float32(cotangent_scale)
    return si.astype(np.float32), z, cot


def perturbation(shape, seed=0, scale=0.02):
    """scale * N(0, 1) in normalised input units, shape [B, 6, D, H, W]: the inputs whose prediction is the inversion
    loss's target. Item b comes from a generator of its own (seed, b), so that any batch item regenerates alone."""
    out = torch.empty(shape, dtype=torch.float32)
    for b in range(shape[0]):
        g = torch.Generator().manual_seed(int(seed) * 1000 + 555 + b)
        out[b] = scale * torch.randn(shape[1:], generator=g, dtype=torch.float32)
    return out


def cotangent(shape, seed=0):
    g = torch.Generator().manual_seed(int(seed) + 777)
    return torch.randn(shape, generator=g, dtype=torch.float32)
