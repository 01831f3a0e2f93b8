"""The production network family on the GPU: width 96 (the reference's final model, docs/technical/ffno-technical.tex:218)
and 1001 vertical modes (z_min_wavelength 0.04 at z_reference_length 40, config.py:294-311), more modes than
levels, so the vertical transform's matrices are 2002 x D and D x 2002 and the LDS-staged kernels fall back to their
partial-staging forms. Forward against the float64 oracle (1e-3 relative in log10 n) and the input and height
gradients against the oracle's autograd (1e-2 relative)."""
import numpy as np
import pytest
import torch

from golden_cases import oracle_config, regenerate_state

pytestmark = pytest.mark.gpu
DEV = "cuda"


def test_width96_with_more_modes_than_levels():
    from ffnoml_amd import synthetic
    from ffnoml_amd.atoms import species
    from ffnoml_amd.model import FFNO3D
    from oracle import ffno_oracle as O
    config = dict(in_channels=6, out_channels=12, width=96, n_layers=2, hidden_spectral=128, hidden_vertical=128,
                  freq_mix_rank=8, z_reference_length=40.0, z_min_wavelength=0.04, z_feature_wavelength=0.3,
                  species=species(["H", "CA"]))
    x, z, mx, sx = synthetic.atmosphere(1, 40, 12, 10, seed=0)
    # heights with distinct spacings: tied spacings make the taper's median spacing non-differentiable, and the
    # float64 oracle and the float32 engine may then send its sensitivity to different intervals
    col = torch.linspace(14.0, -1.0, 40, dtype=torch.float64)
    col = col + 0.04 * torch.sin(1.7 * torch.arange(40, dtype=torch.float64))
    z = col.float()[None, :, None, None].expand(1, 40, 12, 10).contiguous()
    state = regenerate_state(config, 0, mx, sx, np.zeros(12), np.ones(12))
    model = FFNO3D(**config)
    model.load_state_dict(state, strict=True)
    model = model.eval().to(DEV)
    assert model.blocks[0].vertical.modes == 1001
    ct = synthetic.cotangent((1, 12, 40, 12, 10), 3)
    print("[width96 M1001] forward + vjp", flush=True)
    xg = x.to(DEV).requires_grad_(True)
    zg = z.to(DEV).requires_grad_(True)
    y = model(xg, zg, 48e3, 48e3)
    xb, zb = torch.autograd.grad(y, (xg, zg), grad_outputs=ct.to(DEV))
    torch.cuda.synchronize()
    # the per-column path of the oracle gives the per-column height gradient the engine returns
    ref_y, ref_xb, ref_zb = O.vjp(O.to_dtype(state, torch.float64), oracle_config(config), x.double(), z.double(),
                                  48e3, 48e3, ct.double(), force_per_column=True)
    rel = float(((y.detach().cpu().double() - ref_y).abs() / ref_y.abs().clamp_min(1.0)).max())
    gx = float((xb.cpu().double() - ref_xb).norm() / ref_xb.norm())
    gz = float((zb.cpu().double() - ref_zb).norm() / ref_zb.norm())
    print(f"[width96 M1001] forward rel {rel:.3e}, x_bar {gx:.3e}, z_bar {gz:.3e}", flush=True)
    assert rel <= 1e-3
    assert gx <= 1e-2 and gz <= 1e-2
