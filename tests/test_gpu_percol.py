"""Per-column height grids on the GPU (the reference's per-column vertical path, models/ffno_model.py:277-283,
:304-305): every column's own Filon analysis and synthesis matrices. Forward and input gradient against the
reference's own float64 outputs for golden case B (0.01 Mm jitter per column: forward, input and height gradients),
and forward, input and height gradients and the tangent against the float64 oracle at width 64 with two samples of
different jittered columns (the height gradient through every column's own Filon weights, phases, taper, depth
weights and shift).

Tolerances: 1e-3 relative in log10 n, 1e-2 relative for derivatives."""
import numpy as np
import pytest
import torch

from golden_cases import load, oracle_config, regenerate_state

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _model_from(config, state):
    from ffnoml_amd.model import FFNO3D
    m = FFNO3D(**config)
    m.load_state_dict(state, strict=True)
    return m.eval().to(DEV)


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_jitter_golden_forward_and_input_gradient():
    data, config, state = load("case_b_jitter64")
    model = _model_from(config, state)
    x = torch.from_numpy(data["x"]).float().to(DEV).requires_grad_(True)
    z = torch.from_numpy(data["z"]).float().to(DEV).requires_grad_(True)
    ct = torch.from_numpy(data["cotangent"]).float().to(DEV)
    print("[jitter] forward + vjp", flush=True)
    y = model(x, z, 48e3, 48e3)
    assert model.engine().plan(*z.shape, z.detach(), 48e3, 48e3).percol
    xb, zb = torch.autograd.grad(y, (x, z), grad_outputs=ct)
    torch.cuda.synchronize()
    ref = data["y"]
    sy, my = state["std_Y"].numpy().reshape(1, -1, 1, 1, 1), state["mean_Y"].numpy().reshape(1, -1, 1, 1, 1)
    a, b = y.detach().cpu().numpy() * sy + my, ref * sy + my
    rel = float((np.abs(a - b) / np.maximum(np.abs(b), 1.0)).max())
    gx = _rel(xb.cpu().numpy(), data["xbar"])
    gz = _rel(zb.cpu().numpy(), data["zbar"])
    print(f"[jitter] forward rel {rel:.3e}, x_bar rel {gx:.3e}, z_bar rel {gz:.3e}", flush=True)
    assert rel <= 1e-3
    assert gx <= 1e-2 and gz <= 1e-2


def test_per_column_width64_against_the_oracle():
    from ffnoml_amd import synthetic
    from ffnoml_amd.atoms import species
    from oracle import ffno_oracle as O
    config = dict(in_channels=6, out_channels=12, width=64, n_layers=2, hidden_spectral=64, hidden_vertical=64,
                  freq_mix_rank=8, z_reference_length=40.0, z_min_wavelength=1.28, z_feature_wavelength=0.3,
                  species=species(["H", "CA"]))
    B, D, H, W = 2, 24, 10, 12
    x, z, mx, sx = synthetic.atmosphere(B, D, H, W, seed=1)
    rng = np.random.default_rng(7)
    z = (z.double() + torch.from_numpy(0.05 * rng.standard_normal((B, D, H, W)))).float()   # per-column heights
    state = regenerate_state(config, 0, mx, sx, np.zeros(12), np.ones(12))
    model = _model_from(config, state)
    ct = synthetic.cotangent((B, 12, D, H, W), 9)
    t = torch.randn(x.shape, generator=torch.Generator().manual_seed(3))
    print("[percol64] forward + vjp + jvp", flush=True)
    xg = x.to(DEV).requires_grad_(True)
    zg = z.to(DEV).requires_grad_(True)
    y = model(xg, zg, 48e3, 48e3)
    xb, zb = torch.autograd.grad(y, (xg, zg), grad_outputs=ct.to(DEV))
    _, yd = model.engine().jvp(x.to(DEV), z.to(DEV), 48e3, 48e3, t.to(DEV))
    torch.cuda.synchronize()
    sd = O.to_dtype(state, torch.float64)
    cfg = oracle_config(config)
    ry, rxb, rzb = O.vjp(sd, cfg, x.double(), z.double(), 48e3, 48e3, ct.double())
    _, ryd = torch.func.jvp(lambda xx: O.forward(sd, cfg, xx, z.double(), 48e3, 48e3), (x.double(),),
                            (t.double(),))
    rel = float(((y.detach().cpu().double() - ry).abs() / ry.abs().clamp_min(1.0)).max())
    gx = _rel(xb.cpu().numpy(), rxb.numpy())
    gt = _rel(yd.cpu().numpy(), ryd.numpy())
    gz = _rel(zb.cpu().numpy(), rzb.numpy())
    print(f"[percol64] forward rel {rel:.3e}, x_bar {gx:.3e}, z_bar {gz:.3e}, tangent {gt:.3e}", flush=True)
    assert rel <= 1e-3
    assert gx <= 1e-2 and gz <= 1e-2 and gt <= 1e-2
