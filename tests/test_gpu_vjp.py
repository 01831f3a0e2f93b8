"""Adjoint (input VJP) parity on the GPU: the hand-written backward behind torch.autograd against the reference's own
float64 autograd gradients (golden vectors), against the float64 oracle's autograd, and against float64 central
finite differences of the oracle.

Tolerance (the north-star bar for gradients): relative error <= 1e-2, measured as ||g - g_ref|| / ||g_ref|| over the
whole gradient and as max |g - g_ref| / max |g_ref|."""
import numpy as np
import pytest
import torch

from golden_cases import load, oracle_config, regenerate_state

pytestmark = pytest.mark.gpu
DEV = "cuda"
TOL = 1e-2


def _errs(g, ref):
    g = np.asarray(g, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.linalg.norm(g - ref) / np.linalg.norm(ref)), float(np.abs(g - ref).max() / np.abs(ref).max())


def _model_from(config, state):
    from ffno_amd.model import FFNO3D
    m = FFNO3D(**config)
    m.load_state_dict(state, strict=True)
    return m.eval().to(DEV)


@pytest.mark.parametrize("name", ["case_a_shared64", "case_c_rank0", "case_e_tall"])
def test_input_gradient_matches_reference_golden(name):
    data, config, state = load(name)
    model = _model_from(config, state)
    x = torch.from_numpy(data["x"]).float().to(DEV).requires_grad_(True)
    z = torch.from_numpy(data["z"]).float().to(DEV)
    ct = torch.from_numpy(data["cotangent"]).float().to(DEV)
    print(f"[{name}] vjp", flush=True)
    y = model(x, z, 48e3, 48e3)
    (xb,) = torch.autograd.grad(y, x, grad_outputs=ct)
    torch.cuda.synchronize()
    nrm, mx = _errs(xb.cpu().numpy(), data["xbar"])
    print(f"[{name}] x_bar rel norm {nrm:.3e} rel max {mx:.3e}", flush=True)
    assert nrm <= TOL and mx <= TOL


def test_inversion_loss_gradient_matches_oracle_and_finite_differences():
    """The inversion's loss: sum_i (log10 n_i - target_i)^2 with the target the prediction at inputs perturbed by
    0.02 sigma (normalised units). Width 64, 2 layers, B=1, D=24, 12 x 12, 32 vertical modes. The HIP gradient is
    compared with the float64 oracle's autograd gradient over the whole input, and with float64 central finite
    differences of the oracle (step 1e-4) at 64 random input entries."""
    from ffno_amd import synthetic
    from ffno_amd.atoms import species
    from oracle import ffno_oracle as O
    config = dict(in_channels=6, out_channels=12, width=64, n_layers=2, hidden_spectral=64, hidden_vertical=64,
                  freq_mix_rank=8, z_reference_length=40.0, z_min_wavelength=1.28, z_feature_wavelength=0.3,
                  species=species(["H", "CA"]))
    x, z, mx, sx = synthetic.atmosphere(1, 24, 12, 12, seed=0)
    state = regenerate_state(config, 0, mx, sx, np.zeros(12), np.ones(12))
    cfg = oracle_config(config)
    sd = O.to_dtype(state, torch.float64)
    g = torch.Generator().manual_seed(1)
    xt = x.double() + 0.02 * torch.randn(x.shape, generator=g, dtype=torch.float64)
    with torch.no_grad():
        target = O.forward(sd, cfg, xt, z.double(), 48e3, 48e3)

    def loss64(xx):
        return ((O.forward(sd, cfg, xx, z.double(), 48e3, 48e3) - target) ** 2).sum()

    xo = x.double().clone().requires_grad_(True)
    (g_oracle,) = torch.autograd.grad(loss64(xo), xo)
    model = _model_from(config, state)
    xg = x.to(DEV).requires_grad_(True)
    lg = ((model(xg, z.to(DEV), 48e3, 48e3) - target.float().to(DEV)) ** 2).sum()
    (g_hip,) = torch.autograd.grad(lg, xg)
    torch.cuda.synchronize()
    g_hip = g_hip.double().cpu()
    nrm, mxe = _errs(g_hip.numpy(), g_oracle.numpy())
    print(f"[loss] HIP vs oracle autograd: rel norm {nrm:.3e} rel max {mxe:.3e}", flush=True)
    assert nrm <= TOL and mxe <= TOL
    rng = np.random.default_rng(3)
    flat = x.numel()
    idx = rng.choice(flat, 64, replace=False)
    fd = []
    for k in idx:
        e = torch.zeros(flat, dtype=torch.float64)
        e[k] = 1e-4
        e = e.reshape(x.shape)
        with torch.no_grad():
            fd.append(float((loss64(x.double() + e) - loss64(x.double() - e)) / 2e-4))
    fd = np.array(fd)
    got = g_hip.reshape(-1).numpy()[idx]
    oracle_at = g_oracle.reshape(-1).numpy()[idx]
    nrm_fd, mx_fd = _errs(got, fd)
    print(f"[loss] HIP vs FD at 64 entries: rel norm {nrm_fd:.3e} rel max {mx_fd:.3e}; oracle vs FD "
          f"{_errs(oracle_at, fd)[0]:.3e}", flush=True)
    assert nrm_fd <= TOL and mx_fd <= TOL
