"""FFNO3D.head_output on the GPU (ffno_head_output, ffno_head_output_bwd) against the float64 oracle's head_output
(models/ffno_model.py:730-749), whose LTE weights are pinned to an independent MULTI3D transcription in
tests/test_oracle.py. Tolerances: 2e-5 of the output range forward (float32 LTE and log-softmax), 1e-4 relative for
the gradients; the reference test's reproduction of an arbitrary target (tests/test_level_fraction_head.py:53-73) to
1e-4 dex."""
import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _model():
    from ffnoml_amd import synthetic
    from ffnoml_amd.atoms import species
    config = dict(in_channels=6, out_channels=12, width=16, n_layers=1, hidden_spectral=16, hidden_vertical=16,
                  freq_mix_rank=2, species=species(["H", "CA"]))
    model = synthetic.random_model(3, **config)
    model.set_normalization(np.array([3.8, 0, 0, 0, 16, -7], np.float32), np.array([0.4, 3, 3, 3, 1.2, 2.2], np.float32),
                            np.linspace(8.0, 14.0, 12), np.linspace(1.0, 2.0, 12))
    return model


def _oracle(model):
    from oracle import ffno_oracle as O
    sd = O.to_dtype({k: v.detach().cpu() for k, v in model.state_dict().items()}, torch.float64)
    cfg = O.model_config(width=model.width, species=model.species, out_channels=model.out_channels)
    return O, sd, cfg


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def test_head_output_forward_and_gradients():
    model = _model()
    O, sd, cfg = _oracle(model)
    g = torch.Generator().manual_seed(0)
    head = 3.0 * torch.randn(2, 14, 9, 5, 7, generator=g)
    x_in = torch.randn(2, 6, 9, 5, 7, generator=g)
    model = model.to(DEV)
    hg = head.to(DEV).requires_grad_(True)
    xg = x_in.to(DEV).requires_grad_(True)
    y = model.head_output(hg, xg)
    yb = torch.randn(y.shape, generator=g)
    hb, xb = torch.autograd.grad(y, (hg, xg), grad_outputs=yb.to(DEV))
    h64 = head.double().requires_grad_(True)
    x64 = x_in.double().requires_grad_(True)
    y64 = O.head_output(sd, h64, x64, cfg)
    hb64, xb64 = torch.autograd.grad(y64, (h64, x64), grad_outputs=yb.double())
    e_y = _rel(y.detach().cpu(), y64.detach())
    e_h = _rel(hb.cpu(), hb64)
    e_x = _rel(xb.cpu(), xb64)
    print(f"head_output: forward {e_y:.2e}, head gradient {e_h:.2e}, input gradient {e_x:.2e}", flush=True)
    assert e_y < 2e-5
    assert e_h < 1e-4
    assert e_x < 1e-4
    # only the channels the head reads carry an input gradient
    others = [c for c in range(6) if c not in (model.temperature_channel, model.ne_channel, model.rho_channel)]
    assert float(xb[:, others].abs().max()) == 0.0


def test_head_output_reproduces_any_target():
    """Head channels built from target levels (ratios of level populations, and a total within the bound) are mapped
    back onto the target by the head (the reference's construction, tests/test_level_fraction_head.py:53-73)."""
    model = _model()
    model.set_normalization(model.mean_X, model.std_X, np.zeros(12), np.ones(12))
    O, sd, cfg = _oracle(model)
    g = torch.Generator().manual_seed(11)
    x = torch.randn(1, 6, 5, 4, 3, generator=g, dtype=torch.float64)
    mx, sx = sd["mean_X"], sd["std_X"]
    lte = O.lte_log_weights(x[:, 0] * sx[0] + mx[0], x[:, 4] * sx[4] + mx[4], O.lte_tables(cfg["species"]))
    head = torch.empty(1, 14, 5, 4, 3, dtype=torch.float64)
    target = torch.empty(1, 12, 5, 4, 3, dtype=torch.float64)
    for sp in (0, 1):
        ratios = torch.randn(1, 6, 5, 4, 3, generator=g, dtype=torch.float64) * 6.0
        frac = ratios - torch.logsumexp(ratios * math.log(10.0), 1, keepdim=True) / math.log(10.0)
        s = torch.tensor([-3.16, -0.5, 0.0, 1.93, 5.4], dtype=torch.float64)[
            torch.randint(0, 5, (1, 5, 4, 3), generator=g)]
        ref = x[:, 5] * sx[5] + mx[5] - sd["log10_particle_mass"][sp]
        target[:, 6 * sp:6 * sp + 6] = ref[:, None] + s[:, None] + frac
        head[:, 7 * sp:7 * sp + 6] = ratios * math.log(10.0) - lte[:, 6 * sp:6 * sp + 6]
        head[:, 7 * sp + 6] = 6.0 * torch.atanh(s / 6.0)
    model = model.to(DEV)
    with torch.no_grad():
        y = model.head_output(head.float().to(DEV), x.float().to(DEV))
    err = float((y.double().cpu() - target).abs().max())
    print(f"head_output reproduces the target to {err:.2e} dex", flush=True)
    assert err < 1e-4


def test_head_output_rejects_bad_shapes():
    model = _model().to(DEV)
    with pytest.raises(ValueError):
        model.head_output(torch.zeros(1, 13, 2, 2, 2, device=DEV), torch.zeros(1, 6, 2, 2, 2, device=DEV))
    with pytest.raises(ValueError):
        model.head_output(torch.zeros(1, 14, 2, 2, 2, device=DEV), torch.zeros(1, 6, 2, 2, 3, device=DEV))
