"""The service path's SI chain (ffno_fsdp_runtime.py:48-58, :265-292, :76-78), pinned on the CPU.

1. The oracle's model_inputs_si and predict_si against closed forms, against the reference's own float64 outputs
   (golden case C, whose output normalisation is not the identity) carried into SI units, and against the clamp's
   known answers in the reference's tests (tests/test_batched_backend.py:57-70).
2. The product's backend (ffnoml_amd.backend.GPUBackend), whose SI chain rule around the engine's adjoint and tangent
   is written out by hand, against torch autograd through the oracle's predict_si: the backend runs on the CPU here
   with a stand-in engine that evaluates the float64 oracle (the HIP engine itself is checked on the GPU against the
   oracle in tests/test_gpu_vjp.py and tests/test_gpu_jvp.py).
"""
import math

import numpy as np
import pytest
import torch

from golden_cases import load, oracle_config
from oracle import ffno_oracle as O


def _si_from_normalised(x, mean_x, std_x):
    """Normalised training features [B, 6, ...] -> SI (T [K], v [m/s], log10 ne, log10 rho)."""
    f = x * std_x.view(1, -1, 1, 1, 1) + mean_x.view(1, -1, 1, 1, 1)
    return torch.cat((10.0 ** f[:, 0:1], f[:, 1:4] * 1e3, f[:, 4:6]), dim=1)


def test_model_inputs_si_closed_forms():
    """10^3.7 K -> 3.7; 3000 and -1500 m/s -> 3 and -1.5 km/s; log10 ne and log10 rho unchanged; 1.4e7 m -> 14 Mm."""
    feats = torch.tensor([10.0 ** 3.7, 3000.0, -1500.0, 0.0, 16.5, -7.0], dtype=torch.float64).view(1, 6, 1, 1, 1)
    f, z = O.model_inputs_si(feats, torch.full((1, 1, 1, 1), 1.4e7, dtype=torch.float64))
    assert torch.allclose(f.flatten(), torch.tensor([3.7, 3.0, -1.5, 0.0, 16.5, -7.0], dtype=torch.float64),
                          rtol=0, atol=1e-12)
    assert abs(float(z) - 14.0) < 1e-12
    rng = np.random.default_rng(0)
    t = 10.0 ** rng.uniform(3.3, 6.5, (1, 1, 3, 4, 5))
    v = rng.normal(0, 5e3, (1, 3, 3, 4, 5))
    rest = rng.normal(0, 10, (1, 2, 3, 4, 5))
    f, _ = O.model_inputs_si(torch.from_numpy(np.concatenate([t, v, rest], 1)), torch.zeros(1, 3, 4, 5))
    assert np.abs(f[:, 0:1].numpy() - np.log10(t)).max() < 1e-13
    assert np.abs(f[:, 1:4].numpy() - v / 1000.0).max() < 1e-13
    assert np.array_equal(f[:, 4:6].numpy(), rest)


def test_predict_si_reproduces_the_reference_in_si_units():
    """Golden case C (mean_Y 10..15, std_Y 1..2): the reference's float64 output y, de-normalised and raised to
    10^y, equals the oracle's predict_si on the same inputs given in SI units (1e-9 relative in log10 n)."""
    data, config, state = load("case_c_rank0")
    sd = O.to_dtype(state, torch.float64)
    cfg = oracle_config(config)
    x = torch.from_numpy(data["x"]).double()
    z = torch.from_numpy(data["z"]).double()
    si = _si_from_normalised(x, sd["mean_X"], sd["std_X"])
    n = O.predict_si(sd, cfg, si, z * 1e6, 48e3, 48e3)
    log_ref = torch.from_numpy(data["y"]) * sd["std_Y"].view(1, -1, 1, 1, 1) + sd["mean_Y"].view(1, -1, 1, 1, 1)
    inside = log_ref.abs() < 37.0
    assert bool(inside.any()) and bool((~inside).any())        # random inputs: both sides of the +-37 clamp occur
    err = (torch.log10(n) - log_ref.clamp(-37.0, 37.0)).abs().max() / 37.0
    assert float(err) < 1e-9


def test_population_clamp_known_answers(monkeypatch):
    """10^clamp(log10 n, -37, 37): 44.9 -> 37, 41.9 -> 37, -50 -> -37, 10 -> 10 (tests/test_batched_backend.py:57-70),
    in the oracle and in the backend."""
    from ffnoml_amd.backend import GPUBackend
    log10 = torch.full((1, 6, 2, 3, 2), 10.0, dtype=torch.float64)
    log10[0, 4, 1, 2, 1], log10[0, 5, 1, 2, 1], log10[0, 0, 0, 2, 1] = 44.9, 41.9, -50.0
    cfg = {"n_layers": 0}
    sd = dict(mean_X=torch.zeros(6, dtype=torch.float64), std_X=torch.ones(6, dtype=torch.float64),
              mean_Y=torch.zeros(6, dtype=torch.float64), std_Y=torch.ones(6, dtype=torch.float64))
    monkeypatch.setattr(O, "forward", lambda *a, **k: log10.clone())
    feats = torch.ones(1, 6, 2, 3, 2, dtype=torch.float64)
    for n in (O.predict_si(sd, cfg, feats, torch.zeros(1, 2, 3, 2, dtype=torch.float64), 48e3, 48e3),
              GPUBackend.linear(log10)):
        assert bool(torch.isfinite(n).all()) and bool((n > 0).all())
        assert abs(math.log10(float(n[0, 4, 1, 2, 1])) - 37.0) < 1e-12
        assert abs(math.log10(float(n[0, 5, 1, 2, 1])) - 37.0) < 1e-12
        assert abs(math.log10(float(n[0, 0, 0, 2, 1])) + 37.0) < 1e-12
        assert abs(math.log10(float(n[0, 1, 0, 0, 0])) - 10.0) < 1e-12
    # float32, as the service holds populations: 10^37 is finite, 10^44.9 would not be
    assert bool(torch.isfinite(GPUBackend.linear(log10.float())).all())


# ---------------------------------------------------------------------------------------------- the backend's chain
class _OracleEngine:
    """Stands in for ffnoml_amd.engine.Engine on the CPU: the same methods the backend calls, evaluated by the float64
    oracle (forward; adjoint by reverse-mode autograd; tangent by forward-mode autograd)."""

    def __init__(self, sd, cfg):
        self.sd, self.cfg = sd, cfg

    def sync(self, device):
        pass

    def _f(self, x, z, dx, dy):
        return O.forward(self.sd, self.cfg, x, z, dx, dy, force_per_column=True)

    def forward(self, x, z, dx, dy, plan=None):
        return self._f(x, z, dx, dy)

    def forward_saved(self, x, z, dx, dy, plan=None, stride=None, want_z=True, keep=0):
        return self._f(x, z, dx, dy), None

    def vjp(self, x, z, dx, dy, ybar, saved, want_z=False, plan=None):
        xg, zg = x.detach().requires_grad_(True), z.detach().requires_grad_(True)
        xb, zb = torch.autograd.grad(self._f(xg, zg, dx, dy), (xg, zg), grad_outputs=ybar)
        return xb, (zb if want_z else None)

    def jvp(self, x, z, dx, dy, xdot, plan=None):
        return torch.func.jvp(lambda t: self._f(t, z, dx, dy), (x,), (xdot,))


def _backend_and_inputs(rho_shift=0.0):
    from ffnoml_amd.backend import GPUBackend
    from ffnoml_amd.model import FFNO3D
    data, config, state = load("case_c_rank0")
    model = FFNO3D(**config)
    model.load_state_dict(state, strict=True)
    backend = GPUBackend(None, model=model, level_names=[f"L{i}" for i in range(config["out_channels"])],
                         device="cpu")
    sd = O.to_dtype(state, torch.float64)
    cfg = oracle_config(config)
    model.engine = lambda: _OracleEngine(sd, cfg)
    x = torch.from_numpy(data["x"]).double()[:1]
    z = torch.from_numpy(data["z"]).double()[:1]
    si = _si_from_normalised(x, sd["mean_X"], sd["std_X"])
    si[:, 5] += rho_shift
    return backend, sd, cfg, si, z * 1e6


def _rel(a, b):
    return float((a - b).abs().max() / b.abs().max().clamp_min(1e-300))


@pytest.mark.parametrize("rho_shift", [0.0, 30.0])
def test_backend_vjp_chain_matches_autograd_through_the_si_chain(rho_shift):
    """GPUBackend.vjp_tensor (the clamp, 10^, output and input normalisation, log10 T, km/s and Mm written out) equals
    torch autograd through the oracle's predict_si, per K, per m/s and per m (1e-10). Part of the outputs lie beyond
    the clamp, where it passes no gradient: below 10^-37 as the case stands, above 10^37 with log10 rho raised 30 dex."""
    backend, sd, cfg, si, z_m = _backend_and_inputs(rho_shift)
    rng = np.random.default_rng(1)
    cot = torch.from_numpy(rng.standard_normal((1, 6) + tuple(si.shape[2:])))
    n, fb, zb = backend.vjp_tensor(si, z_m, 48e3, 48e3, cot)
    fg, zg = si.clone().requires_grad_(True), z_m.clone().requires_grad_(True)
    # the per-column path gives every column's heights their own gradient, as the engine does on any grid (the
    # reference's shared-grid path sends the vertical transform's part to one column, models/ffno_model.py:270-276)
    ref = O.predict_si(sd, cfg, fg, zg, 48e3, 48e3, force_per_column=True)
    ref_f, ref_z = torch.autograd.grad(ref, (fg, zg), grad_outputs=cot)
    assert _rel(n, ref.detach()) < 1e-12
    log_n = torch.log10(ref.detach())
    assert bool((log_n > 37.0 - 1e-9).any()) == (rho_shift > 0) and bool((log_n < -37.0 + 1e-9).any())
    for c in range(6):
        assert _rel(fb[:, c], ref_f[:, c]) < 1e-10, c
    assert _rel(zb, ref_z) < 1e-10


def test_backend_jvp_chain_matches_forward_mode_through_the_si_chain():
    """GPUBackend.jvp_tensor equals torch forward-mode AD through the oracle's predict_si along an SI feature
    tangent (1e-10)."""
    backend, sd, cfg, si, z_m = _backend_and_inputs()
    rng = np.random.default_rng(2)
    scale = torch.tensor([100.0, 1e3, 1e3, 1e3, 0.1, 0.1], dtype=torch.float64).view(1, 6, 1, 1, 1)
    tangent = torch.from_numpy(rng.standard_normal(si.shape)) * scale
    n, nd = backend.jvp_tensor(si, z_m, 48e3, 48e3, tangent)
    ref, ref_d = torch.func.jvp(lambda f: O.predict_si(sd, cfg, f, z_m, 48e3, 48e3), (si,), (tangent,))
    assert _rel(n, ref) < 1e-12
    assert _rel(nd, ref_d) < 1e-10
