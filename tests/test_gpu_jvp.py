"""Tangent (forward-mode JVP) parity on the GPU: the hand-written tangent kernels, reached through torch.func.jvp on
ffnoml_amd.FFNO3D, against forward-mode autograd through the float64 oracle; the tangent against the adjoint by the
dot-product identity <u, J t> = <J^T u, t> at the benchmark's width and depth; and GPUBackend.jvp_local (SI units)
against the oracle's SI chain.

Tolerance (the accuracy target for derivatives): 1e-2 relative, ||g - g_ref|| / ||g_ref||."""
import numpy as np
import pytest
import torch

from golden_cases import load, oracle_config, regenerate_state

pytestmark = pytest.mark.gpu
DEV = "cuda"
TOL = 1e-2


def _model_from(config, state):
    from ffnoml_amd.model import FFNO3D
    m = FFNO3D(**config)
    m.load_state_dict(state, strict=True)
    return m.eval().to(DEV)


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("name", ["case_a_shared64", "case_c_rank0", "case_e_tall"])
def test_jvp_matches_oracle_forward_mode(name):
    from oracle import ffno_oracle as O
    data, config, state = load(name)
    model = _model_from(config, state)
    x = torch.from_numpy(data["x"]).float()
    z = torch.from_numpy(data["z"]).float()
    t = torch.randn(x.shape, generator=torch.Generator().manual_seed(11))
    print(f"[{name}] jvp {tuple(x.shape)}", flush=True)
    y, yd = torch.func.jvp(lambda xx: model(xx, z.to(DEV), 48e3, 48e3), (x.to(DEV),), (t.to(DEV),))
    torch.cuda.synchronize()
    sd = O.to_dtype(state, torch.float64)
    cfg = oracle_config(config)
    ry, ryd = torch.func.jvp(lambda xx: O.forward(sd, cfg, xx, z.double(), 48e3, 48e3), (x.double(),),
                             (t.double(),))
    e_y = float(((y.cpu().double() - ry).abs() / ry.abs().clamp_min(1.0)).max())
    e_t = _rel(yd.cpu().numpy(), ryd.numpy())
    print(f"[{name}] forward rel {e_y:.2e}, tangent rel {e_t:.2e}", flush=True)
    assert e_y <= 1e-3
    assert e_t <= TOL


def test_jvp_is_the_transpose_of_the_vjp():
    """<u, J t> (tangent kernels) against <J^T u, t> (adjoint kernels) at width 64, 8 layers, 32 vertical modes,
    B=2, D=48, 24 x 20: both sides from the HIP engine, agreeing to 1e-3 relative."""
    from ffnoml_amd import synthetic
    from ffnoml_amd.atoms import species
    config = dict(in_channels=6, out_channels=12, width=64, n_layers=8, hidden_spectral=128, hidden_vertical=128,
                  freq_mix_rank=8, z_reference_length=40.0, z_min_wavelength=1.28, z_feature_wavelength=0.3,
                  species=species(["H", "CA"]))
    x, z, mx, sx = synthetic.atmosphere(2, 48, 24, 20, seed=0)
    state = regenerate_state(config, 0, mx, sx, np.zeros(12), np.ones(12))
    model = _model_from(config, state)
    eng = model.engine()
    x, z = x.to(DEV), z.to(DEV)
    g = torch.Generator(device="cpu").manual_seed(5)
    t = torch.randn(x.shape, generator=g).to(DEV)
    u = torch.randn((2, 12) + tuple(x.shape[2:]), generator=g).to(DEV)
    print("[transpose] jvp + vjp", flush=True)
    _, yd = eng.jvp(x, z, 48e3, 48e3, t)
    xg = x.clone().requires_grad_(True)
    y = model(xg, z, 48e3, 48e3)
    (xb,) = torch.autograd.grad(y, xg, grad_outputs=u)
    lhs = float((u.double() * yd.double()).sum())
    rhs = float((xb.double() * t.double()).sum())
    print(f"[transpose] <u, Jt> {lhs:.8e}  <J^T u, t> {rhs:.8e}", flush=True)
    assert abs(lhs - rhs) <= 1e-3 * abs(rhs)


def test_backend_jvp_local_matches_si_forward_mode():
    """GPUBackend.jvp_local (SI feature tangent, linear populations out) against forward-mode autograd through the
    float64 oracle's SI chain (ffno_fsdp_runtime.py:48-58, :76-78, :350-372)."""
    from ffnoml_amd import synthetic
    from ffnoml_amd.backend import GPUBackend
    from ffnoml_amd.atoms import species
    from oracle import ffno_oracle as O
    config = dict(in_channels=6, out_channels=6, width=32, n_layers=2, hidden_spectral=32, hidden_vertical=32,
                  freq_mix_rank=4, z_reference_length=40.0, z_min_wavelength=1.28, z_feature_wavelength=0.3,
                  species=species(["H"]))
    feats, zc = synthetic.atmosphere_features(1, 20, 10, 12, seed=3)
    mean = feats.mean(axis=(0, 2, 3, 4)).astype(np.float32)
    std = feats.std(axis=(0, 2, 3, 4)).astype(np.float32)
    si = feats[0].copy()
    si[0] = 10.0 ** si[0]
    si[1:4] *= 1e3
    si = si.astype(np.float32)
    z_m = np.broadcast_to(zc[:, None, None] * 1e6, si.shape[1:]).astype(np.float32).copy()
    model = synthetic.random_model(4, **config)
    model.set_normalization(mean, std, np.zeros(6), np.ones(6))
    state = {k: v.detach().clone() for k, v in model.state_dict().items()}
    backend = GPUBackend(None, model=model, level_names=[f"H{i}" for i in range(6)])
    rng = np.random.default_rng(2)
    tangent = (rng.standard_normal(si.shape) * np.array([50.0, 300.0, 300.0, 300.0, 0.02, 0.02])[:, None, None, None]
               ).astype(np.float32)
    got = backend.jvp_local(si, z_m, 48e3, 48e3, tangent)
    sd = O.to_dtype(state, torch.float64)
    cfg = oracle_config(config)
    f64 = torch.from_numpy(si).double()[None]
    z64 = torch.from_numpy(z_m).double()[None]
    _, nd = torch.func.jvp(lambda f: O.predict_si(sd, cfg, f, z64, 48e3, 48e3), (f64,),
                           (torch.from_numpy(tangent).double()[None],))
    ref = nd[0].permute(1, 2, 3, 0).numpy()
    e = _rel(got, ref)
    print(f"[backend] jvp_local rel {e:.2e}", flush=True)
    assert e <= TOL
