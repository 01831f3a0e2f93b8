"""Parity at the benchmark workloads' own sizes, against float64-oracle fixtures written by
tests/golden/make_workload_fixtures.py (inputs and weights regenerate from their seeds; each fixture's weight digest
is checked first).

Bars (the accuracy targets): <= 1e-3 relative in log10 n, max |d log10 n| / max(|log10 n|, 1) over the stored subsample;
<= 1e-2 relative in gradients, ||g - g_ref|| / ||g_ref|| over the stored subsample, and against the float64 central
differences max_i |g_i - fd_i| / max_i |fd_i|. Sums over the whole output check the entries between the subsample."""
import hashlib
import json
import os

import numpy as np
import pytest
import torch

from golden_cases import GOLDEN

pytestmark = pytest.mark.gpu
DEV = "cuda"
LOG_TOL, GRAD_TOL = 1e-3, 1e-2


def _fixture(name):
    path = os.path.join(GOLDEN, name + ".npz")
    if not os.path.exists(path):
        pytest.skip(f"{name}.npz not written yet (python tests/golden/make_workload_fixtures.py)")
    return np.load(path)


def _digest(state):
    h = hashlib.sha256()
    for k in sorted(state):
        h.update(k.encode())
        h.update(np.ascontiguousarray(state[k].detach().double().cpu().numpy()).tobytes())
    return h.hexdigest()


def _model(data, mx, sx):
    from ffnoml_amd import synthetic
    from ffnoml_amd.atoms import species
    config = json.loads(str(data["config_json"]))
    config["species"] = species(config["species"])
    model = synthetic.random_model(0, **config)
    model.set_normalization(mx, sx, np.zeros(12), np.ones(12))
    assert _digest(model.state_dict()) == str(data["state_sha256"]), "regenerated weights differ from the fixture's"
    return model.to(DEV)


def _sub(a, stride):
    return np.asarray(a)[..., ::stride, ::stride, ::stride]


def _log_err(y, ref):
    d = np.abs(np.asarray(y, np.float64) - np.asarray(ref, np.float64))
    return float((d / np.maximum(np.abs(ref), 1.0)).max())


def _grad_err(g, ref):
    g, ref = np.asarray(g, np.float64), np.asarray(ref, np.float64)
    return float(np.linalg.norm(g - ref) / np.linalg.norm(ref))


def _cfg1_batch():
    from ffnoml_amd import synthetic
    x, z, mx, sx = synthetic.atmosphere(8, 128, 64, 64, seed=0)
    return x, z, mx, sx


def test_cfg1_forward_batch():
    """B=8, D=128, 64 x 64, width 64, 8 layers, 32 vertical modes (the benchmark step's forward): batch item 0
    against the float64 oracle; every item finite."""
    data = _fixture("wl_cfg1_forward")
    x, z, mx, sx = _cfg1_batch()
    model = _model(data, mx, sx)
    print("[cfg1] forward B=8 D=128 64x64", flush=True)
    with torch.no_grad():
        y = model(x.to(DEV), z.to(DEV), 48e3, 48e3)
    torch.cuda.synchronize()
    assert bool(torch.isfinite(y).all())
    y0 = y[0].double().cpu().numpy()
    err = _log_err(_sub(y0, int(data["stride"])), data["y"])
    dsum = abs(float(y0.sum()) - float(data["y:sum"])) / float(data["y:abssum"])
    print(f"[cfg1] rel log10 n {err:.3e}; sum {dsum:.2e}", flush=True)
    assert err <= LOG_TOL
    assert dsum <= 1e-5


def test_cfg2_loss_gradient():
    """The inversion loss sum (log10 n - t)^2 with t the prediction at x + 0.02 N(0, 1), at the benchmark shape: its
    input gradient through the hand-written adjoint (torch autograd over FFNO3D), batch item 0 against the float64
    oracle's autograd gradient and its float64 central differences at 64 entries."""
    from ffnoml_amd import synthetic
    data = _fixture("wl_cfg2_gradient")
    x, z, mx, sx = _cfg1_batch()
    model = _model(data, mx, sx)
    xd, zd = x.to(DEV), z.to(DEV)
    noise = synthetic.perturbation(tuple(x.shape), seed=0).to(DEV)
    print("[cfg2] target and loss gradient, B=8", flush=True)
    with torch.no_grad():
        target = model(xd + noise, zd, 48e3, 48e3)
    xg = xd.clone().requires_grad_(True)
    loss = ((model(xg, zd, 48e3, 48e3) - target) ** 2).sum()
    (grad,) = torch.autograd.grad(loss, xg)
    torch.cuda.synchronize()
    g0 = grad[0].double().cpu().numpy()
    stride = int(data["stride"])
    err = _grad_err(_sub(g0, stride), data["grad"])
    idx = data["fd_index"]
    got = np.array([g0[tuple(i)] for i in idx])
    fd = data["fd_value"]
    fd_err = float(np.abs(got - fd).max() / np.abs(fd).max())
    print(f"[cfg2] loss {float(loss) / 8:.4e} (item mean; oracle item 0 {float(data['loss']):.4e}); gradient rel "
          f"{err:.3e}; vs central differences {fd_err:.3e}", flush=True)
    assert err <= GRAD_TOL
    assert fd_err <= GRAD_TOL


def test_tile504_forward_and_adjoint():
    """One 504 x 504 tile (B=1, D=4, width 64, 2 layers, 6 vertical modes): the factorised 504-point DFTs along both
    axes in the whole network, forward and adjoint (x and z gradients) against the float64 oracle. The height
    gradient is compared per level summed over columns, where the engine's per-column attribution and the
    reference's shared-grid one agree."""
    from ffnoml_amd import synthetic
    data = _fixture("wl_tile504")
    B, D, H, W = (int(v) for v in data["shape"])
    x, z, mx, sx = synthetic.atmosphere(B, D, H, W, seed=0)
    model = _model(data, mx, sx)
    ct = synthetic.cotangent((B, 12, D, H, W), 0).to(DEV)
    print(f"[tile504] forward + adjoint {B}x{D}x{H}x{W}", flush=True)
    xg = x.to(DEV).requires_grad_(True)
    zg = z.to(DEV).requires_grad_(True)
    y = model(xg, zg, 48e3, 48e3)
    xb, zb = torch.autograd.grad(y, (xg, zg), grad_outputs=ct)
    torch.cuda.synchronize()
    stride = int(data["stride"])
    y_err = _log_err(_sub(y.detach().double().cpu().numpy(), stride), data["y"])
    x_err = _grad_err(_sub(xb.double().cpu().numpy(), stride), data["xbar"])
    z_err = _grad_err(zb.double().cpu().numpy().sum(axis=(2, 3)), data["zbar_level"])
    print(f"[tile504] rel log10 n {y_err:.3e}; x_bar {x_err:.3e}; z_bar per level {z_err:.3e}", flush=True)
    assert y_err <= LOG_TOL
    assert x_err <= GRAD_TOL
    assert z_err <= GRAD_TOL


def test_volume_proxy_252():
    """The whole-volume network (width 64, 8 layers, 32 vertical modes) on a quarter-area proxy of the 504 x 504 x 256
    volume, B=1, D=256, 252 x 252 (factorised 252-point DFTs along both axes, 128-level blocks of the Filon
    matrices), against the streamed float64 oracle: full-size accuracy rests on this and on the 504-point tile."""
    from ffnoml_amd import synthetic
    data = _fixture("wl_vol252")
    B, D, H, W = (int(v) for v in data["shape"])
    x, z, mx, sx = synthetic.atmosphere(B, D, H, W, seed=0)
    assert np.array_equal(mx, data["mean_x"]) and np.array_equal(sx, data["std_x"])
    model = _model(data, mx, sx)
    print(f"[vol252] forward {B}x{D}x{H}x{W}", flush=True)
    with torch.no_grad():
        y = model(x.to(DEV), z.to(DEV), 48e3, 48e3)
    torch.cuda.synchronize()
    y0 = y.double().cpu().numpy()
    sd, sh, sw = (int(v) for v in data["strides"])
    err = _log_err(y0[..., ::sd, ::sh, ::sw], data["y"])
    dsum = abs(float(y0.sum()) - float(data["y:sum"])) / float(data["y:abssum"])
    print(f"[vol252] rel log10 n {err:.3e}; sum {dsum:.2e}", flush=True)
    assert err <= LOG_TOL
    assert dsum <= 1e-5
