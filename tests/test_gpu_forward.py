"""End-to-end forward parity on the GPU: ffnoml_amd.FFNO3D (HIP engine) against the reference's own float64 outputs
(golden vectors) and against the float64 oracle.

Tolerance (the accuracy target): max over every output of |log10 n - log10 n_ref| / max(|log10 n_ref|, 1) <= 1e-3,
on de-normalised log10 populations."""
import json
import os

import numpy as np
import pytest
import torch

from golden_cases import GOLDEN, load, oracle_config, regenerate_state

pytestmark = pytest.mark.gpu
DEV = "cuda"
TOL = 1e-3


def log10_err(y, ref, mean_y, std_y):
    """max |d log10 n| / max(|log10 n|, 1) on de-normalised outputs, and the max |d log10 n| in dex."""
    y = np.asarray(y, np.float64)
    ref = np.asarray(ref, np.float64)
    shape = (1, -1) + (1,) * (y.ndim - 2)
    my, sy = np.asarray(mean_y, np.float64).reshape(shape), np.asarray(std_y, np.float64).reshape(shape)
    a, b = y * sy + my, ref * sy + my
    d = np.abs(a - b)
    return float((d / np.maximum(np.abs(b), 1.0)).max()), float(d.max())


def _model_from(config, state):
    from ffnoml_amd.model import FFNO3D
    m = FFNO3D(**config)
    m.load_state_dict(state, strict=True)
    return m.eval().to(DEV)


@pytest.mark.parametrize("name", ["case_a_shared64", "case_c_rank0", "case_e_tall"])
def test_forward_matches_reference_golden(name):
    data, config, state = load(name)
    model = _model_from(config, state)
    x = torch.from_numpy(data["x"]).float().to(DEV)
    z = torch.from_numpy(data["z"]).float().to(DEV)
    print(f"[{name}] x {tuple(x.shape)}", flush=True)
    with torch.no_grad():
        y = model(x, z, 48e3, 48e3)
    torch.cuda.synchronize()
    rel, dex = log10_err(y.cpu().numpy(), data["y"], state["mean_Y"].numpy(), state["std_Y"].numpy())
    print(f"[{name}] rel {rel:.3e} max dex {dex:.3e}", flush=True)
    assert rel <= TOL


def test_forward_matches_oracle_width64():
    """Width 64, two layers, H + Ca II, B=2, D=32, 16 x 16, 32 vertical modes: the HIP engine against the float64
    oracle on the same weights and synthetic atmosphere."""
    from ffnoml_amd import synthetic
    from ffnoml_amd.atoms import species
    from oracle import ffno_oracle as O
    config = dict(in_channels=6, out_channels=12, width=64, n_layers=2, hidden_spectral=128, hidden_vertical=128,
                  freq_mix_rank=8, z_reference_length=40.0, z_min_wavelength=1.28, z_feature_wavelength=0.3,
                  species=species(["H", "CA"]))
    x, z, mx, sx = synthetic.atmosphere(2, 32, 16, 16, seed=0)
    state = regenerate_state(config, 0, mx, sx, np.zeros(12), np.ones(12))
    model = _model_from(config, state)
    print("[width64] forward", flush=True)
    with torch.no_grad():
        y = model(x.to(DEV), z.to(DEV), 48e3, 48e3)
    ref = O.forward(O.to_dtype(state, torch.float64), oracle_config(config), x.double(), z.double(), 48e3, 48e3)
    rel, dex = log10_err(y.cpu().numpy(), ref.numpy(), np.zeros(12), np.ones(12))
    print(f"[width64] rel {rel:.3e} max dex {dex:.3e}", flush=True)
    assert rel <= TOL


def test_forward_cpu_workload_matches_golden():
    """The CPU workload (width 32, 4 layers, 64 x 32 x 32, 12 vertical modes) against the reference's float64
    output subsample."""
    from ffnoml_amd import synthetic
    data = np.load(os.path.join(GOLDEN, "case_f_cfg0.npz"))
    config = json.loads(str(data["config_json"]))
    x, z, mx, sx = synthetic.atmosphere(1, 64, 32, 32, seed=0)
    state = regenerate_state(config, 0, mx, sx, np.zeros(12), np.ones(12))
    model = _model_from(config, state)
    with torch.no_grad():
        y = model(x.to(DEV), z.to(DEV), 48e3, 48e3)
    sub = y.cpu().numpy()[:, :, ::4, ::4, ::4]
    rel, dex = log10_err(sub, data["y"], np.zeros(12), np.ones(12))
    print(f"[cfg0] rel {rel:.3e} max dex {dex:.3e}", flush=True)
    assert rel <= TOL
