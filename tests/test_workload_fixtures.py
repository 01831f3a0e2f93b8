"""The benchmark-workload fixtures of tests/golden/wl_*.npz (float64 oracle outputs the GPU tests compare against,
written by tests/golden/make_workload_fixtures.py): their weights and inputs regenerate from the seeds, and at the
benchmark shape the oracle's autograd gradient agrees with its own float64 central differences."""
import hashlib
import json
import os

import numpy as np
import pytest

from golden_cases import GOLDEN

NAMES = ["wl_cfg1_forward", "wl_cfg2_gradient", "wl_tile504", "wl_service", "wl_vol252"]


def _load(name):
    path = os.path.join(GOLDEN, name + ".npz")
    if not os.path.exists(path):
        pytest.skip(f"{name}.npz not written yet (python tests/golden/make_workload_fixtures.py)")
    return np.load(path)


@pytest.mark.parametrize("name", NAMES)
def test_weights_regenerate(name):
    from ffnoml_amd import synthetic
    from ffnoml_amd.atoms import species
    data = _load(name)
    config = json.loads(str(data["config_json"]))
    config["species"] = species(config["species"])
    if name in ("wl_tile504", "wl_vol252"):
        B, D, H, W = (int(v) for v in data["shape"])
        _, _, mx, sx = synthetic.atmosphere(B, D, H, W, seed=0)
    else:
        _, _, mx, sx = synthetic.atmosphere(8, 128, 64, 64, seed=0)
    model = synthetic.random_model(0, **config)
    model.set_normalization(mx, sx, np.zeros(12), np.ones(12))
    h = hashlib.sha256()
    for k, v in sorted(model.state_dict().items()):
        h.update(k.encode())
        h.update(np.ascontiguousarray(v.double().numpy()).tobytes())
    assert h.hexdigest() == str(data["state_sha256"])
    if "mean_x" in data.files:
        assert np.array_equal(mx, data["mean_x"]) and np.array_equal(sx, data["std_x"])


def test_cfg2_autograd_matches_central_differences():
    """At the benchmark shape (B=1 of the batch, D=128, 64 x 64, width 64, 8 layers) the float64 oracle's autograd
    gradient of the inversion loss equals its float64 central differences (h = 1e-4) at 64 random entries to 1e-5 of
    their range (the differences' O(h^2) truncation: 1.05e-6 of the range at the largest entry, 7.6e-8 of each entry
    at the median): the oracle's gradient, against which the GPU adjoint is checked, is the derivative."""
    data = _load("wl_cfg2_gradient")
    fd, ad = data["fd_value"], data["fd_autograd"]
    assert np.isfinite(fd).all() and len(fd) == 64
    assert float(np.abs(fd - ad).max() / np.abs(fd).max()) < 1e-5
    # the stored entries are the gradient's, at the indices stored (the subsample holds every 4th entry)
    stride = int(data["stride"])
    for i, (c, d, h, w) in enumerate(data["fd_index"]):
        if d % stride == 0 and h % stride == 0 and w % stride == 0:
            assert abs(float(data["grad"][c, d // stride, h // stride, w // stride]) - ad[i]) <= 1e-6 * abs(ad[i])
