"""Write the golden vectors under tests/golden/ from the reference implementation itself.

Run in a container that holds the reference tree (read-only) at /root/reference:
    python tests/golden/make_golden.py
It imports the reference's models/ffno_model.py (torch and math only) and config.py (for the atom data), builds a
reference FFNO3D, loads into it with strict=True the state_dict of this repository's ffnoml_amd.model.FFNO3D built
with the same arguments and seed, and records inputs, weights, outputs and input gradients (autograd) as .npz.
Nothing here is used at run time; the tests only read the .npz files.
"""
import hashlib
import json
import os
import sys

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REFERENCE = "/root/reference"
sys.path.insert(0, REPO)

from ffnoml_amd import synthetic  # noqa: E402
from ffnoml_amd.atoms import species as species_of  # noqa: E402


def reference_modules():
    sys.path.insert(0, REFERENCE)
    import config as ref_config  # noqa: F401
    from models import ffno_model as ref_model
    return ref_model, ref_config


def stretched(depth):
    s = np.linspace(0.0, 1.0, depth)
    return (-1.0 + 15.0 * (0.4 * s + 0.6 * s ** 2))[::-1].copy()


def state_digest(sd):
    h = hashlib.sha256()
    for k in sorted(sd):
        h.update(k.encode())
        h.update(np.ascontiguousarray(sd[k].detach().double().cpu().numpy()).tobytes())
    return h.hexdigest()


SOLAR = (np.array([3.8, 0.0, 0.0, 0.0, 16.0, -7.0], np.float32), np.array([0.4, 5.0, 5.0, 5.0, 2.0, 2.5], np.float32))


def run_case(ref_model, name, config, x, z, dtype, seed, hooks=(), store_state=True, subsample=None, grads=True,
             norm=None):
    torch.set_default_dtype(torch.float32)        # the same weights whatever dtype the case runs in
    ours = synthetic.random_model(seed=seed, **config)
    out_ch = config["out_channels"]
    mx, sx, my, sy = norm if norm is not None else (*SOLAR, np.zeros(out_ch), np.ones(out_ch))
    ours.set_normalization(mx, sx, my, sy)
    sd = ours.state_dict()
    torch.set_default_dtype(dtype)                # the reference's frequency grids follow the default dtype
    ref = ref_model.FFNO3D(**config)
    ref.load_state_dict(sd, strict=True)
    ref = ref.to(dtype).eval()
    captured = {}
    handles = []
    for path in hooks:
        module = ref.get_submodule(path)
        handles.append(module.register_forward_hook(
            lambda m, a, out, path=path: captured.__setitem__(path, out.detach().float().numpy())))
    xin = x.to(dtype).clone().requires_grad_(grads)
    zin = z.to(dtype).clone().requires_grad_(grads)
    y = ref(xin, zin, 48e3, 48e3)
    out = dict(x=x.numpy(), z=z.numpy(), y=y.detach().numpy())
    if grads:
        ct = synthetic.cotangent(tuple(y.shape), seed).to(dtype)
        xb, zb = torch.autograd.grad(y, (xin, zin), grad_outputs=ct)
        out.update(cotangent=ct.numpy(), xbar=xb.numpy(), zbar=zb.numpy())
    for h in handles:
        h.remove()
    for path, value in captured.items():
        out["sub:" + path] = value
    if store_state:
        for k, v in sd.items():
            out["w:" + k] = v.detach().numpy()
    out["state_sha256"] = np.array(state_digest(sd))
    out["config_json"] = np.array(json.dumps({k: v for k, v in config.items()}))
    out["seed"] = np.array(seed)
    out["dtype"] = np.array(str(dtype))
    if subsample is not None:
        for key in list(out):
            if key in ("y", "xbar", "zbar", "cotangent"):
                full = out[key]
                out[key + ":sum"] = np.array(float(np.sum(full, dtype=np.float64)))
                out[key + ":abssum"] = np.array(float(np.sum(np.abs(full), dtype=np.float64)))
                out[key] = full[subsample(full.ndim)]
        out.pop("x")
        out.pop("z")
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    print(name, {k: getattr(v, "shape", None) for k, v in out.items() if not k.startswith("w:")}, flush=True)
    torch.set_default_dtype(torch.float32)


def main():
    ref_model, ref_config = reference_modules()
    # the atom data the reference's config gives, to pin ffnoml_amd.atoms
    atoms = {a: {k: ref_config.ATOM_CONFIG[a][k] for k in ("levels", "log10_particle_mass", "multi3d_lte")}
             for a in ("H", "CA")}
    with open(os.path.join(HERE, "atoms.json"), "w") as f:
        json.dump(atoms, f, indent=1, sort_keys=True)

    small = dict(in_channels=6, width=16, n_layers=2, hidden_spectral=16, hidden_vertical=16, freq_mix_rank=4,
                 z_reference_length=40.0, z_min_wavelength=0.3, z_feature_wavelength=0.3)
    hca = dict(small, out_channels=12, species=species_of(["H", "CA"]))
    g = torch.Generator().manual_seed(5)

    # A: shared stretched grid, H + Ca II, odd W
    D, H, W = 24, 11, 9
    x = torch.randn(1, 6, D, H, W, generator=g)
    z = torch.from_numpy(stretched(D))[None, :, None, None].expand(1, D, H, W).float().contiguous()
    hooks = ("lift", "blocks.0.spec", "blocks.0.vertical", "blocks.0", "proj2")
    run_case(ref_model, "case_a_shared64", hca, x, z, torch.float64, 1, hooks)
    run_case(ref_model, "case_a_shared32", hca, x, z, torch.float32, 1, (), grads=False)
    # A': the same with one height moved by 1e-6 Mm: the per-column path of the vertical branch
    zf = z.clone()
    zf[0, 12, 5, 4] += 1e-6
    run_case(ref_model, "case_a_percol64", hca, x, zf, torch.float64, 1, ())
    # B: per-column heights, 0.01 Mm jitter
    zb = z + 0.01 * torch.randn(z.shape, generator=g)
    run_case(ref_model, "case_b_jitter64", hca, x, zb, torch.float64, 1, ())
    # C: rank 0, H only, batch 2, even W (Nyquist column), uniform grid
    D, H, W = 16, 8, 12
    x = torch.randn(2, 6, D, H, W, generator=g)
    z = torch.linspace(14.0, -1.0, D)[None, :, None, None].expand(2, D, H, W).contiguous()
    run_case(ref_model, "case_c_rank0", dict(small, freq_mix_rank=0, out_channels=6, species=species_of(["H"])),
             x, z, torch.float64, 2, (), norm=(*SOLAR, np.linspace(10.0, 15.0, 6), np.linspace(1.0, 2.0, 6)))
    # E: columns taller than half the reference length: a longer periodic box
    D, H, W = 20, 6, 8
    x = torch.randn(1, 6, D, H, W, generator=g)
    z = torch.linspace(31.0, 1.0, D)[None, :, None, None].expand(1, D, H, W).contiguous()
    run_case(ref_model, "case_e_tall", dict(small, out_channels=6, species=species_of(["H"])), x, z, torch.float64, 3,
             ())
    # F: the CPU workload (B=1, width 32, 4 layers, D=64, 32x32, 12 vertical modes, H + Ca II), synthetic atmosphere
    # with seed 0 and weights with seed 0; stored as a subsample plus full sums (inputs and weights regenerate)
    cfg0 = dict(in_channels=6, out_channels=12, width=32, n_layers=4, hidden_spectral=128, hidden_vertical=128,
                freq_mix_rank=8, z_reference_length=40.0, z_min_wavelength=3.5, z_feature_wavelength=0.3,
                species=species_of(["H", "CA"]))
    xs, zs, mean_x, std_x = synthetic.atmosphere(1, 64, 32, 32, seed=0)
    run_case(ref_model, "case_f_cfg0", cfg0, xs, zs, torch.float64, 0, (), store_state=False,
             subsample=lambda nd: (slice(None),) * (nd - 3) + (slice(None, None, 4),) * 3,
             norm=(mean_x, std_x, np.zeros(12), np.ones(12)))


if __name__ == "__main__":
    main()
