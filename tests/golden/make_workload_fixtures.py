"""Write the float64-oracle fixtures of the benchmark workloads under tests/golden/ (wl_*.npz).

The GPU tests compare the HIP path with these at the workloads' own sizes, where the float64 oracle takes minutes to
hours on a CPU and cannot run inside a test. Inputs and weights are not stored: they regenerate from their seeds
(ffnoml_amd.synthetic), and every fixture carries the sha256 of the weights it was made with, which the tests check
first. Outputs are stored as a strided subsample (float32) plus sums over the whole tensor in float64.

    python tests/golden/make_workload_fixtures.py [--only NAME ...] [--threads N] [--scratch DIR]

Cases:
  wl_cfg1_forward   B=8, D=128, 64 x 64, width 64, 8 layers, 32 vertical modes, H + Ca II: the forward of batch
                    item 0 (the batch is z-scored as a whole; GroupNorm statistics are per item, so item 0 of the
                    batch equals item 0 alone).
  wl_cfg2_gradient  the same model and item: the gradient of the inversion loss sum (log10 n - t)^2, t the
                    prediction at x + 0.02 N(0, 1) (normalised units), by autograd over the oracle, and by float64
                    central differences (h = 1e-4) at 64 random input entries. The central differences take two
                    forwards each; they are kept per entry in --scratch, so an interrupted run resumes.
  wl_tile504        B=1, D=4, 504 x 504, width 64, 2 layers, 6 vertical modes: forward, x-gradient and per-level sums of the
                    z-gradient for a fixed cotangent (the factorised 504-point DFTs along both axes).
  wl_service        the cfg1 model in SI units (the service path) on the smallest serving tile, 96 x 32 x 32:
                    linear populations and the feature and height gradients of a fixed population cotangent.
  wl_vol252         the whole-volume network (width 64, 8 layers, 32 vertical modes) on a quarter-area proxy of the
                    504 x 504 x 256 volume, B=1, D=256, 252 x 252: the forward, by the streamed float64 oracle
                    (oracle/ffno_streamed.py; the whole-volume float64 intermediates would not fit host RAM).

Only the oracle (oracle/ffno_oracle.py) runs here; nothing is imported from the reference.
"""
import argparse
import hashlib
import json
import os
import sys
import tempfile
import time

import numpy as np
import torch
from torch.utils.checkpoint import checkpoint

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

from ffnoml_amd import synthetic  # noqa: E402
from ffnoml_amd.atoms import species as species_of  # noqa: E402
from oracle import ffno_oracle as O  # noqa: E402

DX = DY = 48e3
CFG1 = dict(in_channels=6, out_channels=12, width=64, n_layers=8, hidden_spectral=128, hidden_vertical=128,
            freq_mix_rank=8, z_reference_length=40.0, z_min_wavelength=1.28, z_feature_wavelength=0.3,
            species=["H", "CA"])
SHAPE1 = (8, 128, 64, 64)
TILE504 = dict(CFG1, n_layers=2, z_min_wavelength=8.0)      # 6 vertical modes: D=4 levels resolve 4
SHAPE504 = (1, 4, 504, 504)
SERVICE_TILE = (96, 32, 32)
SHAPE252 = (1, 256, 252, 252)
FD_ENTRIES, FD_STEP = 64, 1e-4
STRIDE = 4


def model_config(c):
    return dict(c, species=species_of(c["species"]))


def state_digest(sd):
    h = hashlib.sha256()
    for k in sorted(sd):
        h.update(k.encode())
        h.update(np.ascontiguousarray(sd[k].detach().double().cpu().numpy()).tobytes())
    return h.hexdigest()


def oracle_cfg(c):
    keys = ("width", "n_layers", "hidden_spectral", "hidden_vertical", "freq_mix_rank", "z_reference_length",
            "z_min_wavelength", "z_feature_wavelength", "species", "in_channels", "out_channels")
    return O.model_config(**{k: c[k] for k in keys})


def weights(config, mx, sx):
    m = synthetic.random_model(0, **config)
    m.set_normalization(mx, sx, np.zeros(12), np.ones(12))
    return {k: v.detach().clone() for k, v in m.state_dict().items()}


def forward_ckpt(sd, cfg, x, z):
    """O.forward with every block under non-reentrant checkpointing (the same arithmetic; the autograd graph of one
    block at a time instead of all of them)."""
    w = O.depth_weights(z)
    h = O.gelu(O.weighted_group_norm(O.conv1x1(x, sd["lift.0.weight"]), w, sd["lift.1.weight"], sd["lift.1.bias"],
                                     cfg["groups"]))
    for i in range(cfg["n_layers"]):
        h = checkpoint(lambda t, i=i: O.block(sd, i, t, z, DX, DY, cfg), h, use_reentrant=False)
    head = O.conv1x1(O.gelu(O.conv1x1(h, sd["proj1.weight"])), sd["proj2.weight"], sd["proj2.bias"])
    return O.head_output(sd, head, x, cfg)


def sub(a):
    """Every STRIDE-th entry along the last three axes, as float32."""
    return np.ascontiguousarray(np.asarray(a)[..., ::STRIDE, ::STRIDE, ::STRIDE]).astype(np.float32)


def sums(prefix, a):
    a = np.asarray(a, np.float64)
    return {prefix + ":sum": np.array(a.sum()), prefix + ":abssum": np.array(np.abs(a).sum()),
            prefix + ":sqsum": np.array((a * a).sum())}


def save(name, **arrays):
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **arrays)
    print(f"{name}: {os.path.getsize(path) / 1e3:.0f} kB", flush=True)


def cfg1_inputs():
    x, z, mx, sx = synthetic.atmosphere(*SHAPE1, seed=0)
    return x[:1].double(), z[:1].double(), mx, sx


def make_cfg1():
    x0, z0, mx, sx = cfg1_inputs()
    state = weights(model_config(CFG1), mx, sx)
    sd, cfg = O.to_dtype(state, torch.float64), oracle_cfg(model_config(CFG1))
    t0 = time.time()
    with torch.no_grad():
        y = O.forward(sd, cfg, x0, z0, DX, DY).numpy()
    print(f"cfg1 forward {time.time() - t0:.0f} s", flush=True)
    save("wl_cfg1_forward", y=sub(y), **sums("y", y), state_sha256=np.array(state_digest(state)),
         config_json=np.array(json.dumps(CFG1)), shape=np.array(SHAPE1), item=np.array(0), stride=np.array(STRIDE),
         mean_x=mx, std_x=sx)
    return y


def cfg2_loss_parts(sd, cfg, x0, z0):
    noise = synthetic.perturbation((1, 6) + SHAPE1[1:], seed=0)
    with torch.no_grad():
        target = O.forward(sd, cfg, x0 + noise.double(), z0, DX, DY)
    return target


def make_cfg2(scratch):
    x0, z0, mx, sx = cfg1_inputs()
    state = weights(model_config(CFG1), mx, sx)
    sd, cfg = O.to_dtype(state, torch.float64), oracle_cfg(model_config(CFG1))
    t0 = time.time()
    target = cfg2_loss_parts(sd, cfg, x0, z0)
    xg = x0.clone().requires_grad_(True)
    y = forward_ckpt(sd, cfg, xg, z0)
    loss = ((y - target) ** 2).sum()
    (grad,) = torch.autograd.grad(loss, xg)
    grad = grad.numpy()
    print(f"cfg2 loss {float(loss):.6e}, gradient {time.time() - t0:.0f} s", flush=True)
    del y, xg
    rng = np.random.default_rng(2)
    flat = rng.choice(x0.numel(), FD_ENTRIES, replace=False)
    idx = np.stack(np.unravel_index(flat, x0.shape[1:]), 1)          # (c, d, h, w) of batch item 0
    part = os.path.join(scratch, "wl_cfg2_fd_partial.npz")
    done = dict(np.load(part)) if os.path.exists(part) else {}
    fd = np.full(FD_ENTRIES, np.nan)
    for j in range(FD_ENTRIES):
        key = str(int(flat[j]))
        if key in done:
            fd[j] = float(done[key])
            continue
        t1 = time.time()
        vals = []
        for sgn in (1.0, -1.0):
            xp = x0.clone()
            xp[(0,) + tuple(int(v) for v in idx[j])] += sgn * FD_STEP
            with torch.no_grad():
                vals.append(float(((O.forward(sd, cfg, xp, z0, DX, DY) - target) ** 2).sum()))
        fd[j] = (vals[0] - vals[1]) / (2 * FD_STEP)
        done[key] = np.array(fd[j])
        np.savez(part, **done)
        ad = grad[(0,) + tuple(idx[j])]
        print(f"fd {j + 1}/{FD_ENTRIES} {tuple(int(v) for v in idx[j])}: fd {fd[j]:.8e} autograd {ad:.8e} "
              f"({time.time() - t1:.0f} s)", flush=True)
    ad = np.array([grad[(0,) + tuple(i)] for i in idx])
    save("wl_cfg2_gradient", grad=sub(grad[0]), **sums("grad", grad), fd_index=idx, fd_value=fd, fd_autograd=ad,
         fd_step=np.array(FD_STEP), loss=np.array(float(loss)), state_sha256=np.array(state_digest(state)),
         config_json=np.array(json.dumps(CFG1)), shape=np.array(SHAPE1), item=np.array(0),
         stride=np.array(STRIDE))


def make_tile504():
    B, D, H, W = SHAPE504
    x, z, mx, sx = synthetic.atmosphere(B, D, H, W, seed=0)
    state = weights(model_config(TILE504), mx, sx)
    sd, cfg = O.to_dtype(state, torch.float64), oracle_cfg(model_config(TILE504))
    ct = synthetic.cotangent((B, 12, D, H, W), 0).double()
    t0 = time.time()
    xg = x.double().requires_grad_(True)
    zg = z.double().requires_grad_(True)
    y = forward_ckpt(sd, cfg, xg, zg)
    xb, zb = torch.autograd.grad(y, (xg, zg), grad_outputs=ct)
    y, xb, zb = y.detach().numpy(), xb.numpy(), zb.numpy()
    print(f"tile504 {time.time() - t0:.0f} s", flush=True)
    save("wl_tile504", y=sub(y), **sums("y", y), xbar=sub(xb), **sums("xbar", xb),
         zbar_level=zb.sum(axis=(2, 3)), **sums("zbar", zb), state_sha256=np.array(state_digest(state)),
         config_json=np.array(json.dumps(TILE504)), shape=np.array(SHAPE504), stride=np.array(STRIDE))


def make_service():
    _, _, mx, sx = cfg1_inputs()
    state = weights(model_config(CFG1), mx, sx)
    sd, cfg = O.to_dtype(state, torch.float64), oracle_cfg(model_config(CFG1))
    f, z, cot = synthetic.si_request(*SERVICE_TILE)
    t0 = time.time()
    fg = torch.from_numpy(f).double()[None].requires_grad_(True)
    zg = torch.from_numpy(z).double()[None].requires_grad_(True)
    n = O.predict_si(sd, cfg, fg, zg, DX, DY)                           # [1, 12, nz, nx, ny]
    c = torch.from_numpy(cot).double().permute(3, 0, 1, 2)[None]
    fb, zb = torch.autograd.grad(n, (fg, zg), grad_outputs=c)
    log_n = torch.log10(n.detach())[0].numpy()
    print(f"service tile {time.time() - t0:.0f} s", flush=True)
    save("wl_service", log10_n=sub(log_n), **sums("log10_n", log_n), feature_bar=sub(fb[0].numpy()),
         **sums("feature_bar", fb[0].numpy()), z_bar_level=zb[0].numpy().sum(axis=(1, 2)),
         **sums("z_bar", zb[0].numpy()), state_sha256=np.array(state_digest(state)),
         config_json=np.array(json.dumps(CFG1)), tile=np.array(SERVICE_TILE), stride=np.array(STRIDE))


def make_vol252():
    from oracle import ffno_streamed as S
    B, D, H, W = SHAPE252
    x, z, mx, sx = synthetic.atmosphere(B, D, H, W, seed=0)
    state = weights(model_config(CFG1), mx, sx)
    sd, cfg = O.to_dtype(state, torch.float64), oracle_cfg(model_config(CFG1))
    t0 = time.time()
    y = S.forward(sd, cfg, x.double(), z.double(), DX, DY, plane_chunk=8, row_chunk=16,
                  log=lambda m: print(f"vol252 {m} ({time.time() - t0:.0f} s)", flush=True)).numpy()
    # every 4th level and every 8th column of each axis (2.7 MB): the sums cover the rest
    save("wl_vol252", y=np.ascontiguousarray(y[..., ::STRIDE, ::2 * STRIDE, ::2 * STRIDE]).astype(np.float32),
         **sums("y", y), state_sha256=np.array(state_digest(state)), config_json=np.array(json.dumps(CFG1)),
         shape=np.array(SHAPE252), strides=np.array([STRIDE, 2 * STRIDE, 2 * STRIDE]), mean_x=mx, std_x=sx)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", nargs="*", default=["service", "cfg1", "tile504", "cfg2"])
    ap.add_argument("--threads", type=int, default=len(os.sched_getaffinity(0)))
    ap.add_argument("--scratch", default=tempfile.gettempdir())
    args = ap.parse_args()
    torch.set_num_threads(args.threads)
    for name in args.only:
        {"cfg1": make_cfg1, "cfg2": lambda: make_cfg2(args.scratch), "tile504": make_tile504,
         "service": make_service, "vol252": make_vol252}[name]()


if __name__ == "__main__":
    main()
