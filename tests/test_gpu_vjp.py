"""Adjoint (input VJP) parity on the GPU: the hand-written backward behind torch.autograd against the reference's own
float64 autograd gradients (golden vectors), against the float64 oracle's autograd, and against float64 central
finite differences of the oracle.

Tolerance (the accuracy target for gradients): relative error <= 1e-2, measured as ||g - g_ref|| / ||g_ref|| over the
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
    from ffnoml_amd.model import FFNO3D
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
    from ffnoml_amd import synthetic
    from ffnoml_amd.atoms import species
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


@pytest.mark.parametrize("name", ["case_a_shared64", "case_c_rank0", "case_e_tall"])
def test_height_gradient_matches_per_column_autograd(name):
    """z_bar from the HIP adjoint against the float64 oracle's autograd with the per-column vertical path (every
    column's transform built from its own heights: the mathematically exact per-column gradient; the oracle's
    per-column path is pinned to the reference's by golden case A'). Case C has modes inside the taper band, so its
    height gradient includes the taper term."""
    from oracle import ffno_oracle as O
    data, config, state = load(name)
    model = _model_from(config, state)
    x = torch.from_numpy(data["x"]).float().to(DEV).requires_grad_(True)
    z = torch.from_numpy(data["z"]).float().to(DEV).requires_grad_(True)
    ct = torch.from_numpy(data["cotangent"]).float().to(DEV)
    y = model(x, z, 48e3, 48e3)
    xb, zb = torch.autograd.grad(y, (x, z), grad_outputs=ct)
    torch.cuda.synchronize()
    _, oxb, ozb = O.vjp(O.to_dtype(state, torch.float64), oracle_config(config), torch.from_numpy(data["x"]).double(),
                        torch.from_numpy(data["z"]).double(), 48e3, 48e3, torch.from_numpy(data["cotangent"]).double(),
                        force_per_column=True)
    nrm, mx = _errs(zb.cpu().numpy(), ozb.numpy())
    print(f"[{name}] z_bar rel norm {nrm:.3e} rel max {mx:.3e}", flush=True)
    assert nrm <= TOL and mx <= TOL
    assert _errs(xb.cpu().numpy(), oxb.numpy())[0] <= TOL
    # the reference's own autograd (its shared-grid path sends the vertical transform's height sensitivity to one
    # column of the first sample): per level, summed over every sample's columns, the engine's z_bar is the reference's
    lv, ref_lv = zb.double().cpu().numpy().sum(axis=(0, 2, 3)), data["zbar"].sum(axis=(0, 2, 3))
    assert _errs(lv, ref_lv)[0] <= TOL


def test_backend_vjp_local_matches_si_autograd():
    """GPUBackend.vjp_local (SI features and heights in metres, linear-population cotangent) against float64
    autograd through the oracle's SI chain (ffno_fsdp_runtime.py:48-58, :76-78, :265-281) with per-column heights;
    predict_local against the oracle's populations (1e-3 relative in log10 n)."""
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
    si[0] = 10.0 ** si[0]                       # T [K]
    si[1:4] *= 1e3                              # m/s
    # distinct spacings: modes lie in the taper band, whose median spacing is not differentiable where spacings
    # tie, and float32 metres -> Mm can break a tie differently from the float64 reference
    zc = zc + 0.03 * np.sin(1.7 * np.arange(len(zc)))
    z_m = np.broadcast_to(zc[:, None, None] * 1e6, si.shape[1:]).astype(np.float32).copy()
    model = synthetic.random_model(4, **config)
    model.set_normalization(mean, std, np.zeros(6), np.ones(6))
    state = {k: v.detach().clone() for k, v in model.state_dict().items()}
    backend = GPUBackend(None, model=model, level_names=[f"H{i}" for i in range(6)])
    pops = backend.predict_local(si.astype(np.float32), z_m, 48e3, 48e3)
    sd = O.to_dtype(state, torch.float64)
    cfg = oracle_config(config)
    f64 = torch.from_numpy(si.astype(np.float32)).double()[None].requires_grad_(True)
    z64 = torch.from_numpy(z_m).double()[None].requires_grad_(True)
    n_ref = O.predict_si(sd, cfg, f64, z64, 48e3, 48e3)
    ref_log = np.log10(n_ref.detach().numpy()[0].transpose(1, 2, 3, 0))
    rel = float((np.abs(np.log10(pops) - ref_log) / np.maximum(np.abs(ref_log), 1.0)).max())
    print(f"[backend] predict rel log10 n {rel:.3e}", flush=True)
    assert rel <= 1e-3
    rng = np.random.default_rng(0)
    cot = (rng.standard_normal(pops.shape) / pops).astype(np.float32)      # O(1) in relative terms
    out = backend.vjp_local(si.astype(np.float32), z_m, 48e3, 48e3, cot)
    # float64 reference through the per-column vertical path
    f2 = torch.from_numpy(si.astype(np.float32)).double()[None].requires_grad_(True)
    z2 = torch.from_numpy(z_m).double()[None].requires_grad_(True)
    fm, zm = O.model_inputs_si(f2, z2)
    mx_, sx_ = sd["mean_X"][None, :, None, None, None], sd["std_X"][None, :, None, None, None]
    y = O.forward(sd, cfg, (fm - mx_) / sx_, zm, 48e3, 48e3, force_per_column=True)
    n = torch.pow(10.0, (y * sd["std_Y"][None, :, None, None, None] + sd["mean_Y"][None, :, None, None, None])
                  .clamp(-37, 37))
    gf, gz = torch.autograd.grad(n, (f2, z2), grad_outputs=torch.from_numpy(cot).double().permute(3, 0, 1, 2)[None])
    for key, ref in (("features", gf[0].numpy()), ("z_scale", gz[0].numpy())):
        nrm, mx = _errs(out[key], ref)
        print(f"[backend] vjp {key}: rel norm {nrm:.3e} rel max {mx:.3e}", flush=True)
        assert nrm <= TOL


def test_checkpoint_strides_agree():
    """The adjoint that keeps every block input (stride 1) and the ones that keep every second or third and
    recompute the others from the nearest kept input (strides 2, 3: what whole-volume adjoints use when memory is
    short) give the same x_bar and z_bar bit for bit: the recomputed inputs are the same kernels on the same data."""
    data, config, state = load("case_b_jitter64")
    config = dict(config, n_layers=4) if config.get("n_layers", 0) < 4 else config
    from ffnoml_amd import synthetic
    model = synthetic.random_model(3, **config).to(DEV).eval()
    x, z, mx, sx = synthetic.atmosphere(2, 16, 12, 10, seed=4)
    x, z = x.to(DEV), z.to(DEV)
    ct = synthetic.cotangent((2, config["out_channels"], 16, 12, 10), 5).to(DEV)
    eng = model.engine()
    eng.two_pass_fuse = False              # the same fusion adjoint at every stride
    results = {}
    for stride in (1, 2, 3):
        print(f"stride {stride}", flush=True)
        y, saved = eng.forward_saved(x, z, 48e3, 48e3, stride=stride)
        assert len(saved.inputs) == -(-config["n_layers"] // stride) and (saved.head is not None) == (stride == 1)
        xb, zb = eng.vjp(x, z, 48e3, 48e3, ct, saved, want_z=True)
        torch.cuda.synchronize()
        results[stride] = (y.cpu(), xb.cpu(), zb.cpu())
    for stride in (2, 3):
        for a, b, name in zip(results[stride], results[1], ("y", "x_bar", "z_bar")):
            err = float((a - b).abs().max() / b.abs().max())
            print(f"stride {stride} {name} rel {err:.2e}", flush=True)
            assert err == 0.0


def test_two_pass_fusion_adjoint_matches_fused():
    """The fusion adjoint in two LDS-staged passes over the gates the recompute stored, against the single kernel
    that recomputes them: the same arithmetic except the order in which the gated-branch term joins the gate
    network's adjoint (one ulp). x_bar agrees to 1e-6 relative; z_bar, whose per-column terms cancel (each column's
    shift subtracts its total) and whose own error against the float64 oracle is 3e-5 to 7e-5, to 1e-4."""
    data, config, state = load("case_b_jitter64")
    config = dict(config, n_layers=3, width=64, hidden_spectral=64, hidden_vertical=64)
    from ffnoml_amd import synthetic
    model = synthetic.random_model(4, **config).to(DEV).eval()
    x, z, mx, sx = synthetic.atmosphere(2, 16, 12, 10, seed=6)
    x, z = x.to(DEV), z.to(DEV)
    ct = synthetic.cotangent((2, config["out_channels"], 16, 12, 10), 7).to(DEV)
    eng = model.engine()
    out = {}
    for two_pass in (False, True):
        eng.two_pass_fuse = two_pass
        y, saved = eng.forward_saved(x, z, 48e3, 48e3, stride=1)
        assert saved.two_pass == two_pass
        out[two_pass] = eng.vjp(x, z, 48e3, 48e3, ct, saved, want_z=True)
        torch.cuda.synchronize()
    for a, b, name, tol in zip(out[True], out[False], ("x_bar", "z_bar"), (1e-6, 1e-4)):
        err = float((a - b).abs().max() / b.abs().max())
        print(f"two-pass vs fused {name} rel {err:.2e}", flush=True)
        assert err <= tol


def test_chunked_fusion_adjoint_matches_two_pass():
    """One sample at checkpoint stride 2, where a whole block's gate buffer is not kept: the two-pass fusion adjoint
    run chunk by chunk over the voxels (208-voxel chunks and a ragged last one, each chunk's gates recomputed by the
    fusion forward) against the two-pass adjoint over the stored gates (stride 1) -- the same per-voxel arithmetic,
    so x_bar agrees to 1e-6 relative -- and against the one-pass kernel at stride 2 (x_bar 1e-6, z_bar 1e-4, as the
    two-pass against the one-pass)."""
    data, config, state = load("case_b_jitter64")
    config = dict(config, n_layers=3, width=64, hidden_spectral=64, hidden_vertical=64)
    from ffnoml_amd import synthetic
    model = synthetic.random_model(5, **config).to(DEV).eval()
    x, z, mx, sx = synthetic.atmosphere(1, 16, 12, 10, seed=8)
    x, z = x.to(DEV), z.to(DEV)
    ct = synthetic.cotangent((1, config["out_channels"], 16, 12, 10), 4).to(DEV)
    eng = model.engine()
    out = {}
    for name, stride, two_pass, chunk in (("two-pass", 1, True, None), ("chunked", 2, False, 208),
                                          ("one-pass", 2, False, False)):
        print(f"[fuse] {name}", flush=True)
        eng.two_pass_fuse, eng.fuse_chunk = two_pass, chunk
        y, saved = eng.forward_saved(x, z, 48e3, 48e3, stride=stride)
        assert saved.two_pass == two_pass
        out[name] = eng.vjp(x, z, 48e3, 48e3, ct, saved, want_z=True)
        torch.cuda.synchronize()
    eng.two_pass_fuse, eng.fuse_chunk = None, None
    for other, tols in (("two-pass", (1e-6, 1e-4)), ("one-pass", (1e-6, 1e-4))):
        for a, b, name, tol in zip(out["chunked"], out[other], ("x_bar", "z_bar"), tols):
            err = float((a - b).abs().max() / b.abs().max())
            print(f"chunked vs {other} {name} rel {err:.2e}", flush=True)
            assert err <= tol


def test_adjoint_keeping_every_intermediate_matches_recompute():
    """The adjoint that reads every intermediate the forward kept (no recompute), and the ones that keep the last one
    or two blocks whole and recompute the others, against the one that recomputes each block from its saved input,
    all with the two-pass fusion adjoint: bit-identical x_bar and z_bar (the same kernels on the same data, only
    where the data waited differs)."""
    data, config, state = load("case_b_jitter64")
    config = dict(config, n_layers=3, width=32, hidden_spectral=32, hidden_vertical=32)
    from ffnoml_amd import synthetic
    model = synthetic.random_model(8, **config).to(DEV).eval()
    x, z, mx, sx = synthetic.atmosphere(2, 12, 10, 14, seed=9)
    x, z = x.to(DEV), z.to(DEV)
    ct = synthetic.cotangent((2, config["out_channels"], 12, 10, 14), 3).to(DEV)
    eng = model.engine()
    eng.two_pass_fuse = True
    out = {}
    for mode in ("full", "stride1", "keep1", "keep2"):
        stride, keep = (0, 0) if mode == "full" else (1, int(mode[4:]) if mode.startswith("keep") else 0)
        y, saved = eng.forward_saved(x, z, 48e3, 48e3, stride=stride, keep=keep)
        assert saved.full == (mode == "full") and len(saved.blocks) == (3 if mode == "full" else keep)
        out[mode] = (y,) + tuple(eng.vjp(x, z, 48e3, 48e3, ct, saved, want_z=True))
        assert not saved.blocks                          # each kept block released by its own adjoint
        torch.cuda.synchronize()
    for mode in ("full", "keep1", "keep2"):
        for a, b, name in zip(out[mode], out["stride1"], ("y", "x_bar", "z_bar")):
            err = float((a - b).abs().max() / b.abs().max())
            print(f"{mode} vs recompute {name} rel {err:.2e}", flush=True)
            assert err == 0.0


def test_adjoint_memory_within_admission_footprint():
    """The device memory one forward-and-adjoint allocates beyond what was allocated before it (plans and tables
    already built by a first call) stays within Engine.footprint_bytes, the serving admission's count, for the
    stride-1 adjoint, for 3 of 8 blocks kept whole and for every intermediate kept."""
    config = dict(in_channels=6, out_channels=12, width=64, n_layers=8, hidden_spectral=128, hidden_vertical=128,
                  freq_mix_rank=8, z_reference_length=40.0, z_min_wavelength=1.28, z_feature_wavelength=0.3)
    from ffnoml_amd import synthetic
    from ffnoml_amd.atoms import species
    model = synthetic.random_model(4, species=species(["H", "CA"]), **config).to(DEV).eval()
    B, D, H, W = 1, 48, 64, 64
    x, z, mx, sx = synthetic.atmosphere(B, D, H, W, seed=2)
    x, z = x.to(DEV), z.to(DEV)
    ct = synthetic.cotangent((B, 12, D, H, W), 1).to(DEV)
    eng = model.engine()
    span = float((z.amax(1) - z.amin(1)).max())
    for stride, keep in ((1, 0), (1, 3), (0, 0)):
        for rep in range(2):                             # the first call builds the plan's tables
            torch.cuda.synchronize()
            base = torch.cuda.memory_allocated()
            torch.cuda.reset_peak_memory_stats()
            y, saved = eng.forward_saved(x, z, 48e3, 48e3, stride=stride, keep=keep)
            xb, zb = eng.vjp(x, z, 48e3, 48e3, ct, saved, want_z=True)
            del saved
            torch.cuda.synchronize()
            peak = torch.cuda.max_memory_allocated() - base
            del y, xb, zb
        need = eng.footprint_bytes(B, D, H, W, vjp=True, span=span, full=stride == 0, keep=keep)
        print(f"stride {stride} keep {keep}: peak {peak / 2**20:.0f} MiB, footprint {need / 2**20:.0f} MiB",
              flush=True)
        assert peak <= need


def test_factorised_transforms_match_oracle():
    """The whole network with its horizontal DFTs factorised along H and W (the dense-size limit lowered to 32 so
    that a small tile takes those paths: H = 36 = 6 x 6, W = 40 = 5 x 8, D = 40): forward, x_bar
    and z_bar of the HIP engine against the float64 oracle on the same weights (1e-3 relative in log10 n, 1e-2 in
    the gradients; z_bar against the oracle's per-column path)."""
    from ffnoml_amd import engine as E, synthetic
    from ffnoml_amd.atoms import species
    from ffnoml_amd.model import FFNO3D
    from oracle import ffno_oracle as O
    config = dict(in_channels=6, out_channels=12, width=32, n_layers=2, hidden_spectral=32, hidden_vertical=32,
                  freq_mix_rank=4, z_reference_length=40.0, z_min_wavelength=1.28, z_feature_wavelength=0.3,
                  species=species(["H", "CA"]))
    B, D, H, W = 1, 40, 36, 40
    x, z, mx, sx = synthetic.atmosphere(B, D, H, W, seed=2)
    state = regenerate_state(config, 0, mx, sx, np.zeros(12), np.ones(12))
    model = FFNO3D(**config)
    model.load_state_dict(state, strict=True)
    model = model.eval().to(DEV)
    ct = synthetic.cotangent((B, 12, D, H, W), 6)
    saved = E.DENSE_DFT_MAX
    E.DENSE_DFT_MAX = 32
    try:
        print("[factorised] forward + vjp", flush=True)
        xg = x.to(DEV).requires_grad_(True)
        zg = z.to(DEV).requires_grad_(True)
        y = model(xg, zg, 48e3, 48e3)
        plan = model.engine().plan(B, D, H, W, zg.detach(), 48e3, 48e3)
        assert plan.hfac is not None and plan.wfac is not None
        xb, zb = torch.autograd.grad(y, (xg, zg), grad_outputs=ct.to(DEV))
        torch.cuda.synchronize()
    finally:
        E.DENSE_DFT_MAX = saved
    print("[factorised] float64 oracle", flush=True)
    # the per-column vertical path gives the exact per-column height gradient the engine computes (the shared path's
    # autograd attributes the transform's height dependence to one column, as the reference's does)
    ry, rxb, rzb = O.vjp(O.to_dtype(state, torch.float64), oracle_config(config), x.double(), z.double(), 48e3, 48e3,
                         ct.double(), force_per_column=True)
    rel = float(((y.detach().cpu().double() - ry).abs() / ry.abs().clamp_min(1.0)).max())
    gx = float((xb.cpu().double() - rxb).norm() / rxb.norm())
    gz = float((zb.cpu().double() - rzb).norm() / rzb.norm())
    print(f"[factorised] forward rel {rel:.3e}, x_bar {gx:.3e}, z_bar {gz:.3e}", flush=True)
    assert rel <= 1e-3 and gx <= TOL and gz <= TOL
