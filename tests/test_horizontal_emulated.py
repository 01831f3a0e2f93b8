"""The engine's horizontal spectral pipeline (rfft2 -> per-mode factor -> irfft2 and its adjoint), launch sequence
and index maps as the engine issues them, run on the CPU through a float64 model of the GEMM entry point
(tests/tw_emulator.py): dense transforms, and the factorised DFTs along H, along W and along both (the dense-size
limit lowered so that small shapes take them), against numpy.fft: 1e-5 of range (float32 buffers between stages,
bf16-split matrices summed exactly); the adjoint as the transpose: <T x, y> = <x, T^T y> to 1e-6. The vertical
pipeline with more levels than one LDS-staged GEMM takes (its matrices in row and column blocks) against the dense
float64 matrices."""
import numpy as np
import pytest
import torch

from tw_emulator import Emulator


def _plan(H, W, dense_max, B=1, D=2, C=16, column=None):
    from ffnoml_amd import _lib, engine as E, synthetic
    from ffnoml_amd.atoms import species
    model = synthetic.random_model(0, width=C, n_layers=1, hidden_spectral=16, hidden_vertical=16, freq_mix_rank=4,
                                   z_min_wavelength=1.28, out_channels=12, species=species(["H", "CA"]))
    eng = E.Engine(model)
    cpu = torch.device("cpu")
    eng.device = cpu
    eng.weights = E.PackedWeights(model, cpu)
    eng.weights.adjoint()
    saved = E.DENSE_DFT_MAX
    E.DENSE_DFT_MAX = dense_max
    try:
        col = np.linspace(14.0, -1.0, D) if column is None else column
        plan = E.Plan(eng, B, D, H, W, col - col.min(), float(col.max() - col.min()), 48e3, 48e3)
    finally:
        E.DENSE_DFT_MAX = saved
    return E, _lib, eng, plan


def _tensors(plan, extra):
    out = list(extra)
    for group in (plan.tw, plan.adjoint_tables(), plan.hfac or {}, plan.wfac or {}):
        for v in group.values():
            stack = [v]
            while stack:
                e = stack.pop()
                if isinstance(e, torch.Tensor):
                    out.append(e)
                elif isinstance(e, (tuple, list)):
                    stack.extend(e)
    return out


@pytest.mark.parametrize("dense_max,H,W,fact", [
    (128, 10, 12, (False, False)),
    (16, 24, 20, (True, True)),
    (48, 20, 96, (False, True)),
    (48, 30, 20, (True, False)),
    (16, 21, 15, (True, True)),
    (128, 12, 130, (False, True)),
    (128, 66, 200, (True, True)),
])
def test_horizontal_pipeline_matches_fft(dense_max, H, W, fact):
    E, _lib, eng, plan = _plan(H, W, dense_max)
    assert ((plan.hfac is not None), (plan.wfac is not None)) == fact
    B, D, C, wf = plan.B, plan.D, plan.C, plan.wf
    rng = np.random.default_rng(H * W)
    factor = rng.standard_normal(H * wf) + 1j * rng.standard_normal(H * wf)
    x = torch.from_numpy(rng.standard_normal((B, D, H, W, C)).astype(np.float32))
    y = torch.from_numpy(rng.standard_normal((B, D, H, W, C)).astype(np.float32))
    tx, ty = torch.empty_like(x), torch.empty_like(y)
    scr = dict(s1=torch.empty(plan.h_chunk * plan.plane), s2=torch.empty(plan.h_chunk * plan.plane2))
    tables = {"spec": {"gate": torch.zeros(H * wf, 2), "v": None, "u": None}}
    em = Emulator(_tensors(plan, [x, y, tx, ty, scr["s1"], scr["s2"]]), factor)
    saved = _lib.call, E.Engine._stream
    _lib.call = em
    E.Engine._stream = staticmethod(lambda: None)
    wc = torch.zeros(8)
    eng.weights.blocks[0]["spec.wc"] = wc
    eng.weights.adjoint()["blocks"][0]["spec.wc"] = wc
    plan.tables = lambda i, adjoint=False: tables
    try:
        eng.horizontal(plan, x.view(-1, C), tx.view(-1, C), 0, scr, tables)
        eng.horizontal(plan, y.view(-1, C), ty.view(-1, C), 0, scr, tables, adjoint=True)
    finally:
        _lib.call, E.Engine._stream = saved
    assert "ffno_tw_gemm" in em.calls and "ffno_mix" in em.calls
    xs = np.moveaxis(x.double().numpy(), -1, 2)                    # [B, D, C, H, W]
    spec = np.fft.rfft2(xs) * factor.reshape(H, wf)
    ref = np.moveaxis(np.fft.irfft2(spec, s=(H, W)), 2, -1)
    err = np.abs(tx.double().numpy() - ref).max() / np.abs(ref).max()
    print(f"H {H} W {W} factorised {fact}: rel err {err:.2e}")
    assert err < 1e-5
    lhs = float((tx.double() * y.double()).sum())
    rhs = float((x.double() * ty.double()).sum())
    assert abs(lhs - rhs) <= 1e-6 * float(tx.double().norm() * y.double().norm())


@pytest.mark.parametrize("D,dense_max", [(200, 128), (40, 16)])
def test_vertical_pipeline_in_blocks(D, dense_max):
    """Filon analysis [2M x D] and synthesis [D x 2M] with D beyond the LDS-staged size (column blocks accumulated,
    row blocks), forward and adjoint, against the dense float64 matrices with the per-mode factor between them."""
    E, _lib, eng, plan = _plan(3, 4, 128, B=2, D=D)
    B, C, M = plan.B, plan.C, plan.modes
    hw = plan.H * plan.W
    rng = np.random.default_rng(D)
    factor = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    x = torch.from_numpy(rng.standard_normal((B, D, hw, C)).astype(np.float32))
    out, out_t = torch.empty_like(x), torch.empty_like(x)
    v1 = torch.empty(B * 2 * M * hw * C)
    tables = {"vertical": {"gate": torch.zeros(M, 2), "v": None, "u": None}}
    em = Emulator(_tensors(plan, [x, out, out_t, v1]), factor)
    saved = _lib.call, E.Engine._stream, E.DENSE_DFT_MAX
    _lib.call = em
    E.Engine._stream = staticmethod(lambda: None)
    E.DENSE_DFT_MAX = dense_max
    wc = torch.zeros(8)
    eng.weights.blocks[0]["vertical.wc"] = wc
    eng.weights.adjoint()["blocks"][0]["vertical.wc"] = wc
    plan.tables = lambda i, adjoint=False: tables
    try:
        eng.vertical(plan, x.view(-1, C), v1, out.view(-1, C), 0, tables)
        eng.vertical(plan, x.view(-1, C), v1, out_t.view(-1, C), 0, tables, adjoint=True)
    finally:
        _lib.call, E.Engine._stream, E.DENSE_DFT_MAX = saved
    vf, vi = plan._mats["v_fwd"], plan._mats["v_inv"]                    # [2M, D], [D, 2M]
    blk = np.block([[np.diag(factor.real), -np.diag(factor.imag)], [np.diag(factor.imag), np.diag(factor.real)]])
    T_ = vi @ blk @ vf                                                   # [D, D] per column and channel
    xd = x.double().numpy()
    for got, mat in ((out, T_), (out_t, T_.T)):
        ref = np.einsum("ed,bdnc->benc", mat, xd)
        err = np.abs(got.double().numpy() - ref).max() / np.abs(ref).max()
        print(f"D {D}: rel err {err:.2e}")
        assert err < 1e-5
