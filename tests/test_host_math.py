"""Host-side pieces of the HIP engine, checked on the CPU: the transform matrices against numpy.fft and the oracle,
the weight packing against the kernels' lane-layout contraction, the model's state_dict layout, and the C ABI's
exported symbols."""
import ctypes
import math
import os
import re

import numpy as np
import pytest
from ffnoml_amd import engine as E
import torch

from ffnoml_amd import transforms as T
from ffnoml_amd.engine import pack_mlp_weight


@pytest.mark.parametrize("H,W", [(8, 12), (11, 9), (6, 5), (64, 64)])
def test_horizontal_matrices_are_rfft2_and_irfft2(H, W):
    """rfft2 then irfft2(s=(H, W)) as the four real matrices, on data and on a NON-Hermitian spectrum (as the mix
    leaves it): 1e-12 of numpy.fft's result."""
    rng = np.random.default_rng(0)
    m = T.horizontal_matrices(H, W)
    wf = W // 2 + 1
    x = rng.standard_normal((H, W))
    s1 = m["w_fwd"] @ x.T                                   # [2Wf, H]: rows (part, kx), columns h
    re, im = s1[:wf].T, s1[wf:].T                           # [H, Wf]
    s2 = m["h_fwd"] @ np.concatenate([re, im], 0)           # rows (part, ky)
    ref = np.fft.rfft2(x)
    got = s2[:H] + 1j * s2[H:]
    assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()
    z = rng.standard_normal((H, wf)) + 1j * rng.standard_normal((H, wf))   # not Hermitian
    v = m["h_inv"] @ np.concatenate([z.real, z.imag], 0)    # [2H, Wf]
    y = (m["w_inv"] @ np.concatenate([v[:H].T, v[H:].T], 0)).T     # [H, W]
    yref = np.fft.irfft2(z, s=(H, W))
    assert np.abs(y - yref).max() < 1e-12 * np.abs(yref).max()
    # torch's irfft2 (what the reference calls) agrees on the same spectrum
    yt = torch.fft.irfft2(torch.from_numpy(z), s=(H, W)).numpy()
    assert np.abs(y - yt).max() < 1e-12 * np.abs(yref).max()


@pytest.mark.parametrize("n", [72, 96, 130, 256, 504, 1000])
def test_factorised_dft_stages_are_the_fft(n):
    """The two stages of the factorised DFT along H (stage A with its twiddle rows laid out as the kernel writes
    them, stage B), applied as the engine's index maps apply them, against numpy.fft for both signs and the four
    (sign, scale) pairs the engine uses, and against the dense matrices they replace: 1e-7 of range (the twiddle
    table is float32). Each stage stays within the LDS-staged kernel's limits."""
    rng = np.random.default_rng(n)
    n1, n2 = T.dft_factors(n)
    assert n1 * n2 == n and 2 * T.round_up(n1, 16) <= 128 and 2 * n2 <= 128
    dense = T.horizontal_matrices(n, 4)
    x = rng.standard_normal((n, 3)) + 1j * rng.standard_normal((n, 3))            # 3 columns
    for name, sign, scale in (("h_fwd", -1, 1.0), ("h_inv", 1, 1.0 / n), ("h_inv_T", -1, 1.0 / n),
                              ("h_fwd_T", 1, 1.0)):
        a, tw, b = T.dft_stages(n, n1, n2, sign, scale)
        r = T.round_up(n1, 16)
        assert a.shape == (r + n1, 2 * n1) and np.all(a[n1:r] == 0)
        # stage A: batch over j2 columns; input row (part, j1) at h = n2 j1 + j2
        xin = np.concatenate([x.real, x.imag], 0).reshape(2, n1, n2, 3)
        ya = np.einsum("mk,kjc->mjc", a, xin.reshape(2 * n1, n2, 3))                # [R + n1, j2, c]
        y = (ya[:n1] + 1j * ya[r:]) * (tw[..., 0] + 1j * tw[..., 1])[:, :, None]    # [k1, j2, c]
        # stage B: batch over k1; input row (part, j2); output row (part, k2) at k = k1 + n1 k2
        yb = np.einsum("mk,bkc->bmc", b, np.concatenate([y.real, y.imag], 1))       # [k1, 2 n2, c]
        got = (yb[:, :n2] + 1j * yb[:, n2:]).transpose(1, 0, 2).reshape(n, 3)       # [k2, k1] -> k
        ref = scale * (np.fft.fft(x, axis=0) if sign < 0 else n * np.fft.ifft(x, axis=0))
        assert np.abs(got - ref).max() < 1e-7 * np.abs(ref).max(), name
        if name in dense:
            d = dense[name] @ np.concatenate([x.real, x.imag], 0)
            assert np.abs(d[:n] + 1j * d[n:] - got).max() < 1e-7 * np.abs(ref).max(), name
        else:
            d = dense[name[:-2]].T @ np.concatenate([x.real, x.imag], 0)
            assert np.abs(d[:n] + 1j * d[n:] - got).max() < 1e-7 * np.abs(ref).max(), name
    assert T.dft_factors(127) is None and T.dft_factors(64 * 67) is None


def test_filon_matrices_match_oracle():
    from oracle import ffno_oracle as O
    z = np.linspace(15.0, 0.0, 37) ** 1.1 - 1.0
    z = z - z.min()
    k = 2 * np.pi * np.arange(40) / 40.0
    q_re, q_im = T.filon_weights(z, k)
    o_re, o_im = O.height_fourier_weights(torch.from_numpy(z), torch.from_numpy(k))
    assert np.abs(q_re - o_re.numpy()).max() < 1e-13
    assert np.abs(q_im - o_im.numpy()).max() < 1e-13
    taper = T.column_taper(z, k)
    assert np.abs(taper - O.column_taper(torch.from_numpy(z), torch.from_numpy(k)).numpy()).max() < 1e-13


def test_split_bf16_recovers_float32_accuracy():
    rng = np.random.default_rng(1)
    a = rng.standard_normal((40, 70))
    hi, lo, m, k, mp, kp = T.split_bf16(a, "cpu")
    back = hi.view(torch.bfloat16).double() + lo.view(torch.bfloat16).double()
    assert (mp, kp) == (48, 96)
    assert np.abs(back.numpy()[:m, :k] - a).max() < 2e-5 * np.abs(a).max()
    assert float(back[m:].abs().max()) == 0.0 and float(back[:, k:].abs().max()) == 0.0


def _kernel_contraction(wp, act):
    """What the kernels compute from a packed weight: output o = sum over k-steps s, lane groups q and slots j of
    wp[o, 32s + 8q + j] * act[feature(s, q, j)], where slot j < 4 is feature 32s + 4q + j (tile 2s) and slot
    j >= 4 is feature 32s + 16 + 4q + j - 4 (tile 2s + 1)."""
    npad, kp = wp.shape
    out = np.zeros(npad)
    for s in range(kp // 32):
        for q in range(4):
            for j in range(8):
                f = 32 * s + 4 * q + j if j < 4 else 32 * s + 16 + 4 * q + (j - 4)
                out += wp[:, 32 * s + 8 * q + j] * act[f]
    return out


@pytest.mark.parametrize("N,segments", [(32, [16]), (64, [64]), (96, [96, 96, 96]), (14, [128]), (48, [16, 16])])
def test_packing_matches_kernel_contraction(N, segments):
    rng = np.random.default_rng(2)
    K = sum(segments)
    w = torch.from_numpy(rng.standard_normal((N, K))).float()
    packed = pack_mlp_weight(w, segments).view(torch.bfloat16).double()
    kp = sum(T.round_up(s, 32) for s in segments)
    assert packed.shape == (2, T.round_up(N, 16), kp + E.WPAD)
    assert not packed[..., kp:].any()                          # the row padding the kernels' stride skips
    wp = (packed[0] + packed[1]).numpy()                       # hi + lo
    act_parts = [rng.standard_normal(s) for s in segments]
    act = np.concatenate([np.pad(a, (0, T.round_up(len(a), 32) - len(a))) for a in act_parts])
    got = _kernel_contraction(wp, act)[:N]
    expected = w.double().numpy() @ np.concatenate(act_parts)
    assert np.abs(got - expected).max() < 2e-5 * max(1.0, np.abs(expected).max())
    assert not np.any(_kernel_contraction(wp, act)[N:])


def test_float32_packing_equals_the_float64_split():
    """A float32 weight is split in float32 (its remainders are exact there): the planes are bit-identical to those
    of its float64 copy, over a wide dynamic range, for two and three planes."""
    torch.manual_seed(0)
    for shape, parts in (((64, 96), 3), ((5, 16, 64), 3), ((3, 64, 32), 2)):
        w = torch.randn(*shape) * torch.logspace(-6, 3, shape[-1])
        assert torch.equal(pack_mlp_weight(w, parts=parts), pack_mlp_weight(w.double(), parts=parts))


def test_three_plane_packing_is_exact_and_extends_the_two_plane_split():
    """parts=3: the three bf16 planes add up to the float32 weight exactly, and the first two planes are the two-way
    split (what the adjoint kernels read from the same buffer)."""
    rng = np.random.default_rng(3)
    w = torch.from_numpy(rng.standard_normal((48, 80)) * 10.0 ** rng.uniform(-4, 2, (48, 80))).float()
    three = pack_mlp_weight(w, parts=3).view(torch.bfloat16).double()
    two = pack_mlp_weight(w).view(torch.bfloat16).double()
    assert three.shape[0] == 3 and two.shape[0] == 2
    assert torch.equal(three[:2], two)
    dense = three.sum(0)
    # undo the padding and the k permutation by contracting against unit vectors
    eye = np.eye(T.round_up(80, 32))
    got = np.stack([_kernel_contraction(dense.numpy(), eye[k]) for k in range(80)], 1)[:48]
    assert np.array_equal(got, w.double().numpy())


def test_state_dict_is_the_references():
    """ffnoml_amd.model.FFNO3D has exactly the keys and shapes of the reference checkpoint (golden case A was loaded
    into the reference with strict=True)."""
    from golden_cases import load
    from ffnoml_amd.model import FFNO3D
    data, config, state = load("case_a_shared64")
    model = FFNO3D(**config)
    ours = model.state_dict()
    assert sorted(ours) == sorted(state)
    for k in ours:
        assert tuple(ours[k].shape) == tuple(state[k].shape), k
    model.load_state_dict(state, strict=True)


def test_c_abi_exports_every_declared_symbol():
    """libffno_hip.so loads (no GPU needed) and exports every function include/ffno_hip.h declares."""
    from ffnoml_amd import _lib
    header = open(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                               "ffno_hip.h")).read()
    names = set(re.findall(r"^\s*(?:const char\*|int)\s+(ffno_\w+)\s*\(", header, re.M))
    assert len(names) >= 10
    handle = ctypes.CDLL(_lib.LIB_PATH)
    for n in sorted(names):
        assert hasattr(handle, n), n
    assert set(_lib.SIGNATURES) | {"ffno_last_error", "ffno_version"} >= names


def test_cpu_tensors_are_refused():
    """No CPU path: the module raises on CPU tensors instead of computing something else."""
    from ffnoml_amd import synthetic
    from ffnoml_amd.atoms import species
    model = synthetic.random_model(0, width=16, n_layers=1, hidden_spectral=8, hidden_vertical=8, freq_mix_rank=2,
                                   out_channels=6, species=species(["H"]))
    with pytest.raises(RuntimeError):
        model(torch.zeros(1, 6, 4, 4, 4), torch.zeros(1, 4, 4, 4), 48e3, 48e3)


def test_plain_torch_baseline_matches_oracle():
    """The plain FP32 comparison path (baseline/plain_fp32.py) equals the float64 oracle on golden case A to
    float32 precision (1e-4 of range), so the bench's baseline computes the same network."""
    from golden_cases import load, oracle_config
    from oracle import ffno_oracle as O
    from ffnoml_amd.model import FFNO3D
    from baseline.plain_fp32 import PlainFFNO3D
    data, config, state = load("case_a_shared64")
    model = FFNO3D(**config)
    model.load_state_dict(state, strict=True)
    plain = PlainFFNO3D(model)
    x = torch.from_numpy(data["x"]).float()
    z = torch.from_numpy(data["z"]).float()
    with torch.no_grad():
        y = plain(x, z, 48e3, 48e3)
    assert np.abs(y.double().numpy() - data["y"]).max() / np.abs(data["y"]).max() < 1e-4


def test_height_bands_are_thread_safe():
    """Plans of different shapes are built concurrently by the serving workers: the Filon height Jacobian (forward-mode
    AD inside) must give the same bands from several threads at once as alone."""
    import concurrent.futures as cf
    cols = [np.linspace(14.0, -1.0, d) for d in (24, 40, 56, 72)]
    k = 2 * np.pi * np.arange(12) / 40.0
    alone = [T.filon_height_bands(c, k) for c in cols]
    with cf.ThreadPoolExecutor(8) as ex:
        got = list(ex.map(lambda c: T.filon_height_bands(c, k), cols * 4))
    for i, g in enumerate(got):
        for d in (-1, 0, 1):
            assert np.array_equal(g[d], alone[i % 4][d])
