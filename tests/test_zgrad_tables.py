"""The per-column height-gradient kernel's host tables (engine.zgrad_tables), checked on the CPU: the kernel's
accumulation order (every level d' of a column's products P = xv X_bar^T, Q = yv_bar Y^T, each delta adding
tab[delta + 1][:, d'] . P[d'] to level d' - delta, padded to 16-row tiles) against the formulation it replaces --
the three Filon-band GEMMs G_delta X_bar and the phase GEMM V_dz Y, each followed by the channel dot product with
the level-shifted vertical input (or the branch output's gradient), terms whose shifted level leaves the column
dropped (models/ffno_model.py:272-308)."""
import numpy as np
import pytest

from ffnoml_amd import transforms as T
from ffnoml_amd.engine import zgrad_tables


def band_gemm_path(G, Vdz, xv, yv, Xb, Y):
    """zbar[d] = sum_delta sum_c xv[d + delta, c] (G_delta Xb)[d, c] + sum_c yv[d, c] (V_dz Y)[d, c]."""
    D = xv.shape[0]
    z = np.zeros(D)
    for delta in (-1, 0, 1):
        x1 = G[delta] @ Xb
        for d in range(D):
            if 0 <= d + delta < D:
                z[d] += xv[d + delta] @ x1[d]
    return z + np.sum(yv * (Vdz @ Y), axis=1)


def kernel_order(tab, xv, yv, Xb, Y):
    """The kernel's arithmetic on padded tiles: P, Q [D_pad, R_pad], the column's LDS row zw[-1 .. D_pad]."""
    _, rp, dp = tab.shape
    D, R = xv.shape[0], Xb.shape[0]
    P = np.zeros((dp, rp))
    Q = np.zeros((dp, rp))
    P[:D, :R] = xv @ Xb.T
    Q[:D, :R] = yv @ Y.T
    zw = np.zeros(dp + 2)                       # zw[i] holds level i - 1
    for d1 in range(dp):
        for k, delta in enumerate((-1, 0, 1)):
            zw[d1 - delta + 1] += tab[k, :, d1] @ P[d1]
        zw[d1 + 1] += tab[3, :, d1] @ Q[d1]
    assert abs(zw[0]) < 1e-12 and np.all(np.abs(zw[D + 1:]) < 1e-12), "a level outside the column received a term"
    return zw[1:D + 1]


@pytest.mark.parametrize("D,M,C", [(20, 6, 16), (16, 8, 32), (33, 17, 64)])
def test_tables_reproduce_the_band_gemm_path_on_random_matrices(D, M, C):
    rng = np.random.default_rng(D + M)
    R = 2 * M
    G = {d: rng.standard_normal((D, R)) for d in (-1, 0, 1)}
    Vdz = rng.standard_normal((D, R))
    xv, yv = rng.standard_normal((D, C)), rng.standard_normal((D, C))
    Xb, Y = rng.standard_normal((R, C)), rng.standard_normal((R, C))
    tab = zgrad_tables(G, Vdz)
    assert tab.shape == (4, T.round_up(R, 16), T.round_up(D, 16))
    ref = band_gemm_path(G, Vdz, xv, yv, Xb, Y)
    got = kernel_order(tab, xv, yv, Xb, Y)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


def test_tables_of_a_real_column_and_a_planted_shift_mistake():
    """The Filon bands and synthesis derivative of a stratified column (transforms.filon_height_bands,
    vertical_height_terms); a table built with the band shift reversed fails the same comparison."""
    rng = np.random.default_rng(3)
    col = np.cumsum(0.1 + 0.05 * rng.random(24))[::-1].copy()
    col -= col.min()
    L, k = 40.0, 2 * np.pi * np.arange(9) / 40.0
    bands = T.filon_height_bands(col, k)
    v_dz, _, _, _, _ = T.vertical_height_terms(col, k, np.ones_like(k), L)
    D, R, C = len(col), v_dz.shape[1], 32
    xv, yv = rng.standard_normal((D, C)), rng.standard_normal((D, C))
    Xb, Y = rng.standard_normal((R, C)), rng.standard_normal((R, C))
    ref = band_gemm_path(bands, v_dz, xv, yv, Xb, Y)
    got = kernel_order(zgrad_tables(bands, v_dz), xv, yv, Xb, Y)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    wrong = zgrad_tables({-1: bands[1], 0: bands[0], 1: bands[-1]}, v_dz)
    bad = kernel_order(wrong, xv, yv, Xb, Y)
    assert np.abs(bad - ref).max() > 1e-3 * np.abs(ref).max()
