"""The GroupNorm adjoint's per-channel constant form, checked on the CPU against the direct form it rearranges
(both in float64): t_bar = P act u_bar - w (Q + R (t - M)) and the voxel's depth-weight gradient
w_bar = K - sum_c d (Q + R d / 2), d = t - M, with P = rstd scale gamma, Q = rstd S1 / W, R = rstd^2 S2 / W,
K = sum_c (S2 / 2W)(1 - eps rstd^2), against rstd (g_hat - (w / W)(S1 + x_hat S2)) and
-(1/W) sum_c (S1 x_hat + S2 (x_hat^2 - rstd^2 var) / 2), x_hat = (t - mean) rstd, rstd^2 var = 1 - eps rstd^2
(the weighted GroupNorm of models/ffno_model.py:77-95 differentiated by hand)."""
import numpy as np


def test_constant_form_equals_the_direct_form():
    rng = np.random.default_rng(0)
    G, CPG, N = 8, 8, 50
    C = G * CPG
    mean, rstd = rng.standard_normal(G), 0.5 + rng.random(G)
    Wt = 100.0 + rng.random(G)
    S1, S2 = rng.standard_normal(G), rng.standard_normal(G)
    gamma, scale = rng.standard_normal(C), 0.7
    g = np.repeat(np.arange(G), CPG)
    t = rng.standard_normal((N, C)) * 2 + mean[g]
    ub = rng.standard_normal((N, C))
    act = 0.5 + rng.random((N, C))
    w = rng.random(N) * 2
    # direct
    xh = (t - mean[g]) * rstd[g]
    gh = scale * ub * gamma * act
    tb_ref = rstd[g] * (gh - (w[:, None] / Wt[g]) * (S1[g] + xh * S2[g]))
    r2var = 1 - 1e-5 * rstd[g] ** 2
    wb_ref = -((S1[g] * xh + 0.5 * S2[g] * (xh * xh - r2var)) / Wt[g]).sum(1)
    # per-channel constants
    P, Q, R, M = rstd[g] * scale * gamma, rstd[g] * S1[g] / Wt[g], rstd[g] ** 2 * S2[g] / Wt[g], mean[g]
    K = sum(CPG * 0.5 * S2[j] / Wt[j] * (1 - 1e-5 * rstd[j] ** 2) for j in range(G))
    d = t - M
    tb = P * act * ub - w[:, None] * (Q + R * d)
    wb = K - (d * (Q + 0.5 * R * d)).sum(1)
    assert np.abs(tb - tb_ref).max() <= 1e-12 * np.abs(tb_ref).max()
    assert np.abs(wb - wb_ref).max() <= 1e-12 * np.abs(wb_ref).max()
    # a planted mistake (the half dropped from R d / 2) is seen
    wb_bad = K - (d * (Q + R * d)).sum(1)
    assert np.abs(wb_bad - wb_ref).max() > 1e-3 * np.abs(wb_ref).max()
