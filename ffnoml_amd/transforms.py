"""Transform matrices of the spectral branches, built once per plan in float64 on the host (CPU) and uploaded in
split-bf16: no float64 arithmetic runs on the GPU.

Horizontal (models/ffno_model.py:441, :448): torch.fft.rfft2 over (H, W), then irfft2 with s=(H, W), written as four
real matrices acting on spectra held as separate real/imaginary planes:
  W_fwd  [2Wf x W]    rows kx: cos(2 pi kx w / W); rows Wf+kx: -sin(...)             (rfft along W)
  H_fwd  [2H x 2H]    complex DFT along H, exp(-2 pi i ky h / H)
  H_inv  [2H x 2H]    complex inverse DFT along H with 1/H
  W_inv  [W x 2Wf]    the c2r inverse along W with 1/W: bin 0 (and W/2 for even W) count once and contribute only
                      their real part, every other bin counts twice -- the Hermitian convention irfft applies to a
                      spectrum that is not Hermitian (it is not, after the per-mode mix).
Vertical (models/ffno_model.py:37-64, :239-259, :270-276, :301-305) on one shared column grid:
  V_fwd  [2M x D]     Filon weights q_re, q_im of the column's (shifted) heights
  V_inv  [D x 2M]     f_m cos(k_m z_d) / L and -f_m sin(k_m z_d) / L, f = mode factor * taper
"""
import math
import threading

import numpy as np
import torch


# torch's forward-mode AD keeps its dual levels in process-global state, so concurrent torch.func.jvp calls from
# two threads (two serving workers planning different shapes) can pop each other's level
_FORWARD_AD_LOCK = threading.Lock()


def round_up(n, m):
    return (n + m - 1) // m * m


def horizontal_matrices(H, W):
    wf = W // 2 + 1
    w = np.arange(W)
    kx = np.arange(wf)
    th = 2 * np.pi * np.outer(kx, w) / W                       # [Wf, W]
    w_fwd = np.concatenate([np.cos(th), -np.sin(th)], 0)       # [2Wf, W]
    h = np.arange(H)
    ph = 2 * np.pi * np.outer(h, h) / H                        # [ky, h] (symmetric)
    c, s = np.cos(ph), np.sin(ph)
    h_fwd = np.block([[c, s], [-s, c]])                        # re: c xr + s xi ; im: -s xr + c xi
    h_inv = np.block([[c, -s], [s, c]]) / H                    # re: c zr - s zi ; im: s zr + c zi
    weight = np.full(wf, 2.0)
    weight[0] = 1.0
    if W % 2 == 0:
        weight[-1] = 1.0
    thi = 2 * np.pi * np.outer(w, kx) / W                      # [W, Wf]
    w_inv = np.concatenate([weight * np.cos(thi), -weight * np.sin(thi)], 1) / W   # [W, 2Wf]
    return dict(w_fwd=w_fwd, h_fwd=h_fwd, h_inv=h_inv, w_inv=w_inv)


def vertical_box(span, reference_length, min_wavelength, modes):
    """(L, k [M] as the reference's float32 values in float64, mode factors [M]) (models/ffno_model.py:248-259,
    :226-229): L = reference_length unless twice the tallest column is longer."""
    if 2 * span <= reference_length:
        k = (2 * math.pi * torch.arange(modes, dtype=torch.float32) / reference_length).double().numpy()
        length = reference_length
    else:
        length = 2 * span
        modes = int(math.floor(length / min_wavelength)) + 1
        k = (2 * math.pi * torch.arange(modes, dtype=torch.float32) / length).double().numpy()
    factor = np.full(modes, 2.0)
    factor[0] = 1.0
    return length, k, factor


def dft_factors(n, largest=64):
    """(n1, n2), n1 * n2 = n, both in [2, largest], that needs the fewest MFMA tile steps over the two stages of the
    factorised DFT (dft_stages), or None when n has no such split."""
    best = None
    for n1 in range(2, largest + 1):
        n2 = n // n1
        if n % n1 or not 2 <= n2 <= largest:
            continue
        r = round_up(n1, 16)
        cost = (2 * r // 16) * (round_up(2 * n1, 32) // 32) + (round_up(2 * n2, 16) // 16) * (round_up(2 * n2, 32) // 32)
        if best is None or cost < best[0]:
            best = (cost, n1, n2)
    return None if best is None else best[1:]


def _cmul_block(theta, sign):
    """Real block matrix of the complex products exp(sign i theta) z for a complex vector held as (real parts,
    imaginary parts): [[cos, -sign sin], [sign sin, cos]]."""
    c, s = np.cos(theta), np.sin(theta)
    return np.block([[c, -sign * s], [sign * s, c]])


def dft_stages(n, n1, n2, sign, scale):
    """The complex DFT X[k] = scale sum_h x[h] exp(sign 2 pi i h k / n) in two stages (Cooley-Tukey, n = n1 n2,
    h = n2 j1 + j2, k = k1 + n1 k2):
      stage A  y[k1, j2] = exp(sign 2 pi i j2 k1 / n) sum_j1 exp(sign 2 pi i j1 k1 / n1) x[n2 j1 + j2]
      stage B  X[k1 + n1 k2] = scale sum_j2 exp(sign 2 pi i j2 k2 / n2) y[k1, j2]
    Returns (A [R + n1, 2 n1] with real-part rows 0..n1-1 and imaginary-part rows R..R+n1-1, R = n1 rounded up to
    16; twiddle [n1, n2, 2] float32 (cos, sin of the stage-A factor); B [2 n2, 2 n2]), float64 matrices."""
    r = round_up(n1, 16)
    j = np.arange(n1)
    blk = _cmul_block(2 * np.pi * np.outer(j, j) / n1, sign)
    a = np.zeros((r + n1, 2 * n1))
    a[:n1] = blk[:n1]
    a[r:] = blk[n1:]
    ph = 2 * np.pi * np.outer(np.arange(n1), np.arange(n2)) / n
    tw = np.stack([np.cos(ph), sign * np.sin(ph)], -1).astype(np.float32)
    j2 = np.arange(n2)
    b = scale * _cmul_block(2 * np.pi * np.outer(j2, j2) / n2, sign)
    return a, tw, b


def _runs(keys):
    """[(start, count, key)] of the runs of equal consecutive keys."""
    out = []
    for i, key in enumerate(keys):
        if out and out[-1][2] == key:
            out[-1][1] += 1
        else:
            out.append([i, 1, key])
    return [tuple(r) for r in out]


def _irfft_weight(k, n):
    return 1.0 if k == 0 or 2 * k == n else 2.0


def rdft_stages(n, n1, n2, scale, weighted):
    """Real rows of length n -> half spectrum (n // 2 + 1 bins): X[k] = scale w(k) sum_x x[j] exp(-2 pi i j k / n),
    w(k) = 1, or with weighted the irfft bin weights (1 for bin 0 and, n even, bin n/2; 2 otherwise), in two
    stages: j = n2 j1 + j2, k = k1 + n1 k2.
      stage A  y[k1, j2] = exp(-2 pi i j2 k1 / n) sum_j1 exp(-2 pi i j1 k1 / n1) x[n2 j1 + j2]: A [R + n1, n1]
               (real-part rows 0..n1-1, imaginary-part rows R..R+n1-1, R = n1 rounded up to 16), twiddle
               [n1, n2, 2] float32
      stage B  X[k1 + n1 k2] = scale w sum_j2 exp(-2 pi i j2 k2 / n2) y[k1, j2] for the bins below n // 2 + 1:
               runs of k1 [(k1 start, count, B [2 K2, 2 n2])] with K2 the bins each k1 of the run reaches
    Returns (A, twiddle, runs), float64 matrices."""
    wf = n // 2 + 1
    r = round_up(n1, 16)
    j = np.arange(n1)
    th = 2 * np.pi * np.outer(j, j) / n1                                   # [k1, j1]
    a = np.zeros((r + n1, n1))
    a[:n1], a[r:] = np.cos(th), -np.sin(th)
    ph = 2 * np.pi * np.outer(np.arange(n1), np.arange(n2)) / n
    tw = np.stack([np.cos(ph), -np.sin(ph)], -1).astype(np.float32)
    j2 = np.arange(n2)
    runs = []
    keys = []
    for k1 in range(n1):
        K2 = -(-(wf - k1) // n1)
        wts = tuple(_irfft_weight(k1 + n1 * k2, n) if weighted else 1.0 for k2 in range(K2))
        keys.append((K2, wts))
    for start, count, (K2, wts) in _runs(keys):
        blk = _cmul_block(2 * np.pi * np.outer(np.arange(K2), j2) / n2, -1)   # [2 K2, 2 n2]
        wv = scale * np.asarray(wts)
        runs.append((start, count, blk * np.concatenate([wv, wv])[:, None]))
    return a, tw, runs


def irdft_stages(n, n1, n2, scale, weighted):
    """Half spectrum (n // 2 + 1 bins) -> real rows of length n: x[j] = scale sum_k w(k) Re(Z[k] exp(2 pi i j k / n)),
    w(k) = 1, or with weighted the irfft bin weights (so that weighted, scale 1/n is irfft for any spectrum, Hermitian
    or not), in two stages: k = n2 j1 + j2, j = k1 + n1 k2.
      stage A  y[k1, j2] = exp(2 pi i j2 k1 / n) sum_j1 w exp(2 pi i j1 k1 / n1) Z[n2 j1 + j2] over the bins below
               n // 2 + 1: runs of j2 [(j2 start, count, A [R + n1, 2 J1])] with J1 the bins each j2 of the run
               reaches (real-part rows 0..n1-1, imaginary-part rows R..R+n1-1; columns real parts, then imaginary),
               twiddle [n1, n2, 2] float32 (indexed by the run's j2)
      stage B  x[k1 + n1 k2] = scale Re sum_j2 exp(2 pi i j2 k2 / n2) y[k1, j2]: B [n2, 2 n2]
    Returns (runs, twiddle, B), float64 matrices."""
    wf = n // 2 + 1
    r = round_up(n1, 16)
    k1 = np.arange(n1)
    keys = []
    for j2 in range(n2):
        J1 = -(-(wf - j2) // n2)
        keys.append((J1, tuple(_irfft_weight(n2 * j1 + j2, n) if weighted else 1.0 for j1 in range(J1))))
    runs = []
    for start, count, (J1, wts) in _runs(keys):
        blk = _cmul_block(2 * np.pi * np.outer(k1, np.arange(J1)) / n1, 1)   # [2 n1, 2 J1]
        blk = blk * np.concatenate([wts, wts])[None, :]
        a = np.zeros((r + n1, 2 * J1))
        a[:n1], a[r:] = blk[:n1], blk[n1:]
        runs.append((start, count, a))
    ph = 2 * np.pi * np.outer(np.arange(n1), np.arange(n2)) / n
    tw = np.stack([np.cos(ph), np.sin(ph)], -1).astype(np.float32)
    j2 = np.arange(n2)
    th = 2 * np.pi * np.outer(j2, j2) / n2                                  # [k2, j2]
    b = scale * np.concatenate([np.cos(th), -np.sin(th)], 1)                # Re(exp(i th) y)
    return runs, tw, b


def split_bf16(a, device, parts=2):
    """(hi, lo, m, k, m_pad, k_pad[, l3]) of a float64 matrix [m, k] as bf16 planes [m_pad, k_pad], zero padded:
    hi = bf16(a), lo = bf16(a - hi), and with parts=3 l3 = bf16(a - hi - lo) (the forward transforms' three-way
    products)."""
    a = np.asarray(a, np.float64)
    m, k = a.shape
    mp, kp = round_up(m, 16), round_up(k, 32)
    full = np.zeros((mp, kp))
    full[:m, :k] = a
    rest = torch.from_numpy(full)
    planes = []
    for _ in range(parts):
        p = rest.to(torch.bfloat16)
        planes.append(p.view(torch.int16).to(device))
        rest = rest - p.double()
    return (planes[0], planes[1], m, k, mp, kp) + tuple(planes[2:])


# ----------------------------------------------------------------------------------------------------------------
# vertical (height) transforms, on one shared column [D] or on every column of a per-column grid [n, D] (the
# reference's per-column path, models/ffno_model.py:277-283, :304-305): one torch implementation over any leading
# dimensions, evaluated in float64 on the host
# ----------------------------------------------------------------------------------------------------------------
def _t(a, like=None):
    """a as a float64 tensor (on like's device when given)."""
    t = a if torch.is_tensor(a) else torch.as_tensor(np.asarray(a, np.float64))
    return t if like is None else t.to(like.device)


def filon_weights_t(z, k):
    """Filon quadrature weights (q_re, q_im) [..., D, M] of the heights z [..., D] at the wavenumbers k [M] (torch,
    differentiable in z): sum_j x_j q_jm = int x(z) exp(-i k_m z) dz for x linear between the levels (any order).
    Same quadrature and small-angle series (|theta| < 1e-2) as models/ffno_model.py:37-64."""
    z0, z1 = z[..., :-1, None], z[..., 1:, None]
    low = torch.minimum(z0, z1)
    h = (z1 - z0).abs()
    th = h * k
    small = th.abs() < 1e-2
    tt = torch.where(small, torch.ones_like(th), th)
    c, s = torch.cos(tt), torch.sin(tt)
    # shares of an interval's upper end, int_0^1 u e^{-i th u} du, and of its lower end, int_0^1 (1-u) e^{-i th u} du
    up_re = torch.where(small, 0.5 - th ** 2 / 8, (c + tt * s - 1) / tt ** 2)
    up_im = torch.where(small, -th / 3 + th ** 3 / 30, (tt * c - s) / tt ** 2)
    lo_re = torch.where(small, 0.5 - th ** 2 / 24, s / tt - up_re)
    lo_im = torch.where(small, -th / 6 + th ** 3 / 120, (c - 1) / tt - up_im)
    ph = low * k
    e_re, e_im = h * torch.cos(ph), -h * torch.sin(ph)
    a_re, a_im = e_re * lo_re - e_im * lo_im, e_re * lo_im + e_im * lo_re      # the interval's lower level
    b_re, b_im = e_re * up_re - e_im * up_im, e_re * up_im + e_im * up_re      # its upper level
    asc = z0 <= z1
    pad = torch.nn.functional.pad
    q_re = pad(torch.where(asc, a_re, b_re), (0, 0, 0, 1)) + pad(torch.where(asc, b_re, a_re), (0, 0, 1, 0))
    q_im = pad(torch.where(asc, a_im, b_im), (0, 0, 0, 1)) + pad(torch.where(asc, b_im, a_im), (0, 0, 1, 0))
    return q_re, q_im


def filon_weights(z, k):
    """filon_weights_t of one column (numpy float64 in and out): (q_re, q_im) [D, M]."""
    q_re, q_im = filon_weights_t(_t(z), _t(k))
    return q_re.numpy(), q_im.numpy()


def _median_spacing(z):
    """(index i [...], spacing |z_{i+1} - z_i| [..., 1]) of the lower median of every column's |spacing|: the lower
    middle of a stable sort, the interval CPU torch.median returns for ties (models/ffno_model.py:239-246)."""
    d = (z[..., 1:] - z[..., :-1]).abs()
    i = torch.sort(d, dim=-1, stable=True).indices[..., (d.shape[-1] - 1) // 2]
    return i, d.gather(-1, i[..., None])


def column_taper_t(z, k):
    """Half-cosine taper [..., M] from 2 pi / (3 dz) to pi / dz, dz the lower median |spacing| of each column
    (models/ffno_model.py:239-246); 1 for a single level."""
    if z.shape[-1] < 2:
        return torch.ones(z.shape[:-1] + k.shape, dtype=z.dtype, device=z.device)
    _, dz = _median_spacing(z)
    k_nyq = math.pi / dz
    u = ((k - 2 * k_nyq / 3) / (k_nyq / 3)).clamp(0.0, 1.0)
    return 0.5 * (1 + torch.cos(math.pi * u))


def column_taper(column, k):
    """column_taper_t of one column (numpy float64): [M]."""
    return column_taper_t(_t(column), _t(k)).numpy()


def vertical_matrices_t(z, k, factor, length):
    """V_fwd [..., 2M, D] (Filon weights, real parts then imaginary) and V_inv [..., D, 2M] (f cos(k z) / L and
    -f sin(k z) / L, f = mode factor * taper) of the shifted heights z [..., D]."""
    k, factor = _t(k, z), _t(factor, z)
    q_re, q_im = filon_weights_t(z, k)
    v_fwd = torch.cat([q_re.transpose(-1, -2), q_im.transpose(-1, -2)], -2)
    f = factor * column_taper_t(z, k)
    ph = z[..., :, None] * k
    v_inv = torch.cat([f[..., None, :] * torch.cos(ph), -f[..., None, :] * torch.sin(ph)], -1) / length
    return v_fwd, v_inv


def vertical_matrices(column, k, factor, length):
    """V_fwd [2M, D] and V_inv [D, 2M] of one shared column (numpy float64)."""
    v_fwd, v_inv = vertical_matrices_t(_t(column), k, factor, length)
    return v_fwd.numpy(), v_inv.numpy()


def filon_height_bands_t(z, k):
    """G[delta] [..., D, 2M] for delta = -1, 0, 1: G[delta][..., e, :] = d q[..., e + delta, :] / d z_e (real parts,
    then imaginary), the three non-zero bands of the Filon weights' height Jacobian (the weights of level d depend on
    levels d-1, d, d+1), from forward-mode derivatives along three interleaved tangents."""
    k = _t(k, z)
    D, M = z.shape[-1], k.shape[0]
    bands = {d: torch.zeros(z.shape[:-1] + (D, 2 * M), dtype=z.dtype, device=z.device) for d in (-1, 0, 1)}
    idx = torch.arange(D, device=z.device)
    for color in range(3):
        v = torch.zeros_like(z)
        v[..., color::3] = 1.0
        with _FORWARD_AD_LOCK:           # forward-AD levels are process-global: one plan build at a time here
            _, (dre, dim) = torch.func.jvp(lambda zz: filon_weights_t(zz, k), (z,), (v,))
        for delta in (-1, 0, 1):
            e = idx - delta                        # row d of the derivative belongs to level e = d - delta
            ok = (e >= 0) & (e < D) & (e % 3 == color)
            rows, levels = idx[ok], e[ok]
            bands[delta][..., levels, :M] = dre[..., rows, :]
            bands[delta][..., levels, M:] = dim[..., rows, :]
    return bands


def filon_height_bands(column, k):
    """filon_height_bands_t of one column (numpy float64): {delta: [D, 2M]}."""
    return {d: b.numpy() for d, b in filon_height_bands_t(_t(column), k).items()}


def vertical_height_terms_t(z, k, factor, length):
    """The height derivatives of the synthesis for the heights z [..., D]: V_dz [..., D, 2M] = d V_inv / d z_d
    (-f k sin, -f k cos over L), the unscaled synthesis V0 [..., D, 2M] = (cos, -sin) / L, factor * d taper / d dz
    [..., M] at each column's dz, the index i [...] of its median spacing (dz = |z_{i+1} - z_i|) and the sign of
    that spacing [...]. Spacings that tie at the median make the median non-differentiable; the gradient goes to
    the lower middle of a stable sort, the interval torch.median (models/ffno_model.py:241) returns on the CPU."""
    k, factor = _t(k, z), _t(factor, z)
    ph = z[..., :, None] * k
    f = factor * column_taper_t(z, k)
    v_dz = torch.cat([-(f * k)[..., None, :] * torch.sin(ph), -(f * k)[..., None, :] * torch.cos(ph)], -1) / length
    v0 = torch.cat([torch.cos(ph), -torch.sin(ph)], -1) / length
    if z.shape[-1] < 2:
        zero = torch.zeros(z.shape[:-1], dtype=torch.long, device=z.device)
        return v_dz, v0, torch.zeros(z.shape[:-1] + k.shape, dtype=z.dtype, device=z.device), zero, zero.to(z.dtype)
    i, dz = _median_spacing(z)
    k_nyq = math.pi / dz
    u = (k - 2 * k_nyq / 3) / (k_nyq / 3)
    inside = (u > 0) & (u < 1)
    # taper = (1 + cos(pi u)) / 2 with u = (3 k dz / pi) - 2: d taper / d dz = -(pi/2) sin(pi u) * 3 k / pi
    dtaper = torch.where(inside, -0.5 * torch.sin(math.pi * u) * 3.0 * k, torch.zeros_like(u))
    sgn = torch.sign(z.gather(-1, (i + 1)[..., None]) - z.gather(-1, i[..., None]))[..., 0]
    return v_dz, v0, factor * dtaper, i, sgn


def vertical_height_terms(column, k, factor, length):
    """vertical_height_terms_t of one column (numpy float64): (V_dz, V0, dfac, i, sign) with i an int and sign a
    float."""
    v_dz, v0, dfac, i, sgn = vertical_height_terms_t(_t(column), k, factor, length)
    return v_dz.numpy(), v0.numpy(), dfac.numpy(), int(i), float(sgn)


def split_bf16_columns(a, device, parts=2):
    """(hi, lo, m, k, m_pad, k_pad[, l3]) of per-column matrices a [n, m, k] (host torch float64): bf16 planes
    [n, m_pad, k_pad] on the device, zero padded, as split_bf16 per column (the split is exact arithmetic in float64
    on the host)."""
    n, m, k = a.shape
    mp, kp = round_up(m, 16), round_up(k, 32)
    rest = torch.nn.functional.pad(a, (0, kp - k, 0, mp - m))
    planes = []
    for _ in range(parts):
        p = rest.to(torch.bfloat16)
        planes.append(p.contiguous().view(torch.int16).to(device))
        rest = rest - p.double()
    return (planes[0], planes[1], m, k, mp, kp) + tuple(planes[2:])
