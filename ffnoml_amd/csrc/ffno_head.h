// The level-fraction head's per-voxel arithmetic in float32 (models/ffno_model.py:116-141, :730-749): MULTI3D LTE
// log-weights from T and ne, the log-softmax of head + ln f* per species, the bounded total s = b tanh(u / b), the
// output normalisation -- and the adjoint of all of it. Shared by the head kernels of the forward (k_head), the
// adjoint (k_head_bwd) and the stand-alone head_output entry points.
#pragma once
#include "ffno_device.h"

namespace ffno {

constexpr float LN10 = 2.302585092994046f;

// per voxel: 1/T, conl = ln(ccon ne) - 1.5 ln T and the Debye lowering dxi = 4.98e-4 * 5040 / T * sqrt(ne k T) [eV]
struct Lte {
  float inv_t, conl, dxi;
};

__device__ __forceinline__ Lte lte_of(const FfnoHeadDesc& hd, float log10_t, float log10_ne) {
  const float lt = log10_t * LN10;                 // ln T
  const float lne = (log10_ne - 6.0f) * LN10;      // ln ne [cm^-3]
  return Lte{__expf(-lt), hd.ln_ccon + lne - 1.5f * lt, hd.dxi_coef * __expf(0.5f * (lne + hd.ln_k - lt))};
}

// ln of MULTI3D's unnormalised LTE weight of level li
__device__ __forceinline__ float lte_logw(const FfnoHeadDesc& hd, const Lte& e, int li) {
  return hd.log_g[li] - hd.e_over_k * e.inv_t * (hd.ev[li] - e.dxi * hd.ndxi[li]) - hd.dion[li] * e.conl;
}
// its derivatives with respect to ln T and ln ne
__device__ __forceinline__ float lte_dlt(const FfnoHeadDesc& hd, const Lte& e, int li) {
  return hd.e_over_k * e.inv_t * (hd.ev[li] - 1.5f * e.dxi * hd.ndxi[li]) + 1.5f * hd.dion[li];
}
__device__ __forceinline__ float lte_dlne(const FfnoHeadDesc& hd, const Lte& e, int li) {
  return 0.5f * hd.e_over_k * e.inv_t * e.dxi * hd.ndxi[li] - hd.dion[li];
}

// Normalised log10 n of every output o of one voxel, out(o, value), from its head channels hv(c) and the physical
// log10 T, log10 ne, log10 rho. Species occupy levels + 1 head channels each: the level logits, then the total.
template <class HV, class OUT>
__device__ __forceinline__ void level_fractions(const FfnoHeadDesc& hd, HV hv, float l10t, float l10ne, float l10rho,
                                                OUT out) {
  const Lte e = lte_of(hd, l10t, l10ne);
  int c = 0, k = 0;
  for (int s = 0; s < hd.n_species; ++s) {
    const int nl = hd.levels[s];
    float logit[15];
    float mx = -3.0e38f;
    for (int i = 0; i < nl; ++i) {
      logit[i] = hv(c + i) + lte_logw(hd, e, k + i);
      mx = fmaxf(mx, logit[i]);
    }
    float se = 0.f;
    for (int i = 0; i < nl; ++i) se += __expf(logit[i] - mx);
    const float lse = mx + __logf(se);
    const float stot = hd.total_bound * tanhf(hv(c + nl) / hd.total_bound);
    for (int i = 0; i < nl; ++i) {
      const int o = k + i;
      const float l10n = l10rho - hd.log10_mass[s] + stot + (logit[i] - lse) / LN10;
      out(o, (l10n - hd.mean_y[o]) / hd.std_y[o]);
    }
    c += nl + 1;
    k += nl;
  }
}

// The adjoint of level_fractions for one voxel: from the output gradient yb(o) it writes the gradient of every head
// channel, hbar(c, value) (after all of that species' reads of hv, so hbar may write where hv reads), and adds the
// gradients with respect to ln T, ln ne and log10 rho to g_lt, g_lne, g_rho.
template <class HV, class YB, class HB>
__device__ __forceinline__ void level_fractions_bwd(const FfnoHeadDesc& hd, HV hv, float l10t, float l10ne, YB yb,
                                                    HB hbar, float& g_lt, float& g_lne, float& g_rho) {
  const Lte e = lte_of(hd, l10t, l10ne);
  int c = 0, k = 0;
  for (int s = 0; s < hd.n_species; ++s) {
    const int nl = hd.levels[s];
    float logit[15], a[15];
    float mx = -3.0e38f, asum = 0.f;
    for (int i = 0; i < nl; ++i) {
      const int li = k + i;
      logit[i] = hv(c + i) + lte_logw(hd, e, li);
      mx = fmaxf(mx, logit[i]);
      a[i] = yb(li) / hd.std_y[li];
      asum += a[i];
    }
    const float th = tanhf(hv(c + nl) / hd.total_bound);
    float se = 0.f;
    for (int i = 0; i < nl; ++i) se += __expf(logit[i] - mx);
    const float lse = mx + __logf(se);
    for (int i = 0; i < nl; ++i) {
      const int li = k + i;
      const float lb = (a[i] - __expf(logit[i] - lse) * asum) / LN10;      // gradient of the logit head + ln f*
      hbar(c + i, lb);
      g_lt += lb * lte_dlt(hd, e, li);
      g_lne += lb * lte_dlne(hd, e, li);
    }
    hbar(c + nl, asum * (1.f - th * th));
    g_rho += asum;
    c += nl + 1;
    k += nl;
  }
}

// head and output channel counts of a descriptor, or -1 when it is outside what the kernels hold (1 to 4 species of
// 1 to 15 levels, at most 32 head channels)
inline int head_channels(const FfnoHeadDesc* d, int* n_out) {
  if (!d || d->n_species < 1 || d->n_species > 4) return -1;
  int nh = 0, no = 0;
  for (int s = 0; s < d->n_species; ++s) {
    if (d->levels[s] < 1 || d->levels[s] > 15) return -1;
    nh += d->levels[s] + 1;
    no += d->levels[s];
  }
  if (nh > 32) return -1;
  if (n_out) *n_out = no;
  return nh;
}

}  // namespace ffno
