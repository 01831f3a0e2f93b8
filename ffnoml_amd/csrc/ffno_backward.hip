// FFNO3D adjoint kernels for gfx950: the input VJP of every forward kernel, with the forward recomputed from saved
// block inputs. Same lane layout and split-bf16 MFMA arithmetic as the forward (ffno_device.h).
#include "ffno_device.h"
#include "ffno_head.h"

namespace ffno {
namespace {

// ---------------------------------------------------------------------------------------------------------------
// GroupNorm backward
// ---------------------------------------------------------------------------------------------------------------
// Per (lane, tile t) GroupNorm constants: with at least 4 channels per group a lane's 4 channels of a tile share one
// group, so its statistics and gradient sums are loaded once per tile, not per channel
template <int CT>
__device__ __forceinline__ int gn_group(int c) { return c / ((CT * 16) / G); }

// g_hat and x_hat of one lane's 4 channels of tile t, from u_bar and the GroupNorm input tv of those channels
template <int CT>
__device__ __forceinline__ void gn_ghat_of(const FfnoGnBack& gb, float scale, int b, bool ok, int q, int t, f32x4 ub,
                                           f32x4 tv, f32x4& gh, f32x4& xh) {
  constexpr int C = CT * 16, CPG = C / G;
  const int c0 = 16 * t + 4 * q;
  const f32x4 gm = *(const f32x4*)(gb.gamma + c0);
  f32x4 act = f32x4{1.f, 1.f, 1.f, 1.f};
  if (gb.act_gelu) {
    const f32x4 a = *(const f32x4*)(gb.aff_a + b * C + c0);
    const f32x4 s = *(const f32x4*)(gb.aff_b + b * C + c0);
    act = gelu_grad4(tv * a + s);
  }
  if constexpr (CPG >= 4) {
    const float* st = gb.stats + (b * G + gn_group<CT>(c0)) * 3;
    const float mean = st[0], rstd = st[1];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      xh[j] = ok ? (tv[j] - mean) * rstd : 0.f;
      gh[j] = ok ? scale * ub[j] * gm[j] * act[j] : 0.f;
    }
  } else {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const float* st = gb.stats + (b * G + gn_group<CT>(c0 + j)) * 3;
      xh[j] = ok ? (tv[j] - st[0]) * st[1] : 0.f;
      gh[j] = ok ? scale * ub[j] * gm[j] * act[j] : 0.f;
    }
  }
}

// the same, loading tv from the GroupNorm input
template <int CT>
__device__ __forceinline__ void gn_ghat(const FfnoGnBack& gb, float scale, int64_t row, int b, bool ok, int q, int t,
                                        f32x4 ub, f32x4& gh, f32x4& xh) {
  gn_ghat_of<CT>(gb, scale, b, ok, q, t, ub, ld4(gb.t + row * CT * 16 + 16 * t + 4 * q, ok), gh, xh);
}

// t_bar = rstd * (g_hat - (w / W) * (S1 + x_hat * S2)); with gb.wbar, also the depth-weight gradient of this voxel,
// dL/dw = -(1/W) sum_c (S1 x_hat + S2 (x_hat^2 - rstd^2 var) / 2) with rstd^2 var = 1 - eps rstd^2, accumulated
template <int CT>
__device__ __forceinline__ void gn_back(const FfnoGnBack& gb, float scale, int64_t row, int b, bool ok, int q,
                                        const f32x4 (&ub)[CT], f32x4 (&tb)[CT]) {
  constexpr int C = CT * 16, CPG = C / G;
  const float w = ok ? gb.depth_w[row] : 0.f;
  float wb = 0.f;
#pragma unroll
  for (int t = 0; t < CT; ++t) {
    f32x4 gh, xh;
    gn_ghat<CT>(gb, scale, row, b, ok, q, t, ub[t], gh, xh);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int g = CPG >= 4 ? gn_group<CT>(16 * t + 4 * q) : gn_group<CT>(16 * t + 4 * q + j);
      const float* st = gb.stats + (b * G + g) * 3;
      const float* sm = gb.sums + (b * G + g) * 2;
      const float inv_w = __builtin_amdgcn_rcpf(st[2]);
      tb[t][j] = ok ? st[1] * (gh[j] - (w * inv_w) * (sm[0] + xh[j] * sm[1])) : 0.f;
      if (gb.wbar) {
        const float r2var = 1.f - 1e-5f * st[1] * st[1];
        wb -= (sm[0] * xh[j] + 0.5f * sm[1] * (xh[j] * xh[j] - r2var)) * inv_w;
      }
    }
  }
  if (gb.wbar) {
    wb += __shfl_xor(wb, 16);
    wb += __shfl_xor(wb, 32);
    if (ok && q == 0) gb.wbar[row] += wb;
  }
}

// S1 = sum g_hat and S2 = sum g_hat x_hat of one GroupNorm's backward, accumulated by a kernel that produces that
// GroupNorm's u_bar (so no separate pass re-reads it), into the partials ffno_gn_grad_finalize reduces.
// Same g_hat, x_hat as k_gn_grad_stats. Worth it for a GroupNorm without activation: with a GELU after it, g_hat
// needs the GroupNorm input and a GELU derivative per channel, which cost the producer more than the separate pass.
template <int CT>
struct GradSums : GroupSums<CT> {
  using GroupSums<CT>::SUB;
  __device__ __forceinline__ void add(const FfnoGnBack& gb, float scale, int64_t row, int b, bool ok, int q, int v,
                                      const f32x4 (&ub)[CT]) {
    float a[CT * SUB], c[CT * SUB];
#pragma unroll
    for (int t = 0; t < CT; ++t) {
      f32x4 gh, xh;
      gn_ghat<CT>(gb, scale, row, b, ok, q, t, ub[t], gh, xh);
#pragma unroll
      for (int u = 0; u < SUB; ++u) {
        a[t * SUB + u] = 0.f;
        c[t * SUB + u] = 0.f;
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        a[t * SUB + j / (4 / SUB)] += gh[j];
        c[t * SUB + j / (4 / SUB)] += gh[j] * xh[j];
      }
    }
    this->put(a, c, v);
  }
};

// The GroupNorm backward and an input assembly as per-channel constants of one batch item, computed once per workgroup
// into LDS kc[8][C]: P = rstd scale gamma, Q = rstd S1 / W, R = rstd^2 S2 / W, M = mean (per channel of its group),
// A, B = the GroupNorm's affine (for the GELU after it), XA, XB = the assembly's affine. Then
//   t_bar = P act u_bar - w (Q + R (t - M)),  act = GELU'(A t + B) or 1,
// which is gn_back's rstd (g_hat - (w / W) (S1 + x_hat S2)) rearranged. gn_back's depth-weight gradient of a voxel,
// -(1/W) sum_c (S1 x_hat + S2 (x_hat^2 - rstd^2 var) / 2), is in the same constants with d = t - M
//   w_bar = K - sum_c d (Q + R d / 2),  K = sum_c (S2 / 2W) (1 - eps rstd^2)  (one scalar, written to *ksum).
template <int CT>
__device__ void gn_consts(const FfnoGnBack& gb, float scale, const FfnoAssemble& xin, int b, float* kc,
                          float* ksum = nullptr) {
  constexpr int C = CT * 16, CPG = C / G;
  for (int c = threadIdx.x; c < C; c += blockDim.x) {
    const int g = c / CPG;
    const float* st = gb.stats + (b * G + g) * 3;
    const float* sm = gb.sums + (b * G + g) * 2;
    const float rstd = st[1], inv_w = __builtin_amdgcn_rcpf(st[2]);
    kc[c] = rstd * scale * gb.gamma[c];
    kc[C + c] = rstd * inv_w * sm[0];
    kc[2 * C + c] = rstd * rstd * inv_w * sm[1];
    kc[3 * C + c] = st[0];
    kc[4 * C + c] = gb.act_gelu ? gb.aff_a[b * C + c] : 0.f;
    kc[5 * C + c] = gb.act_gelu ? gb.aff_b[b * C + c] : 0.f;
    kc[6 * C + c] = xin.t ? xin.aff_a[b * C + c] : 0.f;
    kc[7 * C + c] = xin.t ? xin.aff_b[b * C + c] : 0.f;
  }
  if (ksum && threadIdx.x == 0) {
    float k = 0.f;
    for (int g = 0; g < G; ++g) {
      const float* st = gb.stats + (b * G + g) * 3;
      const float* sm = gb.sums + (b * G + g) * 2;
      k += (float)CPG * 0.5f * sm[1] * __builtin_amdgcn_rcpf(st[2]) * (1.f - 1e-5f * st[1] * st[1]);
    }
    *ksum = k;
  }
}

// t_bar from the constants (rows outside the batch item have u_bar = t = 0 and w = 0, so t_bar = 0 there); the
// GroupNorm input is read per tile t, as gn_back does. WB: also the voxel's depth-weight gradient, added to
// gb.wbar (ksum: the scalar K gn_consts wrote)
template <int CT, bool WB = false>
__device__ __forceinline__ void gn_back_k(const FfnoGnBack& gb, const float* kc, int64_t row, bool ok, int q,
                                          const f32x4 (&ub)[CT], f32x4 (&tb)[CT], const float* ksum = nullptr) {
  constexpr int C = CT * 16;
  const float w = ok ? gb.depth_w[row] : 0.f;
  float wb = 0.f;
#pragma unroll
  for (int t = 0; t < CT; ++t) {
    const int c0 = 16 * t + 4 * q;
    const f32x4 tv = ld4(gb.t + row * C + c0, ok);
    f32x4 g = *(const f32x4*)(kc + c0) * ub[t];
    if (gb.act_gelu) g *= gelu_grad4(tv * *(const f32x4*)(kc + 4 * C + c0) + *(const f32x4*)(kc + 5 * C + c0));
    if constexpr (WB) {
      const f32x4 Q = *(const f32x4*)(kc + C + c0);
      const f32x4 d = tv - *(const f32x4*)(kc + 3 * C + c0);
      const f32x4 rd = *(const f32x4*)(kc + 2 * C + c0) * d;
      tb[t] = g - w * (Q + rd);
      const f32x4 e = d * (Q + 0.5f * rd);
      wb -= (e[0] + e[1]) + (e[2] + e[3]);
    } else {
      const f32x4 qr = *(const f32x4*)(kc + C + c0) + *(const f32x4*)(kc + 2 * C + c0) *
                                                          (tv - *(const f32x4*)(kc + 3 * C + c0));
      tb[t] = g - w * qr;
    }
  }
  if constexpr (WB) {
    wb += __shfl_xor(wb, 16);
    wb += __shfl_xor(wb, 32);
    if (ok && q == 0) gb.wbar[row] += wb + *ksum;
  }
}

// the assembly base + scale act(t XA + XB) from the constants (rows outside the batch item give finite values that
// only reach their own, unstored, voxel column)
template <int CT>
__device__ __forceinline__ void assemble_k(const FfnoAssemble& in, const float* kc, float scale, int64_t row, bool ok,
                                           int q, f32x4 (&x)[CT]) {
  constexpr int C = CT * 16;
#pragma unroll
  for (int t = 0; t < CT; ++t) {
    const int c0 = 16 * t + 4 * q;
    f32x4 base = in.base ? ld4(in.base + row * C + c0, ok) : zero4();
    if (in.t) {
      f32x4 u = ld4(in.t + row * C + c0, ok) * *(const f32x4*)(kc + 6 * C + c0) + *(const f32x4*)(kc + 7 * C + c0);
      if (in.act_gelu) u = gelu4(u);
      base += scale * u;
    }
    x[t] = base;
  }
}

template <int CT, int NW>
constexpr size_t grad_sums_bytes() { return (size_t)NW * 4 * GradSums<CT>::SLOT * sizeof(float); }

// the LDS the MLP adjoint keeps ahead of its weights: kc[8][C], K (padded to 4 floats) and the fused sums' slots
template <int CT, int NW>
constexpr size_t mlp_bwd_head_bytes() { return ((size_t)8 * CT * 16 + 4) * sizeof(float) + grad_sums_bytes<CT, NW>(); }

template <int CT, int NW>
__global__ __launch_bounds__(NW * 64) void k_gn_grad_stats(FfnoRows rows, FfnoGnBack gb, const float* __restrict__ ubar,
                                                           float* __restrict__ partials) {
  constexpr int C = CT * 16, CPG = C / G;
  using S = GradSums<CT>;
  __shared__ float red[NW * 4 * S::SLOT];
  // per-channel constants of this batch item: scale gamma, mean, rstd, and the affine before the GELU
  __shared__ float kc[5 * C];
  Lane L;
  const int b = blockIdx.y;
  const float scale = gb.scale ? *gb.scale : 1.f;
  for (int c = threadIdx.x; c < C; c += blockDim.x) {
    const float* st = gb.stats + (b * G + c / CPG) * 3;
    kc[c] = scale * gb.gamma[c];
    kc[C + c] = st[0];
    kc[2 * C + c] = st[1];
    kc[3 * C + c] = gb.act_gelu ? gb.aff_a[b * C + c] : 0.f;
    kc[4 * C + c] = gb.act_gelu ? gb.aff_b[b * C + c] : 0.f;
  }
  __syncthreads();
  const int64_t nt = (rows.voxels + 15) / 16;
  const int64_t stride = (int64_t)gridDim.x * NW;
  // per lane and channel group sums; the next tile's rows are loaded while this tile's are reduced
  float s1[CT * S::SUB], s2[CT * S::SUB];
#pragma unroll
  for (int i = 0; i < CT * S::SUB; ++i) { s1[i] = 0.f; s2[i] = 0.f; }
  f32x4 ubn[CT], tvn[CT];
  auto fetch = [&](int64_t t) {
    const int64_t vox = t * 16 + L.v;
    const bool ok = vox < rows.voxels;
    const int64_t row = (int64_t)b * rows.voxels + vox;
#pragma unroll
    for (int u = 0; u < CT; ++u) {
      ubn[u] = ld4(ubar + row * C + 16 * u + 4 * L.q, ok);
      tvn[u] = ld4(gb.t + row * C + 16 * u + 4 * L.q, ok);
    }
  };
  int64_t tile = (int64_t)blockIdx.x * NW + L.wave;
  if (tile < nt) fetch(tile);
  for (; tile < nt; tile += stride) {
    f32x4 ub[CT], tv[CT];
#pragma unroll
    for (int u = 0; u < CT; ++u) { ub[u] = ubn[u]; tv[u] = tvn[u]; }
    if (tile + stride < nt) fetch(tile + stride);
    asm volatile("" ::: "memory");        // the constants are read from LDS per tile, not hoisted into registers
#pragma unroll
    for (int t = 0; t < CT; ++t) {
      // rows outside the batch item load u_bar = t = 0, so g_hat = 0 there
      const int c0 = 16 * t + 4 * L.q;
      f32x4 gh = *(const f32x4*)(kc + c0) * ub[t];
      if (gb.act_gelu) gh *= gelu_grad4(tv[t] * *(const f32x4*)(kc + 3 * C + c0) + *(const f32x4*)(kc + 4 * C + c0));
      const f32x4 xh = (tv[t] - *(const f32x4*)(kc + C + c0)) * *(const f32x4*)(kc + 2 * C + c0);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        s1[t * S::SUB + j / (4 / S::SUB)] += gh[j];
        s2[t * S::SUB + j / (4 / S::SUB)] += gh[j] * xh[j];
      }
    }
  }
  // sum over the 16 voxel lanes, then over waves and lane groups per channel group
#pragma unroll
  for (int i = 0; i < CT * S::SUB; ++i) {
    s1[i] = row16_sum(s1[i]);
    s2[i] = row16_sum(s2[i]);
  }
  if (L.v == 0)
#pragma unroll
    for (int i = 0; i < CT * S::SUB; ++i) {
      red[(L.wave * 4 + L.q) * S::SLOT + 2 * i] = s1[i];
      red[(L.wave * 4 + L.q) * S::SLOT + 2 * i + 1] = s2[i];
    }
  S sums;
  sums.template flush<NW>(red, partials + ((int64_t)b * gridDim.x + blockIdx.x) * G * 2);
}

// 256 threads: thread t sums the partials i = t/8, t/8 + 32, ... of group t%8, then one thread per group adds the 32
// slices in order (a fixed order: the result does not depend on scheduling)
__global__ __launch_bounds__(256) void k_gn_grad_finalize(int nwg, const float* __restrict__ partials,
                                                           float* __restrict__ sums) {
  __shared__ float sl[32][G][2];
  const int b = blockIdx.x, g = threadIdx.x % G, slice = threadIdx.x / G;
  float a = 0.f, c = 0.f;
  for (int i = slice; i < nwg; i += 32) {
    const float* p = partials + (((int64_t)b * nwg + i) * G + g) * 2;
    a += p[0];
    c += p[1];
  }
  sl[slice][g][0] = a;
  sl[slice][g][1] = c;
  __syncthreads();
  if (threadIdx.x < G) {
    float ta = 0.f, tc = 0.f;
    for (int k = 0; k < 32; ++k) {
      ta += sl[k][threadIdx.x][0];
      tc += sl[k][threadIdx.x][1];
    }
    sums[(b * G + threadIdx.x) * 2 + 0] = ta;
    sums[(b * G + threadIdx.x) * 2 + 1] = tc;
  }
}

// ---------------------------------------------------------------------------------------------------------------
// MLP behind a GroupNorm, with or without the residual path
// ---------------------------------------------------------------------------------------------------------------
// WB: the GroupNorm's depth-weight gradient too (the height gradient is wanted)
template <int CT, int NW, bool LW, bool WB>
__global__ __launch_bounds__(NW * 64) void k_mlp_bwd(FfnoRows rows, const float* ubar, FfnoGnBack gb, FfnoAssemble xin,
                                                     const uint16_t* w1, const uint16_t* w2t, const uint16_t* w1t,
                                                     int residual, float* xbar, FfnoGnBack next,
                                                     float* next_partials) {
  constexpr int C = CT * 16, KS = ks_of(CT), KP = KS * 32;
  extern __shared__ __align__(16) unsigned char smem[];
  // LDS: the per-channel constants kc[8][C], K and the fused sums' slots first (within reach of the LDS instructions'
  // 16-bit immediate offsets), then the staged weights
  float* kc = reinterpret_cast<float*>(smem);
  float* ksum = kc + 8 * C;
  float* red = ksum + 4;
  bf16* lds = reinterpret_cast<bf16*>(smem + mlp_bwd_head_bytes<CT, NW>());
  const Wm W1 = LW ? stage_w(lds, w1, 2 * C * wld(KP)) : global_w(w1, 2 * C * wld(KP));
  const Wm W2T = LW ? stage_w(lds, w2t, 2 * C * wld(KP)) : global_w(w2t, 2 * C * wld(KP));
  const Wm W1T = LW ? stage_w(lds, w1t, C * wld(2 * C)) : global_w(w1t, C * wld(2 * C));
  __syncthreads();
  Lane L;
  const int b = blockIdx.y;
  const float scale = gb.scale ? *gb.scale : 1.f;
  const float xscale = xin.scale ? *xin.scale : 1.f;
  const float nscale = next_partials && next.scale ? *next.scale : 1.f;
  gn_consts<CT>(gb, scale, xin, b, kc, WB ? ksum : nullptr);
  GradSums<CT> sums;
  if (next_partials)
    sums.template init<NW>(red, L.q, L.wave);           // its barrier also publishes kc
  else
    __syncthreads();
  const int64_t nt = (rows.voxels + 15) / 16;
  for (int64_t tile = (int64_t)blockIdx.x * NW + L.wave; tile < nt; tile += (int64_t)gridDim.x * NW) {
    const int64_t vox = tile * 16 + L.v;
    const bool ok = vox < rows.voxels;
    const int64_t row = (int64_t)b * rows.voxels + vox;
    f32x4 ub[CT], acc[CT];
    load_act<CT>(ubar, row, ok, L.q, ub);
    Frag of[KS], xf[KS];
    {
      f32x4 tb[CT], x[CT];
      gn_back_k<CT, WB>(gb, kc, row, ok, L.q, ub, tb, ksum);
      assemble_k<CT>(xin, kc, xscale, row, ok, L.q, x);
      to_frags<CT>(tb, of);
      to_frags<CT>(x, xf);
    }
#pragma unroll
    for (int t = 0; t < CT; ++t) acc[t] = residual ? ub[t] : zero4();
    mlp2_bwd<KS, 2 * CT, KS, CT>(W1, W2T, W1T, xf, of, acc, L.v, L.q);
    store_act<CT>(xbar, row, ok, L.q, acc);
    if (next_partials) sums.add(next, nscale, row, b, ok, L.q, L.v, acc);
  }
  if (next_partials) sums.template flush<NW>(red, next_partials + ((int64_t)b * gridDim.x + blockIdx.x) * G * 2);
}

// ---------------------------------------------------------------------------------------------------------------
// branch fusion adjoint (weights read from global memory: seven matrices do not fit LDS together)
// ---------------------------------------------------------------------------------------------------------------
template <int CT, int NW>
__global__ __launch_bounds__(NW * 64) void k_fuse_bwd(FfnoRows rows, const float* x1bar, FfnoGnBack gf,
                                                      const float* __restrict__ hin, const float* __restrict__ sp,
                                                      const float* __restrict__ vp, const float* __restrict__ s_a,
                                                      const float* __restrict__ s_b, const float* __restrict__ v_a,
                                                      const float* __restrict__ v_b, const uint16_t* bg1,
                                                      const uint16_t* bg2, const float* __restrict__ bg2_bias,
                                                      const uint16_t* f1, const uint16_t* f2t, const uint16_t* f1t,
                                                      const uint16_t* bg2t, const uint16_t* bg1t,
                                                      float* hbar, float* __restrict__ sbar,
                                                      float* __restrict__ vbar) {
  constexpr int C = CT * 16, KS = ks_of(CT), KP = KS * 32;
  const Wm B1 = global_w(bg1, C * wld(3 * KP)), B2 = global_w(bg2, 2 * C * wld(KP));
  const Wm F1 = global_w(f1, 2 * C * wld(2 * KP)), F2T = global_w(f2t, 2 * C * wld(KP));
  const Wm F1T = global_w(f1t, 2 * C * wld(2 * C));
  const Wm B2T = global_w(bg2t, C * wld(2 * C)), B1T = global_w(bg1t, 3 * C * wld(KP));
  Lane L;
  const int b = blockIdx.y;
  const float scale = gf.scale ? *gf.scale : 1.f;
  const int64_t nt = (rows.voxels + 15) / 16;
  for (int64_t tile = (int64_t)blockIdx.x * NW + L.wave; tile < nt; tile += (int64_t)gridDim.x * NW) {
    const int64_t vox = tile * 16 + L.v;
    const bool ok = vox < rows.voxels;
    const int64_t row = (int64_t)b * rows.voxels + vox;
    f32x4 h[CT], s[CT], v[CT], gt[2 * CT];
    load_act<CT>(hin, row, ok, L.q, h);
    load_act<CT>(sp, row, ok, L.q, s);
    load_act<CT>(vp, row, ok, L.q, v);
#pragma unroll
    for (int t = 0; t < CT; ++t) {
      const int c0 = b * C + 16 * t + 4 * L.q;
      s[t] = ok ? gelu4(s[t] * *(const f32x4*)(s_a + c0) + *(const f32x4*)(s_b + c0)) : zero4();
      v[t] = ok ? gelu4(v[t] * *(const f32x4*)(v_a + c0) + *(const f32x4*)(v_b + c0)) : zero4();
    }
    {  // gates (models/ffno_model.py:525-526)
      Frag cat[3 * KS];
      to_frags<CT>(h, cat);
      to_frags<CT>(s, cat + KS);
      to_frags<CT>(v, cat + 2 * KS);
      mlp2<3 * KS, CT, 2 * CT>(B1, B2, cat, gt, L.v, L.q);
#pragma unroll
      for (int t = 0; t < 2 * CT; ++t) gt[t] = sigm4(gt[t] + *(const f32x4*)(bg2_bias + 16 * t + 4 * L.q));
    }
    // fp_bar from the GroupNorm (GELU after it) and the fusion MLP's adjoint (models/ffno_model.py:531-533)
    f32x4 mb[2 * CT];
    {
      f32x4 x1b[CT], fb[CT];
      load_act<CT>(x1bar, row, ok, L.q, x1b);
      gn_back<CT>(gf, scale, row, b, ok, L.q, x1b, fb);
      Frag fbf[KS], mm[2 * KS];
#pragma unroll
      for (int t = 0; t < CT; ++t) {
        mb[t] = 0.5f * fb[t];
        mb[CT + t] = 0.5f * fb[t];
      }
      to_frags<CT>(fb, fbf);
      f32x4 sm[CT], vm[CT];
#pragma unroll
      for (int t = 0; t < CT; ++t) {
        sm[t] = s[t] * gt[t];
        vm[t] = v[t] * gt[CT + t];
      }
      to_frags<CT>(sm, mm);
      to_frags<CT>(vm, mm + KS);
      mlp2_bwd<2 * KS, 2 * CT, KS, 2 * CT>(F1, F2T, F1T, mm, fbf, mb, L.v, L.q);
    }
    // through the gates: s_bar = sm_bar * gs, gate pre-activation gradient = sm_bar * s * gs (1 - gs), same for v
    f32x4 gp[2 * CT];
#pragma unroll
    for (int t = 0; t < CT; ++t) {
      gp[t] = mb[t] * s[t] * gt[t] * (1.f - gt[t]);
      gp[CT + t] = mb[CT + t] * v[t] * gt[CT + t] * (1.f - gt[CT + t]);
      s[t] = mb[t] * gt[t];
      v[t] = mb[CT + t] * gt[CT + t];
    }
    f32x4 cb[3 * CT];
#pragma unroll
    for (int t = 0; t < 3 * CT; ++t) cb[t] = zero4();
    {
      Frag gpf[CT], cat[3 * KS];
      to_frags<2 * CT>(gp, gpf);
      f32x4 sv[CT];
      load_act<CT>(sp, row, ok, L.q, sv);
      to_frags<CT>(h, cat);
#pragma unroll
      for (int t = 0; t < CT; ++t) {
        const int c0 = b * C + 16 * t + 4 * L.q;
        sv[t] = ok ? gelu4(sv[t] * *(const f32x4*)(s_a + c0) + *(const f32x4*)(s_b + c0)) : zero4();
      }
      to_frags<CT>(sv, cat + KS);
      load_act<CT>(vp, row, ok, L.q, sv);
#pragma unroll
      for (int t = 0; t < CT; ++t) {
        const int c0 = b * C + 16 * t + 4 * L.q;
        sv[t] = ok ? gelu4(sv[t] * *(const f32x4*)(v_a + c0) + *(const f32x4*)(v_b + c0)) : zero4();
      }
      to_frags<CT>(sv, cat + 2 * KS);
      mlp2_bwd<3 * KS, CT, CT, 3 * CT>(B1, B2T, B1T, cat, gpf, cb, L.v, L.q);
    }
    f32x4 x1b[CT];
    load_act<CT>(x1bar, row, ok, L.q, x1b);
#pragma unroll
    for (int t = 0; t < CT; ++t) {
      x1b[t] += cb[t];
      s[t] += cb[CT + t];
      v[t] += cb[2 * CT + t];
    }
    store_act<CT>(hbar, row, ok, L.q, x1b);
    store_act<CT>(sbar, row, ok, L.q, s);
    store_act<CT>(vbar, row, ok, L.q, v);
  }
}

// ---------------------------------------------------------------------------------------------------------------
// branch fusion adjoint in two LDS-staged passes, with the gates the recomputed forward stored
// ---------------------------------------------------------------------------------------------------------------
// pass 1: fp_bar from the GroupNorm (GELU after it), then the fusion MLP's adjoint with its hidden layer recomputed:
// [sm_bar, vm_bar] = 0.5 fp_bar (both halves) + F1^T (gelu'(F1 [s gs, v gv]) * F2^T fp_bar)
// WB: the fusion GroupNorm's depth-weight gradient too (the height gradient is wanted)
template <int CT, int NW, bool LW, bool WB>
__global__ __launch_bounds__(NW * 64) void k_fuse_mlp_bwd(FfnoRows rows, const float* __restrict__ x1bar, FfnoGnBack gf,
                                                          const float* __restrict__ sp, const float* __restrict__ vp,
                                                          const float* __restrict__ s_a, const float* __restrict__ s_b,
                                                          const float* __restrict__ v_a, const float* __restrict__ v_b,
                                                          const float* __restrict__ gates, const uint16_t* f1,
                                                          const uint16_t* f2t, const uint16_t* f1t,
                                                          float* __restrict__ smbar, float* __restrict__ vmbar) {
  constexpr int C = CT * 16, KS = ks_of(CT), KP = KS * 32;
  extern __shared__ __align__(16) unsigned char smem[];
  // LDS: kc[12][C] first (the GroupNorm constants, then the branch affines s_a, s_b, v_a, v_b of this batch item),
  // within reach of 16-bit immediate offsets, and K (padded to 4 floats); then the staged weights
  float* kc = reinterpret_cast<float*>(smem);
  float* ksum = kc + 12 * C;
  bf16* lds = reinterpret_cast<bf16*>(smem + (12 * C + 4) * sizeof(float));
  const Wm F1 = LW ? stage_w(lds, f1, 2 * C * wld(2 * KP)) : global_w(f1, 2 * C * wld(2 * KP));
  const Wm F1T = LW ? stage_w(lds, f1t, 2 * C * wld(2 * C)) : global_w(f1t, 2 * C * wld(2 * C));
  const Wm F2T = global_w(f2t, 2 * C * wld(KP));
  Lane L;
  const int b = blockIdx.y;
  const float scale = gf.scale ? *gf.scale : 1.f;
  gn_consts<CT>(gf, scale, FfnoAssemble{}, b, kc, WB ? ksum : nullptr);
  for (int c = threadIdx.x; c < C; c += blockDim.x) {
    kc[8 * C + c] = s_a[b * C + c];
    kc[9 * C + c] = s_b[b * C + c];
    kc[10 * C + c] = v_a[b * C + c];
    kc[11 * C + c] = v_b[b * C + c];
  }
  __syncthreads();
  const int64_t nt = (rows.voxels + 15) / 16;
  for (int64_t tile = (int64_t)blockIdx.x * NW + L.wave; tile < nt; tile += (int64_t)gridDim.x * NW) {
    const int64_t vox = tile * 16 + L.v;
    const bool ok = vox < rows.voxels;
    const int64_t row = (int64_t)b * rows.voxels + vox;
    // the constants are read from LDS per tile: hoisted out of the loop they would not fit the registers
    asm volatile("" ::: "memory");
    Frag mm[2 * KS];
    {
      f32x4 s[CT], v[CT], gt[2 * CT];
      load_act<CT>(sp, row, ok, L.q, s);
      load_act<CT>(vp, row, ok, L.q, v);
      load_act<2 * CT>(gates, row, ok, L.q, gt);
      // rows outside the batch item load zero gates, so s and v are zero there
#pragma unroll
      for (int t = 0; t < CT; ++t) {
        const int c0 = 16 * t + 4 * L.q;
        s[t] = gelu4(s[t] * *(const f32x4*)(kc + 8 * C + c0) + *(const f32x4*)(kc + 9 * C + c0)) * gt[t];
        v[t] = gelu4(v[t] * *(const f32x4*)(kc + 10 * C + c0) + *(const f32x4*)(kc + 11 * C + c0)) * gt[CT + t];
      }
      to_frags<CT>(s, mm);
      to_frags<CT>(v, mm + KS);
    }
    f32x4 mb[2 * CT];
    Frag fbf[KS];
    {
      f32x4 x1b[CT], fb[CT];
      load_act<CT>(x1bar, row, ok, L.q, x1b);
      gn_back_k<CT, WB>(gf, kc, row, ok, L.q, x1b, fb, ksum);
#pragma unroll
      for (int t = 0; t < CT; ++t) {
        mb[t] = 0.5f * fb[t];
        mb[CT + t] = 0.5f * fb[t];
      }
      to_frags<CT>(fb, fbf);
    }
    mlp2_bwd<2 * KS, 2 * CT, KS, 2 * CT>(F1, F2T, F1T, mm, fbf, mb, L.v, L.q);
    store_act<CT>(smbar, row, ok, L.q, *reinterpret_cast<f32x4(*)[CT]>(mb));
    store_act<CT>(vmbar, row, ok, L.q, *reinterpret_cast<f32x4(*)[CT]>(mb + CT));
  }
}

// pass 2: through the gates and the gate network (hidden layer recomputed):
// gp = [sm_bar s, vm_bar v] gt (1 - gt); [h_bar', s_bar', v_bar'] = B1^T (gelu'(B1 [h, s, v]) * B2^T gp);
// h_bar = x1_bar + h_bar', s_bar = sm_bar gs + s_bar', v_bar = vm_bar gv + v_bar'. In place: hbar over x1bar,
// sbar over smbar, vbar over vmbar. With part_s and part_v, also the gradient sums of the two branch GroupNorms
// (descriptors gs, gv; their u_bar are the s_bar and v_bar written here), so that no pass of their own re-reads them.
template <int CT, int NW, bool LW>
__global__ __launch_bounds__(NW * 64) void k_fuse_gates_bwd(FfnoRows rows, float* xh_bar, const float* __restrict__ h,
                                                            const float* __restrict__ sp, const float* __restrict__ vp,
                                                            const float* __restrict__ s_a, const float* __restrict__ s_b,
                                                            const float* __restrict__ v_a, const float* __restrict__ v_b,
                                                            const float* __restrict__ gates, const uint16_t* bg1,
                                                            const uint16_t* bg2t, const uint16_t* bg1t, float* sv_s,
                                                            float* sv_v, FfnoGnBack gs, float* part_s, FfnoGnBack gv,
                                                            float* part_v) {
  constexpr int C = CT * 16, KS = ks_of(CT), KP = KS * 32;
  extern __shared__ __align__(16) unsigned char smem[];
  // LDS: kc[4][C] first (s_a, s_b, v_a, v_b of this batch item), within reach of 16-bit immediate offsets; then the
  // staged weights
  float* kc = reinterpret_cast<float*>(smem);
  bf16* lds = reinterpret_cast<bf16*>(smem + 4 * C * sizeof(float));
  const Wm B1 = LW ? stage_w(lds, bg1, C * wld(3 * KP)) : global_w(bg1, C * wld(3 * KP));
  const Wm B2T = LW ? stage_w(lds, bg2t, C * wld(2 * C)) : global_w(bg2t, C * wld(2 * C));
  const Wm B1T = LW ? stage_w(lds, bg1t, 3 * C * wld(KP)) : global_w(bg1t, 3 * C * wld(KP));
  // the fused sums' slots after the weights (two GroupNorms)
  float* red_s = reinterpret_cast<float*>(lds);
  float* red_v = red_s + NW * 4 * GradSums<CT>::SLOT;
  Lane L;
  const int b = blockIdx.y;
  for (int c = threadIdx.x; c < C; c += blockDim.x) {
    kc[c] = s_a[b * C + c];
    kc[C + c] = s_b[b * C + c];
    kc[2 * C + c] = v_a[b * C + c];
    kc[3 * C + c] = v_b[b * C + c];
  }
  const bool sums = part_s != nullptr;
  const float sscale = sums && gs.scale ? *gs.scale : 1.f, vscale = sums && gv.scale ? *gv.scale : 1.f;
  GradSums<CT> sum_s, sum_v;
  if (sums) {
    sum_s.template init<NW>(red_s, L.q, L.wave);
    sum_v.template init<NW>(red_v, L.q, L.wave);      // its barrier also publishes kc
  } else {
    __syncthreads();
  }
  const int64_t nt = (rows.voxels + 15) / 16;
  for (int64_t tile = (int64_t)blockIdx.x * NW + L.wave; tile < nt; tile += (int64_t)gridDim.x * NW) {
    const int64_t vox = tile * 16 + L.v;
    const bool ok = vox < rows.voxels;
    const int64_t row = (int64_t)b * rows.voxels + vox;
    // the constants are read from LDS per tile: hoisted out of the loop they would not fit the registers
    asm volatile("" ::: "memory");
    Frag cat[3 * KS];
    f32x4 s[CT], v[CT];
    {
      f32x4 x[CT];
      load_act<CT>(h, row, ok, L.q, x);
      to_frags<CT>(x, cat);
    }
    load_act<CT>(sp, row, ok, L.q, s);
    load_act<CT>(vp, row, ok, L.q, v);
#pragma unroll
    for (int t = 0; t < CT; ++t) {
      const int c0 = 16 * t + 4 * L.q;
      s[t] = ok ? gelu4(s[t] * *(const f32x4*)(kc + c0) + *(const f32x4*)(kc + C + c0)) : zero4();
      v[t] = ok ? gelu4(v[t] * *(const f32x4*)(kc + 2 * C + c0) + *(const f32x4*)(kc + 3 * C + c0)) : zero4();
    }
    to_frags<CT>(s, cat + KS);
    to_frags<CT>(v, cat + 2 * KS);
    // cb = [h_bar', sm_bar gs, vm_bar gv], then += B1^T (...): the gate network's adjoint accumulates into it
    Frag gpf[CT];
    f32x4 cb[3 * CT];
    {
      f32x4 gt[2 * CT], gp[2 * CT];
      load_act<2 * CT>(gates, row, ok, L.q, gt);
      load_act<CT>(sv_s, row, ok, L.q, *reinterpret_cast<f32x4(*)[CT]>(cb + CT));
      load_act<CT>(sv_v, row, ok, L.q, *reinterpret_cast<f32x4(*)[CT]>(cb + 2 * CT));
#pragma unroll
      for (int t = 0; t < CT; ++t) {
        gp[t] = cb[CT + t] * s[t] * gt[t] * (1.f - gt[t]);
        gp[CT + t] = cb[2 * CT + t] * v[t] * gt[CT + t] * (1.f - gt[CT + t]);
        cb[CT + t] *= gt[t];
        cb[2 * CT + t] *= gt[CT + t];
        cb[t] = zero4();
      }
      to_frags<2 * CT>(gp, gpf);
    }
    mlp2_bwd<3 * KS, CT, CT, 3 * CT>(B1, B2T, B1T, cat, gpf, cb, L.v, L.q);
    f32x4 x1b[CT];
    load_act<CT>(xh_bar, row, ok, L.q, x1b);
#pragma unroll
    for (int t = 0; t < CT; ++t) x1b[t] += cb[t];
    store_act<CT>(xh_bar, row, ok, L.q, x1b);
    store_act<CT>(sv_s, row, ok, L.q, *reinterpret_cast<f32x4(*)[CT]>(cb + CT));
    store_act<CT>(sv_v, row, ok, L.q, *reinterpret_cast<f32x4(*)[CT]>(cb + 2 * CT));
    if (sums) {
      // the GroupNorm inputs sp, vp are re-read from cache (this tile loaded them above)
      sum_s.add(gs, sscale, row, b, ok, L.q, L.v, *reinterpret_cast<f32x4(*)[CT]>(cb + CT));
      sum_v.add(gv, vscale, row, b, ok, L.q, L.v, *reinterpret_cast<f32x4(*)[CT]>(cb + 2 * CT));
    }
  }
  if (sums) {
    const int64_t slot = ((int64_t)b * gridDim.x + blockIdx.x) * G * 2;
    sum_s.template flush<NW>(red_s, part_s + slot);
    sum_v.template flush<NW>(red_v, part_v + slot);
  }
}

// ---------------------------------------------------------------------------------------------------------------
// input-gate adjoint
// ---------------------------------------------------------------------------------------------------------------
// one gate xo = h * sigmoid(A2 gelu(A1 h)): hb += xb * sig + A1^T (gelu'(A1 h) * A2^T (xb * h * sig (1 - sig)))
template <int CT>
__device__ __forceinline__ void gate_bwd(Wm A1, Wm A2, Wm A2T, Wm A1T, const Frag* hf, const f32x4 (&h)[CT],
                                         const f32x4 (&xb)[CT], f32x4 (&hb)[CT], int v, int q) {
  constexpr int KS = ks_of(CT);
  // forward: the hidden layer's GELU and GELU' from one evaluation per element (the derivative kept for the
  // adjoint below, so A1 h is formed once), then sg = sigmoid(A2 gelu(A1 h))
  f32x4 sg[CT], gd[CT];
#pragma unroll
  for (int t = 0; t < CT; ++t) sg[t] = zero4();
#pragma unroll 1
  for (int s = 0; s < KS; ++s) {
    f32x4 p0 = tile<KS>(A1, 2 * s, hf, v, q), p1 = zero4();
    if (2 * s + 1 < CT) p1 = tile<KS>(A1, 2 * s + 1, hf, v, q);
    f32x4 g0, g1, d0, d1;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      float a, c;
      gelu_and_grad(p0[j], a, c);
      g0[j] = a;
      d0[j] = c;
      gelu_and_grad(p1[j], a, c);
      g1[j] = a;
      d1[j] = c;
    }
    // (registers indexed by the rolled loop's s live in gd through a select, not a dynamic index)
#pragma unroll
    for (int u = 0; u < KS; ++u)
      if (u == s) {
        gd[2 * u] = d0;
        if (2 * u + 1 < CT) gd[2 * u + 1] = d1;
      }
    const Frag gf = split8(g0, g1);
#pragma unroll
    for (int t = 0; t < CT; ++t) {
      const int o = (16 * t + v) * wld(KS * 32) + 32 * s + 8 * q;
      sg[t] = mfma3(*(const bf16x8*)(A2.h + o), *(const bf16x8*)(A2.l + o), gf, sg[t]);
    }
  }
  f32x4 og[CT];
#pragma unroll
  for (int t = 0; t < CT; ++t) {
    sg[t] = sigm4(sg[t]);
    hb[t] += xb[t] * sg[t];
    og[t] = xb[t] * h[t] * sg[t] * (1.f - sg[t]);
  }
  Frag ogf[KS];
  to_frags<CT>(og, ogf);
  // adjoint: hb += A1^T (gelu'(A1 h) * A2^T og)
#pragma unroll 1
  for (int s = 0; s < KS; ++s) {
    f32x4 d0 = zero4(), d1 = zero4();
#pragma unroll
    for (int u = 0; u < KS; ++u)
      if (u == s) {
        d0 = gd[2 * u];
        if (2 * u + 1 < CT) d1 = gd[2 * u + 1];
      }
    const f32x4 e0 = tile<KS>(A2T, 2 * s, ogf, v, q) * d0;
    f32x4 e1 = zero4();
    if (2 * s + 1 < CT) e1 = tile<KS>(A2T, 2 * s + 1, ogf, v, q) * d1;
    const Frag df = split8(e0, e1);
#pragma unroll
    for (int t = 0; t < CT; ++t) {
      const int o = (16 * t + v) * wld(KS * 32) + 32 * s + 8 * q;
      hb[t] = mfma3(*(const bf16x8*)(A1T.h + o), *(const bf16x8*)(A1T.l + o), df, hb[t]);
    }
  }
}

template <int CT, int NW, bool LW>
__global__ __launch_bounds__(NW * 64) void k_gates_bwd(FfnoRows rows, const float* __restrict__ hin,
                                                       const float* __restrict__ xsbar, const float* __restrict__ xvbar,
                                                       const float* hbar_in, const uint16_t* g1,
                                                       const uint16_t* g2, const uint16_t* g2t, const uint16_t* g1t,
                                                       const uint16_t* v1, const uint16_t* v2, const uint16_t* v2t,
                                                       const uint16_t* v1t, const float* __restrict__ z_proj,
                                                       float* hbar_out, float* __restrict__ zbar) {
  constexpr int C = CT * 16, KS = ks_of(CT), SZ = C * wld(KS * 32);
  extern __shared__ __align__(16) unsigned char smem[];
  bf16* lds = reinterpret_cast<bf16*>(smem);
  const Wm G1 = LW ? stage_w(lds, g1, SZ) : global_w(g1, SZ), G2 = LW ? stage_w(lds, g2, SZ) : global_w(g2, SZ);
  const Wm G2T = LW ? stage_w(lds, g2t, SZ) : global_w(g2t, SZ), G1T = LW ? stage_w(lds, g1t, SZ) : global_w(g1t, SZ);
  const Wm V1 = LW ? stage_w(lds, v1, SZ) : global_w(v1, SZ), V2 = LW ? stage_w(lds, v2, SZ) : global_w(v2, SZ);
  const Wm V2T = LW ? stage_w(lds, v2t, SZ) : global_w(v2t, SZ), V1T = LW ? stage_w(lds, v1t, SZ) : global_w(v1t, SZ);
  __syncthreads();
  Lane L;
  const int b = blockIdx.y;
  const int64_t nt = (rows.voxels + 15) / 16;
  for (int64_t tile = (int64_t)blockIdx.x * NW + L.wave; tile < nt; tile += (int64_t)gridDim.x * NW) {
    const int64_t vox = tile * 16 + L.v;
    const bool ok = vox < rows.voxels;
    const int64_t row = (int64_t)b * rows.voxels + vox;
    f32x4 h[CT], hb[CT], xb[CT];
    load_act<CT>(hin, row, ok, L.q, h);
    load_act<CT>(hbar_in, row, ok, L.q, hb);
    Frag hf[KS];
    to_frags<CT>(h, hf);
    load_act<CT>(xsbar, row, ok, L.q, xb);
    gate_bwd<CT>(G1, G2, G2T, G1T, hf, h, xb, hb, L.v, L.q);
    load_act<CT>(xvbar, row, ok, L.q, xb);
    gate_bwd<CT>(V1, V2, V2T, V1T, hf, h, xb, hb, L.v, L.q);
    float zs = 0.f;
#pragma unroll
    for (int t = 0; t < CT; ++t) {
      const f32x4 pr = xb[t] * *(const f32x4*)(z_proj + 16 * t + 4 * L.q);
      zs += pr[0] + pr[1] + pr[2] + pr[3];
    }
    store_act<CT>(hbar_out, row, ok, L.q, hb);
    zs += __shfl_xor(zs, 16);
    zs += __shfl_xor(zs, 32);
    if (zbar && ok && L.q == 0) zbar[row] += zs;
  }
}

// ---------------------------------------------------------------------------------------------------------------
// head adjoint
// ---------------------------------------------------------------------------------------------------------------

// LW: weights staged in LDS (else read from global: widths whose four matrices exceed a CU's LDS)
template <int CT, int NW, int NTH, bool LW>
__global__ __launch_bounds__(NW * 64) void k_head_bwd(FfnoRows rows, const float* hin, const uint16_t* p1,
                                                      const uint16_t* p2, const uint16_t* p2t, const uint16_t* p1t,
                                                      const float* __restrict__ p2_bias, const float* __restrict__ x_in,
                                                      FfnoHeadDesc hd, const float* __restrict__ ybar,
                                                      float* hbar, float* __restrict__ xbar_in) {
  constexpr int C = CT * 16, KS1 = ks_of(CT);
  extern __shared__ __align__(16) unsigned char smem[];
  bf16* lds = reinterpret_cast<bf16*>(smem);
  const Wm P1 = LW ? stage_w(lds, p1, 2 * C * wld(KS1 * 32)) : global_w(p1, 2 * C * wld(KS1 * 32));
  const Wm P2 = LW ? stage_w(lds, p2, NTH * 16 * wld(2 * C)) : global_w(p2, NTH * 16 * wld(2 * C));
  const Wm P2T = LW ? stage_w(lds, p2t, 2 * C * wld(32)) : global_w(p2t, 2 * C * wld(32));
  const Wm P1T = LW ? stage_w(lds, p1t, C * wld(2 * C)) : global_w(p1t, C * wld(2 * C));
  float* hs = reinterpret_cast<float*>(lds);
  __syncthreads();
  Lane L;
  const int b = blockIdx.y;
  const int64_t nt = (rows.voxels + 15) / 16;
  volatile float* mine = hs + L.wave * 16 * NTH * 16;
  int n_out = 0, n_head = 0;
  for (int s = 0; s < hd.n_species; ++s) {
    n_out += hd.levels[s];
    n_head += hd.levels[s] + 1;
  }
  for (int64_t tile = (int64_t)blockIdx.x * NW + L.wave; tile < nt; tile += (int64_t)gridDim.x * NW) {
    const int64_t vox = tile * 16 + L.v;
    const bool ok = vox < rows.voxels;
    const int64_t row = (int64_t)b * rows.voxels + vox;
    f32x4 x[CT];
    load_act<CT>(hin, row, ok, L.q, x);
    Frag xf[KS1];
    to_frags<CT>(x, xf);
    f32x4 o[NTH];
    mlp2<KS1, 2 * CT, NTH>(P1, P2, xf, o, L.v, L.q);
#pragma unroll
    for (int t = 0; t < NTH; ++t) {
      const f32x4 hvv = o[t] + *(const f32x4*)(p2_bias + 16 * t + 4 * L.q);
#pragma unroll
      for (int j = 0; j < 4; ++j) mine[L.v * (NTH * 16) + 16 * t + 4 * L.q + j] = hvv[j];
    }
    if (L.q == 0) {
      // one lane per voxel: the level-fraction head and its adjoint (same lockstep/volatile ordering as k_head)
      volatile float* hv = mine + L.v * (NTH * 16);
      float gt_ = 0.f, gne = 0.f, grho = 0.f;
      if (ok) {
        auto feat = [&](int c) { return x_in[((int64_t)b * 6 + c) * rows.voxels + vox] * hd.std_x[c] + hd.mean_x[c]; };
        level_fractions_bwd(
            hd, [&](int c) { return hv[c]; }, feat(hd.t_channel), feat(hd.ne_channel),
            [&](int o) { return ybar[((int64_t)b * n_out + o) * rows.voxels + vox]; },
            [&](int c, float val) { hv[c] = val; }, gt_, gne, grho);
        for (int i = n_head; i < NTH * 16; ++i) hv[i] = 0.f;
      } else {
        for (int i = 0; i < NTH * 16; ++i) hv[i] = 0.f;
      }
      if (ok) {
        for (int ch = 0; ch < 6; ++ch) {
          float g = 0.f;
          if (ch == hd.t_channel) g += gt_ * LN10 * hd.std_x[ch];
          if (ch == hd.ne_channel) g += gne * LN10 * hd.std_x[ch];
          if (ch == hd.rho_channel) g += grho * hd.std_x[ch];
          xbar_in[((int64_t)b * 6 + ch) * rows.voxels + vox] = g;
        }
      }
    }
    f32x4 hb[NTH];
#pragma unroll
    for (int t = 0; t < NTH; ++t)
#pragma unroll
      for (int j = 0; j < 4; ++j) hb[t][j] = mine[L.v * (NTH * 16) + 16 * t + 4 * L.q + j];
    Frag hbf[1];
    hbf[0] = split8(hb[0], NTH > 1 ? hb[NTH - 1] : zero4());
    f32x4 acc[CT];
#pragma unroll
    for (int t = 0; t < CT; ++t) acc[t] = zero4();
    mlp2_bwd<KS1, 2 * CT, 1, CT>(P1, P2T, P1T, xf, hbf, acc, L.v, L.q);
    store_act<CT>(hbar, row, ok, L.q, acc);
  }
}

// ---------------------------------------------------------------------------------------------------------------
// lift adjoint
// ---------------------------------------------------------------------------------------------------------------
template <int CT, int NW>
__global__ __launch_bounds__(NW * 64) void k_lift_bwd(FfnoRows rows, const float* __restrict__ ubar, FfnoGnBack gb,
                                                      const float* __restrict__ w_lift, float* __restrict__ xbar_in) {
  constexpr int C = CT * 16;
  __shared__ float wl[C * 6];
  for (int i = threadIdx.x; i < C * 6; i += blockDim.x) wl[i] = w_lift[i];
  __syncthreads();
  Lane L;
  const int b = blockIdx.y;
  const float scale = gb.scale ? *gb.scale : 1.f;
  const int64_t nt = (rows.voxels + 15) / 16;
  for (int64_t tile = (int64_t)blockIdx.x * NW + L.wave; tile < nt; tile += (int64_t)gridDim.x * NW) {
    const int64_t vox = tile * 16 + L.v;
    const bool ok = vox < rows.voxels;
    const int64_t row = (int64_t)b * rows.voxels + vox;
    f32x4 ub[CT], tb[CT];
    load_act<CT>(ubar, row, ok, L.q, ub);
    gn_back<CT>(gb, scale, row, b, ok, L.q, ub, tb);
    float acc[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      float a = 0.f;
#pragma unroll
      for (int t = 0; t < CT; ++t)
#pragma unroll
        for (int j = 0; j < 4; ++j) a += wl[(16 * t + 4 * L.q + j) * 6 + c] * tb[t][j];
      a += __shfl_xor(a, 16);
      a += __shfl_xor(a, 32);
      acc[c] = a;
    }
    if (ok && L.q == 0)
#pragma unroll
      for (int c = 0; c < 6; ++c) xbar_in[((int64_t)b * 6 + c) * rows.voxels + vox] += acc[c];
  }
}

// ---------------------------------------------------------------------------------------------------------------
// height gradient of the vertical branch per column of a shared height grid
// ---------------------------------------------------------------------------------------------------------------
// For one height column (b, hw), with P[d', r] = sum_c xv[d', c] X_bar[r, c] and Q[d, r] = sum_c yv_bar[d, c] Y[r, c]
// (r: the 2M rows of the vertical spectrum),
//   zbar[d] += sum_{delta = -1, 0, 1} sum_r G_delta[d, r] P[d + delta, r] + sum_r V_dz[d, r] Q[d, r],
// what the three Filon-band GEMMs and the synthesis-phase GEMM followed by their channel contractions form, without
// writing the GEMMs' D x HW x C outputs. P and Q are two-way split-bf16 MFMA products over the channels (A: 16 levels,
// B: 16 spectrum rows), each 16 x 16 tile contracted in float32 with the tables tab[k][r][d'] = G_delta[d' - delta, r]
// (k = delta + 1; zero where d' - delta leaves the column) and tab[3][r][d] = V_dz[d, r]: a lane holds rows
// d' = 4q + j of spectrum row r = v, so its table reads are one float4 each. The 16 lanes of a row are summed (DPP) and
// the sums added to the wave's LDS row for its column: for one (delta, j) the four lane groups add to four distinct
// levels, so no two lanes add to one entry at once. One wave per column.
template <int NDELTA, int CT, int RTC>
__device__ __forceinline__ void zgrad_pass(const FfnoZGrad& g, int b, int64_t hw, const float* A, const float* Bm,
                                           const float* tab, float* zw, const Lane& L) {
  constexpr int C = CT * 16, KS = ks_of(CT);
  const int DT = g.d_pad / 16, RT = g.r_pad / 16;
  for (int r0 = 0; r0 < RT; r0 += RTC) {
    // B fragments of this chunk of spectrum rows: column v = row r, k = channels 32 s + 8 q + j
    Frag bf[RTC][KS];
#pragma unroll
    for (int u = 0; u < RTC; ++u) {
      const int r = (r0 + u) * 16 + L.v;
      const bool rok = r0 + u < RT && r < g.R;
      const float* p = Bm + (((int64_t)b * g.R + (rok ? r : 0)) * g.hw + hw) * C;
#pragma unroll
      for (int s = 0; s < KS; ++s) {
        const int c0 = 32 * s + 8 * L.q;
        const bool ok = rok && c0 < C;
        bf[u][s] = split8(ld4(p + c0, ok), ld4(p + c0 + 4, ok));
      }
    }
    for (int dt = 0; dt < DT; ++dt) {
      const int d = dt * 16 + L.v;
      const bool dok = d < g.D;
      const float* pa = A + (((int64_t)b * g.D + (dok ? d : 0)) * g.hw + hw) * C;
      Frag af[KS];
#pragma unroll
      for (int s = 0; s < KS; ++s) {
        const int c0 = 32 * s + 8 * L.q;
        const bool ok = dok && c0 < C;
        af[s] = split8(ld4(pa + c0, ok), ld4(pa + c0 + 4, ok));
      }
      const int d1 = dt * 16 + 4 * L.q;             // this lane's first level d' of the tile
      f32x4 zp[NDELTA];
#pragma unroll
      for (int k = 0; k < NDELTA; ++k) zp[k] = zero4();
#pragma unroll
      for (int u = 0; u < RTC; ++u) {
        if (r0 + u < RT) {
          f32x4 acc = zero4();
#pragma unroll
          for (int s = 0; s < KS; ++s) acc = mfma3(af[s].h, af[s].l, bf[u][s], acc);
          const int r = (r0 + u) * 16 + L.v;
#pragma unroll
          for (int k = 0; k < NDELTA; ++k)
            zp[k] += *(const f32x4*)(tab + ((int64_t)k * g.r_pad + r) * g.d_pad + d1) * acc;
        }
      }
#pragma unroll
      for (int k = 0; k < NDELTA; ++k) {
        const int delta = NDELTA == 1 ? 0 : k - 1;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const float sum = row16_sum(zp[k][j]);
          if (L.v == 0) zw[d1 + j - delta] += sum;
        }
      }
    }
  }
}

template <int CT>
__global__ __launch_bounds__(256) void k_zgrad_cols(FfnoZGrad g) {
  constexpr int KS = ks_of(CT);
  constexpr int RTC = KS >= 3 ? 2 : (KS == 2 ? 4 : 8);   // spectrum-row tiles per chunk: RTC * KS fragments held
  extern __shared__ float zl[];                           // per wave: levels -1 .. d_pad
  Lane L;
  float* zw = zl + L.wave * (g.d_pad + 2) + 1;
  const int64_t cols = (int64_t)g.B * g.hw;
  for (int64_t col = (int64_t)blockIdx.x * 4 + L.wave; col < cols; col += (int64_t)gridDim.x * 4) {
    const int b = (int)(col / g.hw);
    const int64_t hw = col % g.hw;
    for (int i = L.lane; i < g.d_pad + 2; i += 64) zw[i - 1] = 0.f;
    zgrad_pass<3, CT, RTC>(g, b, hw, g.xv, g.xbar_spec, g.tables, zw, L);
    zgrad_pass<1, CT, RTC>(g, b, hw, g.yv_bar, g.y_spec, g.tables + (int64_t)3 * g.r_pad * g.d_pad, zw, L);
    for (int d = L.lane; d < g.D; d += 64) g.zbar[((int64_t)b * g.D + d) * g.hw + hw] += zw[d];
  }
}

// ---------------------------------------------------------------------------------------------------------------
// channel contraction for the height gradient
// ---------------------------------------------------------------------------------------------------------------
template <int CT>
__global__ __launch_bounds__(256) void k_dot_reduce(FfnoRows rows, int D, const float* __restrict__ a, int delta,
                                                    const float* __restrict__ r, float* __restrict__ zbar) {
  constexpr int C = CT * 16;
  Lane L;
  const int b = blockIdx.y;
  const int64_t hw = rows.voxels / D;
  const int64_t nt = (rows.voxels + 15) / 16;
  for (int64_t tile = (int64_t)blockIdx.x * 4 + L.wave; tile < nt; tile += (int64_t)gridDim.x * 4) {
    const int64_t vox = tile * 16 + L.v;
    const int64_t d = vox / hw;
    const bool ok = vox < rows.voxels && d + delta >= 0 && d + delta < D;
    const int64_t row = (int64_t)b * rows.voxels + vox;
    float s = 0.f;
#pragma unroll
    for (int t = 0; t < CT; ++t) {
      const f32x4 av = ld4(a + (row + delta * hw) * C + 16 * t + 4 * L.q, ok);
      const f32x4 rv = ld4(r + row * C + 16 * t + 4 * L.q, ok);
      const f32x4 pr = av * rv;
      s += pr[0] + pr[1] + pr[2] + pr[3];
    }
    s += __shfl_xor(s, 16);
    s += __shfl_xor(s, 32);
    if (ok && L.q == 0) zbar[row] += s;
  }
}

}  // namespace
}  // namespace ffno

using namespace ffno;

namespace {
bool gb_ok(const FfnoGnBack& g) { return g.t && g.stats && g.gamma && g.depth_w && (!g.act_gelu || (g.aff_a && g.aff_b)); }
}  // namespace

extern "C" {

int ffno_gn_grad_stats(FfnoRows rows, FfnoGnBack gn, const float* ubar, float* partials, hipStream_t stream) {
  if (!rows_ok(rows) || !gb_ok(gn) || !ubar || !partials) return fail("ffno_gn_grad_stats: bad arguments");
  dim3 grid(rows.nwg, rows.B);
  // 4-wave workgroups: three fit a CU at this kernel's register count (two 8-wave ones would not)
  FFNO_DISPATCH_CT(rows.C, (k_gn_grad_stats<CT, 4><<<grid, 4 * 64, 0, stream>>>(rows, gn, ubar, partials)));
  return check_launch("ffno_gn_grad_stats");
}

int ffno_gn_grad_finalize(int B, int nwg, const float* partials, float* sums, hipStream_t stream) {
  if (B <= 0 || nwg <= 0 || !partials || !sums) return fail("ffno_gn_grad_finalize: bad arguments");
  k_gn_grad_finalize<<<B, 256, 0, stream>>>(nwg, partials, sums);
  return check_launch("ffno_gn_grad_finalize");
}

int ffno_mlp_bwd(FfnoRows rows, const float* ubar, FfnoGnBack gn, FfnoAssemble x, const uint16_t* w1,
                 const uint16_t* w2t, const uint16_t* w1t, int residual, float* xbar, const FfnoGnBack* next,
                 float* next_partials, hipStream_t stream) {
  if (!rows_ok(rows) || !gb_ok(gn) || !gn.sums || !ubar || (!x.base && !x.t) || (x.t && (!x.aff_a || !x.aff_b)) ||
      !w1 || !w2t || !w1t || !xbar || (next_partials && (!next || !gb_ok(*next))))
    return fail("ffno_mlp_bwd: bad arguments");
  const FfnoGnBack nx = next_partials ? *next : FfnoGnBack{};
  dim3 grid(rows.nwg, rows.B);
  FFNO_DISPATCH_CT(rows.C, {
    constexpr int C = CT * 16, KP = ks_of(CT) * 32;
    const size_t red = mlp_bwd_head_bytes<CT, nw_bwd<CT>()>();
    const size_t lds = ((size_t)2 * C * wld(KP) * 2 + (size_t)C * wld(2 * C)) * 2 * 2;
    const bool lw = lds + red <= 160 * 1024;
    const size_t bytes = (lw ? lds : 0) + red;
    auto launch = [&](auto k) -> int {
      if (set_lds(k, bytes)) return -1;
      k<<<grid, nw_bwd<CT>() * 64, bytes, stream>>>(rows, ubar, gn, x, w1, w2t, w1t, residual, xbar, nx, next_partials);
      return 0;
    };
    int rc;
    if (gn.wbar)
      rc = lw ? launch(k_mlp_bwd<CT, nw_bwd<CT>(), true, true>) : launch(k_mlp_bwd<CT, nw_bwd<CT>(), false, true>);
    else
      rc = lw ? launch(k_mlp_bwd<CT, nw_bwd<CT>(), true, false>) : launch(k_mlp_bwd<CT, nw_bwd<CT>(), false, false>);
    if (rc) return rc;
  });
  return check_launch("ffno_mlp_bwd");
}

int ffno_fuse_bwd(FfnoRows rows, const float* x1bar, FfnoGnBack gnf, const float* h, const float* sp, const float* vp,
                  const float* s_a, const float* s_b, const float* v_a, const float* v_b, const uint16_t* bg1,
                  const uint16_t* bg2, const float* bg2_bias, const uint16_t* f1, const uint16_t* f2t,
                  const uint16_t* f1t, const uint16_t* bg2t, const uint16_t* bg1t, float* hbar, float* sbar,
                  float* vbar, hipStream_t stream) {
  if (!rows_ok(rows) || !gb_ok(gnf) || !gnf.sums || !x1bar || !h || !sp || !vp || !s_a || !s_b || !v_a || !v_b ||
      !bg1 || !bg2 || !bg2_bias || !f1 || !f2t || !f1t || !bg2t || !bg1t || !hbar || !sbar || !vbar)
    return fail("ffno_fuse_bwd: bad arguments");
  dim3 grid(rows.nwg, rows.B);
  FFNO_DISPATCH_CT(rows.C, (k_fuse_bwd<CT, NWV><<<grid, NWV * 64, 0, stream>>>(
                               rows, x1bar, gnf, h, sp, vp, s_a, s_b, v_a, v_b, bg1, bg2, bg2_bias, f1, f2t, f1t, bg2t,
                               bg1t, hbar, sbar, vbar)));
  return check_launch("ffno_fuse_bwd");
}

int ffno_fuse_mlp_bwd(FfnoRows rows, const float* x1bar, FfnoGnBack gnf, const float* sp, const float* vp,
                      const float* s_a, const float* s_b, const float* v_a, const float* v_b, const float* gates,
                      const uint16_t* f1, const uint16_t* f2t, const uint16_t* f1t, float* smbar, float* vmbar,
                      hipStream_t stream) {
  if (!rows_ok(rows) || !gb_ok(gnf) || !gnf.sums || !x1bar || !sp || !vp || !s_a || !s_b || !v_a || !v_b || !gates ||
      !f1 || !f2t || !f1t || !smbar || !vmbar)
    return fail("ffno_fuse_mlp_bwd: bad arguments");
  dim3 grid(rows.nwg, rows.B);
  FFNO_DISPATCH_CT(rows.C, {
    constexpr int C = CT * 16, KP = ks_of(CT) * 32;
    const size_t head = ((size_t)12 * C + 4) * sizeof(float);
    const size_t lds = ((size_t)2 * C * wld(2 * KP) + (size_t)2 * C * wld(2 * C)) * 2 * 2;
    const bool lw = lds + head <= 160 * 1024;
    const size_t bytes = (lw ? lds : 0) + head;
    auto launch = [&](auto k) -> int {
      if (set_lds(k, bytes)) return -1;
      k<<<grid, nw_of<CT>() * 64, bytes, stream>>>(rows, x1bar, gnf, sp, vp, s_a, s_b, v_a, v_b, gates, f1, f2t, f1t,
                                                   smbar, vmbar);
      return 0;
    };
    int rc;
    if (gnf.wbar)
      rc = lw ? launch(k_fuse_mlp_bwd<CT, nw_of<CT>(), true, true>) : launch(k_fuse_mlp_bwd<CT, nw_of<CT>(), false, true>);
    else
      rc = lw ? launch(k_fuse_mlp_bwd<CT, nw_of<CT>(), true, false>) : launch(k_fuse_mlp_bwd<CT, nw_of<CT>(), false, false>);
    if (rc) return rc;
  });
  return check_launch("ffno_fuse_mlp_bwd");
}

int ffno_fuse_gates_bwd(FfnoRows rows, float* x1bar_hbar, const float* h, const float* sp, const float* vp,
                        const float* s_a, const float* s_b, const float* v_a, const float* v_b, const float* gates,
                        const uint16_t* bg1, const uint16_t* bg2t, const uint16_t* bg1t, float* sm_sbar, float* vm_vbar,
                        const FfnoGnBack* gn_s, float* partials_s, const FfnoGnBack* gn_v, float* partials_v,
                        hipStream_t stream) {
  if (!rows_ok(rows) || !x1bar_hbar || !h || !sp || !vp || !s_a || !s_b || !v_a || !v_b || !gates || !bg1 || !bg2t ||
      !bg1t || !sm_sbar || !vm_vbar || (!partials_s != !partials_v) ||
      (partials_s && (!gn_s || !gn_v || !gb_ok(*gn_s) || !gb_ok(*gn_v))))
    return fail("ffno_fuse_gates_bwd: bad arguments");
  const FfnoGnBack gs = partials_s ? *gn_s : FfnoGnBack{}, gv = partials_s ? *gn_v : FfnoGnBack{};
  dim3 grid(rows.nwg, rows.B);
  FFNO_DISPATCH_CT(rows.C, {
    constexpr int C = CT * 16, KP = ks_of(CT) * 32;
    const size_t head = (size_t)4 * C * sizeof(float);
    const size_t red = partials_s ? 2 * grad_sums_bytes<CT, nw_of<CT>()>() : 0;
    const size_t lds = ((size_t)C * wld(3 * KP) + (size_t)C * wld(2 * C) + (size_t)3 * C * wld(KP)) * 2 * 2;
    const bool lw = lds + head + red <= 160 * 1024;
    const size_t bytes = (lw ? lds : 0) + head + red;
    auto launch = [&](auto k) -> int {
      if (set_lds(k, bytes)) return -1;
      k<<<grid, nw_of<CT>() * 64, bytes, stream>>>(rows, x1bar_hbar, h, sp, vp, s_a, s_b, v_a, v_b, gates, bg1, bg2t,
                                                   bg1t, sm_sbar, vm_vbar, gs, partials_s, gv, partials_v);
      return 0;
    };
    if (lw ? launch(k_fuse_gates_bwd<CT, nw_of<CT>(), true>) : launch(k_fuse_gates_bwd<CT, nw_of<CT>(), false>))
      return -1;
  });
  return check_launch("ffno_fuse_gates_bwd");
}

int ffno_gates_bwd(FfnoRows rows, const float* h, const float* xsbar, const float* xvbar, const float* hbar_in,
                   const uint16_t* g1, const uint16_t* g2, const uint16_t* g2t, const uint16_t* g1t,
                   const uint16_t* v1, const uint16_t* v2, const uint16_t* v2t, const uint16_t* v1t,
                   const float* z_proj, float* hbar_out, float* zbar, hipStream_t stream) {
  if (!rows_ok(rows) || !h || !xsbar || !xvbar || !hbar_in || !g1 || !g2 || !g2t || !g1t || !v1 || !v2 || !v2t ||
      !v1t || !z_proj || !hbar_out)
    return fail("ffno_gates_bwd: bad arguments");
  dim3 grid(rows.nwg, rows.B);
  FFNO_DISPATCH_CT(rows.C, {
    const size_t lds = 8 * (size_t)(CT * 16) * wld(ks_of(CT) * 32) * 2 * 2;
    if (lds <= 160 * 1024) {
      auto k = k_gates_bwd<CT, nw_bwd<CT>(), true>;
      if (set_lds(k, lds)) return -1;
      k<<<grid, nw_bwd<CT>() * 64, lds, stream>>>(rows, h, xsbar, xvbar, hbar_in, g1, g2, g2t, g1t, v1, v2, v2t, v1t, z_proj,
                                         hbar_out, zbar);
    } else {
      k_gates_bwd<CT, nw_bwd<CT>(), false><<<grid, nw_bwd<CT>() * 64, 0, stream>>>(rows, h, xsbar, xvbar, hbar_in, g1, g2, g2t, g1t, v1,
                                                                 v2, v2t, v1t, z_proj, hbar_out, zbar);
    }
  });
  return check_launch("ffno_gates_bwd");
}

int ffno_head_bwd(FfnoRows rows, const float* h, const uint16_t* p1, const uint16_t* p2, const uint16_t* p2t,
                  const uint16_t* p1t, const float* p2_bias, const float* x_in, const FfnoHeadDesc* desc,
                  const float* ybar, float* hbar, float* xbar_in, hipStream_t stream) {
  if (!rows_ok(rows) || !h || !p1 || !p2 || !p2t || !p1t || !p2_bias || !x_in || !desc || !ybar || !hbar || !xbar_in)
    return fail("ffno_head_bwd: bad arguments");
  int n_head = 0;
  if (desc->n_species < 1 || desc->n_species > 4) return fail("ffno_head_bwd: 1 to 4 species");
  for (int s = 0; s < desc->n_species; ++s) {
    if (desc->levels[s] < 1 || desc->levels[s] > 15) return fail("ffno_head_bwd: 1 to 15 levels per species");
    n_head += desc->levels[s] + 1;
  }
  if (n_head > 32) return fail("ffno_head_bwd: at most 32 head channels");
  const int nth = (n_head + 15) / 16;
  dim3 grid(rows.nwg, rows.B);
  FFNO_DISPATCH_CT(rows.C, {
    constexpr int C = CT * 16;
    const size_t w_bytes = ((size_t)2 * C * wld(ks_of(CT) * 32) + (size_t)nth * 16 * wld(2 * C) + (size_t)2 * C * wld(32) +
                            (size_t)C * wld(2 * C)) * 2 * 2;
    const size_t hs = (size_t)nw_of<CT>() * 16 * nth * 16 * sizeof(float);
    const bool lw = w_bytes + hs <= 160 * 1024;
    const size_t lds = (lw ? w_bytes : 0) + hs;
    auto launch = [&](auto kernel) -> int {
      if (set_lds(kernel, lds)) return -1;
      kernel<<<grid, nw_of<CT>() * 64, lds, stream>>>(rows, h, p1, p2, p2t, p1t, p2_bias, x_in, *desc, ybar, hbar,
                                                      xbar_in);
      return 0;
    };
    int rc;
    if (nth == 1)
      rc = lw ? launch(k_head_bwd<CT, nw_of<CT>(), 1, true>) : launch(k_head_bwd<CT, nw_of<CT>(), 1, false>);
    else
      rc = lw ? launch(k_head_bwd<CT, nw_of<CT>(), 2, true>) : launch(k_head_bwd<CT, nw_of<CT>(), 2, false>);
    if (rc) return rc;
  });
  return check_launch("ffno_head_bwd");
}

int ffno_zgrad_columns(const FfnoZGrad* g, hipStream_t stream) {
  if (!g || g->B <= 0 || g->hw <= 0 || g->D <= 0 || g->R <= 0 || g->C % 16 || g->d_pad < g->D || g->d_pad % 16 ||
      g->r_pad < g->R || g->r_pad % 16 || !g->xv || !g->xbar_spec || !g->yv_bar || !g->y_spec || !g->tables ||
      !g->zbar)
    return fail("ffno_zgrad_columns: bad arguments");
  const int64_t cols = (int64_t)g->B * g->hw;
  const int64_t bx = (cols + 3) / 4;
  const size_t lds = (size_t)4 * (g->d_pad + 2) * sizeof(float);
  if (lds > 160 * 1024) return fail("ffno_zgrad_columns: column too deep");
  FFNO_DISPATCH_CT(g->C, {
    auto k = k_zgrad_cols<CT>;
    if (set_lds(k, lds)) return -1;
    k<<<(unsigned)(bx < 8192 ? bx : 8192), 256, lds, stream>>>(*g);
  });
  return check_launch("ffno_zgrad_columns");
}

int ffno_dot_reduce(FfnoRows rows, int D, const float* a, int delta, const float* r, float* zbar, hipStream_t stream) {
  if (!rows_ok(rows) || D <= 0 || rows.voxels % D || !a || !r || !zbar) return fail("ffno_dot_reduce: bad arguments");
  const int64_t tiles = (rows.voxels + 15) / 16;
  const int64_t bx = (tiles + 3) / 4;
  dim3 grid((unsigned)(bx < 1024 ? bx : 1024), rows.B);
  FFNO_DISPATCH_CT(rows.C, (k_dot_reduce<CT><<<grid, 256, 0, stream>>>(rows, D, a, delta, r, zbar)));
  return check_launch("ffno_dot_reduce");
}

int ffno_lift_bwd(FfnoRows rows, const float* ubar, FfnoGnBack gn, const float* w_lift, float* xbar_in,
                  hipStream_t stream) {
  if (!rows_ok(rows) || !gb_ok(gn) || !gn.sums || !ubar || !w_lift || !xbar_in) return fail("ffno_lift_bwd: bad arguments");
  dim3 grid(rows.nwg, rows.B);
  FFNO_DISPATCH_CT(rows.C, (k_lift_bwd<CT, nw_of<CT>()><<<grid, nw_of<CT>() * 64, 0, stream>>>(rows, ubar, gn, w_lift, xbar_in)));
  return check_launch("ffno_lift_bwd");
}

}  // extern "C"
