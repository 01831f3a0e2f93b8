// FFNO3D tangent (forward-mode, JVP) kernels for gfx950: the directional derivative of every forward layer along a
// tangent of the input features, with the primal quantities each needs recomputed from, or read from, the primal
// forward of the same block. Linear layers (the spectral transforms and mix) reuse the forward kernels on the
// tangent. Products are two-way split-bf16 (the tangent's tolerance is that of the gradients, 1e-2 relative).
// Lane layout and weight packing as in ffno_device.h.
#include "ffno_device.h"
#include "ffno_head.h"

namespace ffno {
namespace {

// out = d/de W2 gelu(W1 (x + e xd)) at e = 0 = W2 (gelu'(W1 x) * (W1 xd)); in, ind: two-way fragments of x and xd
template <int KS1, int NH, int NT2>
__device__ __forceinline__ void mlp2_t(Wm W1, Wm W2, const Frag* in, const Frag* ind, f32x4 (&out)[NT2], int v,
                                       int q) {
  constexpr int KS2 = ks_of(NH);
#pragma unroll
  for (int t = 0; t < NT2; ++t) out[t] = zero4();
#pragma unroll 1
  for (int s = 0; s < KS2; ++s) {
    const f32x4 d0 = gelu_grad4(tile<KS1>(W1, 2 * s, in, v, q)) * tile<KS1>(W1, 2 * s, ind, v, q);
    f32x4 d1 = zero4();
    if (2 * s + 1 < NH) d1 = gelu_grad4(tile<KS1>(W1, 2 * s + 1, in, v, q)) * tile<KS1>(W1, 2 * s + 1, ind, v, q);
    const Frag df = split8(d0, d1);
#pragma unroll
    for (int t = 0; t < NT2; ++t) {
      const int o = (16 * t + v) * wld(KS2 * 32) + 32 * s + 8 * q;
      out[t] = mfma3(*(const bf16x8*)(W2.h + o), *(const bf16x8*)(W2.l + o), df, out[t]);
    }
  }
}

// ---------------------------------------------------------------------------------------------------------------
// GroupNorm tangent: partial sums S1 = sum w yd, S2 = sum w (y - mean) yd per (sample, group), then the per-channel
// coefficients of n_dot = a yd + tan_b y + tan_c (a = gamma rstd, the forward affine)
// ---------------------------------------------------------------------------------------------------------------
template <int CT, int NW>
__global__ __launch_bounds__(NW * 64) void k_gn_tstats(FfnoRows rows, const float* __restrict__ y,
                                                       const float* __restrict__ yd, const float* __restrict__ stats,
                                                       const float* __restrict__ depth_w, float* __restrict__ partials) {
  constexpr int C = CT * 16, CPG = C / G;
  __shared__ float red[NW][4][CT * 4][2];
  Lane L;
  const int b = blockIdx.y;
  const int64_t nt = (rows.voxels + 15) / 16;
  float s1[CT * 4], s2[CT * 4];
#pragma unroll
  for (int i = 0; i < CT * 4; ++i) { s1[i] = 0.f; s2[i] = 0.f; }
  for (int64_t tile = (int64_t)blockIdx.x * NW + L.wave; tile < nt; tile += (int64_t)gridDim.x * NW) {
    const int64_t vox = tile * 16 + L.v;
    const bool ok = vox < rows.voxels;
    const int64_t row = (int64_t)b * rows.voxels + vox;
    const float w = ok ? depth_w[row] : 0.f;
#pragma unroll
    for (int t = 0; t < CT; ++t) {
      const int c0 = 16 * t + 4 * L.q;
      const f32x4 yv = ld4(y + row * C + c0, ok), dv = ld4(yd + row * C + c0, ok);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const float mean = stats[(b * G + (c0 + j) / CPG) * 3];
        s1[t * 4 + j] += w * dv[j];
        s2[t * 4 + j] += w * (yv[j] - mean) * dv[j];
      }
    }
  }
#pragma unroll
  for (int i = 0; i < CT * 4; ++i)
#pragma unroll
    for (int off = 1; off < 16; off <<= 1) {
      s1[i] += __shfl_xor(s1[i], off);
      s2[i] += __shfl_xor(s2[i], off);
    }
  if (L.v == 0)
#pragma unroll
    for (int i = 0; i < CT * 4; ++i) {
      red[L.wave][L.q][i][0] = s1[i];
      red[L.wave][L.q][i][1] = s2[i];
    }
  __syncthreads();
  if (threadIdx.x < G) {
    const int g = threadIdx.x;
    float a = 0.f, c = 0.f;
    for (int w = 0; w < NW; ++w)
      for (int qq = 0; qq < 4; ++qq)
        for (int i = 0; i < CT * 4; ++i) {
          const int ch = 16 * (i / 4) + 4 * qq + (i % 4);
          if (ch / CPG != g) continue;
          a += red[w][qq][i][0];
          c += red[w][qq][i][1];
        }
    float* p = partials + (((int64_t)b * gridDim.x + blockIdx.x) * G + g) * 2;
    p[0] = a;
    p[1] = c;
  }
}

// The same sums accumulated by the kernel that produces yd (its output tiles in lane layout), with y read per tile
// and the group means of this batch item copied per channel into LDS (read per tile, not held in registers)
template <int CT>
struct TanSums : GroupSums<CT> {
  using GroupSums<CT>::SUB;
  using GroupSums<CT>::SLOT;
  const float* mean;                                   // [C] in LDS
  // every thread calls this once, before its first tile; red holds tan_sums_floats<CT, NW>() floats
  template <int NW>
  __device__ void init(float* red, const FfnoTanStatsOut& ts, int b, int q, int wave) {
    constexpr int C = CT * 16, CPG = C / G;
    float* m = red + NW * 4 * SLOT;
    for (int c = threadIdx.x; c < C; c += blockDim.x) m[c] = ts.stats[(b * G + c / CPG) * 3];
    mean = m;
    GroupSums<CT>::template init<NW>(red, q, wave);   // its barrier also publishes the means
  }
  __device__ __forceinline__ void add(const FfnoTanStatsOut& ts, int64_t row, bool ok, int q, int v,
                                      const f32x4 (&yd)[CT]) {
    constexpr int C = CT * 16;
    const float w = ok ? ts.depth_w[row] : 0.f;
    float a[CT * SUB], c[CT * SUB];
#pragma unroll
    for (int t = 0; t < CT; ++t) {
      const int c0 = 16 * t + 4 * q;
      const f32x4 dy = ld4(ts.y + row * C + c0, ok) - *(const f32x4*)(mean + c0);
#pragma unroll
      for (int u = 0; u < SUB; ++u) {
        a[t * SUB + u] = 0.f;
        c[t * SUB + u] = 0.f;
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const float wd = ok ? w * yd[t][j] : 0.f;
        a[t * SUB + j / (4 / SUB)] += wd;
        c[t * SUB + j / (4 / SUB)] += dy[j] * wd;
      }
    }
    this->put(a, c, v);
  }
};

template <int CT, int NW>
constexpr int tan_sums_floats() { return NW * 4 * GroupSums<CT>::SLOT + CT * 16; }

// mean_dot = S1 / W, c = S2 / W, rstd_dot = -rstd^3 c:
// n_dot = gamma (rstd (yd - mean_dot) - (y - mean) rstd^3 c) = a yd + tan_b y + tan_c
__global__ __launch_bounds__(256) void k_gn_tfinalize(int nwg, int C, const float* __restrict__ partials,
                                                      const float* __restrict__ stats, const float* __restrict__ gamma,
                                                      float* __restrict__ tan_b, float* __restrict__ tan_c) {
  __shared__ float sl[32][G][2];
  __shared__ float s1s[G], s2s[G];
  const int b = blockIdx.x;
  {
    const int g = threadIdx.x % G, slice = threadIdx.x / G;
    float a = 0.f, c = 0.f;
    for (int i = slice; i < nwg; i += 32) {
      const float* p = partials + (((int64_t)b * nwg + i) * G + g) * 2;
      a += p[0];
      c += p[1];
    }
    sl[slice][g][0] = a;
    sl[slice][g][1] = c;
  }
  __syncthreads();
  if (threadIdx.x < G) {
    float a = 0.f, c = 0.f;
    for (int k = 0; k < 32; ++k) {
      a += sl[k][threadIdx.x][0];
      c += sl[k][threadIdx.x][1];
    }
    s1s[threadIdx.x] = a;
    s2s[threadIdx.x] = c;
  }
  __syncthreads();
  const int cpg = C / G;
  for (int c = threadIdx.x; c < C; c += blockDim.x) {
    const int g = c / cpg;
    const float* st = stats + (b * G + g) * 3;
    const float mean = st[0], r = st[1], inv_w = 1.f / st[2];
    const float md = s1s[g] * inv_w, cd = s2s[g] * inv_w;
    const float r3 = r * r * r;
    tan_b[b * C + c] = -gamma[c] * r3 * cd;
    tan_c[b * C + c] = -gamma[c] * r * md + gamma[c] * r3 * cd * mean;
  }
}

// out = base_dot + scale * act'(a t + b) * (a t_dot + tan_b t + tan_c), elementwise over [N, C]
__global__ __launch_bounds__(256) void k_assemble_t(int64_t n4, int C, int64_t voxels, FfnoAssembleT a,
                                                    float* __restrict__ out) {
  const float scale = a.scale ? *a.scale : 1.f;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t e = i * 4;
    const int64_t row = e / C;
    const int c0 = (int)(e % C);
    const int b = (int)(row / voxels);
    const f32x4 tv = *(const f32x4*)(a.t + e), dv = *(const f32x4*)(a.t_dot + e);
    const f32x4 ka = *(const f32x4*)(a.aff_a + b * C + c0), kb = *(const f32x4*)(a.aff_b + b * C + c0);
    const f32x4 tb = *(const f32x4*)(a.tan_b + b * C + c0), tc = *(const f32x4*)(a.tan_c + b * C + c0);
    f32x4 n = ka * dv + tb * tv + tc;
    if (a.act_gelu) n = n * gelu_grad4(ka * tv + kb);
    f32x4 r = scale * n;
    if (a.base_dot) r += *(const f32x4*)(a.base_dot + e);
    *(f32x4*)(out + e) = r;
  }
}

// ---------------------------------------------------------------------------------------------------------------
// lift: yd = W_lift x_in_dot (linear)
// ---------------------------------------------------------------------------------------------------------------
template <int CT, int NW>
__global__ __launch_bounds__(NW * 64) void k_lift_t(FfnoRows rows, const float* __restrict__ xd_in,
                                                    const float* __restrict__ w_lift, float* __restrict__ yd) {
  constexpr int C = CT * 16;
  __shared__ float wl[C * 6];
  for (int i = threadIdx.x; i < C * 6; i += blockDim.x) wl[i] = w_lift[i];
  __syncthreads();
  Lane L;
  const int b = blockIdx.y;
  const int64_t nt = (rows.voxels + 15) / 16;
  for (int64_t tile = (int64_t)blockIdx.x * NW + L.wave; tile < nt; tile += (int64_t)gridDim.x * NW) {
    const int64_t vox = tile * 16 + L.v;
    const bool ok = vox < rows.voxels;
    const int64_t row = (int64_t)b * rows.voxels + vox;
    float xin[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) xin[c] = ok ? xd_in[((int64_t)b * 6 + c) * rows.voxels + vox] : 0.f;
    f32x4 y[CT];
#pragma unroll
    for (int t = 0; t < CT; ++t)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int ch = 16 * t + 4 * L.q + j;
        float acc = 0.f;
#pragma unroll
        for (int c = 0; c < 6; ++c) acc += wl[ch * 6 + c] * xin[c];
        y[t][j] = acc;
      }
    store_act<CT>(yd, row, ok, L.q, y);
  }
}

// ---------------------------------------------------------------------------------------------------------------
// input gates: xs = h sigmoid(G2 gelu(G1 h)), xv = h sigmoid(V2 gelu(V1 h)) + z_proj z; with g = sigmoid(a),
// xs_dot = hd g + h g (1 - g) a_dot, a_dot = G2 (gelu'(G1 h) * G1 hd) (the heights carry no tangent)
// ---------------------------------------------------------------------------------------------------------------
template <int CT, int NW, bool LW>
__global__ __launch_bounds__(NW * 64) void k_gates_t(FfnoRows rows, const float* __restrict__ h,
                                                     const float* __restrict__ hd, const uint16_t* g1,
                                                     const uint16_t* g2, const uint16_t* v1, const uint16_t* v2,
                                                     float* __restrict__ xsd, float* __restrict__ xvd) {
  constexpr int C = CT * 16, KS = ks_of(CT), SZ = C * wld(KS * 32);
  extern __shared__ __align__(16) unsigned char smem[];
  bf16* lds = reinterpret_cast<bf16*>(smem);
  const Wm G1 = LW ? stage_w(lds, g1, SZ) : global_w(g1, SZ);
  const Wm G2 = LW ? stage_w(lds, g2, SZ) : global_w(g2, SZ);
  const Wm V1 = LW ? stage_w(lds, v1, SZ) : global_w(v1, SZ);
  const Wm V2 = LW ? stage_w(lds, v2, SZ) : global_w(v2, SZ);
  __syncthreads();
  Lane L;
  const int b = blockIdx.y;
  const int64_t nt = (rows.voxels + 15) / 16;
  for (int64_t tile = (int64_t)blockIdx.x * NW + L.wave; tile < nt; tile += (int64_t)gridDim.x * NW) {
    const int64_t vox = tile * 16 + L.v;
    const bool ok = vox < rows.voxels;
    const int64_t row = (int64_t)b * rows.voxels + vox;
    f32x4 x[CT], xd[CT];
    load_act<CT>(h, row, ok, L.q, x);
    load_act<CT>(hd, row, ok, L.q, xd);
    Frag xf[KS], xdf[KS];
    to_frags<CT>(x, xf);
    to_frags<CT>(xd, xdf);
#pragma unroll 1
    for (int gate = 0; gate < 2; ++gate) {
      const Wm A1 = gate ? V1 : G1, A2 = gate ? V2 : G2;
      f32x4 a[CT], ad[CT];
      mlp2<KS, CT, CT>(A1, A2, xf, a, L.v, L.q);
      mlp2_t<KS, CT, CT>(A1, A2, xf, xdf, ad, L.v, L.q);
#pragma unroll
      for (int t = 0; t < CT; ++t) {
        const f32x4 g = sigm4(a[t]);
        a[t] = xd[t] * g + x[t] * g * (1.f - g) * ad[t];
      }
      store_act<CT>(gate ? xvd : xsd, row, ok, L.q, a);
    }
  }
}

// ---------------------------------------------------------------------------------------------------------------
// a two-layer MLP y = W2 gelu(W1 x) along a tangent: x assembled as the forward does, xd given
// ---------------------------------------------------------------------------------------------------------------
template <int CT, int NW, bool LW>
__global__ __launch_bounds__(NW * 64) void k_mlp_t(FfnoRows rows, FfnoAssemble in, const float* __restrict__ xd,
                                                   const uint16_t* w1, const uint16_t* w2, float* yd,
                                                   FfnoTanStatsOut ts) {
  constexpr int C = CT * 16, KS1 = ks_of(CT);
  __shared__ float red[tan_sums_floats<CT, NW>()];
  extern __shared__ __align__(16) unsigned char smem[];
  bf16* lds = reinterpret_cast<bf16*>(smem);
  const Wm W1 = LW ? stage_w(lds, w1, 2 * C * wld(KS1 * 32)) : global_w(w1, 2 * C * wld(KS1 * 32));
  const Wm W2 = LW ? stage_w(lds, w2, C * wld(2 * C)) : global_w(w2, C * wld(2 * C));
  Lane L;
  const int b = blockIdx.y;
  TanSums<CT> sums;
  if (ts.partials)
    sums.template init<NW>(red, ts, b, L.q, L.wave);    // its barrier also publishes the staged weights
  else
    __syncthreads();
  const float scale = in.scale ? *in.scale : 1.f;
  const int64_t nt = (rows.voxels + 15) / 16;
  for (int64_t tile = (int64_t)blockIdx.x * NW + L.wave; tile < nt; tile += (int64_t)gridDim.x * NW) {
    const int64_t vox = tile * 16 + L.v;
    const bool ok = vox < rows.voxels;
    const int64_t row = (int64_t)b * rows.voxels + vox;
    Frag xf[KS1], xdf[KS1];
    {
      f32x4 x[CT];
      assemble<CT>(in, scale, row, b, ok, L.q, x);
      to_frags<CT>(x, xf);
      load_act<CT>(xd, row, ok, L.q, x);
      to_frags<CT>(x, xdf);
    }
    f32x4 o[CT];
    mlp2_t<KS1, 2 * CT, CT>(W1, W2, xf, xdf, o, L.v, L.q);
    store_act<CT>(yd, row, ok, L.q, o);
    if (ts.partials) sums.add(ts, row, ok, L.q, L.v, o);
  }
  if (ts.partials) sums.template flush<NW>(red, ts.partials + ((int64_t)b * gridDim.x + blockIdx.x) * G * 2);
}

// ---------------------------------------------------------------------------------------------------------------
// branch fusion along a tangent (models/ffno_model.py:524-534, :541-543): s = gelu(GN_s(sp)), v = gelu(GN_v(vp)),
// g = sigmoid(BG2 gelu(BG1 [h, s, v]) + b) (read from the forward), sm = s gs, vm = v gv,
// fp = F2 gelu(F1 [sm, vm]) + 0.5 (sm + vm); tangents hd given, sd and vd formed here from the branch GroupNorms'
// tangent affines (s and gelu' from one tail evaluation); weights from global (L2)
// ---------------------------------------------------------------------------------------------------------------
// x = gelu(a t + b) and its tangent gelu'(a t + b) (a t_dot + tan_b t + tan_c) for one lane's 4 channels of tile t
// (on entry x = t, xd = t_dot); k = the assembly's per-channel constants [a, b, tan_b, tan_c][C] in LDS
template <int CT>
__device__ __forceinline__ void gn_gelu_t(const float* k, bool ok, int q, int t, f32x4& x, f32x4& xd) {
  constexpr int C = CT * 16;
  const int c0 = 16 * t + 4 * q;
  const f32x4 ka = *(const f32x4*)(k + c0), kb = *(const f32x4*)(k + C + c0);
  const f32x4 n = ka * xd + *(const f32x4*)(k + 2 * C + c0) * x + *(const f32x4*)(k + 3 * C + c0);
  const f32x4 u = x * ka + kb;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    float g, d;
    gelu_and_grad(u[j], g, d);
    x[j] = ok ? g : 0.f;
    xd[j] = ok ? n[j] * d : 0.f;
  }
}

template <int CT, int NW>
__global__ __launch_bounds__(NW * 64) __attribute__((amdgpu_waves_per_eu(CT <= 4 ? 2 : 1))) void k_fuse_t(FfnoRows rows, const float* __restrict__ h,
                                                    const float* __restrict__ hd, FfnoAssembleT st, FfnoAssembleT vt,
                                                    const float* __restrict__ gates, const uint16_t* bg1,
                                                    const uint16_t* bg2, const uint16_t* f1, const uint16_t* f2,
                                                    float* __restrict__ fpd, FfnoTanStatsOut ts) {
  constexpr int C = CT * 16, KS = ks_of(CT), KP = KS * 32;
  // the two branch assemblies' per-channel constants of this batch item, then the tangent sums' slots
  __shared__ __align__(16) float kc[8 * C];
  __shared__ __align__(16) float red[tan_sums_floats<CT, NW>()];
  extern __shared__ __align__(16) unsigned char smem[];   // the waves' stashes: [NW][4 CT][64 lanes] float4
  const Wm B1 = global_w(bg1, C * wld(3 * KP)), B2 = global_w(bg2, 2 * C * wld(KP));
  const Wm F1 = global_w(f1, 2 * C * wld(2 * KP)), F2 = global_w(f2, C * wld(2 * C));
  Lane L;
  f32x4* stash = reinterpret_cast<f32x4*>(smem) + (int64_t)L.wave * 4 * CT * 64 + L.lane;
  const int b = blockIdx.y;
  for (int c = threadIdx.x; c < C; c += blockDim.x) {
    const int e = b * C + c;
    kc[c] = st.aff_a[e];
    kc[C + c] = st.aff_b[e];
    kc[2 * C + c] = st.tan_b[e];
    kc[3 * C + c] = st.tan_c[e];
    kc[4 * C + c] = vt.aff_a[e];
    kc[5 * C + c] = vt.aff_b[e];
    kc[6 * C + c] = vt.tan_b[e];
    kc[7 * C + c] = vt.tan_c[e];
  }
  TanSums<CT> sums;
  if (ts.partials)
    sums.template init<NW>(red, ts, b, L.q, L.wave);    // its barrier also publishes kc
  else
    __syncthreads();
  const int64_t nt = (rows.voxels + 15) / 16;
  for (int64_t tile = (int64_t)blockIdx.x * NW + L.wave; tile < nt; tile += (int64_t)gridDim.x * NW) {
    const int64_t vox = tile * 16 + L.v;
    const bool ok = vox < rows.voxels;
    const int64_t row = (int64_t)b * rows.voxels + vox;
    asm volatile("" ::: "memory");        // the constants are read from LDS per tile, not hoisted into registers
    f32x4 s[CT], v[CT], ds[CT], dv[CT];
    load_act<CT>(st.t, row, ok, L.q, s);
    load_act<CT>(vt.t, row, ok, L.q, v);
    load_act<CT>(st.t_dot, row, ok, L.q, ds);
    load_act<CT>(vt.t_dot, row, ok, L.q, dv);
#pragma unroll
    for (int t = 0; t < CT; ++t) {
      gn_gelu_t<CT>(kc, ok, L.q, t, s[t], ds[t]);
      gn_gelu_t<CT>(kc + 4 * C, ok, L.q, t, v[t], dv[t]);
    }
    // gate pre-activation tangent: BG2 (gelu'(BG1 [h, s, v]) * BG1 [hd, sd, vd]); s, v, sd, vd wait in the wave's
    // LDS stash meanwhile (held in registers beside the six input fragments they would leave one wave per SIMD)
    f32x4 gd[2 * CT];
    {
      Frag cat[3 * KS], catd[3 * KS];
      f32x4 x[CT];
      load_act<CT>(h, row, ok, L.q, x);
      to_frags<CT>(x, cat);
      load_act<CT>(hd, row, ok, L.q, x);
      to_frags<CT>(x, catd);
      to_frags<CT>(s, cat + KS);
      to_frags<CT>(ds, catd + KS);
      to_frags<CT>(v, cat + 2 * KS);
      to_frags<CT>(dv, catd + 2 * KS);
#pragma unroll
      for (int t = 0; t < CT; ++t) {
        stash[t * 64] = s[t];
        stash[(CT + t) * 64] = v[t];
        stash[(2 * CT + t) * 64] = ds[t];
        stash[(3 * CT + t) * 64] = dv[t];
      }
      asm volatile("" ::: "memory");
      mlp2_t<3 * KS, CT, 2 * CT>(B1, B2, cat, catd, gd, L.v, L.q);
    }
    asm volatile("" ::: "memory");
#pragma unroll
    for (int t = 0; t < CT; ++t) {
      s[t] = stash[t * 64];
      v[t] = stash[(CT + t) * 64];
      ds[t] = stash[(2 * CT + t) * 64];
      dv[t] = stash[(3 * CT + t) * 64];
    }
    // sm_dot = sd gs + s gs (1 - gs) gd_s, likewise vm; then the primal sm, vm
#pragma unroll
    for (int t = 0; t < CT; ++t) {
      const f32x4 gs = ld4(gates + row * 2 * C + 16 * t + 4 * L.q, ok);
      const f32x4 gv = ld4(gates + row * 2 * C + C + 16 * t + 4 * L.q, ok);
      ds[t] = ds[t] * gs + s[t] * gs * (1.f - gs) * gd[t];
      dv[t] = dv[t] * gv + v[t] * gv * (1.f - gv) * gd[CT + t];
      s[t] = s[t] * gs;
      v[t] = v[t] * gv;
    }
    f32x4 o[CT];
    {
      Frag cat[2 * KS], catd[2 * KS];
      to_frags<CT>(s, cat);
      to_frags<CT>(v, cat + KS);
      to_frags<CT>(ds, catd);
      to_frags<CT>(dv, catd + KS);
      mlp2_t<2 * KS, 2 * CT, CT>(F1, F2, cat, catd, o, L.v, L.q);
    }
#pragma unroll
    for (int t = 0; t < CT; ++t) o[t] = o[t] + 0.5f * (ds[t] + dv[t]);
    store_act<CT>(fpd, row, ok, L.q, o);
    if (ts.partials) sums.add(ts, row, ok, L.q, L.v, o);
  }
  if (ts.partials) sums.template flush<NW>(red, ts.partials + ((int64_t)b * gridDim.x + blockIdx.x) * G * 2);
}

// ---------------------------------------------------------------------------------------------------------------
// head along a tangent (models/ffno_model.py:730-749, :790-798): head = P2 gelu(P1 x) + bias and its tangent; per
// species d log10 n_i = d log10 rho + (1 - tanh^2(u/6)) u_dot + (l_dot_i - sum_j p_j l_dot_j) / ln 10 with
// l_i = head_i + ln f*_i(T, ne) the logits and p their softmax; normalised by std_y
// ---------------------------------------------------------------------------------------------------------------
template <int CT, int NW, int NTH, bool LW>
__global__ __launch_bounds__(NW * 64) void k_head_t(FfnoRows rows, FfnoAssemble in, const float* __restrict__ xd,
                                                    const uint16_t* p1, const uint16_t* p2,
                                                    const float* __restrict__ p2_bias, const float* __restrict__ x_in,
                                                    const float* __restrict__ xd_in, FfnoHeadDesc hd,
                                                    float* __restrict__ yd) {
  constexpr int C = CT * 16, KS1 = ks_of(CT);
  extern __shared__ __align__(16) unsigned char smem[];
  bf16* lds = reinterpret_cast<bf16*>(smem);
  const Wm P1 = LW ? stage_w(lds, p1, 2 * C * wld(KS1 * 32)) : global_w(p1, 2 * C * wld(KS1 * 32));
  const Wm P2 = LW ? stage_w(lds, p2, NTH * 16 * wld(2 * C)) : global_w(p2, NTH * 16 * wld(2 * C));
  float* hs = reinterpret_cast<float*>(lds);   // [NW][16 voxels][2][NTH*16]: head, then its tangent
  __syncthreads();
  Lane L;
  const int b = blockIdx.y;
  const float scale = in.scale ? *in.scale : 1.f;
  const int64_t nt = (rows.voxels + 15) / 16;
  volatile float* mine = hs + L.wave * 16 * 2 * NTH * 16;
  int n_out = 0;
  for (int s = 0; s < hd.n_species; ++s) n_out += hd.levels[s];
  for (int64_t tile = (int64_t)blockIdx.x * NW + L.wave; tile < nt; tile += (int64_t)gridDim.x * NW) {
    const int64_t vox = tile * 16 + L.v;
    const bool ok = vox < rows.voxels;
    const int64_t row = (int64_t)b * rows.voxels + vox;
    Frag xf[KS1], xdf[KS1];
    {
      f32x4 x[CT];
      assemble<CT>(in, scale, row, b, ok, L.q, x);
      to_frags<CT>(x, xf);
      load_act<CT>(xd, row, ok, L.q, x);
      to_frags<CT>(x, xdf);
    }
    f32x4 o[NTH], od[NTH];
    mlp2<KS1, 2 * CT, NTH>(P1, P2, xf, o, L.v, L.q);
    mlp2_t<KS1, 2 * CT, NTH>(P1, P2, xf, xdf, od, L.v, L.q);
#pragma unroll
    for (int t = 0; t < NTH; ++t) {
      const f32x4 hv = o[t] + *(const f32x4*)(p2_bias + 16 * t + 4 * L.q);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        mine[L.v * (2 * NTH * 16) + 16 * t + 4 * L.q + j] = hv[j];
        mine[L.v * (2 * NTH * 16) + NTH * 16 + 16 * t + 4 * L.q + j] = od[t][j];
      }
    }
    if (L.q == 0 && ok) {
      // lanes of one wave run in lockstep and its LDS operations complete in order: lane q = 0 of each voxel reads
      // what lanes q = 1..3 of that voxel stored just above (volatile keeps the order), as the forward head does
      const volatile float* hv = mine + L.v * (2 * NTH * 16);
      const volatile float* hdv = hv + NTH * 16;
      auto feat = [&](int c) { return x_in[((int64_t)b * 6 + c) * rows.voxels + vox] * hd.std_x[c] + hd.mean_x[c]; };
      auto dfeat = [&](int c) { return xd_in[((int64_t)b * 6 + c) * rows.voxels + vox] * hd.std_x[c]; };
      const float dlt = dfeat(hd.t_channel) * LN10, dlne = dfeat(hd.ne_channel) * LN10;
      const float dl10rho = dfeat(hd.rho_channel);
      const Lte e = lte_of(hd, feat(hd.t_channel), feat(hd.ne_channel));
      int c = 0, k = 0;
      for (int s = 0; s < hd.n_species; ++s) {
        const int nl = hd.levels[s];
        float logit[15], dl[15];
        float mx = -3.0e38f;
        for (int i = 0; i < nl; ++i) {
          const int li = k + i;
          logit[i] = hv[c + i] + lte_logw(hd, e, li);
          dl[i] = hdv[c + i] + dlt * lte_dlt(hd, e, li) + dlne * lte_dlne(hd, e, li);
          mx = fmaxf(mx, logit[i]);
        }
        float se = 0.f;
        for (int i = 0; i < nl; ++i) se += __expf(logit[i] - mx);
        const float lse = mx + __logf(se);
        float mean_dl = 0.f;
        for (int i = 0; i < nl; ++i) mean_dl += __expf(logit[i] - lse) * dl[i];
        const float th = tanhf(hv[c + nl] / hd.total_bound);
        const float dtot = (1.f - th * th) * hdv[c + nl];
        for (int i = 0; i < nl; ++i) {
          const int o2 = k + i;
          yd[((int64_t)b * n_out + o2) * rows.voxels + vox] =
              (dl10rho + dtot + (dl[i] - mean_dl) / LN10) / hd.std_y[o2];
        }
        c += nl + 1;
        k += nl;
      }
    }
  }
}

}  // namespace
}  // namespace ffno

using namespace ffno;

extern "C" {

int ffno_gn_tangent_stats(FfnoRows rows, const float* y, const float* yd, const float* stats, const float* depth_w,
                          float* partials, hipStream_t stream) {
  if (!rows_ok(rows) || !y || !yd || !stats || !depth_w || !partials) return fail("ffno_gn_tangent_stats: bad arguments");
  dim3 grid(rows.nwg, rows.B);
  FFNO_DISPATCH_CT(rows.C, (k_gn_tstats<CT, nw_of<CT>()><<<grid, nw_of<CT>() * 64, 0, stream>>>(rows, y, yd, stats,
                                                                                             depth_w, partials)));
  return check_launch("ffno_gn_tangent_stats");
}

int ffno_gn_tangent_finalize(int B, int nwg, int C, const float* partials, const float* stats, const float* gamma,
                             float* tan_b, float* tan_c, hipStream_t stream) {
  if (B <= 0 || nwg <= 0 || C % G || !partials || !stats || !gamma || !tan_b || !tan_c)
    return fail("ffno_gn_tangent_finalize: bad arguments");
  k_gn_tfinalize<<<B, 256, 0, stream>>>(nwg, C, partials, stats, gamma, tan_b, tan_c);
  return check_launch("ffno_gn_tangent_finalize");
}

int ffno_assemble_t(FfnoRows rows, FfnoAssembleT a, float* out, hipStream_t stream) {
  if (!rows_ok(rows) || !a.t || !a.t_dot || !a.aff_a || !a.aff_b || !a.tan_b || !a.tan_c || !out)
    return fail("ffno_assemble_t: bad arguments");
  const int64_t n4 = (int64_t)rows.B * rows.voxels * rows.C / 4;
  const int64_t blocks = (n4 + 255) / 256;
  k_assemble_t<<<(int)(blocks < 8192 ? blocks : 8192), 256, 0, stream>>>(n4, rows.C, rows.voxels, a, out);
  return check_launch("ffno_assemble_t");
}

int ffno_lift_t(FfnoRows rows, const float* xd_in, const float* w_lift, float* yd, hipStream_t stream) {
  if (!rows_ok(rows) || !xd_in || !w_lift || !yd) return fail("ffno_lift_t: bad arguments");
  dim3 grid(rows.nwg, rows.B);
  FFNO_DISPATCH_CT(rows.C, (k_lift_t<CT, nw_of<CT>()><<<grid, nw_of<CT>() * 64, 0, stream>>>(rows, xd_in, w_lift, yd)));
  return check_launch("ffno_lift_t");
}

int ffno_gates_t(FfnoRows rows, const float* h, const float* hd, const uint16_t* g1, const uint16_t* g2,
                 const uint16_t* v1, const uint16_t* v2, float* xsd, float* xvd, hipStream_t stream) {
  if (!rows_ok(rows) || !h || !hd || !g1 || !g2 || !v1 || !v2 || !xsd || !xvd) return fail("ffno_gates_t: bad arguments");
  dim3 grid(rows.nwg, rows.B);
  FFNO_DISPATCH_CT(rows.C, {
    const size_t lds = 4 * (size_t)(CT * 16) * wld(ks_of(CT) * 32) * 2 * 2;
    if (lds <= 160 * 1024) {
      auto k = k_gates_t<CT, nw_of<CT>(), true>;
      if (set_lds(k, lds)) return -1;
      k<<<grid, nw_of<CT>() * 64, lds, stream>>>(rows, h, hd, g1, g2, v1, v2, xsd, xvd);
    } else {
      k_gates_t<CT, nw_of<CT>(), false><<<grid, nw_of<CT>() * 64, 0, stream>>>(rows, h, hd, g1, g2, v1, v2, xsd, xvd);
    }
  });
  return check_launch("ffno_gates_t");
}

int ffno_mlp_t(FfnoRows rows, FfnoAssemble in, const float* xd, const uint16_t* w1, const uint16_t* w2, float* yd,
               FfnoTanStatsOut ts, hipStream_t stream) {
  if (!rows_ok(rows) || (!in.base && !in.t) || (in.t && (!in.aff_a || !in.aff_b)) || !xd || !w1 || !w2 || !yd ||
      (ts.partials && (!ts.y || !ts.stats || !ts.depth_w)))
    return fail("ffno_mlp_t: bad arguments");
  dim3 grid(rows.nwg, rows.B);
  FFNO_DISPATCH_CT(rows.C, {
    constexpr int NW = nw_of<CT>();
    const size_t lds = ((size_t)2 * CT * 16 * wld(ks_of(CT) * 32) + (size_t)CT * 16 * wld(2 * CT * 16)) * 2 * 2;
    if (lds + tan_sums_floats<CT, NW>() * sizeof(float) <= 160 * 1024) {
      auto k = k_mlp_t<CT, NW, true>;
      if (set_lds(k, lds)) return -1;
      k<<<grid, NW * 64, lds, stream>>>(rows, in, xd, w1, w2, yd, ts);
    } else {
      k_mlp_t<CT, NW, false><<<grid, NW * 64, 0, stream>>>(rows, in, xd, w1, w2, yd, ts);
    }
  });
  return check_launch("ffno_mlp_t");
}

static bool assemble_t_ok(const FfnoAssembleT& e) {
  return e.t && e.t_dot && e.aff_a && e.aff_b && e.tan_b && e.tan_c && e.act_gelu == 1 && !e.base_dot && !e.scale;
}

int ffno_fuse_t(FfnoRows rows, const float* h, const float* hd, FfnoAssembleT st, FfnoAssembleT vt, const float* gates,
                const uint16_t* bg1, const uint16_t* bg2, const uint16_t* f1, const uint16_t* f2, float* fpd,
                FfnoTanStatsOut ts, hipStream_t stream) {
  if (!rows_ok(rows) || !h || !hd || !assemble_t_ok(st) || !assemble_t_ok(vt) || !gates || !bg1 || !bg2 || !f1 ||
      !f2 || !fpd || (ts.partials && (!ts.y || !ts.stats || !ts.depth_w)))
    return fail("ffno_fuse_t: bad arguments");
  dim3 grid(rows.nwg, rows.B);
  FFNO_DISPATCH_CT(rows.C, {
    const size_t lds = (size_t)NWV * 4 * CT * 64 * sizeof(f32x4);
    auto k = k_fuse_t<CT, NWV>;
    if (set_lds(k, lds)) return -1;
    k<<<grid, NWV * 64, lds, stream>>>(rows, h, hd, st, vt, gates, bg1, bg2, f1, f2, fpd, ts);
  });
  return check_launch("ffno_fuse_t");
}

int ffno_head_t(FfnoRows rows, FfnoAssemble in, const float* xd, const uint16_t* p1, const uint16_t* p2,
                const float* p2_bias, const float* x_in, const float* xd_in, const FfnoHeadDesc* desc, float* yd,
                hipStream_t stream) {
  if (!rows_ok(rows) || (!in.base && !in.t) || !xd || !p1 || !p2 || !p2_bias || !x_in || !xd_in || !desc || !yd)
    return fail("ffno_head_t: bad arguments");
  int n_head = 0;
  if (desc->n_species < 1 || desc->n_species > 4) return fail("ffno_head_t: 1 to 4 species");
  for (int s = 0; s < desc->n_species; ++s) {
    if (desc->levels[s] < 1 || desc->levels[s] > 15) return fail("ffno_head_t: 1 to 15 levels per species");
    n_head += desc->levels[s] + 1;
  }
  if (n_head > 32) return fail("ffno_head_t: at most 32 head channels");
  const int nth = (n_head + 15) / 16;
  dim3 grid(rows.nwg, rows.B);
  FFNO_DISPATCH_CT(rows.C, {
    const size_t w_bytes = ((size_t)2 * CT * 16 * wld(ks_of(CT) * 32) + (size_t)nth * 16 * wld(2 * CT * 16)) * 2 * 2;
    const size_t hs = (size_t)nw_of<CT>() * 16 * 2 * nth * 16 * sizeof(float);
    const bool lw = w_bytes + hs <= 160 * 1024;
    const size_t lds = (lw ? w_bytes : 0) + hs;
    auto launch = [&](auto kernel) -> int {
      if (set_lds(kernel, lds)) return -1;
      kernel<<<grid, nw_of<CT>() * 64, lds, stream>>>(rows, in, xd, p1, p2, p2_bias, x_in, xd_in, *desc, yd);
      return 0;
    };
    int rc;
    if (nth == 1)
      rc = lw ? launch(k_head_t<CT, nw_of<CT>(), 1, true>) : launch(k_head_t<CT, nw_of<CT>(), 1, false>);
    else
      rc = lw ? launch(k_head_t<CT, nw_of<CT>(), 2, true>) : launch(k_head_t<CT, nw_of<CT>(), 2, false>);
    if (rc) return rc;
  });
  return check_launch("ffno_head_t");
}

}  // extern "C"
