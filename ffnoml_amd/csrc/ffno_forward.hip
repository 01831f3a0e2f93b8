// FFNO3D forward kernels for gfx950 (MI355X): register-chained split-bf16 MFMA MLPs with fused GroupNorm
// statistics, the split-bf16 twiddle GEMM for the DFTs and the vertical Filon transform, and the per-mode
// spectral channel mix. Conventions are in ffno_device.h.
#include <mutex>

#include "ffno_device.h"
#include "ffno_head.h"

// Partial GroupNorm statistics of the allocation-free entry point ffno_post_gn_noalloc: [B][nwg][8][3] floats of one
// call, at most FFNO_NOALLOC_MAX_SLOTS (B * nwg) slots (384 KiB), held by every GPU the library runs on until the
// process ends. Host code never uses this name as an address: the entry point looks up each device's own copy with
// hipGetSymbolAddress once and passes that pointer to the kernels.
static __device__ __align__(256) float scratch[FFNO_NOALLOC_MAX_SLOTS * 8 * 3];

namespace ffno {

thread_local char g_err[512] = "";

int fail(const char* msg) {
  snprintf(g_err, sizeof(g_err), "%s", msg);
  return -1;
}

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    snprintf(g_err, sizeof(g_err), "%s: %s", what, hipGetErrorString(e));
    return -2;
  }
  return 0;
}

namespace {

// ---------------------------------------------------------------------------------------------------------------
// lift: Conv3d 6 -> C (fp32) + GroupNorm statistics
// ---------------------------------------------------------------------------------------------------------------
template <int CT, int NW>
__global__ __launch_bounds__(NW * 64) void k_lift(FfnoRows rows, const float* __restrict__ x_in,
                                                  const float* __restrict__ w_lift, float* __restrict__ out,
                                                  FfnoStatsOut st) {
  constexpr int C = CT * 16;
  __shared__ float wl[C * 6];
  __shared__ float red[NW * 4 * CT * 2 * 3];
  for (int i = threadIdx.x; i < C * 6; i += blockDim.x) wl[i] = w_lift[i];
  __syncthreads();
  Lane L;
  const int b = blockIdx.y;
  const int64_t nt = (rows.voxels + 15) / 16;
  GnAcc<CT> acc;
  acc.init();
  for (int64_t tile = (int64_t)blockIdx.x * NW + L.wave; tile < nt; tile += (int64_t)gridDim.x * NW) {
    const int64_t vox = tile * 16 + L.v;
    const bool ok = vox < rows.voxels;
    const int64_t row = (int64_t)b * rows.voxels + vox;
    float xin[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) xin[c] = ok ? x_in[((int64_t)b * 6 + c) * rows.voxels + vox] : 0.f;
    f32x4 y[CT];
#pragma unroll
    for (int t = 0; t < CT; ++t)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int ch = 16 * t + 4 * L.q + j;
        float a = 0.f;
#pragma unroll
        for (int c = 0; c < 6; ++c) a += wl[ch * 6 + c] * xin[c];
        y[t][j] = a;
      }
    store_act<CT>(out, row, ok, L.q, y);
    if (st.partials) acc.add(y, ok ? st.depth_w[row] : 0.f, ok);
  }
  if (st.partials)
    acc.template flush<NW>(red, st.partials + ((int64_t)b * gridDim.x + blockIdx.x) * G * 3, L.v, L.q, L.wave);
}

// ---------------------------------------------------------------------------------------------------------------
// GroupNorm finalize
// ---------------------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_gn_finalize(int nwg, int C, const float* __restrict__ partials, const float* __restrict__ gamma,
                              const float* __restrict__ beta, float* __restrict__ stats, float* __restrict__ aff_a,
                              float* __restrict__ aff_b) {
  // 256 threads: thread t merges the partials i = t/8, t/8 + 32, ... of group t%8, then one thread per group merges
  // the 32 slices in order (fixed order: deterministic)
  __shared__ float mean_s[G], rstd_s[G];
  __shared__ float sl[32][G][3];
  const int b = blockIdx.x;
  {
    const int g = threadIdx.x % G, slice = threadIdx.x / G;
    float W = 0.f, m = 0.f, M2 = 0.f;
    for (int i = slice; i < nwg; i += 32) {
      const float* p = partials + (((int64_t)b * nwg + i) * G + g) * 3;
      chan(W, m, M2, p[0], p[1], p[2]);
    }
    sl[slice][g][0] = W;
    sl[slice][g][1] = m;
    sl[slice][g][2] = M2;
  }
  __syncthreads();
  if (threadIdx.x < G) {
    const int g = threadIdx.x;
    float W = 0.f, m = 0.f, M2 = 0.f;
    for (int k = 0; k < 32; ++k) chan(W, m, M2, sl[k][g][0], sl[k][g][1], sl[k][g][2]);
    float var = W > 0.f ? fmaxf(M2 / W, 0.f) : 0.f;
    float r = rsqrtf(var + 1e-5f);
    mean_s[g] = m;
    rstd_s[g] = r;
    if (stats) {
      stats[(b * G + g) * 3 + 0] = m;
      stats[(b * G + g) * 3 + 1] = r;
      stats[(b * G + g) * 3 + 2] = W;
    }
  }
  __syncthreads();
  const int cpg = C / G;
  for (int c = threadIdx.x; c < C; c += blockDim.x) {
    const int g = c / cpg;
    const float a = rstd_s[g] * gamma[c];
    aff_a[b * C + c] = a;
    aff_b[b * C + c] = beta[c] - mean_s[g] * a;
  }
}

// ---------------------------------------------------------------------------------------------------------------
// block entry: residual assembly + both input gates
// ---------------------------------------------------------------------------------------------------------------
template <int CT, int NW, int NL>
__global__ __launch_bounds__(NW * 64) void k_block_in(FfnoRows rows, FfnoAssemble in, float* x_out,
                                                      const uint16_t* g1, const uint16_t* g2, const uint16_t* v1,
                                                      const uint16_t* v2, const float* __restrict__ z_proj,
                                                      const float* __restrict__ z, float* xs, float* xv) {
  constexpr int C = CT * 16, KS = ks_of(CT), SZ = C * wld(KS * 32);
  extern __shared__ __align__(16) unsigned char smem[];
  // LDS: the assembly's affine kc[2][C] first (within reach of 16-bit immediate offsets), then the weights
  float* kc = reinterpret_cast<float*>(smem);
  bf16* lds = reinterpret_cast<bf16*>(smem + 2 * C * sizeof(float));
  const Wm3 G1 = stage_w3<NL>(lds, g1, SZ), G2 = stage_w3<NL>(lds, g2, SZ);
  const Wm3 V1 = stage_w3<NL>(lds, v1, SZ), V2 = stage_w3<NL>(lds, v2, SZ);
  Lane L;
  const int b = blockIdx.y;
  stage_affine<CT>(in, b, kc);
  __syncthreads();
  const float scale = in.scale ? *in.scale : 1.f;
  const int64_t nt = (rows.voxels + 15) / 16;
  for (int64_t tile = (int64_t)blockIdx.x * NW + L.wave; tile < nt; tile += (int64_t)gridDim.x * NW) {
    const int64_t vox = tile * 16 + L.v;
    const bool ok = vox < rows.voxels;
    const int64_t row = (int64_t)b * rows.voxels + vox;
    // the constants are read from LDS per tile: hoisted out of the loop they would not fit the registers
    asm volatile("" ::: "memory");
    f32x4 x[CT];
    assemble_c<CT>(in, kc, scale, row, ok, L.q, x);
    if (x_out) store_act<CT>(x_out, row, ok, L.q, x);
    Frag3 xf[KS];
    to_frags<CT>(x, xf);
    f32x4 g[CT];
    // horizontal gate (models/ffno_model.py:439-440); xs NULL: the vertical branch's input alone
    if (xs) {
      mlp2<KS, CT, CT>(G1, G2, xf, g, L.v, L.q);
#pragma unroll
      for (int t = 0; t < CT; ++t) g[t] = x[t] * sigm4(g[t]);
      store_act<CT>(xs, row, ok, L.q, g);
    }
    // vertical gate and height projection (models/ffno_model.py:266)
    mlp2<KS, CT, CT>(V1, V2, xf, g, L.v, L.q);
    const float zv = ok ? z[row] : 0.f;
#pragma unroll
    for (int t = 0; t < CT; ++t) {
      const f32x4 zp = *(const f32x4*)(z_proj + 16 * t + 4 * L.q);
      g[t] = x[t] * sigm4(g[t]) + zp * zv;
    }
    store_act<CT>(xv, row, ok, L.q, g);
  }
}

// ---------------------------------------------------------------------------------------------------------------
// post MLP C -> 2C -> C with statistics
// ---------------------------------------------------------------------------------------------------------------
template <int CT, int NW, int NL>
__global__ __launch_bounds__(NW * 64) void k_post(FfnoRows rows, const float* u, const uint16_t* w1,
                                                  const uint16_t* w2, float* y, FfnoStatsOut st) {
  constexpr int C = CT * 16, KS1 = ks_of(CT);
  extern __shared__ __align__(16) unsigned char smem[];
  bf16* lds = reinterpret_cast<bf16*>(smem);
  const Wm3 W1 = stage_w3<NL>(lds, w1, 2 * C * wld(KS1 * 32)), W2 = stage_w3<NL>(lds, w2, C * wld(2 * C));
  float* red = reinterpret_cast<float*>(lds);
  __syncthreads();
  Lane L;
  const int b = blockIdx.y;
  const int64_t nt = (rows.voxels + 15) / 16;
  GnAcc<CT> acc;
  acc.init();
  // two voxel tiles (2p, 2p + 1) per wave and step: every weight fragment read serves both
  const int64_t np = (nt + 1) / 2;
  for (int64_t p = (int64_t)blockIdx.x * NW + L.wave; p < np; p += (int64_t)gridDim.x * NW) {
    const int64_t vox0 = p * 32 + L.v, vox1 = vox0 + 16;
    const bool ok0 = vox0 < rows.voxels, ok1 = vox1 < rows.voxels;
    const int64_t row0 = (int64_t)b * rows.voxels + vox0, row1 = row0 + 16;
    Frag3 xf0[KS1], xf1[KS1];
    {
      f32x4 x[CT];
      load_act<CT>(u, row0, ok0, L.q, x);
      to_frags<CT>(x, xf0);
      load_act<CT>(u, row1, ok1, L.q, x);
      to_frags<CT>(x, xf1);
    }
    f32x4 o0[CT], o1[CT];
    mlp2x2<KS1, 2 * CT, CT>(W1, W2, xf0, xf1, o0, o1, L.v, L.q);
    store_act<CT>(y, row0, ok0, L.q, o0);
    store_act<CT>(y, row1, ok1, L.q, o1);
    if (st.partials) {
      acc.add(o0, ok0 ? st.depth_w[row0] : 0.f, ok0);
      acc.add(o1, ok1 ? st.depth_w[row1] : 0.f, ok1);
    }
  }
  if (st.partials)
    acc.template flush<NW>(red, st.partials + ((int64_t)b * gridDim.x + blockIdx.x) * G * 3, L.v, L.q, L.wave);
}

// ---------------------------------------------------------------------------------------------------------------
// branch fusion; LW: weights in LDS (all but the lo half of the first branch-gate layer), else all from global
// ---------------------------------------------------------------------------------------------------------------
template <int CT, int NW, bool LW>
__global__ __launch_bounds__(NW * 64) void k_fuse(FfnoRows rows, const float* __restrict__ xin,
                                                  const float* sp, const float* __restrict__ vp,
                                                  const float* __restrict__ s_a, const float* __restrict__ s_b,
                                                  const float* __restrict__ v_a, const float* __restrict__ v_b,
                                                  const uint16_t* bg1, const uint16_t* bg2,
                                                  const float* __restrict__ bg2_bias, const uint16_t* f1,
                                                  const uint16_t* f2, float* fp, FfnoStatsOut st,
                                                  float* __restrict__ gates_out) {
  constexpr int C = CT * 16, KS = ks_of(CT), KP = KS * 32;
  constexpr int N1 = C * wld(3 * KP), N2 = 2 * C * wld(KP), N3 = 2 * C * wld(2 * KP), N4 = C * wld(2 * C);
  extern __shared__ __align__(16) unsigned char smem[];
  // LDS: kc[6][C] first (s_a, s_b, v_a, v_b of this batch item, then the gate bias [2C]), within reach of 16-bit
  // immediate offsets; then the staged weights and the statistics scratch
  float* kc = reinterpret_cast<float*>(smem);
  bf16* lds = reinterpret_cast<bf16*>(smem + 6 * C * sizeof(float));
  Wm B1, B2, F1, F2;
  if (LW) {
    // B1 and F2: hi plane staged, lo plane read from global memory (everything staged would pass 160 KB)
    stage(lds, bg1, N1);
    B1 = Wm{lds, reinterpret_cast<const bf16*>(bg1) + N1};
    lds += N1;
    B2 = stage_w(lds, bg2, N2);
    F1 = stage_w(lds, f1, N3);
    stage(lds, f2, N4);
    F2 = Wm{lds, reinterpret_cast<const bf16*>(f2) + N4};
    lds += N4;
  } else {
    B1 = global_w(bg1, N1);
    B2 = global_w(bg2, N2);
    F1 = global_w(f1, N3);
    F2 = global_w(f2, N4);
  }
  float* red = reinterpret_cast<float*>(lds);
  Lane L;
  const int b = blockIdx.y;
  for (int c = threadIdx.x; c < C; c += blockDim.x) {
    kc[c] = s_a[b * C + c];
    kc[C + c] = s_b[b * C + c];
    kc[2 * C + c] = v_a[b * C + c];
    kc[3 * C + c] = v_b[b * C + c];
    kc[4 * C + c] = bg2_bias[c];
    kc[5 * C + c] = bg2_bias[C + c];
  }
  __syncthreads();
  const int64_t nt = (rows.voxels + 15) / 16;
  GnAcc<CT> acc;
  acc.init();
  for (int64_t tile = (int64_t)blockIdx.x * NW + L.wave; tile < nt; tile += (int64_t)gridDim.x * NW) {
    const int64_t vox = tile * 16 + L.v;
    const bool ok = vox < rows.voxels;
    const int64_t row = (int64_t)b * rows.voxels + vox;
    // the constants are read from LDS per tile: hoisted out of the loop they would not fit the registers
    asm volatile("" ::: "memory");
    f32x4 s[CT], v[CT];
    Frag cat[3 * KS];
    {
      f32x4 x[CT];
      load_act<CT>(xin, row, ok, L.q, x);
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
    f32x4 g[2 * CT];
    mlp2<3 * KS, CT, 2 * CT>(B1, B2, cat, g, L.v, L.q);
#pragma unroll
    for (int t = 0; t < 2 * CT; ++t) g[t] = sigm4(g[t] + *(const f32x4*)(kc + 4 * C + 16 * t + 4 * L.q));
    if (gates_out) store_act<2 * CT>(gates_out, row, ok, L.q, g);
#pragma unroll
    for (int t = 0; t < CT; ++t) {
      s[t] = s[t] * g[t];
      v[t] = v[t] * g[CT + t];
    }
    to_frags<CT>(s, cat);
    to_frags<CT>(v, cat + KS);
    f32x4 o[CT];
    mlp2<2 * KS, 2 * CT, CT>(F1, F2, cat, o, L.v, L.q);
#pragma unroll
    for (int t = 0; t < CT; ++t) o[t] = o[t] + 0.5f * (s[t] + v[t]);
    store_act<CT>(fp, row, ok, L.q, o);
    if (st.partials) acc.add(o, ok ? st.depth_w[row] : 0.f, ok);
  }
  if (st.partials)
    acc.template flush<NW>(red, st.partials + ((int64_t)b * gridDim.x + blockIdx.x) * G * 3, L.v, L.q, L.wave);
}

// ---------------------------------------------------------------------------------------------------------------
// residual pointwise path
// ---------------------------------------------------------------------------------------------------------------
template <int CT, int NW, int NL>
__global__ __launch_bounds__(NW * 64) void k_resmlp(FfnoRows rows, FfnoAssemble in, float* x_out,
                                                    const uint16_t* w1, const uint16_t* w2, float* y,
                                                    FfnoStatsOut st) {
  constexpr int C = CT * 16, KS1 = ks_of(CT);
  extern __shared__ __align__(16) unsigned char smem[];
  // LDS: the assembly's affine kc[2][C] first (within reach of 16-bit immediate offsets), then the weights and the
  // statistics scratch
  float* kc = reinterpret_cast<float*>(smem);
  bf16* lds = reinterpret_cast<bf16*>(smem + 2 * C * sizeof(float));
  const Wm3 W1 = stage_w3<NL>(lds, w1, 2 * C * wld(KS1 * 32)), W2 = stage_w3<NL>(lds, w2, C * wld(2 * C));
  float* red = reinterpret_cast<float*>(lds);
  Lane L;
  const int b = blockIdx.y;
  stage_affine<CT>(in, b, kc);
  __syncthreads();
  const float scale = in.scale ? *in.scale : 1.f;
  const int64_t nt = (rows.voxels + 15) / 16;
  GnAcc<CT> acc;
  acc.init();
  // two voxel tiles (2p, 2p + 1) per wave and step: every weight fragment read serves both
  const int64_t np = (nt + 1) / 2;
  for (int64_t p = (int64_t)blockIdx.x * NW + L.wave; p < np; p += (int64_t)gridDim.x * NW) {
    const int64_t vox0 = p * 32 + L.v, vox1 = vox0 + 16;
    const bool ok0 = vox0 < rows.voxels, ok1 = vox1 < rows.voxels;
    const int64_t row0 = (int64_t)b * rows.voxels + vox0, row1 = row0 + 16;
    // the constants are read from LDS per step: hoisted out of the loop they would not fit the registers
    asm volatile("" ::: "memory");
    Frag3 xf0[KS1], xf1[KS1];
    {
      f32x4 x[CT];
      assemble_c<CT>(in, kc, scale, row0, ok0, L.q, x);
      if (x_out) store_act<CT>(x_out, row0, ok0, L.q, x);
      to_frags<CT>(x, xf0);
      assemble_c<CT>(in, kc, scale, row1, ok1, L.q, x);
      if (x_out) store_act<CT>(x_out, row1, ok1, L.q, x);
      to_frags<CT>(x, xf1);
    }
    f32x4 o0[CT], o1[CT];
    mlp2x2<KS1, 2 * CT, CT>(W1, W2, xf0, xf1, o0, o1, L.v, L.q);
    store_act<CT>(y, row0, ok0, L.q, o0);
    store_act<CT>(y, row1, ok1, L.q, o1);
    if (st.partials) {
      acc.add(o0, ok0 ? st.depth_w[row0] : 0.f, ok0);
      acc.add(o1, ok1 ? st.depth_w[row1] : 0.f, ok1);
    }
  }
  if (st.partials)
    acc.template flush<NW>(red, st.partials + ((int64_t)b * gridDim.x + blockIdx.x) * G * 3, L.v, L.q, L.wave);
}

// ---------------------------------------------------------------------------------------------------------------
// head: projection MLP + level-fraction output, channel-first store
// ---------------------------------------------------------------------------------------------------------------
template <int CT, int NW, int NTH, int NL>
__global__ __launch_bounds__(NW * 64) void k_head(FfnoRows rows, FfnoAssemble in, float* x_out,
                                                  const uint16_t* p1,
                                                  const uint16_t* p2, const float* __restrict__ p2_bias,
                                                  const float* __restrict__ x_in, FfnoHeadDesc hd,
                                                  float* __restrict__ y) {
  constexpr int C = CT * 16, KS1 = ks_of(CT);
  extern __shared__ __align__(16) unsigned char smem[];
  bf16* lds = reinterpret_cast<bf16*>(smem);
  const Wm3 P1 = stage_w3<NL>(lds, p1, 2 * C * wld(KS1 * 32)), P2 = stage_w3<NL>(lds, p2, NTH * 16 * wld(2 * C));
  float* hs = reinterpret_cast<float*>(lds);   // [NW][16 voxels][NTH*16]
  __syncthreads();
  Lane L;
  const int b = blockIdx.y;
  const float scale = in.scale ? *in.scale : 1.f;
  const int64_t nt = (rows.voxels + 15) / 16;
  volatile float* mine = hs + L.wave * 16 * NTH * 16;
  int n_out = 0;
  for (int s = 0; s < hd.n_species; ++s) n_out += hd.levels[s];
  for (int64_t tile = (int64_t)blockIdx.x * NW + L.wave; tile < nt; tile += (int64_t)gridDim.x * NW) {
    const int64_t vox = tile * 16 + L.v;
    const bool ok = vox < rows.voxels;
    const int64_t row = (int64_t)b * rows.voxels + vox;
    f32x4 x[CT];
    assemble<CT>(in, scale, row, b, ok, L.q, x);
    if (x_out) store_act<CT>(x_out, row, ok, L.q, x);
    if (!y) continue;                            // only the head input wanted (uniform: no barrier in this loop)
    Frag3 xf[KS1];
    to_frags<CT>(x, xf);
    f32x4 o[NTH];
    mlp2<KS1, 2 * CT, NTH>(P1, P2, xf, o, L.v, L.q);
#pragma unroll
    for (int t = 0; t < NTH; ++t) {
      const f32x4 hvv = o[t] + *(const f32x4*)(p2_bias + 16 * t + 4 * L.q);
#pragma unroll
      for (int j = 0; j < 4; ++j) mine[L.v * (NTH * 16) + 16 * t + 4 * L.q + j] = hvv[j];
    }
    if (L.q == 0 && ok) {
      // lanes of one wave run in lockstep and a wave's LDS operations complete in order: the q = 0 lane of each
      // voxel reads what the lanes q = 1..3 of that voxel stored just above (volatile keeps the order)
      const volatile float* hv = mine + L.v * (NTH * 16);
      auto feat = [&](int c) { return x_in[((int64_t)b * 6 + c) * rows.voxels + vox] * hd.std_x[c] + hd.mean_x[c]; };
      level_fractions(hd, [&](int c) { return hv[c]; }, feat(hd.t_channel), feat(hd.ne_channel), feat(hd.rho_channel),
                      [&](int o, float val) { y[((int64_t)b * n_out + o) * rows.voxels + vox] = val; });
    }
  }
}

// ---------------------------------------------------------------------------------------------------------------
// split-bf16 twiddle GEMM
// ---------------------------------------------------------------------------------------------------------------
typedef float f32x2 __attribute__((ext_vector_type(2)));
// a lane's NC consecutive columns (NC = 4 or 2) as one vector load
template <int NC>
struct ColVec;
template <>
struct ColVec<4> {
  typedef f32x4 T;
};
template <>
struct ColVec<2> {
  typedef f32x2 T;
};

// split of a column group: 8 k-values of the lane's NC columns into 2 or 3 bf16 planes
template <bool THREE, int NC = 4>
__device__ __forceinline__ void split_cols(const typename ColVec<NC>::T (&xv)[8], bf16x8 (&bh)[NC], bf16x8 (&bm)[NC],
                                           bf16x8 (&bl)[NC]) {
  const NegPair n;
#pragma unroll
  for (int i = 0; i < NC; ++i)
#pragma unroll
    for (int j = 0; j < 8; j += 2) {
      float x0 = xv[j][i], x1 = xv[j + 1][i];
      const bf16x2 h = split_pair(x0, x1, n);           // remainders x - h, exact
      bh[i][j] = h[0];
      bh[i][j + 1] = h[1];
      if (THREE) {
        const bf16x2 m = split_pair(x0, x1, n);         // x - h - m, exact
        bm[i][j] = m[0];
        bm[i][j + 1] = m[1];
        bl[i][j] = (bf16)x0;
        bl[i][j + 1] = (bf16)x1;
      } else {
        bm[i][j] = (bf16)x0;
        bm[i][j + 1] = (bf16)x1;
      }
    }
}

// acc += A-fragment (ah, am[, al]) x the lane's NC column fragments
template <bool THREE, int NC = 4>
__device__ __forceinline__ void tw_mac(bf16x8 ah, bf16x8 am, bf16x8 al, const bf16x8 (&bh)[NC], const bf16x8 (&bm)[NC],
                                       const bf16x8 (&bl)[NC], f32x4 (&acc)[NC]) {
#pragma unroll
  for (int i = 0; i < NC; ++i) {
    if (THREE) {
      acc[i] = mfma(al, bh[i], acc[i]);
      acc[i] = mfma(am, bm[i], acc[i]);
      acc[i] = mfma(ah, bl[i], acc[i]);
    }
    acc[i] = mfma(ah, bm[i], acc[i]);
    acc[i] = mfma(am, bh[i], acc[i]);
    acc[i] = mfma(ah, bh[i], acc[i]);
  }
}

template <bool THREE>
__global__ __launch_bounds__(256) void k_tw_gemm(FfnoTwGemm g) {
  Lane L;
  const int64_t ntn = (g.n + 63) / 64;                 // 64-column tasks per batch item
  const int64_t tasks = (int64_t)g.batch * ntn;
  const int nmb = (g.m_pad + 63) / 64;
  for (int64_t task = (int64_t)blockIdx.x * 4 + L.wave; task < tasks; task += (int64_t)gridDim.x * 4) {
    const int64_t bi = task / ntn;
    const int n0 = (int)(task % ntn) * 64 + 4 * L.v;   // this lane's 4 columns
    const bool nok = n0 < g.n;
    const float* xb = g.x + imap(bi, g.b_split, g.x_b_hi, g.x_b_lo) + n0;
    float* yb = g.y + imap(bi, g.b_split, g.y_b_hi, g.y_b_lo) + n0;
    for (int mb = 0; mb < nmb; ++mb) {
      f32x4 acc[4][4];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[a][i] = zero4();
      for (int s = 0; s < g.k_pad / 32; ++s) {
        f32x4 xv[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const int k = 32 * s + 8 * L.q + j;
          xv[j] = (nok && k < g.k) ? *(const f32x4*)(xb + imap(k, g.k_split, g.x_k_hi, g.x_k_lo)) : zero4();
        }
        bf16x8 bh[4], bm[4], bl[4];
        split_cols<THREE>(xv, bh, bm, bl);
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          const int mt = mb * 4 + a;
          if (mt * 16 >= g.m_pad) break;
          const int64_t ao = (int64_t)(mt * 16 + L.v) * g.k_pad + 32 * s + 8 * L.q;
          const bf16x8 ah = *(const bf16x8*)(g.a_hi + ao);
          const bf16x8 am = *(const bf16x8*)(g.a_lo + ao);
          const bf16x8 al = THREE ? *(const bf16x8*)(g.a_l3 + ao) : ah;
          tw_mac<THREE>(ah, am, al, bh, bm, bl, acc[a]);
        }
      }
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        const int mt = mb * 4 + a;
        if (mt * 16 >= g.m_pad) break;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int m = mt * 16 + 4 * L.q + j;
          if (m < g.m && nok) {
            float* p = yb + imap(m, g.m_split, g.y_m_hi, g.y_m_lo);
            f32x4 r = f32x4{acc[a][0][j], acc[a][1][j], acc[a][2][j], acc[a][3][j]};
            if (g.accumulate) r += *(const f32x4*)p;
            *(f32x4*)p = r;
          }
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------------------------------------------
// per-column vertical transform: every height column has its own matrix A_j [m x k] (the Filon weights of its own
// heights, or the synthesis at them); Y_j [m][C] = A_j X_j with X [B][k][HW][C], Y [B][m][HW][C]. A wave takes one
// column and 64 of its channels; its A_j fragments come from global memory (each column's matrix is read once per
// 64-row block).
// ---------------------------------------------------------------------------------------------------------------
template <bool THREE>
__global__ __launch_bounds__(256) void k_col_gemm(FfnoColGemm g) {
  Lane L;
  const int nchunk = (g.C + 63) / 64;
  const int64_t cols = (int64_t)g.B * g.hw;
  const int64_t tasks = cols * nchunk;
  const int nmb = (g.m_pad + 63) / 64;
  const int64_t rs = (int64_t)g.hw * g.C;               // row stride of X and Y
  const int64_t asz = (int64_t)g.m_pad * g.k_pad;
  for (int64_t task = (int64_t)blockIdx.x * 4 + L.wave; task < tasks; task += (int64_t)gridDim.x * 4) {
    const int64_t j = task / nchunk;
    const int b = (int)(j / g.hw);
    const int64_t hw = j % g.hw;
    const int n0 = (int)(task % nchunk) * 64 + 4 * L.v;
    const bool nok = n0 < g.C;
    const float* xb = g.x + ((int64_t)b * g.k * g.hw + hw) * g.C + n0;
    float* yb = g.y + ((int64_t)b * g.m * g.hw + hw) * g.C + n0;
    const bf16* ah = reinterpret_cast<const bf16*>(g.a_hi) + j * asz;
    const bf16* am = reinterpret_cast<const bf16*>(g.a_lo) + j * asz;
    const bf16* al = THREE ? reinterpret_cast<const bf16*>(g.a_l3) + j * asz : ah;
    for (int mb = 0; mb < nmb; ++mb) {
      f32x4 acc[4][4];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[a][i] = zero4();
      for (int s = 0; s < g.k_pad / 32; ++s) {
        f32x4 xv[8];
#pragma unroll
        for (int jj = 0; jj < 8; ++jj) {
          const int k = 32 * s + 8 * L.q + jj;
          xv[jj] = (nok && k < g.k) ? *(const f32x4*)(xb + k * rs) : zero4();
        }
        bf16x8 bh[4], bm[4], bl[4];
        split_cols<THREE>(xv, bh, bm, bl);
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          const int mt = mb * 4 + a;
          if (mt * 16 >= g.m_pad) break;
          const int64_t ao = (int64_t)(mt * 16 + L.v) * g.k_pad + 32 * s + 8 * L.q;
          const bf16x8 fh = *(const bf16x8*)(ah + ao);
          const bf16x8 fm = *(const bf16x8*)(am + ao);
          const bf16x8 fl = THREE ? *(const bf16x8*)(al + ao) : fh;
          tw_mac<THREE>(fh, fm, fl, bh, bm, bl, acc[a]);
        }
      }
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        const int mt = mb * 4 + a;
        if (mt * 16 >= g.m_pad) break;
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) {
          const int m = mt * 16 + 4 * L.q + jj;
          if (m < g.m && nok)
            *(f32x4*)(yb + m * rs) = f32x4{acc[a][0][jj], acc[a][1][jj], acc[a][2][jj], acc[a][3][jj]};
        }
      }
    }
  }
}

// twiddle matrices of at most 128 x 128: A staged in LDS once per persistent workgroup (rows padded by 16 bytes so
// the lanes' 16-byte reads fall in distinct banks), the row offsets of both index maps tabulated in LDS, all MT
// m-tiles of a task of 16 * NC columns accumulated in registers (the task's input is read once), and the next
// k-step's input loaded while the current one's MFMAs run. NC = 2 halves the accumulators of the larger tiles so two
// waves fit each SIMD.
constexpr int TW_PAD = 8;
// column groups per lane of the small twiddle GEMM: 2 (32-column tasks, two waves per SIMD) from 4 m-tiles up, except
// the three-way 8-tile case, whose 64-column tasks at one wave per SIMD measured faster (462 against 548 us on the
// 128-row vertical inverse of a 128 x 64 x 64 x 64 batch of 8)
constexpr int tw_cols(int mt, bool three) { return mt >= 4 && !(mt == 8 && three) ? 2 : 4; }

// TWID: the first stage of a factorised complex DFT; output rows k1 (real part, tile a) and m_split + k1 (imaginary
// part, tile a + MT/2) sit in the same lane and register slot, so the twiddle multiply is lane-local.
template <int MT, bool THREE, int NC, bool TWID = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NC == 2 ? 2 : 1))) void k_tw_small(FfnoTwGemm g) {
  static_assert(!TWID || MT % 2 == 0, "twiddle rows come in real/imaginary tile pairs");
  typedef typename ColVec<NC>::T V;
  extern __shared__ __align__(16) unsigned char smem[];
  const int kp = g.k_pad, ldk = kp + TW_PAD;
  int64_t* koff = reinterpret_cast<int64_t*>(smem);
  int64_t* moff = koff + kp;
  bf16* ah = reinterpret_cast<bf16*>(moff + MT * 16);
  bf16* am = ah + MT * 16 * ldk;
  bf16* al = am + MT * 16 * ldk;          // third plane (THREE only)
  for (int i = threadIdx.x; i < MT * 16 * (kp / 8); i += blockDim.x) {
    const int r = i / (kp / 8), c = (i % (kp / 8)) * 8;
    uint4 h = {0, 0, 0, 0}, m = {0, 0, 0, 0}, l = {0, 0, 0, 0};
    if (r < g.m_pad) {
      h = *reinterpret_cast<const uint4*>(g.a_hi + (int64_t)r * kp + c);
      m = *reinterpret_cast<const uint4*>(g.a_lo + (int64_t)r * kp + c);
      if (THREE) l = *reinterpret_cast<const uint4*>(g.a_l3 + (int64_t)r * kp + c);
    }
    *reinterpret_cast<uint4*>(ah + r * ldk + c) = h;
    *reinterpret_cast<uint4*>(am + r * ldk + c) = m;
    if (THREE) *reinterpret_cast<uint4*>(al + r * ldk + c) = l;
  }
  for (int k = threadIdx.x; k < kp; k += blockDim.x) koff[k] = k < g.k ? imap(k, g.k_split, g.x_k_hi, g.x_k_lo) : -1;
  for (int m = threadIdx.x; m < MT * 16; m += blockDim.x)
    moff[m] = m < g.m && (!TWID || m % g.m_split < g.tw_n1) ? imap(m, g.m_split, g.y_m_hi, g.y_m_lo) : -1;
  __syncthreads();
  Lane L;
  const int64_t ntn = (g.n + 16 * NC - 1) / (16 * NC);
  const int64_t tasks = (int64_t)g.batch * ntn;
  const int ks = kp / 32;
  const int64_t stride = (int64_t)gridDim.x * 4;
  // the columns of task t and the input pointer of its batch item
  auto cols = [&](int64_t t, int& n0, bool& nok, const float*& xb) {
    n0 = (int)(t % ntn) * (16 * NC) + NC * L.v;
    nok = n0 < g.n;
    xb = g.x + imap(t / ntn, g.b_split, g.x_b_hi, g.x_b_lo) + n0;
  };
  // k-step s of one lane's 4 columns (koff < 0: padding rows beyond K)
  auto load = [&](const float* xb, bool nok, int s, V (&xv)[8]) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int64_t o = koff[32 * s + 8 * L.q + j];
      xv[j] = (nok && o >= 0) ? *(const V*)(xb + o) : V{};
    }
  };
  int64_t task = (int64_t)blockIdx.x * 4 + L.wave;
  if (task >= tasks) return;
  int n0;
  bool nok;
  const float* xb;
  cols(task, n0, nok, xb);
  V xv[8];
  load(xb, nok, 0, xv);
  // every wave runs its tasks with the next k-step's loads in flight across task boundaries: the first k-step of
  // the next task is fetched during the last k-step of this one
  while (true) {
    float* yb = g.y + imap(task / ntn, g.b_split, g.y_b_hi, g.y_b_lo) + n0;
    const bool out_ok = nok;
    const int n_cur = n0;
    const int64_t next = task + stride;
    f32x4 acc[MT][NC];
#pragma unroll
    for (int a = 0; a < MT; ++a)
#pragma unroll
      for (int i = 0; i < NC; ++i) acc[a][i] = zero4();
    for (int s = 0; s < ks; ++s) {
      bf16x8 bh[NC], bm[NC], bl[NC];
      split_cols<THREE, NC>(xv, bh, bm, bl);
      if (s + 1 < ks) {
        load(xb, nok, s + 1, xv);
      } else if (next < tasks) {
        cols(next, n0, nok, xb);
        load(xb, nok, 0, xv);
      }
#pragma unroll
      for (int a = 0; a < MT; ++a) {
        const int ao = (a * 16 + L.v) * ldk + 32 * s + 8 * L.q;
        const bf16x8 fh = *reinterpret_cast<const bf16x8*>(ah + ao);
        const bf16x8 fm = *reinterpret_cast<const bf16x8*>(am + ao);
        const bf16x8 fl = THREE ? *reinterpret_cast<const bf16x8*>(al + ao) : fh;
        tw_mac<THREE, NC>(fh, fm, fl, bh, bm, bl, acc[a]);
      }
    }
    if constexpr (TWID) {
      constexpr int RT = MT / 2;
      const int64_t tcol = g.tw_cols > 0 ? (out_ok ? n_cur / g.tw_cols : 0) : task / ntn % g.tw_bsplit;
      const float2* tw = reinterpret_cast<const float2*>(g.twiddle) + tcol;
#pragma unroll
      for (int a = 0; a < RT; ++a)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int k1 = a * 16 + 4 * L.q + j;
          const int64_t o = moff[k1];
          if (o >= 0 && out_ok) {
            const float2 w = tw[(int64_t)k1 * g.tw_n2];
            V re, im;
#pragma unroll
            for (int i = 0; i < NC; ++i) {
              const float xr = acc[a][i][j], xi = acc[a + RT][i][j];
              re[i] = xr * w.x - xi * w.y;
              im[i] = xr * w.y + xi * w.x;
            }
            *(V*)(yb + o) = re;
            *(V*)(yb + moff[RT * 16 + k1]) = im;
          }
        }
    } else {
#pragma unroll
      for (int a = 0; a < MT; ++a)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int64_t o = moff[a * 16 + 4 * L.q + j];
          if (o >= 0 && out_ok) {
            float* p = yb + o;
            V r;
#pragma unroll
            for (int i = 0; i < NC; ++i) r[i] = acc[a][i][j];
            if (g.accumulate) r += *(const V*)p;
            *(V*)p = r;
          }
        }
    }
    task = next;
    if (task >= tasks) break;
  }
}

// ---------------------------------------------------------------------------------------------------------------
// per-mode complex channel mix, in place
// ---------------------------------------------------------------------------------------------------------------
constexpr int MIX_TILES = 8;   // row tiles of 16 per work item of the mix

// NL: planes of wc staged in LDS (0 or 2 for the two-way form, 0, 2 or 3 for the three-way one)
template <int CT, int NL, bool THREE, int NWM>
__global__ __launch_bounds__(NWM * 64) void k_mix(FfnoMix mx) {
  using S = SplitT<THREE>;
  using F = typename S::F;
  constexpr int C = CT * 16, KS = ks_of(CT), KP = KS * 32, NWC = 2 * C * wld(2 * KP), P = S::P;
  extern __shared__ __align__(16) unsigned char smem[];
  bf16* lds = reinterpret_cast<bf16*>(smem);
  typename S::W Wc;
  if constexpr (THREE)
    Wc = stage_w3<NL>(lds, mx.wc, NWC);
  else
    Wc = NL ? stage_w(lds, mx.wc, NWC) : global_w(mx.wc, NWC);
  __syncthreads();
  Lane L;
  // work items are (mode, chunk of MIX_TILES row tiles), mode-major so that neighbouring waves share a mode's tables
  const int rtiles = (mx.rows + 15) / 16;
  const int nchunk = (rtiles + MIX_TILES - 1) / MIX_TILES;
  const int64_t items = (int64_t)mx.modes * nchunk;
  for (int64_t item = (int64_t)blockIdx.x * NWM + L.wave; item < items; item += (int64_t)gridDim.x * NWM) {
    const int mode = (int)(item / nchunk);
    const int rt0 = (int)(item % nchunk) * MIX_TILES;
    const int rt1 = rt0 + MIX_TILES < rtiles ? rt0 + MIX_TILES : rtiles;
    const float gr = mx.gate[2 * mode], gi = mx.gate[2 * mode + 1];
    const float fac = mx.factor ? mx.factor[mode] : 1.f;
    // per-mode low-rank tables: V [16][wld(KP)] and U [C][wld(32)], P planes of each mode's block
    const auto V = S::global(mx.v_p ? mx.v_p + (int64_t)mode * P * 16 * wld(KP) : mx.wc, 16 * wld(KP));
    const auto U = S::global(mx.u_p ? mx.u_p + (int64_t)mode * P * C * wld(32) : mx.wc, C * wld(32));
    // row tile rt's offset; its loads are issued one tile ahead, into registers the split has just freed, so that
    // they are in flight while the previous tile's GEMMs run
    auto row_of = [&](int rt, bool& ok) {
      const int r = rt * 16 + L.v;
      ok = r < mx.rows;
      return imap(ok ? r : 0, mx.r_split, mx.r_hi, mx.r_lo) + (int64_t)mode * mx.mode_stride;
    };
    f32x4 yr[CT], yi[CT];
    bool nok;
    int64_t nbase = row_of(rt0, nok);
#pragma unroll
    for (int t = 0; t < CT; ++t) {
      yr[t] = ld4(mx.re + nbase + 16 * t + 4 * L.q, nok);
      yi[t] = ld4(mx.im + nbase + 16 * t + 4 * L.q, nok);
    }
    for (int rt = rt0; rt < rt1; ++rt) {
      const bool ok = nok;
      const int64_t base = nbase;
      F cat[2 * KS];
      to_frags<CT>(yr, cat);
      to_frags<CT>(yi, cat + KS);
      if (rt + 1 < rt1) {
        nbase = row_of(rt + 1, nok);
#pragma unroll
        for (int t = 0; t < CT; ++t) {
          yr[t] = ld4(mx.re + nbase + 16 * t + 4 * L.q, nok);
          yi[t] = ld4(mx.im + nbase + 16 * t + 4 * L.q, nok);
        }
      }
      f32x4 zr[CT], zi[CT];
#pragma unroll
      for (int t = 0; t < CT; ++t) {
        const f32x4 ar = tile<2 * KS>(Wc, t, cat, L.v, L.q);
        const f32x4 ai = tile<2 * KS>(Wc, CT + t, cat, L.v, L.q);
        zr[t] = gr * ar - gi * ai;
        zi[t] = gi * ar + gr * ai;
        __builtin_amdgcn_sched_barrier(0);   // keep the next tile's weight loads from being hoisted here
      }
      if (mx.rank) {
        const f32x4 lr = tile<KS>(V, 0, cat, L.v, L.q);
        const f32x4 li = tile<KS>(V, 0, cat + KS, L.v, L.q);
        const F lrf = S::split(lr, zero4()), lif = S::split(li, zero4());
#pragma unroll
        for (int t = 0; t < CT; ++t) {
          zr[t] += tile<1>(U, t, &lrf, L.v, L.q);
          zi[t] += tile<1>(U, t, &lif, L.v, L.q);
        }
      }
#pragma unroll
      for (int t = 0; t < CT; ++t) {
        st4(mx.re + base + 16 * t + 4 * L.q, fac * zr[t], ok);
        st4(mx.im + base + 16 * t + 4 * L.q, fac * zi[t], ok);
      }
    }
  }
}

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[MAX_DEVICES] = {};     // the device address of `scratch` on each GPU, looked up once

// The current device's copy of `scratch`: hipGetSymbolAddress on the first call on a device (with that device
// current), the cached pointer afterwards.
int scratch_of_current_device(float** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES)
    return fail("ffno_post_gn_noalloc: no current HIP device (or its index is above 63)");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      snprintf(g_err, sizeof(g_err), "ffno_post_gn_noalloc: hipGetSymbolAddress(scratch) failed: %s",
               hipGetErrorString(e));
      return -1;
    }
    g_scratch_dev[dev] = p;
  }
  *out = static_cast<float*>(g_scratch_dev[dev]);
  return 0;
}

}  // namespace
}  // namespace ffno

using namespace ffno;

extern "C" {

const char* ffno_last_error(void) { return g_err; }
int ffno_version(void) { return 1; }

int ffno_lift(FfnoRows rows, const float* x_in, const float* w_lift, float* out, FfnoStatsOut stats,
              hipStream_t stream) {
  if (!rows_ok(rows) || !x_in || !w_lift || !out) return fail("ffno_lift: bad arguments");
  dim3 grid(rows.nwg, rows.B);
  FFNO_DISPATCH_CT(rows.C, (k_lift<CT, nw_of<CT>()><<<grid, nw_of<CT>() * 64, 0, stream>>>(rows, x_in, w_lift, out, stats)));
  return check_launch("ffno_lift");
}

int ffno_gn_finalize(int B, int nwg, int C, const float* partials, const float* gamma, const float* beta,
                     float* stats, float* aff_a, float* aff_b, hipStream_t stream) {
  if (B <= 0 || nwg <= 0 || C % G || !partials || !gamma || !beta || !aff_a || !aff_b)
    return fail("ffno_gn_finalize: bad arguments");
  k_gn_finalize<<<B, 256, 0, stream>>>(nwg, C, partials, gamma, beta, stats, aff_a, aff_b);
  return check_launch("ffno_gn_finalize");
}

int ffno_block_in(FfnoRows rows, FfnoAssemble in, float* x_out, const uint16_t* g1, const uint16_t* g2,
                  const uint16_t* v1, const uint16_t* v2, const float* z_proj, const float* z, float* xs, float* xv,
                  hipStream_t stream) {
  if (!rows_ok(rows) || !in.t || !in.aff_a || !in.aff_b || !g1 || !g2 || !v1 || !v2 || !z_proj || !z || !xv)
    return fail("ffno_block_in: bad arguments");
  dim3 grid(rows.nwg, rows.B);
  FFNO_DISPATCH_CT(rows.C, {
    const size_t elems = 4 * (size_t)(CT * 16) * wld(ks_of(CT) * 32);
    const size_t head = (size_t)2 * CT * 16 * sizeof(float);   // the assembly's affine
    const int nl = lds_planes(elems, head);
    FFNO_DISPATCH_NL(nl, {
      const size_t lds = NL * elems * 2 + head;
      auto k = k_block_in<CT, nw_fwd<CT>(), NL>;
      if (set_lds(k, lds)) return -1;
      k<<<grid, nw_fwd<CT>() * 64, lds, stream>>>(rows, in, x_out, g1, g2, v1, v2, z_proj, z, xs, xv);
    });
  });
  return check_launch("ffno_block_in");
}

int ffno_post(FfnoRows rows, const float* u, const uint16_t* w1, const uint16_t* w2, float* y, FfnoStatsOut stats,
              hipStream_t stream) {
  if (!rows_ok(rows) || !u || !w1 || !w2 || !y) return fail("ffno_post: bad arguments");
  dim3 grid(rows.nwg, rows.B);
  FFNO_DISPATCH_CT(rows.C, {
    constexpr int NWP = nw_bwd<CT>();   // 12 waves at CT 4: two tiles per wave need a 3-wave SIMD's registers
    const size_t elems = (size_t)2 * CT * 16 * wld(ks_of(CT) * 32) + (size_t)CT * 16 * wld(2 * CT * 16);
    const size_t red = red_bytes_nw<CT, NWP>();
    const int nl = lds_planes(elems, red);
    FFNO_DISPATCH_NL(nl, {
      const size_t lds = NL * elems * 2 + red;
      auto k = k_post<CT, NWP, NL>;
      if (set_lds(k, lds)) return -1;
      k<<<grid, NWP * 64, lds, stream>>>(rows, u, w1, w2, y, stats);
    });
  });
  return check_launch("ffno_post");
}

int ffno_post_gn_noalloc(FfnoRows rows, const float* u, const uint16_t* w1, const uint16_t* w2, float* y,
                         const float* depth_w, const float* gamma, const float* beta, float* stats, float* aff_a,
                         float* aff_b, hipStream_t stream) {
  if (!rows_ok(rows) || !u || !w1 || !w2 || !y || !depth_w || !gamma || !beta || !aff_a || !aff_b)
    return fail("ffno_post_gn_noalloc: bad arguments");
  if ((int64_t)rows.B * rows.nwg > FFNO_NOALLOC_MAX_SLOTS) {
    snprintf(g_err, sizeof(g_err),
             "ffno_post_gn_noalloc: B * nwg = %lld partial-statistics slots exceed the static scratch's %d",
             (long long)rows.B * rows.nwg, FFNO_NOALLOC_MAX_SLOTS);
    return -1;
  }
  float* part = nullptr;
  if (scratch_of_current_device(&part)) return -1;
  // every workgroup of the post kernel writes all G slots of its own [b][wg] entry before the finalize reads them:
  // nothing in the scratch has to start at zero
  const int rc = ffno_post(rows, u, w1, w2, y, FfnoStatsOut{depth_w, part}, stream);
  if (rc) return rc;
  return ffno_gn_finalize(rows.B, rows.nwg, rows.C, part, gamma, beta, stats, aff_a, aff_b, stream);
}

int ffno_fuse(FfnoRows rows, const float* x, const float* sp, const float* vp, const float* s_a, const float* s_b,
              const float* v_a, const float* v_b, const uint16_t* bg1, const uint16_t* bg2, const float* bg2_bias,
              const uint16_t* f1, const uint16_t* f2, float* fp, FfnoStatsOut stats, float* gates_out,
              hipStream_t stream) {
  if (!rows_ok(rows) || !x || !sp || !vp || !s_a || !s_b || !v_a || !v_b || !bg1 || !bg2 || !bg2_bias || !f1 ||
      !f2 || !fp)
    return fail("ffno_fuse: bad arguments");
  dim3 grid(rows.nwg, rows.B);
  FFNO_DISPATCH_CT(rows.C, {
    constexpr int C = CT * 16, KP = ks_of(CT) * 32;
    const size_t lds = ((size_t)C * wld(3 * KP) + 2 * ((size_t)2 * C * wld(KP) + 2 * C * wld(2 * KP)) +
                        (size_t)C * wld(2 * C)) * 2 +
                       red_bytes_nw<CT, nw_of<CT>()>() + 6 * C * sizeof(float);
    if (lds <= 160 * 1024) {
      auto k = k_fuse<CT, nw_of<CT>(), true>;
      if (set_lds(k, lds)) return -1;
      k<<<grid, nw_of<CT>() * 64, lds, stream>>>(rows, x, sp, vp, s_a, s_b, v_a, v_b, bg1, bg2, bg2_bias, f1, f2, fp, stats,
                                                 gates_out);
    } else {
      auto k = k_fuse<CT, NWV, false>;
      const size_t bytes = red_bytes<CT>() + 6 * C * sizeof(float);
      if (set_lds(k, bytes)) return -1;
      k<<<grid, NWV * 64, bytes, stream>>>(rows, x, sp, vp, s_a, s_b, v_a, v_b, bg1, bg2, bg2_bias, f1, f2, fp, stats,
                                           gates_out);
    }
  });
  return check_launch("ffno_fuse");
}

int ffno_resmlp(FfnoRows rows, FfnoAssemble in, float* x_out, const uint16_t* w1, const uint16_t* w2, float* y,
                FfnoStatsOut stats, hipStream_t stream) {
  if (!rows_ok(rows) || !in.t || !in.aff_a || !in.aff_b || !w1 || !w2 || !y) return fail("ffno_resmlp: bad arguments");
  dim3 grid(rows.nwg, rows.B);
  FFNO_DISPATCH_CT(rows.C, {
    const size_t elems = (size_t)2 * CT * 16 * wld(ks_of(CT) * 32) + (size_t)CT * 16 * wld(2 * CT * 16);
    // 12 waves at CT 4: the next tile's two input rows held in flight need the registers of a 3-wave SIMD
    constexpr int NWR = nw_bwd<CT>();
    const size_t red = red_bytes_nw<CT, NWR>() + (size_t)2 * CT * 16 * sizeof(float);   // + the assembly's affine
    const int nl = lds_planes(elems, red);
    FFNO_DISPATCH_NL(nl, {
      const size_t lds = NL * elems * 2 + red;
      auto k = k_resmlp<CT, NWR, NL>;
      if (set_lds(k, lds)) return -1;
      k<<<grid, NWR * 64, lds, stream>>>(rows, in, x_out, w1, w2, y, stats);
    });
  });
  return check_launch("ffno_resmlp");
}

int ffno_head(FfnoRows rows, FfnoAssemble in, float* x_out, const uint16_t* p1, const uint16_t* p2,
              const float* p2_bias, const float* x_in, const FfnoHeadDesc* desc, float* y, hipStream_t stream) {
  if (!rows_ok(rows) || !in.t || !p1 || !p2 || !p2_bias || !x_in || !desc || (!y && !x_out))
    return fail("ffno_head: bad arguments");
  int n_head = 0, n_out = 0;
  if (desc->n_species < 1 || desc->n_species > 4) return fail("ffno_head: 1 to 4 species");
  for (int s = 0; s < desc->n_species; ++s) {
    if (desc->levels[s] < 1 || desc->levels[s] > 15) return fail("ffno_head: 1 to 15 levels per species");
    n_head += desc->levels[s] + 1;
    n_out += desc->levels[s];
  }
  if (n_head > 32 || n_out > 32) return fail("ffno_head: at most 32 head channels");
  dim3 grid(rows.nwg, rows.B);
  const int nth = (n_head + 15) / 16;
  FFNO_DISPATCH_CT(rows.C, {
    const size_t elems = (size_t)2 * CT * 16 * wld(ks_of(CT) * 32) + (size_t)nth * 16 * wld(2 * CT * 16);
    const size_t hs = (size_t)nw_fwd<CT>() * 16 * nth * 16 * sizeof(float);
    const int nl = lds_planes(elems, hs);
    FFNO_DISPATCH_NL(nl, {
      const size_t lds = NL * elems * 2 + hs;
      if (nth == 1) {
        auto k = k_head<CT, nw_fwd<CT>(), 1, NL>;
        if (set_lds(k, lds)) return -1;
        k<<<grid, nw_fwd<CT>() * 64, lds, stream>>>(rows, in, x_out, p1, p2, p2_bias, x_in, *desc, y);
      } else {
        auto k = k_head<CT, nw_fwd<CT>(), 2, NL>;
        if (set_lds(k, lds)) return -1;
        k<<<grid, nw_fwd<CT>() * 64, lds, stream>>>(rows, in, x_out, p1, p2, p2_bias, x_in, *desc, y);
      }
    });
  });
  return check_launch("ffno_head");
}

int ffno_tw_gemm(const FfnoTwGemm* g, hipStream_t stream) {
  if (!g || !g->a_hi || !g->a_lo || !g->x || !g->y) return fail("ffno_tw_gemm: null pointer");
  if (g->m <= 0 || g->k <= 0 || g->n <= 0 || g->batch <= 0 || g->n % 4 || g->m_pad % 16 || g->k_pad % 32 ||
      g->m_pad < g->m || g->k_pad < g->k || g->b_split <= 0 || g->k_split <= 0 || g->m_split <= 0)
    return fail("ffno_tw_gemm: bad sizes");
  const int64_t tasks = (int64_t)g->batch * ((g->n + 63) / 64);
  const int64_t blocks = (tasks + 3) / 4;
  const bool three = g->a_l3 != nullptr;
  if (g->twiddle) {
    const int mt = g->m_pad / 16;
    if (g->m_pad > 128 || g->k_pad > 128 || mt % 2 || g->m_split != g->m_pad / 2 || g->accumulate ||
        g->tw_n1 <= 0 || g->tw_n1 > g->m_split || g->m != g->m_split + g->tw_n1 || g->tw_n2 <= 0 ||
        g->tw_cols < 0 || g->tw_cols % 4 || (g->tw_cols > 0 && (g->n + g->tw_cols - 1) / g->tw_cols > g->tw_n2) ||
        (g->tw_cols == 0 && (g->tw_bsplit <= 0 || g->tw_bsplit > g->tw_n2)))
      return fail("ffno_tw_gemm: bad twiddle stage");
    const size_t lds = (size_t)(g->k_pad + mt * 16) * sizeof(int64_t) +
                       (size_t)(three ? 3 : 2) * mt * 16 * (g->k_pad + TW_PAD) * sizeof(uint16_t);
    const int nc = tw_cols(mt, three);
    const int64_t sblocks = ((int64_t)g->batch * ((g->n + 16 * nc - 1) / (16 * nc)) + 3) / 4;
    const int grid = (int)(sblocks < 512 ? sblocks : 512);
    switch (mt) {
#define FFNO_TWID_CASE(M)                                                                                 \
  case M: {                                                                                               \
    auto k = three ? k_tw_small<M, true, tw_cols(M, true), true> : k_tw_small<M, false, tw_cols(M, false), true>; \
    if (set_lds(k, lds)) return -1;                                                                       \
    k<<<grid, 256, lds, stream>>>(*g);                                                                    \
    break;                                                                                                \
  }
      FFNO_TWID_CASE(2) FFNO_TWID_CASE(4) FFNO_TWID_CASE(6) FFNO_TWID_CASE(8)
#undef FFNO_TWID_CASE
      default: return fail("ffno_tw_gemm: bad tile count");
    }
    return check_launch("ffno_tw_gemm");
  }
  if (g->m_pad <= 128 && g->k_pad <= 128) {
    const int mt = (g->m_pad + 15) / 16;
    const size_t lds = (size_t)(g->k_pad + mt * 16) * sizeof(int64_t) +
                       (size_t)(three ? 3 : 2) * mt * 16 * (g->k_pad + TW_PAD) * sizeof(uint16_t);
    // persistent: A is staged once per workgroup; 4 tiles and more take 32-column tasks (NC = 2)
    const int nc = tw_cols(mt, three);
    const int64_t sblocks = ((int64_t)g->batch * ((g->n + 16 * nc - 1) / (16 * nc)) + 3) / 4;
    const int grid = (int)(sblocks < 512 ? sblocks : 512);
    switch (mt) {
#define FFNO_TW_CASE(M)                                                                                   \
  case M: {                                                                                               \
    auto k = three ? k_tw_small<M, true, tw_cols(M, true)> : k_tw_small<M, false, tw_cols(M, false)>;    \
    if (set_lds(k, lds)) return -1;                                                                       \
    k<<<grid, 256, lds, stream>>>(*g);                                                                    \
    break;                                                                                                \
  }
      FFNO_TW_CASE(1) FFNO_TW_CASE(2) FFNO_TW_CASE(3) FFNO_TW_CASE(4)
      FFNO_TW_CASE(5) FFNO_TW_CASE(6) FFNO_TW_CASE(7) FFNO_TW_CASE(8)
#undef FFNO_TW_CASE
      default: return fail("ffno_tw_gemm: bad tile count");
    }
    return check_launch("ffno_tw_gemm");
  }
  const int grid = (int)(blocks < 4096 ? blocks : 4096);
  if (three)
    k_tw_gemm<true><<<grid, 256, 0, stream>>>(*g);
  else
    k_tw_gemm<false><<<grid, 256, 0, stream>>>(*g);
  return check_launch("ffno_tw_gemm");
}

int ffno_col_gemm(const FfnoColGemm* g, hipStream_t stream) {
  if (!g || !g->a_hi || !g->a_lo || !g->x || !g->y) return fail("ffno_col_gemm: null pointer");
  if (g->m <= 0 || g->k <= 0 || g->B <= 0 || g->hw <= 0 || g->C <= 0 || g->C % 4 || g->m_pad % 16 ||
      g->k_pad % 32 || g->m_pad < g->m || g->k_pad < g->k)
    return fail("ffno_col_gemm: bad sizes");
  const int64_t tasks = (int64_t)g->B * g->hw * ((g->C + 63) / 64);
  const int64_t blocks = (tasks + 3) / 4;
  const int grid = (int)(blocks < 8192 ? blocks : 8192);
  if (g->a_l3)
    k_col_gemm<true><<<grid, 256, 0, stream>>>(*g);
  else
    k_col_gemm<false><<<grid, 256, 0, stream>>>(*g);
  return check_launch("ffno_col_gemm");
}

int ffno_mix(const FfnoMix* m, hipStream_t stream) {
  if (!m || !m->re || !m->im || !m->wc || !m->gate) return fail("ffno_mix: null pointer");
  if (m->rank && (!m->v_p || !m->u_p || m->rank > 16)) return fail("ffno_mix: low-rank tables missing or rank > 16");
  if (m->rows <= 0 || m->modes <= 0 || m->r_split <= 0) return fail("ffno_mix: bad sizes");
  const int64_t items = (int64_t)m->modes * (((m->rows + 15) / 16 + MIX_TILES - 1) / MIX_TILES);
  if (m->parts != 2 && m->parts != 3) return fail("ffno_mix: parts must be 2 or 3");
  // the staged wc allows one workgroup per CU: its waves are the occupancy (8 for the three-way form's registers,
  // 12 for the two-way one's, 4 at widths whose wc is read from global memory)
  auto launch = [&](auto k, int nw, size_t lds) -> int {
    if (set_lds(k, lds)) return -1;
    const int64_t blocks = (items + nw - 1) / nw;
    const int grid = (int)(blocks < 2048 ? blocks : 2048);
    k<<<grid, nw * 64, lds, stream>>>(*m);
    return 0;
  };
  FFNO_DISPATCH_CT(m->C, {
    constexpr int NW3 = CT <= 4 ? 8 : 4, NW2 = CT <= 4 ? 12 : 4;
    const size_t elems = (size_t)2 * CT * 16 * wld(2 * ks_of(CT) * 32);
    if (m->parts == 3) {
      FFNO_DISPATCH_NL(lds_planes(elems, 0), {
        if (launch(k_mix<CT, NL, true, NW3>, NW3, NL * elems * 2)) return -1;
      });
    } else if (2 * elems * 2 <= 160 * 1024) {
      if (launch(k_mix<CT, 2, false, NW2>, NW2, 2 * elems * 2)) return -1;
    } else {
      if (launch(k_mix<CT, 0, false, 4>, 4, 0)) return -1;
    }
  });
  return check_launch("ffno_mix");
}

}  // extern "C"
