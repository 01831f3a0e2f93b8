// FFNO3D forward kernels for gfx950 (MI355X): register-chained bf16 MFMA MLPs with fused GroupNorm statistics,
// a split-bf16 twiddle GEMM for the DFTs and the vertical Filon transform, and the per-mode spectral channel mix.
//
// MFMA conventions (v_mfma_f32_16x16x32_bf16, 64 lanes; lane l: v = l & 15, q = l >> 4):
//   A operand (16 x 32): lane l supplies row v and 8 k-values; B operand (32 x 16): lane l supplies column v and the
//   same 8 k-slots; D (16 x 16 f32): lane l holds rows 4q..4q+3 of column v.
// Pointwise layers are computed transposed, out^T = W * in^T: rows are output features, columns are 16 voxels.
// A lane then holds, for every 16-feature tile t, features 16t+4q..16t+4q+3 of voxel v (the "lane layout"), both for
// what it loads and for what an MFMA leaves in it. Two such float4s (tiles 2s, 2s+1) make the B operand of k-step s;
// the weights are packed on the host with the matching k permutation, so the hidden activations of a layer feed the
// next layer straight from registers.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include "ffno_hip.h"

namespace {

typedef __bf16 bf16;
typedef bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int G = 8;               // GroupNorm groups for every supported width (models/ffno_model.py:21-25)

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ f32x4 zero4() { return f32x4{0.f, 0.f, 0.f, 0.f}; }

__device__ __forceinline__ bf16x8 pack8(f32x4 a, f32x4 b) {
  bf16x8 r;
  r[0] = (bf16)a[0]; r[1] = (bf16)a[1]; r[2] = (bf16)a[2]; r[3] = (bf16)a[3];
  r[4] = (bf16)b[0]; r[5] = (bf16)b[1]; r[6] = (bf16)b[2]; r[7] = (bf16)b[3];
  return r;
}

__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float sigm(float x) { return 1.0f / (1.0f + __expf(-x)); }
__device__ __forceinline__ f32x4 gelu4(f32x4 x) { return f32x4{gelu(x[0]), gelu(x[1]), gelu(x[2]), gelu(x[3])}; }
__device__ __forceinline__ f32x4 sigm4(f32x4 x) { return f32x4{sigm(x[0]), sigm(x[1]), sigm(x[2]), sigm(x[3])}; }

__device__ __forceinline__ f32x4 ld4(const float* p, bool ok) { return ok ? *(const f32x4*)p : zero4(); }
__device__ __forceinline__ void st4(float* p, f32x4 v, bool ok) { if (ok) *(f32x4*)p = v; }

constexpr int ks_of(int ntiles) { return (ntiles + 1) / 2; }   // k-steps of 32 for ntiles 16-feature tiles

template <int NT>
__device__ __forceinline__ void to_frags(const f32x4 (&a)[NT], bf16x8* f) {
#pragma unroll
  for (int s = 0; s < ks_of(NT); ++s) f[s] = pack8(a[2 * s], (2 * s + 1 < NT) ? a[2 * s + 1] : zero4());
}

// out^T[NT tiles] = W[NT*16 rows][KS*32] * in^T, W in LDS (row-major, K permuted per 32-block on the host)
template <int NT, int KS>
__device__ __forceinline__ void gemm(const bf16* W, const bf16x8* b, f32x4 (&acc)[NT], int v, int q) {
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    f32x4 c = zero4();
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      bf16x8 a = *(const bf16x8*)(W + (16 * t + v) * (KS * 32) + 32 * s + 8 * q);
      c = mfma(a, b[s], c);
    }
    acc[t] = c;
  }
}

__device__ __forceinline__ void stage(bf16* dst, const uint16_t* src, int n) {   // n bf16, multiple of 8
  const uint4* s = reinterpret_cast<const uint4*>(src);
  uint4* d = reinterpret_cast<uint4*>(dst);
  for (int i = threadIdx.x; i < n / 8; i += blockDim.x) d[i] = s[i];
}

// ---------------------------------------------------------------------------------------------------------------
// GroupNorm statistics, accumulated per lane with a shift, merged (Chan) across lanes, waves and workgroups.
// ---------------------------------------------------------------------------------------------------------------
__device__ __forceinline__ void chan(float& W, float& m, float& M2, float W2, float m2, float M22) {
  float Wt = W + W2;
  if (Wt <= 0.f) return;
  float d = m2 - m;
  m += d * (W2 / Wt);
  M2 += M22 + d * d * (W * (W2 / Wt));
  W = Wt;
}

template <int CT>
struct GnAcc {
  static constexpr int CPG = CT * 16 / G;          // channels per group: 2, 4, 8, 12, 16
  static constexpr int SUB = CPG >= 4 ? 1 : 2;     // sub-chunks per 4-feature chunk
  static constexpr int NF = 4 / SUB;               // features per sub-chunk
  float s1[CT * SUB], s2[CT * SUB], k[CT * SUB];
  float wsum;
  bool seen;

  __device__ void init() {
#pragma unroll
    for (int i = 0; i < CT * SUB; ++i) { s1[i] = 0.f; s2[i] = 0.f; k[i] = 0.f; }
    wsum = 0.f;
    seen = false;
  }
  __device__ __forceinline__ void add(const f32x4 (&y)[CT], float w, bool valid) {
    if (!valid) return;
    if (!seen) {
#pragma unroll
      for (int t = 0; t < CT; ++t)
#pragma unroll
        for (int u = 0; u < SUB; ++u) {
          float a = 0.f;
#pragma unroll
          for (int j = 0; j < NF; ++j) a += y[t][u * NF + j];
          k[t * SUB + u] = a / NF;
        }
      seen = true;
    }
    wsum += w;
#pragma unroll
    for (int t = 0; t < CT; ++t)
#pragma unroll
      for (int u = 0; u < SUB; ++u) {
        float a = 0.f, b = 0.f;
#pragma unroll
        for (int j = 0; j < NF; ++j) {
          float d = y[t][u * NF + j] - k[t * SUB + u];
          a += d;
          b += d * d;
        }
        s1[t * SUB + u] += w * a;
        s2[t * SUB + u] += w * b;
      }
  }
  // merge everything the workgroup saw into partials[g][3]; lds needs NW*4*CT*SUB*3 floats
  template <int NW>
  __device__ void flush(float* lds, float* partials, int v, int q, int wave) {
#pragma unroll
    for (int i = 0; i < CT * SUB; ++i) {
      float W = NF * wsum;
      float m = W > 0.f ? k[i] + s1[i] / W : 0.f;
      float M2 = W > 0.f ? fmaxf(s2[i] - s1[i] * s1[i] / W, 0.f) : 0.f;
#pragma unroll
      for (int off = 1; off < 16; off <<= 1) {
        float W2 = __shfl_xor(W, off), m2 = __shfl_xor(m, off), M22 = __shfl_xor(M2, off);
        chan(W, m, M2, W2, m2, M22);
      }
      if (v == 0) {
        float* e = lds + ((wave * 4 + q) * (CT * SUB) + i) * 3;
        e[0] = W; e[1] = m; e[2] = M2;
      }
    }
    __syncthreads();
    if (threadIdx.x < G) {
      const int g = threadIdx.x;
      float W = 0.f, m = 0.f, M2 = 0.f;
      for (int w = 0; w < NW; ++w)
        for (int qq = 0; qq < 4; ++qq)
          for (int i = 0; i < CT * SUB; ++i) {
            int t = i / SUB, u = i % SUB;
            int grp = (16 * t + 4 * qq + u * NF) / CPG;
            if (grp != g) continue;
            const float* e = lds + ((w * 4 + qq) * (CT * SUB) + i) * 3;
            chan(W, m, M2, e[0], e[1], e[2]);
          }
      partials[g * 3 + 0] = W;
      partials[g * 3 + 1] = m;
      partials[g * 3 + 2] = M2;
    }
  }
};

// ---------------------------------------------------------------------------------------------------------------
// shared pieces of the voxel-row kernels
// ---------------------------------------------------------------------------------------------------------------
struct Lane {
  int wave, lane, v, q;
  __device__ Lane() {
    wave = threadIdx.x >> 6;
    lane = threadIdx.x & 63;
    v = lane & 15;
    q = lane >> 4;
  }
};

template <int CT>
__device__ __forceinline__ void load_act(const float* p, int64_t row, bool ok, int q, f32x4 (&x)[CT]) {
#pragma unroll
  for (int t = 0; t < CT; ++t) x[t] = ld4(p + row * (CT * 16) + 16 * t + 4 * q, ok);
}

template <int CT>
__device__ __forceinline__ void store_act(float* p, int64_t row, bool ok, int q, const f32x4 (&x)[CT]) {
#pragma unroll
  for (int t = 0; t < CT; ++t) st4(p + row * (CT * 16) + 16 * t + 4 * q, x[t], ok);
}

template <int CT>
__device__ __forceinline__ void assemble(const FfnoAssemble& in, float scale, int64_t row, int b, bool ok, int q,
                                         f32x4 (&x)[CT]) {
  constexpr int C = CT * 16;
#pragma unroll
  for (int t = 0; t < CT; ++t) {
    const int c0 = 16 * t + 4 * q;
    f32x4 tv = ld4(in.t + row * C + c0, ok);
    f32x4 a = *(const f32x4*)(in.aff_a + b * C + c0);
    f32x4 s = *(const f32x4*)(in.aff_b + b * C + c0);
    f32x4 u = tv * a + s;
    if (in.act_gelu) u = gelu4(u);
    f32x4 base = in.base ? ld4(in.base + row * C + c0, ok) : zero4();
    x[t] = ok ? base + scale * u : zero4();
  }
}

template <int CT>
__device__ __forceinline__ void gelu_tiles(f32x4 (&x)[CT]) {
#pragma unroll
  for (int t = 0; t < CT; ++t) x[t] = gelu4(x[t]);
}

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
__global__ void k_gn_finalize(int nwg, int C, const float* __restrict__ partials, const float* __restrict__ gamma,
                              const float* __restrict__ beta, float* __restrict__ stats, float* __restrict__ aff_a,
                              float* __restrict__ aff_b) {
  __shared__ float mean_s[G], rstd_s[G];
  const int b = blockIdx.x;
  if (threadIdx.x < G) {
    const int g = threadIdx.x;
    float W = 0.f, m = 0.f, M2 = 0.f;
    for (int i = 0; i < nwg; ++i) {
      const float* p = partials + (((int64_t)b * nwg + i) * G + g) * 3;
      chan(W, m, M2, p[0], p[1], p[2]);
    }
    float var = W > 0.f ? fmaxf(M2 / W, 0.f) : 0.f;
    float r = rsqrtf(var + 1e-5f);
    mean_s[g] = m;
    rstd_s[g] = r;
    if (stats) {
      stats[(b * G + g) * 2 + 0] = m;
      stats[(b * G + g) * 2 + 1] = r;
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
template <int CT, int NW>
__global__ __launch_bounds__(NW * 64) void k_block_in(FfnoRows rows, FfnoAssemble in, float* __restrict__ x_out,
                                                      const uint16_t* g1, const uint16_t* g2, const uint16_t* v1,
                                                      const uint16_t* v2, const float* __restrict__ z_proj,
                                                      const float* __restrict__ z, float* __restrict__ xs,
                                                      float* __restrict__ xv) {
  constexpr int C = CT * 16, KS = ks_of(CT), KP = KS * 32;
  extern __shared__ __align__(16) unsigned char smem[];
  bf16* W = reinterpret_cast<bf16*>(smem);
  constexpr int SZ = C * KP;
  stage(W, g1, SZ);
  stage(W + SZ, g2, SZ);
  stage(W + 2 * SZ, v1, SZ);
  stage(W + 3 * SZ, v2, SZ);
  __syncthreads();
  Lane L;
  const int b = blockIdx.y;
  const float scale = in.scale ? *in.scale : 1.f;
  const int64_t nt = (rows.voxels + 15) / 16;
  for (int64_t tile = (int64_t)blockIdx.x * NW + L.wave; tile < nt; tile += (int64_t)gridDim.x * NW) {
    const int64_t vox = tile * 16 + L.v;
    const bool ok = vox < rows.voxels;
    const int64_t row = (int64_t)b * rows.voxels + vox;
    f32x4 x[CT];
    assemble<CT>(in, scale, row, b, ok, L.q, x);
    if (x_out) store_act<CT>(x_out, row, ok, L.q, x);
    bf16x8 xf[KS];
    to_frags<CT>(x, xf);
    f32x4 h[CT], g[CT];
    bf16x8 hf[KS];
    // horizontal gate (models/ffno_model.py:439-440)
    gemm<CT, KS>(W, xf, h, L.v, L.q);
    gelu_tiles<CT>(h);
    to_frags<CT>(h, hf);
    gemm<CT, KS>(W + SZ, hf, g, L.v, L.q);
#pragma unroll
    for (int t = 0; t < CT; ++t) g[t] = x[t] * sigm4(g[t]);
    store_act<CT>(xs, row, ok, L.q, g);
    // vertical gate and height projection (models/ffno_model.py:266)
    gemm<CT, KS>(W + 2 * SZ, xf, h, L.v, L.q);
    gelu_tiles<CT>(h);
    to_frags<CT>(h, hf);
    gemm<CT, KS>(W + 3 * SZ, hf, g, L.v, L.q);
    const float zv = ok ? z[row] : 0.f;
#pragma unroll
    for (int t = 0; t < CT; ++t) {
      f32x4 zp = *(const f32x4*)(z_proj + 16 * t + 4 * L.q);
      g[t] = x[t] * sigm4(g[t]) + zp * zv;
    }
    store_act<CT>(xv, row, ok, L.q, g);
  }
}

// ---------------------------------------------------------------------------------------------------------------
// post MLP C -> 2C -> C with statistics
// ---------------------------------------------------------------------------------------------------------------
template <int CT, int NW>
__global__ __launch_bounds__(NW * 64) void k_post(FfnoRows rows, const float* __restrict__ u, const uint16_t* w1,
                                                  const uint16_t* w2, float* __restrict__ y, FfnoStatsOut st) {
  constexpr int C = CT * 16, KS1 = ks_of(CT), KS2 = ks_of(2 * CT);
  extern __shared__ __align__(16) unsigned char smem[];
  bf16* W1 = reinterpret_cast<bf16*>(smem);
  bf16* W2 = W1 + 2 * C * KS1 * 32;
  float* red = reinterpret_cast<float*>(W2 + C * KS2 * 32);
  stage(W1, w1, 2 * C * KS1 * 32);
  stage(W2, w2, C * KS2 * 32);
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
    f32x4 x[CT];
    load_act<CT>(u, row, ok, L.q, x);
    bf16x8 xf[KS1];
    to_frags<CT>(x, xf);
    f32x4 h[2 * CT];
    gemm<2 * CT, KS1>(W1, xf, h, L.v, L.q);
    gelu_tiles<2 * CT>(h);
    bf16x8 hf[KS2];
    to_frags<2 * CT>(h, hf);
    f32x4 o[CT];
    gemm<CT, KS2>(W2, hf, o, L.v, L.q);
    store_act<CT>(y, row, ok, L.q, o);
    if (st.partials) acc.add(o, ok ? st.depth_w[row] : 0.f, ok);
  }
  if (st.partials)
    acc.template flush<NW>(red, st.partials + ((int64_t)b * gridDim.x + blockIdx.x) * G * 3, L.v, L.q, L.wave);
}

// ---------------------------------------------------------------------------------------------------------------
// branch fusion
// ---------------------------------------------------------------------------------------------------------------
template <int CT, int NW>
__global__ __launch_bounds__(NW * 64) void k_fuse(FfnoRows rows, const float* __restrict__ xin,
                                                  const float* __restrict__ sp, const float* __restrict__ vp,
                                                  const float* __restrict__ s_a, const float* __restrict__ s_b,
                                                  const float* __restrict__ v_a, const float* __restrict__ v_b,
                                                  const uint16_t* bg1, const uint16_t* bg2,
                                                  const float* __restrict__ bg2_bias, const uint16_t* f1,
                                                  const uint16_t* f2, float* __restrict__ fp, FfnoStatsOut st) {
  constexpr int C = CT * 16, KS = ks_of(CT), KP = KS * 32, KS2 = ks_of(2 * CT);
  extern __shared__ __align__(16) unsigned char smem[];
  bf16* B1 = reinterpret_cast<bf16*>(smem);          // [C][3KP]
  bf16* B2 = B1 + C * 3 * KP;                        // [2C][KP]
  bf16* F1 = B2 + 2 * C * KP;                        // [2C][2KP]
  bf16* F2 = F1 + 2 * C * 2 * KP;                    // [C][KS2*32]
  float* red = reinterpret_cast<float*>(F2 + C * KS2 * 32);
  stage(B1, bg1, C * 3 * KP);
  stage(B2, bg2, 2 * C * KP);
  stage(F1, f1, 2 * C * 2 * KP);
  stage(F2, f2, C * KS2 * 32);
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
    f32x4 x[CT], s[CT], v[CT];
    load_act<CT>(xin, row, ok, L.q, x);
    load_act<CT>(sp, row, ok, L.q, s);
    load_act<CT>(vp, row, ok, L.q, v);
#pragma unroll
    for (int t = 0; t < CT; ++t) {
      const int c0 = b * C + 16 * t + 4 * L.q;
      s[t] = gelu4(s[t] * *(const f32x4*)(s_a + c0) + *(const f32x4*)(s_b + c0));
      v[t] = gelu4(v[t] * *(const f32x4*)(v_a + c0) + *(const f32x4*)(v_b + c0));
      if (!ok) { s[t] = zero4(); v[t] = zero4(); }
    }
    bf16x8 cat[3 * KS];
    to_frags<CT>(x, cat);
    to_frags<CT>(s, cat + KS);
    to_frags<CT>(v, cat + 2 * KS);
    f32x4 h[CT];
    gemm<CT, 3 * KS>(B1, cat, h, L.v, L.q);
    gelu_tiles<CT>(h);
    bf16x8 hf[KS];
    to_frags<CT>(h, hf);
    f32x4 g[2 * CT];
    gemm<2 * CT, KS>(B2, hf, g, L.v, L.q);
#pragma unroll
    for (int t = 0; t < CT; ++t) {
      s[t] = s[t] * sigm4(g[t] + *(const f32x4*)(bg2_bias + 16 * t + 4 * L.q));
      v[t] = v[t] * sigm4(g[CT + t] + *(const f32x4*)(bg2_bias + C + 16 * t + 4 * L.q));
    }
    to_frags<CT>(s, cat);
    to_frags<CT>(v, cat + KS);
    gemm<2 * CT, 2 * KS>(F1, cat, g, L.v, L.q);
    gelu_tiles<2 * CT>(g);
    bf16x8 gf[KS2];
    to_frags<2 * CT>(g, gf);
    f32x4 o[CT];
    gemm<CT, KS2>(F2, gf, o, L.v, L.q);
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
template <int CT, int NW>
__global__ __launch_bounds__(NW * 64) void k_resmlp(FfnoRows rows, FfnoAssemble in, float* __restrict__ x_out,
                                                    const uint16_t* w1, const uint16_t* w2, float* __restrict__ y,
                                                    FfnoStatsOut st) {
  constexpr int C = CT * 16, KS1 = ks_of(CT), KS2 = ks_of(2 * CT);
  extern __shared__ __align__(16) unsigned char smem[];
  bf16* W1 = reinterpret_cast<bf16*>(smem);
  bf16* W2 = W1 + 2 * C * KS1 * 32;
  float* red = reinterpret_cast<float*>(W2 + C * KS2 * 32);
  stage(W1, w1, 2 * C * KS1 * 32);
  stage(W2, w2, C * KS2 * 32);
  __syncthreads();
  Lane L;
  const int b = blockIdx.y;
  const float scale = in.scale ? *in.scale : 1.f;
  const int64_t nt = (rows.voxels + 15) / 16;
  GnAcc<CT> acc;
  acc.init();
  for (int64_t tile = (int64_t)blockIdx.x * NW + L.wave; tile < nt; tile += (int64_t)gridDim.x * NW) {
    const int64_t vox = tile * 16 + L.v;
    const bool ok = vox < rows.voxels;
    const int64_t row = (int64_t)b * rows.voxels + vox;
    f32x4 x[CT];
    assemble<CT>(in, scale, row, b, ok, L.q, x);
    if (x_out) store_act<CT>(x_out, row, ok, L.q, x);
    bf16x8 xf[KS1];
    to_frags<CT>(x, xf);
    f32x4 h[2 * CT];
    gemm<2 * CT, KS1>(W1, xf, h, L.v, L.q);
    gelu_tiles<2 * CT>(h);
    bf16x8 hf[KS2];
    to_frags<2 * CT>(h, hf);
    f32x4 o[CT];
    gemm<CT, KS2>(W2, hf, o, L.v, L.q);
    store_act<CT>(y, row, ok, L.q, o);
    if (st.partials) acc.add(o, ok ? st.depth_w[row] : 0.f, ok);
  }
  if (st.partials)
    acc.template flush<NW>(red, st.partials + ((int64_t)b * gridDim.x + blockIdx.x) * G * 3, L.v, L.q, L.wave);
}

// ---------------------------------------------------------------------------------------------------------------
// head: projection MLP + level-fraction output, channel-first store
// ---------------------------------------------------------------------------------------------------------------
constexpr float LN10 = 2.302585092994046f;

template <int CT, int NW, int NTH>
__global__ __launch_bounds__(NW * 64) void k_head(FfnoRows rows, FfnoAssemble in, const uint16_t* p1,
                                                  const uint16_t* p2, const float* __restrict__ p2_bias,
                                                  const float* __restrict__ x_in, FfnoHeadDesc hd,
                                                  float* __restrict__ y) {
  constexpr int C = CT * 16, KS1 = ks_of(CT), KS2 = ks_of(2 * CT);
  extern __shared__ __align__(16) unsigned char smem[];
  bf16* P1 = reinterpret_cast<bf16*>(smem);
  bf16* P2 = P1 + 2 * C * KS1 * 32;
  float* hs = reinterpret_cast<float*>(P2 + NTH * 16 * KS2 * 32);   // [NW][16 voxels][NTH*16]
  stage(P1, p1, 2 * C * KS1 * 32);
  stage(P2, p2, NTH * 16 * KS2 * 32);
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
    bf16x8 xf[KS1];
    to_frags<CT>(x, xf);
    f32x4 h[2 * CT];
    gemm<2 * CT, KS1>(P1, xf, h, L.v, L.q);
    gelu_tiles<2 * CT>(h);
    bf16x8 hf[KS2];
    to_frags<2 * CT>(h, hf);
    f32x4 o[NTH];
    gemm<NTH, KS2>(P2, hf, o, L.v, L.q);
#pragma unroll
    for (int t = 0; t < NTH; ++t) {
      f32x4 bb = *(const f32x4*)(p2_bias + 16 * t + 4 * L.q);
      const f32x4 hvv = o[t] + bb;
#pragma unroll
      for (int j = 0; j < 4; ++j) mine[L.v * (NTH * 16) + 16 * t + 4 * L.q + j] = hvv[j];
    }
    if (L.q == 0 && ok) {
      // lanes of one wave run in lockstep and a wave's LDS operations complete in order: the q = 0 lane of each
      // voxel reads what the lanes q = 1..3 of that voxel stored just above (volatile keeps the order)
      const volatile float* hv = mine + L.v * (NTH * 16);
      auto feat = [&](int c) { return x_in[((int64_t)b * 6 + c) * rows.voxels + vox] * hd.std_x[c] + hd.mean_x[c]; };
      const float lt = feat(hd.t_channel) * LN10;                 // ln T
      const float lne = (feat(hd.ne_channel) - 6.0f) * LN10;      // ln ne [cm^-3]
      const float l10rho = feat(hd.rho_channel);
      const float inv_t = __expf(-lt);
      // MULTI3D calc_lte_populations: conl = ln(ccon ne) - 1.5 ln T, dxi = 4.98e-4 * 5040 / T * sqrt(ne k T) [eV]
      const float conl = hd.ln_ccon + lne - 1.5f * lt;
      const float dxi = hd.dxi_coef * __expf(0.5f * (lne + hd.ln_k - lt));
      int c = 0, k = 0;
      for (int s = 0; s < hd.n_species; ++s) {
        const int nl = hd.levels[s];
        float logit[15];
        float mx = -3.0e38f;
        for (int i = 0; i < nl; ++i) {
          const int li = k + i;
          const float lte = hd.log_g[li] - hd.e_over_k * inv_t * (hd.ev[li] - dxi * hd.ndxi[li]) - hd.dion[li] * conl;
          logit[i] = hv[c + i] + lte;
          mx = fmaxf(mx, logit[i]);
        }
        float se = 0.f;
        for (int i = 0; i < nl; ++i) se += __expf(logit[i] - mx);
        const float lse = mx + __logf(se);
        const float bound = hd.total_bound;
        const float stot = bound * tanhf(hv[c + nl] / bound);
        for (int i = 0; i < nl; ++i) {
          const int o = k + i;
          const float l10n = l10rho - hd.log10_mass[s] + stot + (logit[i] - lse) / LN10;
          y[((int64_t)b * n_out + o) * rows.voxels + vox] = (l10n - hd.mean_y[o]) / hd.std_y[o];
        }
        c += nl + 1;
        k += nl;
      }
    }
  }
}

// ---------------------------------------------------------------------------------------------------------------
// split-bf16 twiddle GEMM
// ---------------------------------------------------------------------------------------------------------------
__device__ __forceinline__ int64_t imap(int64_t i, int64_t split, int64_t hi, int64_t lo) {
  return (i / split) * hi + (i % split) * lo;
}

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
        bf16x8 bh[4], bl[4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 8; ++j) {
            const bf16 h = (bf16)xv[j][i];
            bh[i][j] = h;
            bl[i][j] = (bf16)(xv[j][i] - (float)h);
          }
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          const int mt = mb * 4 + a;
          if (mt * 16 >= g.m_pad) break;
          const int64_t ao = (int64_t)(mt * 16 + L.v) * g.k_pad + 32 * s + 8 * L.q;
          const bf16x8 ah = *(const bf16x8*)(g.a_hi + ao);
          const bf16x8 al = *(const bf16x8*)(g.a_lo + ao);
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            acc[a][i] = mfma(ah, bh[i], acc[a][i]);
            acc[a][i] = mfma(ah, bl[i], acc[a][i]);
            acc[a][i] = mfma(al, bh[i], acc[a][i]);
          }
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
// per-mode complex channel mix, in place
// ---------------------------------------------------------------------------------------------------------------
template <int CT>
__global__ __launch_bounds__(256) void k_mix(FfnoMix mx) {
  constexpr int C = CT * 16, KS = ks_of(CT), KP = KS * 32;
  extern __shared__ __align__(16) unsigned char smem[];
  bf16* Wc = reinterpret_cast<bf16*>(smem);             // [2C][2KP]
  stage(Wc, mx.wc, 2 * C * 2 * KP);
  __syncthreads();
  Lane L;
  const int rtiles = (mx.rows + 15) / 16;
  for (int mode = blockIdx.x * 4 + L.wave; mode < mx.modes; mode += gridDim.x * 4) {
    const float gr = mx.gate[2 * mode], gi = mx.gate[2 * mode + 1];
    const float fac = mx.factor ? mx.factor[mode] : 1.f;
    bf16x8 va[KS], ua[CT];
    if (mx.rank) {
      const uint16_t* vp = mx.v_p + (int64_t)mode * 16 * KP;
#pragma unroll
      for (int s = 0; s < KS; ++s) va[s] = *(const bf16x8*)(vp + L.v * KP + 32 * s + 8 * L.q);
      const uint16_t* up = mx.u_p + (int64_t)mode * C * 32;
#pragma unroll
      for (int t = 0; t < CT; ++t) ua[t] = *(const bf16x8*)(up + (16 * t + L.v) * 32 + 8 * L.q);
    }
    for (int rt = 0; rt < rtiles; ++rt) {
      const int r = rt * 16 + L.v;
      const bool ok = r < mx.rows;
      const int64_t base = imap(ok ? r : 0, mx.r_split, mx.r_hi, mx.r_lo) + (int64_t)mode * mx.mode_stride;
      f32x4 yr[CT], yi[CT];
#pragma unroll
      for (int t = 0; t < CT; ++t) {
        yr[t] = ld4(mx.re + base + 16 * t + 4 * L.q, ok);
        yi[t] = ld4(mx.im + base + 16 * t + 4 * L.q, ok);
      }
      bf16x8 cat[2 * KS];
      to_frags<CT>(yr, cat);
      to_frags<CT>(yi, cat + KS);
      f32x4 a[2 * CT];
      gemm<2 * CT, 2 * KS>(Wc, cat, a, L.v, L.q);
      f32x4 zr[CT], zi[CT];
#pragma unroll
      for (int t = 0; t < CT; ++t) {
        zr[t] = gr * a[t] - gi * a[CT + t];
        zi[t] = gi * a[t] + gr * a[CT + t];
      }
      if (mx.rank) {
        f32x4 lr = zero4(), li = zero4();
#pragma unroll
        for (int s = 0; s < KS; ++s) {
          lr = mfma(va[s], cat[s], lr);
          li = mfma(va[s], cat[KS + s], li);
        }
        const bf16x8 lrf = pack8(lr, zero4()), lif = pack8(li, zero4());
#pragma unroll
        for (int t = 0; t < CT; ++t) {
          zr[t] = mfma(ua[t], lrf, zr[t]);
          zi[t] = mfma(ua[t], lif, zi[t]);
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

// ---------------------------------------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------------------------------------
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

template <typename K>
int set_lds(K kernel, size_t bytes) {
  if (bytes > 160 * 1024) return fail("kernel needs more LDS than a CU has");
  if (bytes > 64 * 1024) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e != hipSuccess) return fail("hipFuncSetAttribute(MaxDynamicSharedMemorySize) failed");
  }
  return 0;
}

bool rows_ok(const FfnoRows& r) { return r.B > 0 && r.voxels > 0 && r.nwg > 0 && r.C % 16 == 0; }

constexpr int NWV = 4;   // waves per workgroup of the voxel-row kernels

#define FFNO_DISPATCH_CT(C, ...)                          \
  switch ((C) / 16) {                                     \
    case 1: { constexpr int CT = 1; __VA_ARGS__; break; } \
    case 2: { constexpr int CT = 2; __VA_ARGS__; break; } \
    case 4: { constexpr int CT = 4; __VA_ARGS__; break; } \
    case 6: { constexpr int CT = 6; __VA_ARGS__; break; } \
    case 8: { constexpr int CT = 8; __VA_ARGS__; break; } \
    default: return fail("unsupported width (16, 32, 64, 96 or 128)"); \
  }

template <int CT>
constexpr size_t red_bytes() { return (size_t)NWV * 4 * CT * GnAcc<CT>::SUB * 3 * sizeof(float); }

}  // namespace

extern "C" {

const char* ffno_last_error(void) { return g_err; }
int ffno_version(void) { return 1; }

int ffno_lift(FfnoRows rows, const float* x_in, const float* w_lift, float* out, FfnoStatsOut stats,
              hipStream_t stream) {
  if (!rows_ok(rows) || !x_in || !w_lift || !out) return fail("ffno_lift: bad arguments");
  dim3 grid(rows.nwg, rows.B);
  FFNO_DISPATCH_CT(rows.C, (k_lift<CT, NWV><<<grid, NWV * 64, 0, stream>>>(rows, x_in, w_lift, out, stats)));
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
  if (!rows_ok(rows) || !in.t || !in.aff_a || !in.aff_b || !g1 || !g2 || !v1 || !v2 || !z_proj || !z || !xs || !xv)
    return fail("ffno_block_in: bad arguments");
  dim3 grid(rows.nwg, rows.B);
  FFNO_DISPATCH_CT(rows.C, {
    const size_t lds = 4 * (size_t)(CT * 16) * ks_of(CT) * 32 * 2;
    auto k = k_block_in<CT, NWV>;
    if (set_lds(k, lds)) return -1;
    k<<<grid, NWV * 64, lds, stream>>>(rows, in, x_out, g1, g2, v1, v2, z_proj, z, xs, xv);
  });
  return check_launch("ffno_block_in");
}

int ffno_post(FfnoRows rows, const float* u, const uint16_t* w1, const uint16_t* w2, float* y, FfnoStatsOut stats,
              hipStream_t stream) {
  if (!rows_ok(rows) || !u || !w1 || !w2 || !y) return fail("ffno_post: bad arguments");
  dim3 grid(rows.nwg, rows.B);
  FFNO_DISPATCH_CT(rows.C, {
    const size_t lds = ((size_t)2 * CT * 16 * ks_of(CT) * 32 + (size_t)CT * 16 * ks_of(2 * CT) * 32) * 2 +
                       red_bytes<CT>();
    auto k = k_post<CT, NWV>;
    if (set_lds(k, lds)) return -1;
    k<<<grid, NWV * 64, lds, stream>>>(rows, u, w1, w2, y, stats);
  });
  return check_launch("ffno_post");
}

int ffno_fuse(FfnoRows rows, const float* x, const float* sp, const float* vp, const float* s_a, const float* s_b,
              const float* v_a, const float* v_b, const uint16_t* bg1, const uint16_t* bg2, const float* bg2_bias,
              const uint16_t* f1, const uint16_t* f2, float* fp, FfnoStatsOut stats, hipStream_t stream) {
  if (!rows_ok(rows) || !x || !sp || !vp || !s_a || !s_b || !v_a || !v_b || !bg1 || !bg2 || !bg2_bias || !f1 ||
      !f2 || !fp)
    return fail("ffno_fuse: bad arguments");
  dim3 grid(rows.nwg, rows.B);
  FFNO_DISPATCH_CT(rows.C, {
    constexpr int C = CT * 16, KP = ks_of(CT) * 32;
    const size_t lds = ((size_t)C * 3 * KP + 2 * C * KP + 2 * C * 2 * KP + (size_t)C * ks_of(2 * CT) * 32) * 2 +
                       red_bytes<CT>();
    auto k = k_fuse<CT, NWV>;
    if (set_lds(k, lds)) return -1;
    k<<<grid, NWV * 64, lds, stream>>>(rows, x, sp, vp, s_a, s_b, v_a, v_b, bg1, bg2, bg2_bias, f1, f2, fp, stats);
  });
  return check_launch("ffno_fuse");
}

int ffno_resmlp(FfnoRows rows, FfnoAssemble in, float* x_out, const uint16_t* w1, const uint16_t* w2, float* y,
                FfnoStatsOut stats, hipStream_t stream) {
  if (!rows_ok(rows) || !in.t || !in.aff_a || !in.aff_b || !w1 || !w2 || !y) return fail("ffno_resmlp: bad arguments");
  dim3 grid(rows.nwg, rows.B);
  FFNO_DISPATCH_CT(rows.C, {
    const size_t lds = ((size_t)2 * CT * 16 * ks_of(CT) * 32 + (size_t)CT * 16 * ks_of(2 * CT) * 32) * 2 +
                       red_bytes<CT>();
    auto k = k_resmlp<CT, NWV>;
    if (set_lds(k, lds)) return -1;
    k<<<grid, NWV * 64, lds, stream>>>(rows, in, x_out, w1, w2, y, stats);
  });
  return check_launch("ffno_resmlp");
}

int ffno_head(FfnoRows rows, FfnoAssemble in, const uint16_t* p1, const uint16_t* p2, const float* p2_bias,
              const float* x_in, const FfnoHeadDesc* desc, float* y, hipStream_t stream) {
  if (!rows_ok(rows) || !in.t || !p1 || !p2 || !p2_bias || !x_in || !desc || !y) return fail("ffno_head: bad arguments");
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
    const size_t w_bytes = ((size_t)2 * CT * 16 * ks_of(CT) * 32 + (size_t)nth * 16 * ks_of(2 * CT) * 32) * 2;
    const size_t lds = w_bytes + (size_t)NWV * 16 * nth * 16 * sizeof(float);
    if (nth == 1) {
      auto k = k_head<CT, NWV, 1>;
      if (set_lds(k, lds)) return -1;
      k<<<grid, NWV * 64, lds, stream>>>(rows, in, p1, p2, p2_bias, x_in, *desc, y);
    } else {
      auto k = k_head<CT, NWV, 2>;
      if (set_lds(k, lds)) return -1;
      k<<<grid, NWV * 64, lds, stream>>>(rows, in, p1, p2, p2_bias, x_in, *desc, y);
    }
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
  const int grid = (int)(blocks < 4096 ? blocks : 4096);
  k_tw_gemm<<<grid, 256, 0, stream>>>(*g);
  return check_launch("ffno_tw_gemm");
}

int ffno_mix(const FfnoMix* m, hipStream_t stream) {
  if (!m || !m->re || !m->im || !m->wc || !m->gate) return fail("ffno_mix: null pointer");
  if (m->rank && (!m->v_p || !m->u_p || m->rank > 16)) return fail("ffno_mix: low-rank tables missing or rank > 16");
  if (m->rows <= 0 || m->modes <= 0 || m->r_split <= 0) return fail("ffno_mix: bad sizes");
  const int blocks = (m->modes + 3) / 4;
  const int grid = blocks < 4096 ? blocks : 4096;
  FFNO_DISPATCH_CT(m->C, {
    const size_t lds = (size_t)2 * CT * 16 * 2 * ks_of(CT) * 32 * 2;
    auto k = k_mix<CT>;
    if (set_lds(k, lds)) return -1;
    k<<<grid, 256, lds, stream>>>(*m);
  });
  return check_launch("ffno_mix");
}

}  // extern "C"
