// Device helpers shared by the FFNO3D forward and adjoint kernels (gfx950).
//
// MFMA conventions (v_mfma_f32_16x16x32_bf16, 64 lanes; lane l: v = l & 15, q = l >> 4):
//   A operand (16 x 32): lane l supplies row v and 8 k-values; B operand (32 x 16): lane l supplies column v and the
//   same 8 k-slots; D (16 x 16 f32): lane l holds rows 4q..4q+3 of column v.
// Pointwise layers are computed transposed, out^T = W * in^T: rows are output features, columns are 16 voxels.
// A lane then holds, for every 16-feature tile t, features 16t+4q..16t+4q+3 of voxel v (the "lane layout"), both for
// what it loads and for what an MFMA leaves in it. Two such float4s (tiles 2s, 2s+1) make the B operand of k-step s;
// the weights are packed on the host with the matching k permutation, so the hidden activations of a layer feed the
// next layer straight from registers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include "ffno_hip.h"

namespace ffno {


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

// The standard normal CDF Phi(x) = (1 + erf(x / sqrt 2)) / 2 by Abramowitz & Stegun, Handbook of Mathematical
// Functions, formula 7.1.26 (Hastings): for a >= 0, erf(a) = 1 - (a1 t + a2 t^2 + a3 t^3 + a4 t^4 + a5 t^5) e^(-a^2),
// t = 1 / (1 + p a), |error| <= 1.5e-7. The tail Phi(-|x|) = (1 - erf(|x| / sqrt 2)) / 2 is the polynomial term itself
// (no cancellation for negative x). Written for few instructions: with u = k x, k = sqrt(log2(e) / 2), the Gaussian
// E = e^(-x^2 / 2) is 2^(-u^2) (one v_exp_f32), p a = (p / (sqrt 2 k)) |u|, and the 1/2 is folded into the a_i.
// Returns the tail and E (which the GELU derivative reuses).
__device__ __forceinline__ float norm_tail(float x, float& e) {
  const float u = x * 0.84932180028801907f;
  const float t = __builtin_amdgcn_rcpf(fmaf(fabsf(u), 0.27273747324943542f, 1.0f));
  const float poly = t * (0.127414796f + t * (-0.142248368f + t * (0.7107068705f + t * (-0.7265760135f +
                                                                                        t * 0.5307027145f))));
  e = __builtin_amdgcn_exp2f(-(u * u));
  return poly * e;
}
// GELU(x) = x Phi(x) = max(x, 0) - |x| Phi(-|x|): no compare or select
__device__ __forceinline__ float gelu(float x) {
  float e;
  return fmaf(-fabsf(x), norm_tail(x, e), fmaxf(x, 0.0f));
}
// 1 / (1 + e^-x) with the hardware reciprocal (1 ulp); e^-x = inf gives 0
__device__ __forceinline__ float sigm(float x) {
  return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(x * -1.4426950408889634f));
}
__device__ __forceinline__ f32x4 gelu4(f32x4 x) { return f32x4{gelu(x[0]), gelu(x[1]), gelu(x[2]), gelu(x[3])}; }
__device__ __forceinline__ f32x4 sigm4(f32x4 x) { return f32x4{sigm(x[0]), sigm(x[1]), sigm(x[2]), sigm(x[3])}; }

__device__ __forceinline__ f32x4 ld4(const float* p, bool ok) { return ok ? *(const f32x4*)p : zero4(); }
__device__ __forceinline__ void st4(float* p, f32x4 v, bool ok) { if (ok) *(f32x4*)p = v; }

constexpr int ks_of(int ntiles) { return (ntiles + 1) / 2; }   // k-steps of 32 for ntiles 16-feature tiles

// Every packed weight row of K bf16 (K a multiple of 32) is followed by WPAD bf16 of zeros, in global memory and in
// LDS alike. The row stride K + 16 (2K + 32 bytes) is the one that reads A-fragments at the full LDS rate on gfx950:
// timed over 16-row ds_read_b128 fragment reads, pads of 0, 4, 8, 16 and 24 bf16 take 3.5, 14, 1.8, 1 and 1.8 times
// the time of a pad of 16, at K = 64 and K = 128 alike.
constexpr int WPAD = 16;
constexpr int wld(int k) { return k + WPAD; }                   // row stride of a packed weight of K columns

// split-bf16 operands: x = hi + lo with hi = bf16(x), lo = bf16(x - hi); products hi*hi + hi*lo + lo*hi keep about
// 16 mantissa bits, so every MFMA layer is accurate to ~1e-5 relative with float32 accumulation
struct Frag {
  bf16x8 h, l;
};

typedef bf16 bf16x2 __attribute__((ext_vector_type(2)));

// The bf16 pairs (-1, 0) and (0, -1), made in scalar registers by an instruction the compiler cannot fold: given as
// an inline constant -1.0, v_dot2c_f32_bf16 applies it to the other half of the pair on gfx950
struct NegPair {
  bf16x2 lo, hi;
  __device__ __forceinline__ NegPair() {
    unsigned a, b;
    asm("s_mov_b32 %0, 0xbf80" : "=s"(a));
    asm("s_mov_b32 %0, 0xbf800000" : "=s"(b));
    lo = __builtin_bit_cast(bf16x2, a);
    hi = __builtin_bit_cast(bf16x2, b);
  }
};

// p = (bf16(x0), bf16(x1)) (one v_cvt_pk_bf16_f32), and x0, x1 replaced by their remainders x - p, exact in float32,
// one v_dot2 each: x0 + p . (-1, 0) (the pair's other half enters with weight 0)
__device__ __forceinline__ bf16x2 split_pair(float& x0, float& x1, const NegPair& n) {
  bf16x2 p;
  p[0] = (bf16)x0;
  p[1] = (bf16)x1;
  x0 = __builtin_amdgcn_fdot2_f32_bf16(p, n.lo, x0, false);
  x1 = __builtin_amdgcn_fdot2_f32_bf16(p, n.hi, x1, false);
  return p;
}

__device__ __forceinline__ Frag split8(f32x4 a, f32x4 b) {
  const NegPair n;
  float x[8] = {a[0], a[1], a[2], a[3], b[0], b[1], b[2], b[3]};
  Frag f;
#pragma unroll
  for (int j = 0; j < 8; j += 2) {
    const bf16x2 p = split_pair(x[j], x[j + 1], n);
    f.h[j] = p[0];
    f.h[j + 1] = p[1];
    f.l[j] = (bf16)x[j];
    f.l[j + 1] = (bf16)x[j + 1];
  }
  return f;
}

__device__ __forceinline__ f32x4 mfma3(bf16x8 ah, bf16x8 al, const Frag& b, f32x4 c) {
  c = mfma(ah, b.h, c);
  c = mfma(ah, b.l, c);
  return mfma(al, b.h, c);
}

template <int NT>
__device__ __forceinline__ void to_frags(const f32x4 (&a)[NT], Frag* f) {
#pragma unroll
  for (int s = 0; s < ks_of(NT); ++s) f[s] = split8(a[2 * s], (2 * s + 1 < NT) ? a[2 * s + 1] : zero4());
}

// A weight matrix [rows][K] in kernel form (k permuted per 32-block on the host), hi and lo halves; the pointers are
// LDS or global memory
struct Wm {
  const bf16* h;
  const bf16* l;
};

// one 16-row output tile t of W * in^T over KS k-steps
template <int KS>
__device__ __forceinline__ f32x4 tile(Wm W, int t, const Frag* b, int v, int q) {
  f32x4 c = zero4();
#pragma unroll
  for (int s = 0; s < KS; ++s) {
    const int o = (16 * t + v) * wld(KS * 32) + 32 * s + 8 * q;
    c = mfma3(*(const bf16x8*)(W.h + o), *(const bf16x8*)(W.l + o), b[s], c);
  }
  return c;
}

// out^T[NT tiles] = W[NT*16 rows][KS*32] * in^T
template <int NT, int KS>
__device__ __forceinline__ void gemm(Wm W, const Frag* b, f32x4 (&acc)[NT], int v, int q) {
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = tile<KS>(W, t, b, v, q);
}

// d/dx of the exact GELU: Phi(x) + x phi(x), phi(x) = e^(-x^2/2) / sqrt(2 pi)
__device__ __forceinline__ float gelu_grad(float x) {
  float e;
  const float tail = norm_tail(x, e);
  const float c = x >= 0.0f ? 1.0f - tail : tail;
  return fmaf(x * 0.3989422804014327f, e, c);
}
// GELU and its derivative from one tail evaluation
__device__ __forceinline__ void gelu_and_grad(float x, float& g, float& d) {
  float e;
  const float tail = norm_tail(x, e);
  g = fmaf(-fabsf(x), tail, fmaxf(x, 0.0f));
  d = fmaf(x * 0.3989422804014327f, e, x >= 0.0f ? 1.0f - tail : tail);
}
__device__ __forceinline__ f32x4 gelu_grad4(f32x4 x) {
  return f32x4{gelu_grad(x[0]), gelu_grad(x[1]), gelu_grad(x[2]), gelu_grad(x[3])};
}

// Adjoint of out = W2 gelu(W1 in) with respect to in, accumulated into acc (NTI tiles of in):
// acc += W1^T (gelu'(W1 in) * (W2^T og)). W2T and W1T are the transposed weights in kernel form; the hidden
// pre-activations are recomputed two tiles at a time, as in mlp2.
template <int KSI, int NH, int KSO, int NTI>
__device__ __forceinline__ void mlp2_bwd(Wm W1, Wm W2T, Wm W1T, const Frag* in, const Frag* og, f32x4 (&acc)[NTI],
                                         int v, int q) {
  constexpr int KSH = ks_of(NH);
#pragma unroll 1
  for (int s = 0; s < KSH; ++s) {
    f32x4 d0 = tile<KSO>(W2T, 2 * s, og, v, q) * gelu_grad4(tile<KSI>(W1, 2 * s, in, v, q));
    f32x4 d1 = zero4();
    if (2 * s + 1 < NH) d1 = tile<KSO>(W2T, 2 * s + 1, og, v, q) * gelu_grad4(tile<KSI>(W1, 2 * s + 1, in, v, q));
    const Frag df = split8(d0, d1);
#pragma unroll
    for (int t = 0; t < NTI; ++t) {
      const int o = (16 * t + v) * wld(KSH * 32) + 32 * s + 8 * q;
      acc[t] = mfma3(*(const bf16x8*)(W1T.h + o), *(const bf16x8*)(W1T.l + o), df, acc[t]);
    }
  }
}

struct GeluAct {
  __device__ __forceinline__ f32x4 operator()(f32x4 x) const { return gelu4(x); }
};

// out = W2 * gelu(W1 * in): the NH hidden tiles are produced two at a time and folded straight into the NT2
// output accumulators, so the hidden layer is never held whole
template <int KS1, int NH, int NT2>
__device__ __forceinline__ void mlp2(Wm W1, Wm W2, const Frag* in, f32x4 (&out)[NT2], int v, int q) {
  constexpr int KS2 = ks_of(NH);
#pragma unroll
  for (int t = 0; t < NT2; ++t) out[t] = zero4();
  // a rolled loop: unrolled, the scheduler hoists every weight fragment of the layer into registers and spills
#pragma unroll 1
  for (int s = 0; s < KS2; ++s) {
    const f32x4 h0 = gelu4(tile<KS1>(W1, 2 * s, in, v, q));
    const f32x4 h1 = (2 * s + 1 < NH) ? gelu4(tile<KS1>(W1, 2 * s + 1, in, v, q)) : zero4();
    const Frag hf = split8(h0, h1);
#pragma unroll
    for (int t = 0; t < NT2; ++t) {
      const int o = (16 * t + v) * wld(KS2 * 32) + 32 * s + 8 * q;
      out[t] = mfma3(*(const bf16x8*)(W2.h + o), *(const bf16x8*)(W2.l + o), hf, out[t]);
    }
  }
}

// ---------------------------------------------------------------------------------------------------------------
// Three-way split for the forward's pointwise layers: x = h + m + l, h = bf16(x), m = bf16(x - h),
// l = bf16(x - h - m), which holds all 24 bits of a float32. Of the nine partial products the six of order <= 2
// (hh, hm, mh, hl, mm, lh; each split level is 2^-8 smaller) are kept, so a layer is exact to about 2^-24 relative
// (float32 level). The two-way form above drops the order-2 terms and leaves ~2^-17 per layer, which eight blocks
// amplify past 1e-3 in log10 n; the adjoint keeps the two-way form (its tolerance is 1e-2). A weight packed in three
// planes [h, m, l] is read by the two-way kernels through its first two planes, which are its two-way split.
// ---------------------------------------------------------------------------------------------------------------
struct Frag3 {
  bf16x8 h, m, l;
};

struct Wm3 {
  const bf16* h;
  const bf16* m;
  const bf16* l;
};

__device__ __forceinline__ Frag3 split8_3(f32x4 a, f32x4 b) {
  const NegPair n;
  float x[8] = {a[0], a[1], a[2], a[3], b[0], b[1], b[2], b[3]};
  Frag3 f;
#pragma unroll
  for (int j = 0; j < 8; j += 2) {
    const bf16x2 h = split_pair(x[j], x[j + 1], n);     // x - h, exact
    const bf16x2 m = split_pair(x[j], x[j + 1], n);     // x - h - m, exact
    f.h[j] = h[0];
    f.h[j + 1] = h[1];
    f.m[j] = m[0];
    f.m[j + 1] = m[1];
    f.l[j] = (bf16)x[j];
    f.l[j + 1] = (bf16)x[j + 1];
  }
  return f;
}

// the small terms first, so that they meet the accumulator before the large ones
__device__ __forceinline__ f32x4 mfma6(bf16x8 ah, bf16x8 am, bf16x8 al, const Frag3& b, f32x4 c) {
  c = mfma(al, b.h, c);
  c = mfma(am, b.m, c);
  c = mfma(ah, b.l, c);
  c = mfma(am, b.h, c);
  c = mfma(ah, b.m, c);
  return mfma(ah, b.h, c);
}

template <int NT>
__device__ __forceinline__ void to_frags(const f32x4 (&a)[NT], Frag3* f) {
#pragma unroll
  for (int s = 0; s < ks_of(NT); ++s) f[s] = split8_3(a[2 * s], (2 * s + 1 < NT) ? a[2 * s + 1] : zero4());
}

// mlp2 for two voxel tiles at once: every weight fragment read from LDS feeds both tiles' MFMAs, which halves the
// LDS traffic per voxel and gives each wait on a fragment twice the independent MFMA work
template <int KS1, int NH, int NT2>
__device__ __forceinline__ void mlp2x2(Wm3 W1, Wm3 W2, const Frag3* in0, const Frag3* in1, f32x4 (&out0)[NT2],
                                       f32x4 (&out1)[NT2], int v, int q) {
  constexpr int KS2 = ks_of(NH);
#pragma unroll
  for (int t = 0; t < NT2; ++t) {
    out0[t] = zero4();
    out1[t] = zero4();
  }
#pragma unroll 1
  for (int s = 0; s < KS2; ++s) {
    f32x4 a0 = zero4(), a1 = zero4(), b0 = zero4(), b1 = zero4();   // hidden tiles 2s (a) and 2s + 1 (b)
#pragma unroll
    for (int k = 0; k < KS1; ++k) {
      const int oa = (32 * s + v) * wld(KS1 * 32) + 32 * k + 8 * q;
      const bf16x8 ah = *(const bf16x8*)(W1.h + oa), am = *(const bf16x8*)(W1.m + oa);
      const bf16x8 al = *(const bf16x8*)(W1.l + oa);
      a0 = mfma6(ah, am, al, in0[k], a0);
      a1 = mfma6(ah, am, al, in1[k], a1);
      if (2 * s + 1 < NH) {
        const int ob = oa + 16 * wld(KS1 * 32);
        const bf16x8 bh = *(const bf16x8*)(W1.h + ob), bm = *(const bf16x8*)(W1.m + ob);
        const bf16x8 bl = *(const bf16x8*)(W1.l + ob);
        b0 = mfma6(bh, bm, bl, in0[k], b0);
        b1 = mfma6(bh, bm, bl, in1[k], b1);
      }
    }
    const Frag3 hf0 = split8_3(gelu4(a0), gelu4(b0));
    const Frag3 hf1 = split8_3(gelu4(a1), gelu4(b1));
#pragma unroll
    for (int t = 0; t < NT2; ++t) {
      const int o = (16 * t + v) * wld(KS2 * 32) + 32 * s + 8 * q;
      const bf16x8 wh = *(const bf16x8*)(W2.h + o), wm = *(const bf16x8*)(W2.m + o);
      const bf16x8 wl = *(const bf16x8*)(W2.l + o);
      out0[t] = mfma6(wh, wm, wl, hf0, out0[t]);
      out1[t] = mfma6(wh, wm, wl, hf1, out1[t]);
    }
  }
}

template <int KS>
__device__ __forceinline__ f32x4 tile(Wm3 W, int t, const Frag3* b, int v, int q) {
  f32x4 c = zero4();
#pragma unroll
  for (int s = 0; s < KS; ++s) {
    const int o = (16 * t + v) * wld(KS * 32) + 32 * s + 8 * q;
    c = mfma6(*(const bf16x8*)(W.h + o), *(const bf16x8*)(W.m + o), *(const bf16x8*)(W.l + o), b[s], c);
  }
  return c;
}

template <int KS1, int NH, int NT2>
__device__ __forceinline__ void mlp2(Wm3 W1, Wm3 W2, const Frag3* in, f32x4 (&out)[NT2], int v, int q) {
  constexpr int KS2 = ks_of(NH);
#pragma unroll
  for (int t = 0; t < NT2; ++t) out[t] = zero4();
#pragma unroll 1
  for (int s = 0; s < KS2; ++s) {
    const f32x4 h0 = gelu4(tile<KS1>(W1, 2 * s, in, v, q));
    const f32x4 h1 = (2 * s + 1 < NH) ? gelu4(tile<KS1>(W1, 2 * s + 1, in, v, q)) : zero4();
    const Frag3 hf = split8_3(h0, h1);
#pragma unroll
    for (int t = 0; t < NT2; ++t) {
      const int o = (16 * t + v) * wld(KS2 * 32) + 32 * s + 8 * q;
      out[t] = mfma6(*(const bf16x8*)(W2.h + o), *(const bf16x8*)(W2.m + o), *(const bf16x8*)(W2.l + o), hf, out[t]);
    }
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

// sum over the 16 lanes of a row (the 16 voxels of a lane group): four DPP adds, xor 1, xor 2 (quad permutes), then
// the half-row and row mirrors; every lane of the row ends with the row's sum
__device__ __forceinline__ float row16_sum(float x) {
  x += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0xB1, 0xF, 0xF, true));    // quad_perm 1,0,3,2
  x += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0x4E, 0xF, 0xF, true));    // quad_perm 2,3,0,1
  x += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0x141, 0xF, 0xF, true));   // row_half_mirror
  x += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0x140, 0xF, 0xF, true));   // row_mirror
  return x;
}

// A pair of per-(sample, GroupNorm group) sums (S1, S2) accumulated by the kernel that produces one of their factors,
// so that no separate pass re-reads it. Each tile's per-lane sums are reduced over the 16 voxel lanes at once and
// added to the wave's own LDS slot [wave][lane group][channel group of the quad][2] (nothing held in registers across
// tiles); flush() merges the slots into the workgroup's partials [8][2], the layout the finalize kernels reduce.
template <int CT>
struct GroupSums {
  static constexpr int CPG = CT * 16 / G;
  static constexpr int SUB = CPG >= 4 ? 1 : 4 / CPG;   // channel groups per 4-channel quad
  static constexpr int SLOT = CT * SUB * 2;            // floats per (wave, lane group)
  float* slot;                                         // this wave's and lane group's slot in LDS

  // every thread of the workgroup calls this once, before its first tile; red holds NW * 4 * SLOT floats
  template <int NW>
  __device__ void init(float* red, int q, int wave) {
    for (int i = threadIdx.x; i < NW * 4 * SLOT; i += blockDim.x) red[i] = 0.f;
    slot = red + (wave * 4 + q) * SLOT;
    __syncthreads();
  }
  // a, c: this lane's sums over its own 4 channels of each tile, index t * SUB + (channel group within the quad)
  __device__ __forceinline__ void put(float (&a)[CT * SUB], float (&c)[CT * SUB], int v) {
#pragma unroll
    for (int i = 0; i < CT * SUB; ++i) {
      a[i] = row16_sum(a[i]);
      c[i] = row16_sum(c[i]);
    }
    if (v == 0)
#pragma unroll
      for (int i = 0; i < CT * SUB; ++i) {
        slot[2 * i] += a[i];
        slot[2 * i + 1] += c[i];
      }
  }
  // every thread of the workgroup calls this once, after its last tile
  template <int NW>
  __device__ void flush(const float* red, float* partials) {
    __syncthreads();
    if (threadIdx.x < G) {
      const int g = threadIdx.x;
      float s1 = 0.f, s2 = 0.f;
      for (int w = 0; w < NW; ++w)
        for (int qq = 0; qq < 4; ++qq)
          for (int i = 0; i < CT * SUB; ++i) {
            const int ch = 16 * (i / SUB) + 4 * qq + (i % SUB) * CPG;
            if (ch / CPG != g) continue;
            s1 += red[(w * 4 + qq) * SLOT + 2 * i];
            s2 += red[(w * 4 + qq) * SLOT + 2 * i + 1];
          }
      partials[g * 2 + 0] = s1;
      partials[g * 2 + 1] = s2;
    }
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
    f32x4 base = in.base ? ld4(in.base + row * C + c0, ok) : zero4();
    if (in.t) {
      f32x4 tv = ld4(in.t + row * C + c0, ok);
      f32x4 a = *(const f32x4*)(in.aff_a + b * C + c0);
      f32x4 s = *(const f32x4*)(in.aff_b + b * C + c0);
      f32x4 u = tv * a + s;
      if (in.act_gelu) u = gelu4(u);
      base += scale * u;
    }
    x[t] = ok ? base : zero4();
  }
}

// An assembly's affine (a, b of its batch item) staged into LDS kc[2][C] by the whole workgroup (the caller then
// synchronises), and the assembly read from there: x = base + scale act(t a + b). Rows outside the batch item give
// finite values that reach only their own, unstored, voxel column.
template <int CT>
__device__ __forceinline__ void stage_affine(const FfnoAssemble& in, int b, float* kc) {
  constexpr int C = CT * 16;
  for (int c = threadIdx.x; c < C; c += blockDim.x) {
    kc[c] = in.t ? in.aff_a[b * C + c] : 0.f;
    kc[C + c] = in.t ? in.aff_b[b * C + c] : 0.f;
  }
}

template <int CT>
__device__ __forceinline__ void assemble_c(const FfnoAssemble& in, const float* kc, float scale, int64_t row, bool ok,
                                           int q, f32x4 (&x)[CT]) {
  constexpr int C = CT * 16;
#pragma unroll
  for (int t = 0; t < CT; ++t) {
    const int c0 = 16 * t + 4 * q;
    f32x4 base = in.base ? ld4(in.base + row * C + c0, ok) : zero4();
    if (in.t) {
      f32x4 u = ld4(in.t + row * C + c0, ok) * *(const f32x4*)(kc + c0) + *(const f32x4*)(kc + C + c0);
      if (in.act_gelu) u = gelu4(u);
      base += scale * u;
    }
    x[t] = base;
  }
}

template <int CT>
__device__ __forceinline__ void gelu_tiles(f32x4 (&x)[CT]) {
#pragma unroll
  for (int t = 0; t < CT; ++t) x[t] = gelu4(x[t]);
}

// ---------------------------------------------------------------------------------------------------------------
// weights: packed [2][rows][wld(K)] (hi then lo) in global memory; staged into LDS when the kernel's set fits
// ---------------------------------------------------------------------------------------------------------------
__device__ __forceinline__ Wm stage_w(bf16*& lds, const uint16_t* g, int n) {   // n = rows * wld(K) of one half
  stage(lds, g, 2 * n);
  Wm w{lds, lds + n};
  lds += 2 * n;
  return w;
}

__device__ __forceinline__ Wm global_w(const uint16_t* g, int n) {
  return Wm{reinterpret_cast<const bf16*>(g), reinterpret_cast<const bf16*>(g) + n};
}

// a three-plane weight [3][rows][wld(K)]: its first NL planes (0, 2 or 3) staged into LDS, the rest read from global
template <int NL>
__device__ __forceinline__ Wm3 stage_w3(bf16*& lds, const uint16_t* g, int n) {
  const bf16* gb = reinterpret_cast<const bf16*>(g);
  if (NL == 0) return Wm3{gb, gb + n, gb + 2 * n};
  stage(lds, g, NL * n);
  const Wm3 w{lds, lds + n, NL == 3 ? lds + 2 * n : gb + 2 * n};
  lds += NL * n;
  return w;
}

__device__ __forceinline__ Wm3 global_w3(const uint16_t* g, int n) {
  const bf16* gb = reinterpret_cast<const bf16*>(g);
  return Wm3{gb, gb + n, gb + 2 * n};
}

// the two split forms behind one name, for kernels that run either (THREE: the forward's three-way products)
template <bool THREE>
struct SplitT {
  using F = Frag;
  using W = Wm;
  static constexpr int P = 2;
  __device__ static F split(f32x4 a, f32x4 b) { return split8(a, b); }
  __device__ static W global(const uint16_t* g, int n) { return global_w(g, n); }
};
template <>
struct SplitT<true> {
  using F = Frag3;
  using W = Wm3;
  static constexpr int P = 3;
  __device__ static F split(f32x4 a, f32x4 b) { return split8_3(a, b); }
  __device__ static W global(const uint16_t* g, int n) { return global_w3(g, n); }
};

// LDS planes per three-plane weight for a kernel whose weights hold `elems` elements per plane next to `extra`
// bytes of other LDS use: 3 when everything fits a CU's 160 KB, else 2, else 0
inline int lds_planes(size_t elems, size_t extra) {
  if (3 * elems * 2 + extra <= 160 * 1024) return 3;
  if (2 * elems * 2 + extra <= 160 * 1024) return 2;
  return 0;
}

#define FFNO_DISPATCH_NL(nl, ...)                          \
  switch (nl) {                                            \
    case 3: { constexpr int NL = 3; __VA_ARGS__; break; }  \
    case 2: { constexpr int NL = 2; __VA_ARGS__; break; }  \
    default: { constexpr int NL = 0; __VA_ARGS__; break; } \
  }

__device__ __forceinline__ int64_t imap(int64_t i, int64_t split, int64_t hi, int64_t lo) {
  return (i / split) * hi + (i % split) * lo;
}


// error reporting of the C ABI (ffno_forward.hip): a thread-local message and a negative return code
int fail(const char* msg);
int check_launch(const char* what);

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

inline bool rows_ok(const FfnoRows& r) { return r.B > 0 && r.voxels > 0 && r.nwg > 0 && r.C % 16 == 0; }

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

// waves per workgroup of the LDS-staged row kernels: 8 (two per SIMD sharing one copy of the staged weights) while
// a wave's registers fit 256; wider channel counts keep 4
template <int CT>
constexpr int nw_of() { return CT <= 4 ? 8 : 4; }
// the same for kernels whose waves fit 128 VGPRs (the forward's three-way row kernels: 4 waves per SIMD) or 170
// (the MLP and input-gate adjoints: 3 per SIMD); their staged weights allow one workgroup per CU
template <int CT>
constexpr int nw_fwd() { return CT <= 4 ? 16 : 4; }
template <int CT>
constexpr int nw_bwd() { return CT <= 4 ? 12 : 4; }

template <int CT, int NW>
constexpr size_t red_bytes_nw() { return (size_t)NW * 4 * CT * GnAcc<CT>::SUB * 3 * sizeof(float); }

}  // namespace ffno
