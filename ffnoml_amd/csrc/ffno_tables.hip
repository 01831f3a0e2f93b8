// Per-mode low-rank tables of the spectral mix, built on the device: the last two layers of the reference's
// frequency low-rank network, u(k) = freq_u(h(k)) viewed [C][r] and v(k) = freq_v(h(k)) viewed [r][C]
// (models/ffno_model.py:404-429), written straight into the packed split-bf16 form ffno_mix reads. One float32 dot
// product of length `hidden` per table entry (the same arithmetic as a float32 linear layer), then the bf16 planes
// w = p0 + p1 (+ p2), each remainder exact in float32.
#include "ffno_device.h"

namespace ffno {
namespace {

constexpr int TB_MODES = 32;       // modes per workgroup tile (rows of the hidden-feature tile in LDS)
constexpr int TB_THREADS = 256;    // table entries per workgroup tile (one per thread)
constexpr int TB_HIDDEN = 256;     // largest hidden width staged

// column of logical k in a packed row: k = 32 s + 16 h + 4 q + j is stored at 32 s + 8 q + 4 h + j (the MFMA lane
// layout; the host packer does the same)
__device__ __forceinline__ int packed_col(int k) {
  const int s = k >> 5, rem = k & 31;
  return 32 * s + 8 * ((rem & 15) >> 2) + 4 * (rem >> 4) + (rem & 3);
}

// grid (ceil(2 C r / TB_THREADS), ceil(modes / TB_MODES)); thread t of tile (x, y) computes entry
// e = x * TB_THREADS + t (u entries first, then v) for the TB_MODES modes of tile y
__global__ __launch_bounds__(TB_THREADS) void k_lowrank_tables(int64_t modes, int C, int rank, int hidden,
                                                               const float* __restrict__ hid,
                                                               const float* __restrict__ wut,
                                                               const float* __restrict__ bu,
                                                               const float* __restrict__ wvt,
                                                               const float* __restrict__ bv, int parts, int adjoint,
                                                               uint16_t* __restrict__ v_out,
                                                               uint16_t* __restrict__ u_out) {
  __shared__ __align__(16) float hs[TB_MODES][TB_HIDDEN + 4];
  const int64_t m0 = (int64_t)blockIdx.y * TB_MODES;
  const int nm = (int)(modes - m0 < TB_MODES ? modes - m0 : TB_MODES);
  for (int i = threadIdx.x; i < TB_MODES * hidden; i += TB_THREADS) {
    const int mm = i / hidden, k = i % hidden;
    hs[mm][k] = mm < nm ? hid[(m0 + mm) * hidden + k] : 0.f;
  }
  __syncthreads();
  const int cr = C * rank;
  const int e = blockIdx.x * TB_THREADS + threadIdx.x;
  if (e >= 2 * cr) return;
  const bool is_u = e < cr;
  const int idx = is_u ? e : e - cr;
  // weights transposed on the host ([hidden][C r]), so consecutive lanes read consecutive addresses
  const float* wt = is_u ? wut : wvt;
  float acc[TB_MODES];
  const float bias = is_u ? bu[idx] : bv[idx];
#pragma unroll
  for (int mm = 0; mm < TB_MODES; ++mm) acc[mm] = bias;
  // four hidden features per step: one 16-byte LDS read (the same address on every lane) per mode and step
  for (int k = 0; k < hidden; k += 4) {
    const float w0 = wt[(int64_t)k * cr + idx], w1 = wt[(int64_t)(k + 1) * cr + idx];
    const float w2 = wt[(int64_t)(k + 2) * cr + idx], w3 = wt[(int64_t)(k + 3) * cr + idx];
#pragma unroll
    for (int mm = 0; mm < TB_MODES; ++mm) {
      const f32x4 h = *reinterpret_cast<const f32x4*>(&hs[mm][k]);
      acc[mm] = fmaf(h[3], w3, fmaf(h[2], w2, fmaf(h[1], w1, fmaf(h[0], w0, acc[mm]))));
    }
  }
  // u[c][j] (idx = c r + j), v[j][c] (idx = j C + c). The forward packs V-shaped [16][KC + 8] tables from v (row j,
  // column c) and U-shaped [C][32 + 8] ones from u (row c, column j); the adjoint swaps roles: V-shaped from u^T,
  // U-shaped from v^T.
  const int c = is_u ? idx / rank : idx % C;
  const int j = is_u ? idx % rank : idx / C;
  const bool vshape = adjoint ? is_u : !is_u;
  const int kc = ((C + 31) / 32) * 32;
  int64_t off, plane, per_mode;
  if (vshape) {
    plane = (int64_t)16 * (kc + WPAD);
    off = (int64_t)j * (kc + WPAD) + packed_col(c);
  } else {
    plane = (int64_t)C * (32 + WPAD);
    off = (int64_t)c * (32 + WPAD) + packed_col(j);
  }
  per_mode = plane * parts;
  uint16_t* out = vshape ? v_out : u_out;
  for (int mm = 0; mm < nm; ++mm) {
    uint16_t* o = out + (m0 + mm) * per_mode + off;
    float rest = acc[mm];
    for (int p = 0; p < parts; ++p) {
      const bf16 h = (bf16)rest;
      o[p * plane] = *reinterpret_cast<const uint16_t*>(&h);
      rest = rest - (float)h;
    }
  }
}

}  // namespace
}  // namespace ffno

using namespace ffno;

extern "C" {

int ffno_lowrank_tables(int64_t modes, int C, int rank, int hidden, const float* hid, const float* wu_t,
                        const float* bu, const float* wv_t, const float* bv, int parts, int adjoint, uint16_t* v_out,
                        uint16_t* u_out, hipStream_t stream) {
  if (modes < 1 || C < 1 || C % 16 || rank < 1 || rank > 16 || hidden < 4 || hidden % 4 || hidden > TB_HIDDEN ||
      !hid || !wu_t || !bu || !wv_t || !bv || (parts != 2 && parts != 3) || !v_out || !u_out)
    return fail("ffno_lowrank_tables: bad arguments");
  const int64_t tiles = (modes + TB_MODES - 1) / TB_MODES;
  if (tiles > 65535) return fail("ffno_lowrank_tables: more than 2097120 modes");
  dim3 grid((2 * C * rank + TB_THREADS - 1) / TB_THREADS, (unsigned)tiles);
  k_lowrank_tables<<<grid, TB_THREADS, 0, stream>>>(modes, C, rank, hidden, hid, wu_t, bu, wv_t, bv, parts, adjoint,
                                                    v_out, u_out);
  return check_launch("ffno_lowrank_tables");
}

}  // extern "C"
