// The level-fraction output on given head channels, stand alone: FFNO3D.head_output(head, x_in)
// (models/ffno_model.py:730-749) and its adjoint. One thread per voxel; every channel plane is read and written with
// consecutive voxels on consecutive lanes.
#include "ffno_device.h"
#include "ffno_head.h"

namespace ffno {
namespace {

__global__ __launch_bounds__(256) void k_head_output(int64_t voxels, int n_head, int n_out,
                                                     const float* __restrict__ head, const float* __restrict__ x_in,
                                                     FfnoHeadDesc hd, float* __restrict__ y) {
  const int b = blockIdx.y;
  const float* hb = head + (int64_t)b * n_head * voxels;
  const float* xb = x_in + (int64_t)b * 6 * voxels;
  float* yb = y + (int64_t)b * n_out * voxels;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < voxels; v += (int64_t)gridDim.x * blockDim.x) {
    auto feat = [&](int c) { return xb[c * voxels + v] * hd.std_x[c] + hd.mean_x[c]; };
    level_fractions(hd, [&](int c) { return hb[c * voxels + v]; }, feat(hd.t_channel), feat(hd.ne_channel),
                    feat(hd.rho_channel), [&](int o, float val) { yb[o * voxels + v] = val; });
  }
}

__global__ __launch_bounds__(256) void k_head_output_bwd(int64_t voxels, int n_head, int n_out,
                                                         const float* __restrict__ head,
                                                         const float* __restrict__ x_in, FfnoHeadDesc hd,
                                                         const float* __restrict__ ybar, float* __restrict__ head_bar,
                                                         float* __restrict__ x_bar) {
  const int b = blockIdx.y;
  const float* hb = head + (int64_t)b * n_head * voxels;
  const float* xb = x_in + (int64_t)b * 6 * voxels;
  const float* yb = ybar + (int64_t)b * n_out * voxels;
  float* hbb = head_bar + (int64_t)b * n_head * voxels;
  float* xbb = x_bar + (int64_t)b * 6 * voxels;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < voxels; v += (int64_t)gridDim.x * blockDim.x) {
    auto feat = [&](int c) { return xb[c * voxels + v] * hd.std_x[c] + hd.mean_x[c]; };
    float g_lt = 0.f, g_lne = 0.f, g_rho = 0.f;
    level_fractions_bwd(
        hd, [&](int c) { return hb[c * voxels + v]; }, feat(hd.t_channel), feat(hd.ne_channel),
        [&](int o) { return yb[o * voxels + v]; }, [&](int c, float val) { hbb[c * voxels + v] = val; }, g_lt, g_lne,
        g_rho);
    // the normalised inputs: log10 T and log10 ne enter as ln = log10 * ln 10, log10 rho directly
    for (int c = 0; c < 6; ++c) {
      float g = 0.f;
      if (c == hd.t_channel) g += g_lt * LN10;
      if (c == hd.ne_channel) g += g_lne * LN10;
      if (c == hd.rho_channel) g += g_rho;
      xbb[c * voxels + v] = g * hd.std_x[c];
    }
  }
}

dim3 grid_of(int B, int64_t voxels) {
  const int64_t blocks = (voxels + 255) / 256;
  return dim3((unsigned)(blocks < 4096 ? blocks : 4096), (unsigned)B);
}

int check_desc(const FfnoHeadDesc* desc, const char* what) {
  if (desc->t_channel < 0 || desc->t_channel >= 6 || desc->ne_channel < 0 || desc->ne_channel >= 6 ||
      desc->rho_channel < 0 || desc->rho_channel >= 6)
    return fail(what);
  return 0;
}

}  // namespace
}  // namespace ffno

using namespace ffno;

extern "C" {

int ffno_head_output(int B, int64_t voxels, const float* head, const float* x_in, const FfnoHeadDesc* desc, float* y,
                     hipStream_t stream) {
  int n_out = 0;
  const int n_head = head_channels(desc, &n_out);
  if (B < 1 || voxels < 1 || !head || !x_in || !y || n_head < 0) return fail("ffno_head_output: bad arguments");
  if (check_desc(desc, "ffno_head_output: input channel index outside 0..5")) return -1;
  k_head_output<<<grid_of(B, voxels), 256, 0, stream>>>(voxels, n_head, n_out, head, x_in, *desc, y);
  return check_launch("ffno_head_output");
}

int ffno_head_output_bwd(int B, int64_t voxels, const float* head, const float* x_in, const FfnoHeadDesc* desc,
                         const float* ybar, float* head_bar, float* x_bar, hipStream_t stream) {
  int n_out = 0;
  const int n_head = head_channels(desc, &n_out);
  if (B < 1 || voxels < 1 || !head || !x_in || !ybar || !head_bar || !x_bar || n_head < 0)
    return fail("ffno_head_output_bwd: bad arguments");
  if (check_desc(desc, "ffno_head_output_bwd: input channel index outside 0..5")) return -1;
  k_head_output_bwd<<<grid_of(B, voxels), 256, 0, stream>>>(voxels, n_head, n_out, head, x_in, *desc, ybar, head_bar,
                                                            x_bar);
  return check_launch("ffno_head_output_bwd");
}

}  // extern "C"
