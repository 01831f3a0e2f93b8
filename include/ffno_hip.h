/* C ABI of the FFNO3D HIP library (libffno_hip.so), gfx950.
 *
 * The reference has no native code: its FFNO3D forward (models/ffno_model.py:800-832) is a chain of stock torch
 * ops. Each entry point below launches one hand-written kernel that replaces a span of that chain; the comment on
 * each names the reference lines it stands for. Host code (ffnoml_amd/engine.py) owns every device buffer, decides the
 * launch order and passes plain device pointers, sizes and a hipStream_t. Nothing here allocates device memory.
 *
 * Layouts (all float32 unless noted):
 *   activations    channel-last [B, D, H, W, C]  ("voxel rows" of C channels)
 *   model inputs   channel-first [B, 6, D, H, W], outputs channel-first [B, Cout, D, H, W] (the reference's)
 *   packed weights bf16 (uint16_t) [N_pad][K_pad], K permuted within every 32-block so that one 16-byte load is
 *                  one MFMA operand (see ffnoml_amd/engine.py: pack_mlp_weight)
 *   twiddles       bf16 pairs (hi, lo) [M_pad][K_pad], the split-bf16 form of a float64 matrix
 *   GroupNorm      partial statistics [B][nwg][G][3] = (weight, mean, M2); finalized stats [B][G][3] = (mean, rstd, W);
 *                  per-channel affine [B][C] scale and [B][C] shift
 *
 * Every call returns 0 on success or a negative code; ffno_last_error() gives the message (thread-local).
 */
#ifndef FFNO_HIP_H
#define FFNO_HIP_H

#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

const char* ffno_last_error(void);
int ffno_version(void);

/* Voxel-row sizes shared by the pointwise kernels. */
typedef struct {
  int B;              /* samples */
  int64_t voxels;     /* D*H*W per sample */
  int C;              /* width: 16, 32, 64, 96 or 128 */
  int nwg;            /* workgroups per sample (grid = nwg x B) */
} FfnoRows;

/* GroupNorm statistics output of a producer kernel (NULL partials: no statistics). */
typedef struct {
  const float* depth_w;   /* [B, D*H*W] depth_weights (models/ffno_model.py:67-74) */
  float* partials;        /* [B][nwg][8][3] */
} FfnoStatsOut;

/* "base + scale * act(A * t + Bshift)" input assembly from a finalized GroupNorm (affine per (b, c)). */
typedef struct {
  const float* base;      /* [N, C] or NULL */
  const float* t;         /* [N, C] the normalised tensor's input */
  const float* aff_a;     /* [B][C] */
  const float* aff_b;     /* [B][C] */
  const float* scale;     /* device scalar or NULL (1) */
  int act_gelu;           /* 1: GELU after the affine, 0: none */
} FfnoAssemble;           /* x = base + scale * act(aff_a * t + aff_b); base NULL: 0; t NULL: x = base */

/* Lift: Conv3d 6->C then GroupNorm statistics (models/ffno_model.py:686-690, :782-783). */
int ffno_lift(FfnoRows rows, const float* x_in, const float* w_lift /* [C][6] */, float* out, FfnoStatsOut stats,
              hipStream_t stream);

/* GroupNorm finalize: merge partial statistics, produce stats [B][8][3] = (mean, rstd, total weight) and the
 * per-channel affine (models/ffno_model.py:77-95, eps 1e-5). gamma/beta [C]. */
int ffno_gn_finalize(int B, int nwg, int C, const float* partials, const float* gamma, const float* beta,
                     float* stats, float* aff_a, float* aff_b, hipStream_t stream);

/* Block entry (models/ffno_model.py:536-542, :266, :439-440): x = assemble(...); writes x, then both input gates:
 * xs = x * sigmoid(G2 gelu(G1 x)), xv = x * sigmoid(V2 gelu(V1 x)) + z_proj * z. x_out may be NULL, or alias
 * in.base; xs may alias in.t (each row is read before it is written), or be NULL (xv alone). */
int ffno_block_in(FfnoRows rows, FfnoAssemble in, float* x_out, const uint16_t* g1, const uint16_t* g2,
                  const uint16_t* v1, const uint16_t* v2, const float* z_proj /* [C] */, const float* z /* [B,DHW] */,
                  float* xs, float* xv, hipStream_t stream);

/* Two-layer pointwise MLP y = W2 gelu(W1 u) with GroupNorm statistics of y: the post networks of both branches
 * (models/ffno_model.py:221-225, :369-373, :309, :449). y may alias u. */
int ffno_post(FfnoRows rows, const float* u, const uint16_t* w1, const uint16_t* w2, float* y, FfnoStatsOut stats,
              hipStream_t stream);

/* ffno_post followed by ffno_gn_finalize of its output, allocating nothing: the partial statistics live in a static
 * device array of the library (FFNO_NOALLOC_MAX_SLOTS = 4096 slots of [8][3] floats, 384 KiB per GPU, held until
 * the process ends) instead of a caller's workspace. It has no counterpart in the reference; it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Writes y, stats [B][8][3] and the affine aff_a, aff_b [B][C] exactly as ffno_post + ffno_gn_finalize
 * do. Fails when rows.B * rows.nwg exceeds FFNO_NOALLOC_MAX_SLOTS. All calls on one GPU share its array: calls that
 * overlap on one GPU (concurrent streams or threads) race and must not be issued. */
#define FFNO_NOALLOC_MAX_SLOTS 4096
int ffno_post_gn_noalloc(FfnoRows rows, const float* u, const uint16_t* w1, const uint16_t* w2, float* y,
                         const float* depth_w, const float* gamma, const float* beta, float* stats, float* aff_a,
                         float* aff_b, hipStream_t stream);

/* Branch fusion (models/ffno_model.py:524-534, :541-543): s = gelu(GN(sp)), v = gelu(GN(vp)),
 * gates = sigmoid(BG2 gelu(BG1 [x,s,v]) + b), fp = F2 gelu(F1 [s*gs, v*gv]) + 0.5 (s*gs + v*gv), statistics of fp.
 * fp may alias sp. */
int ffno_fuse(FfnoRows rows, const float* x, const float* sp, const float* vp, const float* s_a, const float* s_b,
              const float* v_a, const float* v_b, const uint16_t* bg1, const uint16_t* bg2, const float* bg2_bias,
              const uint16_t* f1, const uint16_t* f2, float* fp, FfnoStatsOut stats,
              float* gates_out /* optional [N, 2C]: (gs, gv) for the two-pass adjoint */, hipStream_t stream);

/* Residual pointwise path (models/ffno_model.py:545-551): x_out = assemble(...); y = W2 gelu(W1 x_out); statistics
 * of y. Used for pw (assemble = x + res_fused gelu(GN(fp))) and mlp (assemble = x1 + res_pw GN(pp)). x_out may be
 * NULL or alias in.base. */
int ffno_resmlp(FfnoRows rows, FfnoAssemble in, float* x_out, const uint16_t* w1, const uint16_t* w2, float* y,
                FfnoStatsOut stats, hipStream_t stream);

/* Head (models/ffno_model.py:790-798, :730-749, :116-141): x = assemble(...); head = P2 gelu(P1 x) + bias;
 * per species log10 n = log10 rho - log10 m + 6 tanh(u/6) + log_softmax(head + ln f*(T, ne)) / ln 10, normalised
 * with mean_Y, std_Y. Writes y channel-first [B, Cout, D, H, W]. */
typedef struct {
  int n_species;
  int levels[4];
  float log10_mass[4];
  float log_g[32], ev[32], ndxi[32], dion[32];   /* per output level, concatenated over species */
  float mean_x[6], std_x[6];
  float mean_y[32], std_y[32];
  int t_channel, ne_channel, rho_channel;
  float total_bound;
  float ln_ccon;      /* ln(0.5 (h / sqrt(2 pi m_e) / sqrt(k))^3), cgs (models/ffno_model.py:113) */
  float ln_k;         /* ln k_B, cgs */
  float e_over_k;     /* eV / k_B [K] */
  float dxi_coef;     /* 4.98e-4 * 5040 */
} FfnoHeadDesc;

/* x_out (optional, may alias in.base) receives the assembled head input; y NULL: only that is done. */
int ffno_head(FfnoRows rows, FfnoAssemble in, float* x_out, const uint16_t* p1, const uint16_t* p2,
              const float* p2_bias, const float* x_in /* [B,6,DHW] */, const FfnoHeadDesc* desc, float* y,
              hipStream_t stream);

/* FFNO3D.head_output(head, x_in) (models/ffno_model.py:730-749) on given head channels: per species the MULTI3D LTE
 * log-weights of T and ne (:116-141), log10 n = log10 rho - log10 m + b tanh(u / b) + log_softmax(head + ln f*) /
 * ln 10, normalised by mean_Y, std_Y; float32. head [B, Cout + S, DHW], x_in [B, 6, DHW] (normalised inputs), y
 * [B, Cout, DHW], all channel-first. */
int ffno_head_output(int B, int64_t voxels, const float* head, const float* x_in, const FfnoHeadDesc* desc, float* y,
                     hipStream_t stream);

/* Adjoint of ffno_head_output (the backward the reference's autograd runs through :730-749): head_bar [B, Cout + S,
 * DHW] and x_bar [B, 6, DHW] (the T, ne and rho channels through the LTE weights and the total; the others 0) from
 * ybar [B, Cout, DHW]. head_bar may alias head. */
int ffno_head_output_bwd(int B, int64_t voxels, const float* head, const float* x_in, const FfnoHeadDesc* desc,
                         const float* ybar, float* head_bar, float* x_bar, hipStream_t stream);

/* Split-bf16 GEMM of a small precomputed real matrix A [m x k] against batched data X [k x n]:
 * Y[b][m][n] = sum_k A[m][k] X[b][k][n]. It carries every DFT (rfft/ifft/irfft along W and H) and the vertical
 * Filon transform and its inverse (models/ffno_model.py:441, :448, :272-275, :303-305). Row offsets come from
 * two-level index maps: offset(i) = (i / split) * hi + (i % split) * lo. Two planes of A (a_hi, a_lo): three
 * products per MFMA pair (~2^-17 relative); with a third plane a_l3, A and X are split three ways and the six
 * products of order <= 2 are summed (float32 level, the forward's transforms). */
typedef struct {
  const uint16_t* a_hi;
  const uint16_t* a_lo;
  int m, k, n;            /* n % 4 == 0 */
  int m_pad, k_pad;       /* multiples of 16 and 32 */
  int batch;
  const float* x;
  float* y;
  int64_t b_split, x_b_hi, x_b_lo, y_b_hi, y_b_lo;
  int64_t k_split, x_k_hi, x_k_lo;
  int64_t m_split, y_m_hi, y_m_lo;
  int accumulate;         /* 1: Y += A X */
  const uint16_t* a_l3;   /* third plane of A or NULL */
  /* Optional twiddle of a factorised DFT (the first of its two stages; NULL: none). A's rows are then [real part
   * rows 0..tw_n1-1, zero rows up to m_split, imaginary part rows m_split..m_split+tw_n1-1], m_split = m_pad / 2 a
   * multiple of 16, m_pad <= 128, k_pad <= 128, no accumulate. Output row pair (k1, m_split + k1) of column j of
   * batch item b is multiplied, as one complex number, by twiddle[k1 * tw_n2 + t] (complex float pairs), t =
   * j / tw_cols when tw_cols > 0, else b % tw_bsplit; rows k1 >= tw_n1 of either part are not written. */
  const float* twiddle;
  int tw_n1, tw_n2;
  int64_t tw_cols;        /* columns per twiddle column group, % 4 == 0, or 0 */
  int64_t tw_bsplit;      /* with tw_cols == 0: batch items per twiddle column cycle */
} FfnoTwGemm;

int ffno_tw_gemm(const FfnoTwGemm* g, hipStream_t stream);

/* The vertical transform on per-column height grids (models/ffno_model.py:277-283, :304-305, the reference's
 * per-column path): each (sample, h, w) column j has its own split-bf16 matrix A_j [m x k] (planes a_hi, a_lo and
 * optionally a_l3, each [B*hw][m_pad][k_pad]); Y[b][m][hw][c] = sum_k A_j[m][k] X[b][k][hw][c]. */
typedef struct {
  const uint16_t* a_hi;
  const uint16_t* a_lo;
  const uint16_t* a_l3;   /* third plane or NULL (two-way) */
  int m, k, m_pad, k_pad;
  int B;
  int64_t hw;             /* columns per sample (H*W) */
  int C;                  /* channels, % 4 == 0 */
  const float* x;         /* [B][k][hw][C] */
  float* y;               /* [B][m][hw][C] */
} FfnoColGemm;

int ffno_col_gemm(const FfnoColGemm* g, hipStream_t stream);

/* Per-mode complex channel mix of a spectrum held as separate real and imaginary planes, in place
 * (models/ffno_model.py:394-429 for the horizontal branch, :286-302 for the vertical one):
 * Z = g(k) (Y (W_r + i W_i)) + U(k) (V(k) Y), times an optional per-mode factor. */
typedef struct {
  float* re;
  float* im;
  int C, rows, modes, rank;
  int64_t mode_stride, r_split, r_hi, r_lo;
  const uint16_t* wc;     /* packed [parts][2C][2C] */
  const float* gate;      /* [modes][2] */
  const uint16_t* v_p;    /* [modes][parts][16][KS*32] or NULL */
  const uint16_t* u_p;    /* [modes][parts][C][32] or NULL */
  const float* factor;    /* [modes] or NULL */
  int parts;              /* planes of wc, v_p, u_p: 2, or 3 for the three-way products of the forward */
} FfnoMix;

int ffno_mix(const FfnoMix* m, hipStream_t stream);

/* The per-mode low-rank tables ffno_mix reads, from the hidden features hid [modes][hidden] of the frequency low-rank
 * network (models/ffno_model.py:404-429): u(k) = freq_u(h) as [C][rank] and v(k) = freq_v(h) as [rank][C] in
 * float32 (wu_t, wv_t: the two Linear weights transposed, [hidden][C rank]; bu, bv their biases), split into
 * `parts` bf16 planes and packed: forward (adjoint = 0, parts 3) v_out [modes][parts][16][KC + 8] from v and
 * u_out [modes][parts][C][40] from u (KC = C rounded up to 32); adjoint (adjoint = 1, parts 2) the same shapes from
 * u^T and v^T. Padding entries are left as they are (the caller zeroes the outputs). hidden <= 256, a multiple
 * of 4; rank <= 16. Replaces the table build of SpectralConv2dFull / HeightSpectralConv's mix (:394-429, :285-296). */
int ffno_lowrank_tables(int64_t modes, int C, int rank, int hidden, const float* hid, const float* wu_t,
                        const float* bu, const float* wv_t, const float* bv, int parts, int adjoint, uint16_t* v_out,
                        uint16_t* u_out, hipStream_t stream);

/* ------------------------------------------------------------------------------------------------------------
 * Adjoint (input VJP). The reference differentiates its forward with torch autograd through non-reentrant
 * checkpoints (models/ffno_model.py:751-798; ffno_fsdp_runtime.py:306-348). Here each block is recomputed from its
 * saved input and differentiated by the kernels below; they return input gradients only.
 * ------------------------------------------------------------------------------------------------------------ */

/* One depth-weighted GroupNorm seen from its output gradient u_bar: g_hat = scale * u_bar * gamma (* GELU'(n) when
 * the GN output went through GELU, n = a*t + b the forward affine), x_hat = (t - mean) * rstd. The backward needs
 * S1 = sum g_hat and S2 = sum g_hat * x_hat per (sample, group), then
 * t_bar = rstd * (g_hat - (w / W) * (S1 + x_hat * S2)), w the voxel's depth weight and W the group's total weight. */
typedef struct {
  const float* t;         /* GN input [N, C] */
  const float* stats;     /* [B][8][3] (mean, rstd, W) from ffno_gn_finalize */
  const float* gamma;     /* [C] */
  const float* aff_a;     /* [B][C] forward affine (used when act_gelu) */
  const float* aff_b;
  const float* scale;     /* device scalar or NULL (1) */
  int act_gelu;
  const float* sums;      /* [B][8][2] (S1, S2) from ffno_gn_grad_finalize; unused by ffno_gn_grad_stats */
  const float* depth_w;   /* [B, D*H*W] */
  float* wbar;            /* [B, D*H*W] or NULL: accumulates the gradient with respect to the depth weights,
                             -(1/W) sum_c (S1 x_hat + S2 (x_hat^2 - rstd^2 var) / 2) */
} FfnoGnBack;

/* S1, S2 partial sums of one GroupNorm's backward: partials [B][nwg][8][2]. */
int ffno_gn_grad_stats(FfnoRows rows, FfnoGnBack gn, const float* ubar, float* partials, hipStream_t stream);
/* partials [B][nwg][8][2] -> sums [B][8][2] (fixed order, deterministic) */
int ffno_gn_grad_finalize(int B, int nwg, const float* partials, float* sums, hipStream_t stream);

/* Adjoint of y = W2 gelu(W1 x) behind a GroupNorm: t_bar = GN backward of u_bar, then
 * x_bar = [u_bar if residual] + W1^T (GELU'(W1 x) * W2^T t_bar) (models/ffno_model.py:547-551 for pw and mlp with
 * residual; :221-225, :369-373 for the branch post networks without). w2t, w1t are the transposed weights packed.
 * The MLP input x is given as an assembly (x = base, or base + scale act(a t + b)), so a residual sum need not be
 * stored. x_bar may alias u_bar or x.base (each row is read before it is written). */
int ffno_mlp_bwd(FfnoRows rows, const float* ubar, FfnoGnBack gn, FfnoAssemble x, const uint16_t* w1,
                 const uint16_t* w2t, const uint16_t* w1t, int residual, float* xbar, const FfnoGnBack* next,
                 float* next_partials, hipStream_t stream);
/* (next, next_partials): when next_partials is given, the kernel also accumulates the S1, S2 partial sums of the
 * GroupNorm whose u_bar is this call's x_bar (next describes that GroupNorm; its sums field is not read) into
 * next_partials [B][nwg][8][2], exactly what ffno_gn_grad_stats(rows, *next, xbar, next_partials) would write, so that
 * pass is skipped (models/ffno_model.py:547-551: x2_bar of the mlp is u_bar of the pw GroupNorm). */

/* Adjoint of the branch fusion (models/ffno_model.py:524-545): from x1_bar (gradient of x1 = x + res_fused * f)
 * to the gradients of the block input (including x1_bar's residual path), of s = gelu(GN_s(sp)) and of
 * v = gelu(GN_v(vp)); gates and hidden layers are recomputed. hbar may alias x1bar. */
int ffno_fuse_bwd(FfnoRows rows, const float* x1bar, FfnoGnBack gnf, const float* h, const float* sp, const float* vp,
                  const float* s_a, const float* s_b, const float* v_a, const float* v_b, const uint16_t* bg1,
                  const uint16_t* bg2, const float* bg2_bias, const uint16_t* f1, const uint16_t* f2t,
                  const uint16_t* f1t, const uint16_t* bg2t, const uint16_t* bg1t, float* hbar, float* sbar,
                  float* vbar, hipStream_t stream);

/* The same adjoint in two passes whose weights fit LDS, given the gates the recomputed ffno_fuse stored
 * (models/ffno_model.py:524-545). Pass 1: fp_bar = GN backward of x1bar (GELU after it), then
 * [smbar, vmbar] = 0.5 fp_bar + F1^T (gelu'(F1 [s gs, v gv]) * F2^T fp_bar). Pass 2 (in place): through the gates
 * and the gate network, x1bar -> hbar, smbar -> sbar, vmbar -> vbar; with partials_s and partials_v (both or
 * neither), also the gradient sums of the two branch GroupNorms gn_s, gn_v whose u_bar are sbar and vbar, as
 * ffno_gn_grad_stats would write them (reduce with ffno_gn_grad_finalize). */
int ffno_fuse_mlp_bwd(FfnoRows rows, const float* x1bar, FfnoGnBack gnf, const float* sp, const float* vp,
                      const float* s_a, const float* s_b, const float* v_a, const float* v_b, const float* gates,
                      const uint16_t* f1, const uint16_t* f2t, const uint16_t* f1t, float* smbar, float* vmbar,
                      hipStream_t stream);
int ffno_fuse_gates_bwd(FfnoRows rows, float* x1bar_hbar, const float* h, const float* sp, const float* vp,
                        const float* s_a, const float* s_b, const float* v_a, const float* v_b, const float* gates,
                        const uint16_t* bg1, const uint16_t* bg2t, const uint16_t* bg1t, float* sm_sbar, float* vm_vbar,
                        const FfnoGnBack* gn_s, float* partials_s, const FfnoGnBack* gn_v, float* partials_v,
                        hipStream_t stream);

/* Adjoint of both input gates (models/ffno_model.py:266, :439-440): hbar_out = hbar_in + d(xs)/dh^T xs_bar +
 * d(xv)/dh^T xv_bar; zbar (optional, accumulated) += sum_c xv_bar * z_proj. hbar_out may alias hbar_in. */
int ffno_gates_bwd(FfnoRows rows, const float* h, const float* xsbar, const float* xvbar, const float* hbar_in,
                   const uint16_t* g1, const uint16_t* g2, const uint16_t* g2t, const uint16_t* g1t,
                   const uint16_t* v1, const uint16_t* v2, const uint16_t* v2t, const uint16_t* v1t,
                   const float* z_proj, float* hbar_out, float* zbar, hipStream_t stream);

/* Adjoint of the head (models/ffno_model.py:730-749, :790-798) from the output gradient ybar [B, Cout, DHW]:
 * writes hbar [N, C] and the direct input gradient xbar_in [B, 6, DHW] (T, ne and rho through the LTE weights
 * and the total; the other channels 0). hbar may alias h. */
int ffno_head_bwd(FfnoRows rows, const float* h, const uint16_t* p1, const uint16_t* p2, const uint16_t* p2t,
                  const uint16_t* p1t, const float* p2_bias, const float* x_in, const FfnoHeadDesc* desc,
                  const float* ybar, float* hbar, float* xbar_in, hipStream_t stream);

/* zbar[b, d, hw] += sum_c a[b, d + delta, hw, c] * r[b, d, hw, c] (terms with d + delta outside [0, D) skipped):
 * the channel contraction of the height gradient of the vertical branch (models/ffno_model.py:272-308). */
int ffno_dot_reduce(FfnoRows rows, int D, const float* a, int delta, const float* r, float* zbar, hipStream_t stream);

/* The same contractions for the three Filon bands and the synthesis phases at once, per column of a shared height
 * grid (models/ffno_model.py:272-308), without the band GEMMs' outputs:
 *   zbar[b, d, hw] += sum_{delta=-1,0,1} sum_r G_delta[d, r] P[d + delta, r] + sum_r V_dz[d, r] Q[d, r],
 *   P[d', r] = sum_c xv[b, d', hw, c] xbar_spec[b, r, hw, c],  Q[d, r] = sum_c yv_bar[b, d, hw, c] y_spec[b, r, hw, c],
 * with tables [4][r_pad][d_pad] float32: tables[delta + 1][r][d'] = G_delta[d' - delta, r] (zero where d' - delta
 * is outside [0, D)), tables[3][r][d] = V_dz[d, r]; r_pad, d_pad multiples of 16, zero beyond R, D. */
typedef struct {
  int B;
  int64_t hw;             /* columns per batch item (H * W) */
  int D;                  /* levels */
  int R;                  /* spectrum rows, 2M */
  int C;
  int r_pad, d_pad;
  const float* xv;        /* [B, D, HW, C] the vertical branch input */
  const float* xbar_spec; /* [B, R, HW, C] the gradient of its Filon spectrum */
  const float* yv_bar;    /* [B, D, HW, C] the gradient of the branch output */
  const float* y_spec;    /* [B, R, HW, C] the spectrum after the mix */
  const float* tables;
  float* zbar;            /* [B, D, HW], accumulated */
} FfnoZGrad;
int ffno_zgrad_columns(const FfnoZGrad* g, hipStream_t stream);

/* Adjoint of the lift (models/ffno_model.py:782-783): xbar_in += W_lift^T (GN backward of ubar). */
int ffno_lift_bwd(FfnoRows rows, const float* ubar, FfnoGnBack gn, const float* w_lift, float* xbar_in,
                  hipStream_t stream);

/* ------------------------------------------------------------------------------------------------------------
 * Tangent (forward-mode JVP along a tangent of the input features). The reference computes it with torch forward-AD
 * duals through its forward with checkpointing bypassed (ffno_fsdp_runtime.py:61-73, :350-372). Here the primal
 * forward of each block runs first (the kernels above, keeping the block's intermediates), then these kernels carry
 * the tangent through the same block; the spectral transforms and the mix are linear and run on the tangent as they
 * are. The heights carry no tangent (jvp_local perturbs the features only).
 * ------------------------------------------------------------------------------------------------------------ */

/* Tangent of "base + scale * act(a t + b)" where t is a depth-weighted GroupNorm's input: with the GroupNorm
 * tangent n_dot = a t_dot + tan_b t + tan_c (ffno_gn_tangent_finalize), out = base_dot + scale * act'(a t + b) *
 * n_dot, elementwise. */
typedef struct {
  const float* base_dot;  /* [N, C] or NULL */
  const float* t;         /* primal GroupNorm input [N, C] */
  const float* t_dot;     /* its tangent [N, C] */
  const float* aff_a;     /* [B][C] the forward affine (gamma rstd) */
  const float* aff_b;     /* [B][C] */
  const float* tan_b;     /* [B][C] */
  const float* tan_c;     /* [B][C] */
  const float* scale;     /* device scalar or NULL (1) */
  int act_gelu;
} FfnoAssembleT;

/* Tangent sums of a GroupNorm (models/ffno_model.py:77-95): partials [B][nwg][8][2] of S1 = sum w yd and
 * S2 = sum w (y - mean) yd, then per channel tan_b = -gamma rstd^3 S2/W and
 * tan_c = -gamma rstd S1/W + gamma rstd^3 (S2/W) mean (stats from ffno_gn_finalize, gamma [C]). */
int ffno_gn_tangent_stats(FfnoRows rows, const float* y, const float* yd, const float* stats, const float* depth_w,
                          float* partials, hipStream_t stream);

/* The same tangent sums accumulated by the kernel that produces yd (ffno_mlp_t, ffno_fuse_t), so that no separate
 * pass re-reads it: y the GroupNorm's primal input, partials laid out as ffno_gn_tangent_stats writes them (NULL: no
 * sums). */
typedef struct {
  const float* y;         /* [N, C] */
  const float* stats;     /* [B][8][3] from ffno_gn_finalize */
  const float* depth_w;   /* [B, D*H*W] */
  float* partials;        /* [B][nwg][8][2] or NULL */
} FfnoTanStatsOut;
int ffno_gn_tangent_finalize(int B, int nwg, int C, const float* partials, const float* stats, const float* gamma,
                             float* tan_b, float* tan_c, hipStream_t stream);
int ffno_assemble_t(FfnoRows rows, FfnoAssembleT a, float* out, hipStream_t stream);

/* Lift tangent (models/ffno_model.py:782-783): yd = W_lift xd_in, xd_in channel-first [B, 6, DHW]. */
int ffno_lift_t(FfnoRows rows, const float* xd_in, const float* w_lift, float* yd, hipStream_t stream);

/* Input-gate tangents (models/ffno_model.py:266, :439-440): from the block input h and its tangent hd,
 * xsd = hd g + h g (1 - g) a_dot with g = sigmoid(a), a = G2 gelu(G1 h), a_dot = G2 (gelu'(G1 h) * G1 hd); xvd the
 * same with V1, V2. */
int ffno_gates_t(FfnoRows rows, const float* h, const float* hd, const uint16_t* g1, const uint16_t* g2,
                 const uint16_t* v1, const uint16_t* v2, float* xsd, float* xvd, hipStream_t stream);

/* Tangent of y = W2 gelu(W1 x) (the post networks, pw and mlp: models/ffno_model.py:221-225, :369-373, :545-551):
 * x assembled as the forward assembles it, xd given: yd = W2 (gelu'(W1 x) * W1 xd); with ts.partials, also the
 * tangent sums of the GroupNorm that y feeds. */
int ffno_mlp_t(FfnoRows rows, FfnoAssemble in, const float* xd, const uint16_t* w1, const uint16_t* w2, float* yd,
               FfnoTanStatsOut ts, hipStream_t stream);

/* Branch-fusion tangent (models/ffno_model.py:524-534, :541-543) from h, hd, the gates the forward stored [N, 2C]
 * and the two branch GroupNorms as tangent assemblies: s = gelu(GN_s(sp)) with its tangent
 * sd = gelu'(a sp + b) (a sp_dot + tan_b sp + tan_c) (st.t = sp, st.t_dot = sp_dot, st.act_gelu = 1, no base or
 * scale), likewise v from vt: fpd = F2 (gelu'(F1 [sm, vm]) * F1 [smd, vmd]) + 0.5 (smd + vmd); with ts.partials, also
 * the tangent sums of the fusion GroupNorm (ts.y = fp). */
int ffno_fuse_t(FfnoRows rows, const float* h, const float* hd, FfnoAssembleT st, FfnoAssembleT vt, const float* gates,
                const uint16_t* bg1, const uint16_t* bg2, const uint16_t* f1, const uint16_t* f2, float* fpd,
                FfnoTanStatsOut ts, hipStream_t stream);

/* Head tangent (models/ffno_model.py:730-749, :790-798): normalised log10 n tangent yd [B, Cout, DHW] from the head
 * input (assembled as the forward does), its tangent xd, the model input x_in and its tangent xd_in (T, ne and rho
 * enter the LTE weights and the total directly). */
int ffno_head_t(FfnoRows rows, FfnoAssemble in, const float* xd, const uint16_t* p1, const uint16_t* p2,
                const float* p2_bias, const float* x_in, const float* xd_in, const FfnoHeadDesc* desc, float* yd,
                hipStream_t stream);

#ifdef __cplusplus
}
#endif
#endif
