// sia_surrogate.hpp -- learned column surrogate of the SIA diffusivity (host-side launch interface).
//
// The surrogate replaces the per-level flow-law evaluation of the SIA column
// integral for the GPBLD flow law with the default enthalpy converter
// (SIAFD.cc:620-721; flowlaws.cc:209-221; enthalpyConverter.cc:56-320).
//
// D of a staggered column is the discrete sum D = C * sum_k G_k A_k of the
// reference's trapezoid rule (SIAFD.cc:677-684), with C = 2 e theta alpha^(n-1),
// G_k = W_k p_k^n and W_k the trapezoid weights of level k. The surrogate
//   * sums the temperate levels exactly (A = A(T0) (1 + c min(omega, limit)) is
//     linear in the water fraction: no transcendental function);
//   * reduces the cold levels to 32 cells in zeta = z / thk: per cell the sums
//     of G, G x, G x^2 with x = T_pa - 263.15 K, and a physics baseline
//     Qb = (sum G) A(T_mean) (1 + (k^2 - 2k/T_mean) Var / 2), the second-order
//     expansion of A = A_0 exp(-Q / (R T)) about the cell mean;
//   * lets an MLP (34 -> 4 x 128 GELU -> 32, BF16 MFMA, FP32 accumulate) predict
//     a log-correction r_m per cell: D = C (S_temperate + sum_m Qb_m exp(r_m)).
// MLP inputs: the 32 cell-mean temperatures (the resampled enthalpy profile),
// the thickness and the surface slope |grad h|.
#pragma once

#include "sia_kernels.hpp"

namespace pism_sia {

constexpr int SG_NCELL = 32;
constexpr int SG_NIN = 34;     // 32 cell temperatures + thickness + slope
constexpr int SG_KIN = 64;     // inputs padded to two MFMA k-chunks
constexpr int SG_WIDTH = 128;
constexpr int SG_NOUT = 32;
constexpr int SG_NLAYER = 5;   // 4 hidden layers (GELU) + linear output
constexpr int SG_FSTRIDE = 36; // floats per feature record (34 used)

// parameter offsets in the flat FP32 master array (row = output unit)
constexpr int SG_W1 = 0;                                  // [128][34]
constexpr int SG_B1 = SG_W1 + SG_WIDTH * SG_NIN;          // [128]
constexpr int SG_W2 = SG_B1 + SG_WIDTH;                   // [128][128]
constexpr int SG_B2 = SG_W2 + SG_WIDTH * SG_WIDTH;
constexpr int SG_W3 = SG_B2 + SG_WIDTH;
constexpr int SG_B3 = SG_W3 + SG_WIDTH * SG_WIDTH;
constexpr int SG_W4 = SG_B3 + SG_WIDTH;
constexpr int SG_B4 = SG_W4 + SG_WIDTH * SG_WIDTH;
constexpr int SG_W5 = SG_B4 + SG_WIDTH;                   // [32][128]
constexpr int SG_B5 = SG_W5 + SG_NOUT * SG_WIDTH;         // [32]
constexpr int SG_NPARAM = SG_B5 + SG_NOUT;                // 58144

// packed inference weights: 120 A-fragments of 64 lanes x 8 bf16
constexpr int SG_NFRAG = 8 * 2 + 3 * 8 * 4 + 2 * 4;

// column status of a surrogate update
enum SgStatus : unsigned char { SG_ZERO = 0, SG_SURR = 1, SG_FALLBACK = 2 };

// model constants used by the device kernels
struct SgNorm {
  float mean[SG_NIN], inv_std[SG_NIN], lo[SG_NIN], hi[SG_NIN];
  float rscale;  // output unit -> log-correction
};

// one training sample (device), 4-byte fields
struct SgSample {
  float f[SG_FSTRIDE];     // raw features (34 used)
  float r[SG_NCELL];       // target log-corrections, 0 where the cell has no cold level
  float share[SG_NCELL];   // cell share of D (C Q_ref,m / D_ref)
  unsigned has;            // bit m: cell m has cold levels
  int col;                 // staggered column index s = 2 (i My + j) + o
  double dref;             // D of the column from the exact per-level sum
};

struct SgColumnArgs {
  const double *z, *E, *thks, *theta, *hx, *hy;
  int *err;
};

// classification + compaction of the staggered columns
struct SgListArgs {
  unsigned char *status;  // [2 Mx My]
  int *block_count;       // [nblocks]
  int *list;              // [2 Mx My] staggered indices of SG_SURR columns, in order
  int *n_list;            // device scalar
  int *counts;            // device [3]: zero, surrogate, fallback (fallback grows in the MLP pass)
  double *flux, *Dstag;   // written for SG_ZERO columns
  int *fb_list;           // [2 Mx My] SG_FALLBACK columns: the classified ones in order, then the MLP's
  int *n_fb;              // device count
  // list order: 0 staggered index order; 1 strips of SG_STRIP map columns
  // along j, each walked down i (row i, then j, then the two faces)
  int strips;
};

int sg_classify_blocks(const DevParams &P);
void launch_sg_list(const DevParams &P, const SgColumnArgs &C, const SgListArgs &L, hipStream_t st);

// features (+ targets) of listed columns
struct SgFeatureArgs {
  const int *list;
  const int *n_list;      // device count
  float *F;               // [slot][SG_FSTRIDE]
  float *CQb;             // [slot][32] C * Qb
  double *CSt;            // [slot] C * S_temperate
  unsigned char *status;  // set to SG_FALLBACK where no surrogate applies
  // targets (training): nullptr in inference
  SgSample *samples;      // appended at *n_samples (capacity cap)
  long long *n_samples;
  long long cap;
  int stride;             // keep every stride-th listed column
};
void launch_sg_features(const DevParams &P, const SgColumnArgs &C, const SgFeatureArgs &F, bool targets,
                        hipStream_t st);

// fused inference
struct SgMlpArgs {
  const int *list, *n_list;
  const float *F, *CQb;
  const double *CSt;
  const unsigned short *frag;  // packed bf16 A-fragments [SG_NFRAG][64][8]
  const float *bias;           // FP32 master array (biases read from it)
  SgNorm norm;
  const double *hx, *hy;
  unsigned char *status;
  double *flux, *Dstag;
  double *dmax_part;           // per block
  int *counts;
  float tau;                   // fallback score threshold (+inf: none)
  float *score;                // per slot fallback score (-1 out of range), or nullptr
  int *fb_list, *n_fb;         // fallback columns are appended here
};
int sg_mlp_blocks();
void launch_sg_mlp(const DevParams &P, const SgMlpArgs &A, hipStream_t st);
void launch_sg_pack(const float *master, unsigned short *frag, hipStream_t st);

// global-address-space views of table-held arguments (see glb)
__device__ __forceinline__ SgColumnArgs global_view(SgColumnArgs a) {
  a.z = glb(a.z);
  a.E = glb(a.E);
  a.thks = glb(a.thks);
  a.theta = glb(a.theta);
  a.hx = glb(a.hx);
  a.hy = glb(a.hy);
  a.err = glb(a.err);
  return a;
}

__device__ __forceinline__ SgListArgs global_view(SgListArgs a) {
  a.status = glb(a.status);
  a.block_count = glb(a.block_count);
  a.list = glb(a.list);
  a.n_list = glb(a.n_list);
  a.counts = glb(a.counts);
  a.flux = glb(a.flux);
  a.Dstag = glb(a.Dstag);
  a.fb_list = glb(a.fb_list);
  a.n_fb = glb(a.n_fb);
  return a;
}

__device__ __forceinline__ SgFeatureArgs global_view(SgFeatureArgs a) {
  a.list = glb(a.list);
  a.n_list = glb(a.n_list);
  a.F = glb(a.F);
  a.CQb = glb(a.CQb);
  a.CSt = glb(a.CSt);
  a.status = glb(a.status);
  a.samples = glb(a.samples);
  a.n_samples = glb(a.n_samples);
  return a;
}

__device__ __forceinline__ SgMlpArgs global_view(const SgMlpArgs &a) {
  SgMlpArgs g = a;
  g.list = glb(a.list);
  g.n_list = glb(a.n_list);
  g.F = glb(a.F);
  g.CQb = glb(a.CQb);
  g.CSt = glb(a.CSt);
  g.frag = glb(a.frag);
  g.bias = glb(a.bias);
  g.hx = glb(a.hx);
  g.hy = glb(a.hy);
  g.status = glb(a.status);
  g.flux = glb(a.flux);
  g.Dstag = glb(a.Dstag);
  g.dmax_part = glb(a.dmax_part);
  g.counts = glb(a.counts);
  g.score = glb(a.score);
  g.fb_list = glb(a.fb_list);
  g.n_fb = glb(a.n_fb);
  return g;
}

// ---- ensembles: the surrogate stages of every member bound to one surrogate,
// one launch per stage. Each member keeps its own list, feature rows and
// status; blocks map to members through block-prefix tables as in the plain
// ensemble stages, and run the single-member kernels' bodies.
struct EnsSgMember {
  DevParams P;
  SgColumnArgs C;
  SgListArgs L;
  SgFeatureArgs F;
  SgMlpArgs M;  // M.dmax_part: the member's partials after its column-stage ones
};
struct EnsSgLaunch {
  const EnsSgMember *table;                        // device [nm]
  const int *cls_start, *feat_start, *mlp_start;   // device [nm + 1]
  int nm, cls_blocks, feat_blocks, mlp_blocks;
  int feat_waves, nm1, max_mz;                     // wavefronts per feature block, n - 1, largest Mz
};
// wavefronts per block of the inference feature pass for columns of Mz levels
int sg_features_waves(int Mz);
void launch_ens_sg(const EnsSgLaunch &L, hipStream_t st);

// training
struct SgTrainBuffers {
  int B;                                   // batch
  unsigned short *X0;                      // [B][64] bf16
  float *Z[SG_NLAYER];                     // pre-activations [B][N_l]
  unsigned short *Abf[SG_NLAYER];          // bf16 activations: Abf[0] = X0, Abf[l] = gelu(Z[l-1])
  float *T, *S, *Hm;                       // [B][32] targets, shares, masks
  float *dZ;                               // [B][128] f32 scratch
  unsigned short *dZbf;                    // [B][128] bf16
  float *dA;                               // [B][128]
  float *grad;                             // [SG_NPARAM]
  float *dWpad;                            // [128][128] scratch for a padded weight gradient
  unsigned short *Wbf;                     // bf16 weights, W1 padded to [128][64]
  double *loss;                            // [1] device
};
void launch_sg_gather(const SgSample *samples, long long n, const SgNorm &norm, unsigned long long seed, long long step,
                      const SgTrainBuffers &T, hipStream_t st);
void launch_sg_set_inputs(const float *Xn, const float *r, const float *share, const float *has, int B,
                          const SgTrainBuffers &T, hipStream_t st);
void launch_sg_weights_bf16(const float *master, unsigned short *Wbf, hipStream_t st);
void launch_sg_forward_backward(const float *master, const SgTrainBuffers &T, float rscale, float lambda,
                                hipStream_t st);
void launch_sg_adam(float *master, float *m, float *v, const float *grad, float lr, float b1, float b2, float eps,
                    long long t, hipStream_t st);
// per-wavefront partial statistics: [nwaves][SG_NIN][4] (sum, sum of squares,
// min, max of each feature) followed by [nwaves][3] (max |r|, sum r^2, count)
// over the cells with cold levels
constexpr int SG_STAT_WAVES = 2048;
void launch_sg_norm_stats(const SgSample *samples, long long n, double *partial, hipStream_t st);

}  // namespace pism_sia
