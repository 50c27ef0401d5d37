// sia_surrogate.hip -- learned column surrogate of the SIA diffusivity for gfx950.
//
// Kernels (the model is described in sia_surrogate.hpp):
//   k_sg_classify / k_sg_scan / k_sg_scatter  status of every staggered column
//       (zero, surrogate, FP64 fallback) and the ordered list of surrogate columns
//   k_sg_features<TARGETS>  one pass over the levels of each listed column: exact
//       temperate sum, 32-cell moments of the cold levels, physics baseline and
//       the MLP inputs; with TARGETS also the exact per-cell cold sums of the
//       plain FP64 flow law (training targets)
//   k_sg_mlp   fused BF16 MFMA inference (v_mfma_f32_16x16x32_bf16, weights in
//       LDS, activations never leave registers) + D, flux, D_max of the column
//   training: k_sg_gather, k_sg_gemm (generic BF16 MFMA GEMM), k_sg_bias_act,
//       k_sg_loss, k_sg_act_bwd, k_sg_colsum, k_sg_adam
#include "sia_surrogate.hpp"

namespace pism_sia {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
// two consecutive levels of a column (columns start at any 8-byte boundary)
typedef double d2u __attribute__((ext_vector_type(2), aligned(8)));
typedef float f2a __attribute__((ext_vector_type(2), aligned(8)));

// ---------------------------------------------------------------- helpers
// FP32 -> BF16 round to nearest even (finite inputs)
__host__ __device__ inline unsigned short f2bf(float f) {
  unsigned u;
  __builtin_memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ inline float bf2f(unsigned short b) {
  const unsigned u = (unsigned)b << 16;
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}

// GELU, tanh form: 0.5 x (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3))), with
// tanh(u) = 1 - 2 / (exp(2u) + 1)
__device__ inline float gelu_f(float x) {
  const float u = 0.7978845608028654f * (x + 0.044715f * x * x * x);
  const float t = __expf(2.0f * u);
  const float th = 1.0f - 2.0f * __builtin_amdgcn_rcpf(t + 1.0f);
  return 0.5f * x * (1.0f + th);
}
__device__ inline float gelu_grad_f(float x) {
  const float u = 0.7978845608028654f * (x + 0.044715f * x * x * x);
  const float t = __expf(2.0f * u);
  const float th = 1.0f - 2.0f * __builtin_amdgcn_rcpf(t + 1.0f);
  const float du = 0.7978845608028654f * (1.0f + 3.0f * 0.044715f * x * x);
  return 0.5f * (1.0f + th) + 0.5f * x * (1.0f - th * th) * du;
}

__device__ inline bf16x8 as_bf16x8(uint4 v) {
  bf16x8 r;
  __builtin_memcpy(&r, &v, 16);
  return r;
}

template <int CTRL, int ROW_MASK, bool BOUND_ZERO>
__device__ __forceinline__ double sg_dpp(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROW_MASK, 0xf, BOUND_ZERO);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROW_MASK, 0xf, BOUND_ZERO);
  return __hiloint2double(hi, lo);
}
// inclusive prefix sum over the 64 lanes (all lanes active)
__device__ __forceinline__ double sg_scan(double v) {
  v += sg_dpp<0x111, 0xf, true>(v);
  v += sg_dpp<0x112, 0xf, true>(v);
  v += sg_dpp<0x114, 0xf, true>(v);
  v += sg_dpp<0x118, 0xf, true>(v);
  v += sg_dpp<0x142, 0xa, false>(v);
  v += sg_dpp<0x143, 0xc, false>(v);
  return v;
}
__device__ __forceinline__ double sg_lane(double v, int l) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double lane_value_d(double v, int l) { return sg_lane(v, l); }
__device__ __forceinline__ int lane_value_i(int v, int l) { return __builtin_amdgcn_readlane(v, l); }

// sum of v over the 32 lanes of this lane's half wave
__device__ __forceinline__ double half_sum(double v, int half) {
  const double s = sg_scan(v);
  const double a = sg_lane(s, 31), b = sg_lane(s, 63);
  return half ? b - a : a;
}

// ---------------------------------------------------------------- classification
constexpr int SG_CLS_BLOCK = 256;

// Strip order: a strip of SG_STRIP map columns contributes 2 SG_STRIP entries
// per row i, so one row of a strip fills about one 32-entry wavefront chunk of
// the feature pass, and the four chunks of its 4-wavefront blocks are four
// consecutive rows on one CU. Staggered column (i, j, 0) reads the enthalpy of
// (i, j) and (i + 1, j): the block's rows share those columns through its L2,
// which the staggered index order leaves to a wavefront a whole row of the
// list away, on another XCD.
constexpr int SG_STRIP = 16;

__host__ __device__ inline int sg_strips(const DevParams &P) { return (P.My + SG_STRIP - 1) / SG_STRIP; }

// enough blocks for either order
int sg_classify_blocks(const DevParams &P) {
  const long rows = 2L * P.Mx * P.My, strips = 2L * SG_STRIP * P.Mx * sg_strips(P);
  return (int)((std::max(rows, strips) + SG_CLS_BLOCK - 1) / SG_CLS_BLOCK);
}

// staggered column of classification slot g in the list order, -1 for none
__device__ __forceinline__ int sg_slot_column(const DevParams &P, long g, bool strips) {
  const long nstag = 2L * P.Mx * P.My;
  if (!strips) return g < nstag ? (int)g : -1;
  const int o = (int)(g & 1), jj = (int)((g >> 1) & (SG_STRIP - 1));
  const long r = g / (2 * SG_STRIP);
  const int strip = (int)(r / P.Mx), i = (int)(r - (long)strip * P.Mx);
  const int j = strip * SG_STRIP + jj;
  if (strip >= sg_strips(P) || j >= P.My) return -1;
  return 2 * (i * P.My + j) + o;
}

// status of staggered column s: SG_ZERO (thk == 0, SIAFD.cc:635-641, or zero
// slope with thk <= Lz: D = 0), SG_FALLBACK (thk above Lz, whatever the slope:
// left to the FP64 kernel, which reports the error as kBelowHeight does for
// any thk != 0, SIAFD.cc:654; or thk below the second level), else SG_SURR.
__device__ __forceinline__ void sg_classify_block(const DevParams &P, const SgColumnArgs &C, const SgListArgs &L,
                                                  int bid) {
  const int s = sg_slot_column(P, (long)bid * SG_CLS_BLOCK + threadIdx.x, L.strips != 0);
  int st = -1;
  if (s >= 0) {
    const int q = s >> 1, o = s & 1;
    const int i = q / P.My, j = q - i * P.My;
    const int q2 = (o == 0) ? wrapi(i + 1, P.Mx) * P.My + j : i * P.My + wrapi(j + 1, P.My);
    const double thk = 0.5 * (C.thks[q] + C.thks[q2]);
    const double hxs = C.hx[s], hys = C.hy[s];
    const double alpha = sqrt(hxs * hxs + hys * hys);
    if (thk > P.Lz + 1.0e-6) {
      st = SG_FALLBACK;
    } else if (thk == 0.0 || alpha == 0.0) {
      st = SG_ZERO;
      const double slope = (o == 0) ? hxs : hys;
      L.flux[s] = (thk == 0.0) ? 0.0 : -0.0 * slope;
      L.Dstag[s] = 0.0;
    } else if (thk < C.z[1]) {
      st = SG_FALLBACK;
    } else {
      st = SG_SURR;
    }
    L.status[s] = (unsigned char)st;
  }
  const int nz = __syncthreads_count(st == SG_ZERO);
  const int ns = __syncthreads_count(st == SG_SURR);
  const int nf = __syncthreads_count(st == SG_FALLBACK);
  if (threadIdx.x == 0) {
    L.block_count[3 * bid + 0] = nz;
    L.block_count[3 * bid + 1] = ns;
    L.block_count[3 * bid + 2] = nf;
  }
}

__global__ __launch_bounds__(SG_CLS_BLOCK) void k_sg_classify(DevParams P, SgColumnArgs C, SgListArgs L) {
  sg_classify_block(P, C, L, blockIdx.x);
}

// exclusive scans of the per-block surrogate and fallback counts (one block of
// 1024); the offsets overwrite block_count[3 b + 1] and [3 b + 2]; totals go
// to counts[0..2], n_list and n_fb. part, partf: 1024 entries of LDS each.
__device__ __forceinline__ void sg_scan_block(int nb, const SgListArgs &L, long long *part, long long *partf,
                                              long long *tot) {
  const int t = threadIdx.x;
  const int per = (nb + 1023) / 1024;
  const int b0 = t * per, b1 = min(nb, b0 + per);
  long long s = 0, z = 0, f = 0;
  // unrolled: the loads of 8 blocks are in flight together
#pragma unroll 8
  for (int b = b0; b < b1; b++) {
    z += L.block_count[3 * b + 0];
    s += L.block_count[3 * b + 1];
    f += L.block_count[3 * b + 2];
  }
  part[t] = s;
  partf[t] = f;
  if (t < 3) tot[t] = 0;
  __syncthreads();
  atomicAdd((unsigned long long *)&tot[0], (unsigned long long)z);
  // Hillis-Steele over the 1024 partial sums of both counts
  for (int off = 1; off < 1024; off <<= 1) {
    const long long v = (t >= off) ? part[t - off] : 0;
    const long long vf = (t >= off) ? partf[t - off] : 0;
    __syncthreads();
    part[t] += v;
    partf[t] += vf;
    __syncthreads();
  }
  long long run = part[t] - s, runf = partf[t] - f;  // exclusive starts of this thread's chunk
#pragma unroll 8
  for (int b = b0; b < b1; b++) {
    const int c = L.block_count[3 * b + 1], cf = L.block_count[3 * b + 2];
    L.block_count[3 * b + 1] = (int)run;
    L.block_count[3 * b + 2] = (int)runf;
    run += c;
    runf += cf;
  }
  if (t == 1023) {
    *L.n_list = (int)part[1023];
    L.counts[1] = (int)part[1023];
    *L.n_fb = (int)partf[1023];
    L.counts[2] = (int)partf[1023];
  }
  __syncthreads();
  if (t == 0) L.counts[0] = (int)tot[0];
}

__global__ __launch_bounds__(1024) void k_sg_scan(int nb, SgListArgs L) {
  __shared__ long long part[1024], partf[1024];
  __shared__ long long tot[3];
  sg_scan_block(nb, L, part, partf, tot);
}

// ordered compaction of the SG_SURR columns into list and of the classified
// SG_FALLBACK columns into fb_list; wsum: 2 * SG_CLS_BLOCK / 64 ints of LDS
__device__ __forceinline__ void sg_scatter_block(const DevParams &P, const SgListArgs &L, int bid, int *wsum) {
  const int s = sg_slot_column(P, (long)bid * SG_CLS_BLOCK + threadIdx.x, L.strips != 0);
  const unsigned char st = (s >= 0) ? L.status[s] : (unsigned char)SG_ZERO;
  const bool flag = st == SG_SURR, fflag = st == SG_FALLBACK;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  constexpr int NW = SG_CLS_BLOCK / 64;
  const unsigned long long m = __ballot(flag), mf = __ballot(fflag);
  const unsigned long long below = (1ull << lane) - 1ull;
  if (lane == 0) {
    wsum[wv] = __popcll(m);
    wsum[NW + wv] = __popcll(mf);
  }
  __syncthreads();
  int base = L.block_count[3 * bid + 1], basef = L.block_count[3 * bid + 2];
  for (int w = 0; w < wv; w++) {
    base += wsum[w];
    basef += wsum[NW + w];
  }
  if (flag) L.list[base + __popcll(m & below)] = s;
  if (fflag) L.fb_list[basef + __popcll(mf & below)] = s;
}

__global__ __launch_bounds__(SG_CLS_BLOCK) void k_sg_scatter(DevParams P, SgListArgs L) {
  __shared__ int wsum[2 * SG_CLS_BLOCK / 64];
  sg_scatter_block(P, L, blockIdx.x, wsum);
}

void launch_sg_list(const DevParams &P, const SgColumnArgs &C, const SgListArgs &L, hipStream_t st) {
  const int nb = sg_classify_blocks(P);
  hipLaunchKernelGGL(k_sg_classify, dim3(nb), dim3(SG_CLS_BLOCK), 0, st, P, C, L);
  hipLaunchKernelGGL(k_sg_scan, dim3(1), dim3(1024), 0, st, nb, L);
  hipLaunchKernelGGL(k_sg_scatter, dim3(nb), dim3(SG_CLS_BLOCK), 0, st, P, L);
}

// ---------------------------------------------------------------- features
// s^(n-1) and p^n for the integer exponents the surrogate supports (n - 1 in {1, 2, 3})
__device__ __forceinline__ double ipow_nm1(const DevParams &P, double s) {
  const double s2 = s * s;
  return P.nm1_int == 1 ? s : (P.nm1_int == 2 ? s2 : s2 * s);
}
__device__ __forceinline__ double pow_n(const DevParams &P, double p) { return ipow_nm1(P, p) * p; }

// Training samples: a wavefront owns 16 consecutive list entries. Lane c
// prepares entry c (2-D inputs, kBelowHeight), so the setup of all 16 is in
// flight at once. The entries are then taken in pairs: all 64 lanes stage the
// averaged enthalpy of both columns in LDS with coalesced loads, then half wave
// hh takes column hh of the pair and lane m its zeta-cell m, walking the
// cell's levels from LDS: the level weight G_k, the exact temperate sum, and
// the moments of x = T_pa - 263.15 over the cold levels. The level weights are
// the reference's:
// W_k = d_k (dz_k [k >= 1] + dz_{k+1} [k < ks]) / 2 + d_ks^2 / 2 [k == ks]
// (SIAFD.cc:677-684 regrouped per delta_k), G_k = W_k p_k^n.
constexpr int SG_FEAT_COLS = 16;
// list entries a wavefront prepares at once in inference (lane c < this: entry c)
constexpr int SG_INFER_COLS = 32;
// 128-level batches of a pair's enthalpy loads issued together (two: with the
// loads in flight the inference pass fits 128 VGPRs, 4 waves per SIMD)
constexpr int SG_STAGE_BATCHES = 2;

// buckets of the level search: bucket b holds the first level with z >= b Lz / SG_NBUCKET
constexpr int SG_NBUCKET = 256;
// flag beside ks of an inference entry: the top level lies below the surface
constexpr int SG_TOP_POS = 1 << 16;

// first level k in [0, ks] with z_k * s32 >= m (ks + 1 if none), i.e.
// cell_of(z_k) >= m for m <= 31: start at the bucket of height m / s32 and
// step to the exact answer (0-2 steps for equal or quadratic spacing)
// kbelow over the interleaved (z, half spacing) table
__device__ __forceinline__ int kbelow_zsk(const double2 *zsk, int Mz, double Lz, double height) {
  if (height < 0.0 - 1.0e-6 || height > Lz + 1.0e-6) return -1;
  int ilo = 0, ihi = Mz - 1;
  while (ihi > ilo + 1) {
    const int i = (ihi + ilo) >> 1;
    if (zsk[i].x > height) ihi = i; else ilo = i;
  }
  return ilo;
}

__device__ __forceinline__ int cell_first_level(const double2 *zsk, const int *bucket, double bscale, int ks,
                                                double s32, int m) {
  const double hgt = (double)m / s32;
  int b = (int)(hgt * bscale);
  b = b < 0 ? 0 : (b > SG_NBUCKET - 1 ? SG_NBUCKET - 1 : b);
  int k = min(bucket[b], ks + 1);
  while (k > 0 && zsk[k - 1].x * s32 >= (double)m) k--;
  while (k <= ks && zsk[k].x * s32 < (double)m) k++;
  return k;
}

// p^N for a compile-time N in {2, 3, 4}
template <int N>
__device__ __forceinline__ double pown(double p) {
  const double p2 = p * p;
  return N == 2 ? p2 : (N == 3 ? p2 * p : p2 * p2);
}

template <int N>
__device__ __forceinline__ float pownf(float p) {
  const float p2 = p * p;
  return N == 2 ? p2 : (N == 3 ? p2 * p : p2 * p2);
}

__device__ __forceinline__ float lane_value_f(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// v of the lane whose byte address (4 x lane) is src, per lane (ds_bpermute)
__device__ __forceinline__ float lane_get_f(float v, int src) {
  return __int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(v)));
}
__device__ __forceinline__ double lane_get_d(double v, int src) {
  const int lo = __builtin_amdgcn_ds_bpermute(src, __double2loint(v));
  const int hi = __builtin_amdgcn_ds_bpermute(src, __double2hiint(v));
  return __hiloint2double(hi, lo);
}

// sum of v over the 32 lanes of this lane's half wave (FP32, five DPP steps)
template <int CTRL, int ROW_MASK, bool BOUND_ZERO>
__device__ __forceinline__ float sg_dppf(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xf, BOUND_ZERO));
}
__device__ __forceinline__ float half_total_f(float v, int half) {
  v += sg_dppf<0x111, 0xf, true>(v);
  v += sg_dppf<0x112, 0xf, true>(v);
  v += sg_dppf<0x114, 0xf, true>(v);
  v += sg_dppf<0x118, 0xf, true>(v);
  v += sg_dppf<0x142, 0xa, false>(v);
  const float a = lane_value_f(v, 31), b = lane_value_f(v, 63);
  return half ? b : a;
}

// a wave-uniform value pinned where it is computed (the compiler would
// otherwise fold selects of such constants into per-lane FP64 divisions)
__device__ __forceinline__ float pinned(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

// kBelowHeight (IceGrid.cc:379-391) clamped to Mz - 2, from the bucket of the
// height: the last level below the bucket's lower edge, then 0-2 steps
__device__ __forceinline__ int kbelow_bucket(const double *zd, const int *bucket, double bscale, int Mz, double Lz,
                                             double height) {
  if (height < 0.0 - 1.0e-6 || height > Lz + 1.0e-6) return -1;
  int b = (int)(height * bscale);
  b = b < 0 ? 0 : (b > SG_NBUCKET - 1 ? SG_NBUCKET - 1 : b);
  int k = bucket[b] - 1;
  k = k < 0 ? 0 : (k > Mz - 2 ? Mz - 2 : k);
  while (k > 0 && zd[k] > height) k--;
  while (k < Mz - 2 && zd[k + 1] <= height) k++;
  return k;
}

// Inference pass (FP32 level arithmetic; the column constants, the cell
// boundaries and the outputs' scales stay FP64). A wavefront owns
// SG_INFER_COLS consecutive list entries; lane c prepares entry c (32 entries:
// the per-entry setup runs on half the lanes, 1.87 -> 1.77 ms against 16 on
// the 1501x2801x401 sheet). The entries are taken
// in pairs: all 64 lanes stage the averaged enthalpy of both columns in LDS as
// y_k = E_k / c_i + T0 - T_ref (up to 2 x 128 levels of both columns in flight),
// then half wave hh takes column hh and lane m its zeta-cell m (the levels with
// floor(32 z_k / thk) = m, level ks in cell 31 at most). Per level, with
// pressures relative to the basal one, ph = p_k / p_b = (p_air + rho g d_k) / p_b:
//   W_k = d_k (dz_k [k >= 1] + dz_{k+1} [k < ks]) / 2 + d_ks^2 / 2 [k = ks]
//   (SIAFD.cc:677-684 regrouped per delta_k), G_k = W_k ph^n,
//   x_k = T_pa - T_ref = y_k + beta p_b ph, cold iff x_k < T_melt - T_ref (E < E_cts)
// g0, g1, g2: sums of G, G x, G x^2 over the cold levels; st: the exact
// temperate sum of G A_T0 (1 + wf_coeff min(omega, wf_limit)). The outputs are
// scaled back by C p_b^n in FP64. The cell epilogue evaluates the physics
// baseline Qb = g0 A(T) (1 + (k^2 - 2k/T) var / 2), T = T_ref + g1/g0,
// k = Q / (R T^2), with the hardware exp2 and reciprocal (about 1 ulp).
template <int NP>
__device__ __forceinline__ void sg_infer_features_block(const DevParams &P, const SgColumnArgs &C,
                                                        const SgFeatureArgs &F, int bid, int nblk, double2 *zsk,
                                                        int *bucket) {
  const int Mz = P.Mz;
  const int Mzp = (Mz + 1) & ~1;  // even: 16-byte aligned LDS regions, 8-byte aligned level pairs
  // FP64 z_k (kBelowHeight, cell boundaries), FP32 (z_k, half spacing) (levels)
  double *zd = reinterpret_cast<double *>(zsk);
  float2 *zf = reinterpret_cast<float2 *>(zd + Mzp);
  for (int k = threadIdx.x; k < Mz; k += blockDim.x) {
    const double zk = C.z[k];
    const double w = 0.5 * (((k >= 1) ? zk - C.z[k - 1] : 0.0) + ((k + 1 < Mz) ? C.z[k + 1] - zk : 0.0));
    zd[k] = zk;
    zf[k] = make_float2((float)zk, (float)w);
  }
  const double bscale = (double)SG_NBUCKET / C.z[Mz - 1];
  for (int b = threadIdx.x; b < SG_NBUCKET; b += blockDim.x) {
    const double lo = (double)b / bscale;  // first level at or above the bucket's lower edge
    int a = 0, c = Mz;
    while (a < c) {
      const int mid = (a + c) >> 1;
      if (C.z[mid] >= lo) c = mid; else a = mid + 1;
    }
    bucket[b] = a;
  }
  __syncthreads();
  const int nwb = blockDim.x >> 6;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  float *ebw = reinterpret_cast<float *>(zf + Mzp) + (long)wv * 4 * Mzp;  // two buffers of a pair's profiles
  const int n = *F.n_list;
  const double rhog = P.rho_i * P.g;
  const float T_ref = 263.15f;
  const double y_off = P.T0 - 263.15;
  const float x_cts = pinned((float)(P.T_melt - 263.15));
  // temperate softness A_T0 + ta min(x - x_cts, tcap), ta = A_T0 wf_coeff c_i / L, tcap = wf_limit L / c_i
  const double ta = P.A_T0 * P.wf_coeff * P.ci * P.inv_L, tcap = P.wf_limit / (P.ci * P.inv_L);
  const float t_a = pinned((float)ta), t_b = pinned((float)(P.A_T0 - ta * (P.T_melt - 263.15)));
  const float t_cap = pinned((float)(P.A_T0 + ta * tcap));
  const float QRc = pinned((float)(P.Q_cold / P.R)), QRw = pinned((float)(P.Q_warm / P.R));
  const float Ac = pinned((float)P.A_cold), Aw = pinned((float)P.A_warm), Tcrit = pinned((float)P.T_crit);
  const int hh = lane >> 5, m = lane & (SG_NCELL - 1);
  for (int t0 = (bid * nwb + wv) * SG_INFER_COLS; t0 < n; t0 += nblk * nwb * SG_INFER_COLS) {
    // ---- setup: lane c < SG_INFER_COLS prepares entry t0 + c (c_kt = -1: no entry)
    // c_kt: ks, plus SG_TOP_POS when z_ks < thk (d_ks > 0 in FP64)
    int c_kt = -1, c_q = 0, c_q2 = 0;
    double c_s32 = 0.0, c_cs = 0.0;
    float c_thk = 0.0f, c_fthk = 0.0f, c_lgs = 0.0f, c_cz = 0.0f, c_pa = 0.0f, c_bp = 0.0f, c_csf = 0.0f;
    if (lane < SG_INFER_COLS && t0 + lane < n) {
      const int s = F.list[t0 + lane];
      const int q = s >> 1, o = s & 1;
      const int i = q / P.My, j = q - i * P.My;
      const int q2 = (o == 0) ? wrapi(i + 1, P.Mx) * P.My + j : i * P.My + wrapi(j + 1, P.My);
      const double thk = 0.5 * (C.thks[q] + C.thks[q2]);
      const double hxs = C.hx[s], hys = C.hy[s];
      const double alpha = sqrt(hxs * hxs + hys * hys);
      c_lgs = (float)(log10(alpha) + 2.0);  // the slope input of the network
      c_fthk = (float)(thk / 1000.0);       // the thickness input
      const double pb = P.p_air + rhog * thk;  // basal pressure
      c_cs = P.e * (0.5 * (C.theta[q] + C.theta[q2])) * 2.0 * ipow_nm1(P, alpha) * pown<NP>(pb);
      c_csf = (float)c_cs;
      const double ipb = 1.0 / pb;
      c_cz = (float)(rhog * ipb);
      c_pa = (float)(P.p_air * ipb);
      c_bp = (float)(P.beta * pb);
      c_s32 = (double)SG_NCELL / thk;
      const int ks = kbelow_bucket(zd, bucket, bscale, Mz, P.Lz, thk);  // >= 1 for listed columns
      c_kt = ks | ((zd[ks] < thk) ? SG_TOP_POS : 0);
      c_thk = (float)thk;
      c_q = q;
      c_q2 = q2;
    }
    // ---- the chunk's pairs (entries 2 pp, 2 pp + 1), software-pipelined: the
    // loads of pair pp + 1 (its first SG_STAGE_BATCHES x 128 levels) are in
    // flight while pair pp walks its levels; they are converted into the other
    // profile buffer before pair pp's epilogue. Lane l stages levels k, k + 1
    // with k = kb + 128 b + 2 l, up to the larger ks of the pair (16-byte loads;
    // k <= ks <= Mz - 2 keeps k + 1 inside the column).
    const int npair = (min(SG_INFER_COLS, n - t0) + 1) / 2;
    const double hci = 0.5 * P.inv_ci;
    struct PairSrc {
      const double *E[2][2];
      int kt0, kt1, ksmax;
    };
    auto pair_src = [&](int pp) {
      PairSrc r;
      const int c = 2 * pp;
      r.kt0 = lane_value_i(c_kt, c);
      r.kt1 = lane_value_i(c_kt, c + 1);
      const int ks0 = r.kt0 & (SG_TOP_POS - 1), ks1 = r.kt1 < 0 ? -1 : (r.kt1 & (SG_TOP_POS - 1));
      r.ksmax = max(ks0, ks1);
      r.E[0][0] = C.E + (long)lane_value_i(c_q, c) * Mz;
      r.E[0][1] = C.E + (long)lane_value_i(c_q2, c) * Mz;
      // a missing second entry stages the first column twice
      r.E[1][0] = r.kt1 < 0 ? r.E[0][0] : C.E + (long)lane_value_i(c_q, c + 1) * Mz;
      r.E[1][1] = r.kt1 < 0 ? r.E[0][1] : C.E + (long)lane_value_i(c_q2, c + 1) * Mz;
      return r;
    };
    d2u ea[SG_STAGE_BATCHES][2], ebv[SG_STAGE_BATCHES][2];
    auto issue = [&](const PairSrc &r, int kb) {
#pragma unroll
      for (int bt = 0; bt < SG_STAGE_BATCHES; bt++) {
        const int k = kb + 128 * bt + 2 * lane;
        if (k <= r.ksmax) {
#pragma unroll
          for (int h = 0; h < 2; h++) {
            ea[bt][h] = *reinterpret_cast<const d2u *>(r.E[h][0] + k);
            ebv[bt][h] = *reinterpret_cast<const d2u *>(r.E[h][1] + k);
          }
        }
      }
    };
    auto convert = [&](const PairSrc &r, int kb, float *buf) {
#pragma unroll
      for (int bt = 0; bt < SG_STAGE_BATCHES; bt++) {
        const int k = kb + 128 * bt + 2 * lane;
        if (k <= r.ksmax) {
#pragma unroll
          for (int h = 0; h < 2; h++) {
            f2a y;
            y.x = (float)__builtin_fma(ea[bt][h].x + ebv[bt][h].x, hci, y_off);
            y.y = (float)__builtin_fma(ea[bt][h].y + ebv[bt][h].y, hci, y_off);
            *reinterpret_cast<f2a *>(buf + h * Mzp + k) = y;
          }
        }
      }
    };
    PairSrc cur = pair_src(0);
    for (int kb = 0; kb <= cur.ksmax; kb += 128 * SG_STAGE_BATCHES) {
      issue(cur, kb);
      convert(cur, kb, ebw);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (int pp = 0; pp < npair; pp++) {
      const int cc = 2 * pp;
      const int kt0 = cur.kt0, kt1 = cur.kt1;
      const int ks0 = kt0 & (SG_TOP_POS - 1), ks1 = kt1 < 0 ? -1 : (kt1 & (SG_TOP_POS - 1));
      float *bcur = ebw + (pp & 1) * 2 * Mzp, *bnext = ebw + ((pp + 1) & 1) * 2 * Mzp;
      const bool more = pp + 1 < npair;
      PairSrc nxt = cur;
      if (more) {
        nxt = pair_src(pp + 1);
        issue(nxt, 0);
      }
      // ---- half wave hh: column cc + hh, lane m: zeta-cell m
      const int src = (cc + hh) * 4;
      const int ks = hh ? ks1 : ks0;
      const bool live = ks >= 0;
      const bool top_pos = ((hh ? kt1 : kt0) & SG_TOP_POS) != 0;
      const double s32 = lane_get_d(c_s32, src);
      const float thkf = lane_get_f(c_thk, src);
      // cell m: levels [klo, khi), the first level with z_k s32 >= m from the
      // bucket of height m thk / 32; khi is the next lane's klo
      int klo = 0;
      if (live) {
        int b = (int)((float)m * thkf * (float)(bscale / SG_NCELL));
        b = b > SG_NBUCKET - 1 ? SG_NBUCKET - 1 : b;
        int k = min(bucket[b], ks + 1);
        while (k > 0 && zd[k - 1] * s32 >= (double)m) k--;
        while (k <= ks && zd[k] * s32 < (double)m) k++;
        klo = k;
      }
      int khi = __builtin_amdgcn_update_dpp(0, klo, 0x130, 0xf, 0xf, true);  // lane l + 1 (wave_shl:1)
      if (m == SG_NCELL - 1) khi = ks + 1;
      if (!live) khi = 0;
      const float cz = lane_get_f(c_cz, src);
      const float pa = lane_get_f(c_pa, src);
      const float bp = lane_get_f(c_bp, src);
      // W_ks = d_ks (dz_ks + d_ks) / 2
      const int kt = live ? ks : 0;
      const float skt = 0.5f * ((kt >= 1 ? zf[kt].x - zf[kt - 1].x : 0.0f) + (thkf - zf[kt].x));
      const float *yb = bcur + hh * Mzp;
      float g0 = 0.0f, g1 = 0.0f, g2 = 0.0f, st = 0.0f;
#pragma unroll 2
      for (int k = klo; k < khi; k++) {
        const float2 zw = zf[k];
        const float d = thkf - zw.x;  // >= 0: rounding to FP32 keeps z_k <= thk
        const float ph = __builtin_fmaf(cz, d, pa);
        const float G = d * (k == ks ? skt : zw.y) * pownf<NP>(ph);
        const float x = __builtin_fmaf(bp, ph, yb[k]);
        const float Gc = (x < x_cts) ? G : 0.0f;
        g0 += Gc;
        const float gx = Gc * x;
        g1 += gx;
        g2 = __builtin_fmaf(gx, x, g2);
        st = __builtin_fmaf(G - Gc, __builtin_fminf(__builtin_fmaf(t_a, x, t_b), t_cap), st);
      }
      // a cold top level with d = thk - z_ks > 0 whose G underflows in FP32 still
      // populates its cell (the FP64 pass sees G > 0 of negligible weight)
      if (live && top_pos && klo <= ks && ks < khi && g0 == 0.0f) {
        const float x = __builtin_fmaf(bp, __builtin_fmaf(cz, thkf - zf[ks].x, pa), yb[ks]);
        if (x < x_cts) {
          g0 = 1.0e-30f;
          g1 = g0 * x;
          g2 = g1 * x;
        }
      }
      // pair pp + 1 into the other buffer (its levels beyond the prefetch too)
      if (more) {
        convert(nxt, 0, bnext);
        for (int kb = 128 * SG_STAGE_BATCHES; kb <= nxt.ksmax; kb += 128 * SG_STAGE_BATCHES) {
          issue(nxt, kb);
          convert(nxt, kb, bnext);
        }
      }
      const float S_t = half_total_f(st, hh);
      const bool has = live && g0 > 0.0f;
      float Qh = 0.0f, feat = 2.0f;
      if (has) {
        const float ig0 = __builtin_amdgcn_rcpf(g0);
        const float mx = g1 * ig0;
        const float Tbar = T_ref + mx;
        const float var = __builtin_fmaxf(__builtin_fmaf(g2, ig0, -mx * mx), 0.0f);
        const float iT = __builtin_amdgcn_rcpf(Tbar);
        const bool cold = Tbar < Tcrit;
        const float QR = cold ? QRc : QRw;
        const float kk = QR * iT * iT;
        // Paterson-Budd softness at the cell mean (flowlaws.cc:90-95)
        const float Ab = (cold ? Ac : Aw) * __builtin_amdgcn_exp2f(-1.4426950408889634f * QR * iT);
        Qh = g0 * Ab * __builtin_fmaf(0.5f * (kk * kk - 2.0f * kk * iT), var, 1.0f);
        feat = (Tbar - 250.0f) * (1.0f / 15.0f);
      }
      const double cs = lane_get_d(c_cs, src);
      const float csf = lane_get_f(c_csf, src);
      const float fthk = lane_get_f(c_fthk, src), lgs = lane_get_f(c_lgs, src);
      if (live) {  // row t0 + cc + hh: F (36 floats), CQb (32 floats), CSt
        const long r = t0 + cc + hh;
        F.F[r * SG_FSTRIDE + m] = feat;
        F.CQb[r * SG_NCELL + m] = csf * Qh;
        if (m < SG_FSTRIDE - SG_NCELL) F.F[r * SG_FSTRIDE + SG_NCELL + m] = m == 0 ? fthk : (m == 1 ? lgs : 0.0f);
        if (m == 0) F.CSt[r] = cs * (double)S_t;
      }
      // pair pp + 1's profiles are complete; pair pp + 2 overwrites this pair's
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      cur = nxt;
    }
  }
}

// Training samples, in FP64 throughout: the features as the inference pass
// computes them and the exact per-cell cold sums of the plain per-level flow
// law (in the reference's order). Block bid of nblk persistent blocks walking
// every stride-th list entry. LDS: zsk (dynamic): (z_k, (dz_k [k >= 1] +
// dz_{k+1}) / 2) per level, then per wavefront the two staged FP64 enthalpy
// profiles; bucket: SG_NBUCKET ints.
__device__ __forceinline__ void sg_sample_block(const DevParams &P, const SgColumnArgs &C, const SgFeatureArgs &F,
                                                int bid, int nblk, double2 *zsk, int *bucket) {
  const int Mz = P.Mz;
  for (int k = threadIdx.x; k < Mz; k += blockDim.x) {
    const double zk = C.z[k];
    zsk[k] = make_double2(zk, 0.5 * (((k >= 1) ? zk - C.z[k - 1] : 0.0) + ((k + 1 < Mz) ? C.z[k + 1] - zk : 0.0)));
  }
  const double bscale = (double)SG_NBUCKET / C.z[Mz - 1];
  for (int b = threadIdx.x; b < SG_NBUCKET; b += blockDim.x) {
    // first level at or above the bucket's lower edge
    const double lo = (double)b / bscale;
    int a = 0, c = Mz;
    while (a < c) {
      const int mid = (a + c) >> 1;
      if (C.z[mid] >= lo) c = mid; else a = mid + 1;
    }
    bucket[b] = a;
  }
  __syncthreads();
  const int nwb = blockDim.x >> 6;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int Mzp = (Mz + 1) & ~1;  // even stride: 8-byte aligned level pairs
  double *ebw = reinterpret_cast<double *>(zsk + Mz) + (long)wv * 2 * Mzp;
  const int n = *F.n_list;
  const int stride = F.stride;
  const int nsel = (n + stride - 1) / stride;
  const long long base_sample = *F.n_samples;
  const double T_ref = 263.15;
  const double rhog = P.rho_i * P.g, T_off = P.T0 + P.T_melt - T_ref;
  const double QR_cold = P.Q_cold / P.R, QR_warm = P.Q_warm / P.R;
  const int hh = lane >> 5, m = lane & (SG_NCELL - 1);
  for (int t0 = (bid * nwb + wv) * SG_FEAT_COLS; t0 < nsel; t0 += nblk * nwb * SG_FEAT_COLS) {
    // ---- setup: lane c < 16 prepares entry t0 + c (c_ks = -1: no entry)
    int c_ks = -1, c_q = 0, c_q2 = 0, c_s = 0;
    double c_thk = 0.0, c_lgs = 0.0, c_Cf = 0.0;
    if (lane < SG_FEAT_COLS && t0 + lane < nsel) {
      const int s = F.list[(t0 + lane) * stride];
      const int q = s >> 1, o = s & 1;
      const int i = q / P.My, j = q - i * P.My;
      const int q2 = (o == 0) ? wrapi(i + 1, P.Mx) * P.My + j : i * P.My + wrapi(j + 1, P.My);
      const double thk = 0.5 * (C.thks[q] + C.thks[q2]);
      const double hxs = C.hx[s], hys = C.hy[s];
      const double alpha = sqrt(hxs * hxs + hys * hys);
      c_lgs = log10(alpha) + 2.0;  // the slope input of the network
      c_Cf = P.e * (0.5 * (C.theta[q] + C.theta[q2])) * 2.0 * ipow_nm1(P, alpha);
      c_ks = kbelow_zsk(zsk, Mz, P.Lz, thk);  // >= 1 for listed columns
      c_thk = thk;
      c_q = q;
      c_q2 = q2;
      c_s = s;
    }
    for (int cc = 0; cc < SG_FEAT_COLS; cc += 2) {
      if (lane_value_i(c_ks, cc) < 0) break;  // entries are consecutive: nothing left
      // ---- stage the averaged enthalpy of columns cc, cc + 1
      int ksp[2];
      const double *Ep[2][2];
#pragma unroll
      for (int h = 0; h < 2; h++) {
        ksp[h] = lane_value_i(c_ks, cc + h);
        Ep[h][0] = C.E + (long)lane_value_i(c_q, cc + h) * Mz;
        Ep[h][1] = C.E + (long)lane_value_i(c_q2, cc + h) * Mz;
      }
      // lane l stages levels k, k + 1 with k = kb + 128 b + 2 l (16-byte loads;
      // k <= ks <= Mz - 2 keeps k + 1 inside the column, level ks + 1 is never
      // read); the loads of SG_STAGE_BATCHES x 128 levels are all in flight
      // before the first is used
      const int ksmax = max(ksp[0], ksp[1]);
      for (int kb = 0; kb <= ksmax; kb += 128 * SG_STAGE_BATCHES) {
        d2u ea[SG_STAGE_BATCHES][2], ebv[SG_STAGE_BATCHES][2];
#pragma unroll
        for (int bt = 0; bt < SG_STAGE_BATCHES; bt++) {
          const int k = kb + 128 * bt + 2 * lane;
#pragma unroll
          for (int h = 0; h < 2; h++) {
            if (k <= ksp[h]) {
              ea[bt][h] = *reinterpret_cast<const d2u *>(Ep[h][0] + k);
              ebv[bt][h] = *reinterpret_cast<const d2u *>(Ep[h][1] + k);
            }
          }
        }
#pragma unroll
        for (int bt = 0; bt < SG_STAGE_BATCHES; bt++) {
          const int k = kb + 128 * bt + 2 * lane;
#pragma unroll
          for (int h = 0; h < 2; h++) {
            if (k <= ksp[h]) {
              d2u e;
              e.x = 0.5 * (ea[bt][h].x + ebv[bt][h].x);
              e.y = 0.5 * (ea[bt][h].y + ebv[bt][h].y);
              *reinterpret_cast<d2u *>(ebw + h * Mzp + k) = e;
            }
          }
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      // ---- half wave hh: column cc + hh, lane m: zeta-cell m
      const int ks = hh ? ksp[1] : ksp[0];
      const bool live = ks >= 0;
      const double thk = hh ? lane_value_d(c_thk, cc + 1) : lane_value_d(c_thk, cc);
      const double lgs = hh ? lane_value_d(c_lgs, cc + 1) : lane_value_d(c_lgs, cc);
      const double Cf = hh ? lane_value_d(c_Cf, cc + 1) : lane_value_d(c_Cf, cc);
      const int s = hh ? lane_value_i(c_s, cc + 1) : lane_value_i(c_s, cc);
      // cell m: levels [klo, khi); khi is the next lane's klo
      const double s32 = (double)SG_NCELL / (live ? thk : 1.0);
      const int klo = live ? cell_first_level(zsk, bucket, bscale, ks, s32, m) : 0;
      int khi = __shfl_down(klo, 1, 64);
      if (m == SG_NCELL - 1) khi = ks + 1;
      if (!live) khi = 0;
      const double *eb = ebw + hh * Mzp;
      double st = 0.0, g0 = 0.0, g1 = 0.0, g2 = 0.0, qref = 0.0, dsum = 0.0, gs = 1.0;
      {
        // G is accumulated relative to thk^2 p_base^n
        gs = live ? thk * thk * pow_n(P, P.p_air + rhog * thk) : 1.0;
        const double inv_gs = 1.0 / gs;
#pragma unroll 4
        for (int k = klo; k < khi; k++) {
          const double zk = zsk[k].x;
          const double d = thk - zk;  // >= 0 up to ks
          const double p = P.p_air + rhog * d;
          const double W = (k < ks) ? d * zsk[k].y : 0.5 * d * (((k >= 1) ? zk - zsk[k - 1].x : 0.0) + d);
          const double G = W * pow_n(P, p) * inv_gs;
          const double Ek = (double)eb[k];
          // default enthalpy converter: T_m = T_melt - beta p, E_cts = c_i (T_m - T0)
          // (enthalpyConverter.cc:128-143)
          const double Tm = P.T_melt - P.beta * p;
          const double Ects = P.ci * (Tm - P.T0);
          if (Ek < Ects) {  // cold branch of GPBLD (flowlaws.cc:212-214): T_pa - 263.15
            const double x = Ek * P.inv_ci + T_off - Tm;
            g0 += G;
            g1 += G * x;
            g2 += G * x * x;
            const double A = softness_pb(P, ec_pressure_adjusted_temperature(P, Ek, p));
            qref += G * A;
            dsum += G * A;
          } else {  // temperate branch (flowlaws.cc:215-219), omega = (E - E_cts) / L
            const double omega = fmin((Ek - Ects) * P.inv_L, P.wf_limit);
            st += G * (P.A_T0 * (1.0 + P.wf_coeff * omega));
          }
        }
      }
      const double S_t = half_sum(st, hh) * gs;
      const bool has = live && g0 > 0.0;
      double Qb = 0.0, feat = 2.0;
      if (has) {
        const double inv_g0 = 1.0 / g0;
        const double mx = g1 * inv_g0;
        const double Tbar = T_ref + mx;
        const double var = fmax(g2 * inv_g0 - mx * mx, 0.0);
        const double inv_T = 1.0 / Tbar;
        const bool cold = Tbar < P.T_crit;
        const double QR = cold ? QR_cold : QR_warm;
        const double kk = QR * inv_T * inv_T;
        // Paterson-Budd softness at the cell mean (flowlaws.cc:90-95)
        const double Ab = (cold ? P.A_cold : P.A_warm) * exp(-QR * inv_T);
        Qb = g0 * gs * Ab * (1.0 + 0.5 * (kk * kk - 2.0 * kk * inv_T) * var);
        feat = (Tbar - 250.0) * (1.0 / 15.0);
      }
      const int t = t0 + cc + hh;
      {
        const double qcell = qref * gs;  // exact cold sum of the cell
        const double Dref = Cf * (S_t + half_sum(dsum, hh) * gs);
        const unsigned long long hm = __ballot(has);
        const unsigned hasbits = (unsigned)(hh ? (hm >> 32) : (hm & 0xffffffffull));
        const long long idx = base_sample + t;
        if (live && idx < F.cap) {
          SgSample &S = F.samples[idx];
          S.f[m] = (float)feat;
          S.r[m] = has ? (float)log(qcell / Qb) : 0.0f;
          S.share[m] = (Dref > 0.0) ? (float)(Cf * qcell / Dref) : 0.0f;
          if (m == 0) {
            S.f[32] = (float)(thk / 1000.0);
            S.f[33] = (float)lgs;
            S.f[34] = 0.0f;
            S.f[35] = 0.0f;
            S.has = hasbits;
            S.col = s;
            S.dref = Dref;
          }
        }
      }
      // the next pair overwrites the enthalpy buffers
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
  }
}

// inference: at most 128 VGPRs (4 waves per SIMD); training samples: no cap
template <bool TARGETS, int NP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(TARGETS ? 1 : 4)))
void k_sg_features(DevParams P, SgColumnArgs C, SgFeatureArgs F) {
  extern __shared__ double2 zsk[];
  __shared__ int bucket[SG_NBUCKET];
  if constexpr (TARGETS) {
    sg_sample_block(P, C, F, blockIdx.x, gridDim.x, zsk, bucket);
  } else {
    sg_infer_features_block<NP>(P, C, F, blockIdx.x, gridDim.x, zsk, bucket);
  }
}

// dynamic LDS of k_sg_features: per level FP64 (z, half spacing) for the
// training samples, FP64 z and FP32 (z, half spacing) in inference; per
// wavefront the staged profiles of a pair of columns (FP64), in inference two
// such buffers (FP32)
static size_t sg_features_lds(int Mz, int waves, bool targets) {
  const size_t Mzp = (size_t)((Mz + 1) & ~1);
  return (targets ? sizeof(double2) * (size_t)Mz : (sizeof(double) + sizeof(float2)) * Mzp) +
         (size_t)waves * 2 * Mzp * (targets ? sizeof(double) : 2 * sizeof(float));
}

void launch_sg_features(const DevParams &P, const SgColumnArgs &C, const SgFeatureArgs &F, bool targets,
                        hipStream_t st) {
  // four wavefronts a block while their enthalpy buffers fit the LDS, else one
  const int waves = sg_features_lds(P.Mz, 4, targets) <= 64 * 1024 ? 4 : 1;
  const int nb = 8192 / waves;  // 8192 wavefronts walk the list in chunks of entries
  const size_t lds = sg_features_lds(P.Mz, waves, targets);
  if (targets) {
    hipLaunchKernelGGL((k_sg_features<true, 3>), dim3(nb), dim3(64 * waves), lds, st, P, C, F);
  } else if (P.nm1_int == 1) {
    hipLaunchKernelGGL((k_sg_features<false, 2>), dim3(nb), dim3(64 * waves), lds, st, P, C, F);
  } else if (P.nm1_int == 2) {
    hipLaunchKernelGGL((k_sg_features<false, 3>), dim3(nb), dim3(64 * waves), lds, st, P, C, F);
  } else {
    hipLaunchKernelGGL((k_sg_features<false, 4>), dim3(nb), dim3(64 * waves), lds, st, P, C, F);
  }
}

// ---------------------------------------------------------------- inference
// A-fragment index of (layer, output tile, k-chunk)
__host__ __device__ constexpr int frag_index(int layer, int t, int c) {
  return layer == 0 ? t * 2 + c : (layer < 4 ? 16 + (layer - 1) * 32 + t * 4 + c : 16 + 96 + t * 4 + c);
}

// input feature of k-slot (g, j) of a chunk: layer 0 reads its inputs from
// memory in natural order; deeper layers take the previous layer's MFMA
// accumulators, whose lane holds rows 4g..4g+3 of two consecutive 16-row tiles
__host__ __device__ constexpr int slot_feature(int layer, int c, int g, int j) {
  return layer == 0 ? 32 * c + 8 * g + j : 32 * c + (j < 4 ? 4 * g + j : 16 + 4 * g + (j - 4));
}

// packs the FP32 master weights into the A-fragments (bf16) of the fused kernel
__global__ void k_sg_pack(const float *__restrict__ w, unsigned short *__restrict__ frag) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;  // element of [SG_NFRAG][64][8]
  if (e >= SG_NFRAG * 512) return;
  const int f = e / 512, l = (e / 8) % 64, j = e % 8;
  int layer, t, c;
  if (f < 16) { layer = 0; t = f / 2; c = f % 2; }
  else if (f < 112) { layer = 1 + (f - 16) / 32; t = ((f - 16) % 32) / 4; c = (f - 16) % 4; }
  else { layer = 4; t = (f - 112) / 4; c = (f - 112) % 4; }
  const int row = 16 * t + (l & 15);
  const int col = slot_feature(layer, c, l >> 4, j);
  float v = 0.0f;
  if (layer == 0) v = (col < SG_NIN) ? w[SG_W1 + row * SG_NIN + col] : 0.0f;
  else if (layer == 1) v = w[SG_W2 + row * SG_WIDTH + col];
  else if (layer == 2) v = w[SG_W3 + row * SG_WIDTH + col];
  else if (layer == 3) v = w[SG_W4 + row * SG_WIDTH + col];
  else v = w[SG_W5 + row * SG_WIDTH + col];
  frag[e] = f2bf(v);
}

void launch_sg_pack(const float *master, unsigned short *frag, hipStream_t st) {
  hipLaunchKernelGGL(k_sg_pack, dim3((SG_NFRAG * 512 + 255) / 256), dim3(256), 0, st, master, frag);
}

constexpr int SG_MLP_BLOCKS = 256;  // one block per CU: the 120 KB of weights fill its LDS
int sg_mlp_blocks() { return SG_MLP_BLOCKS; }

__device__ __forceinline__ bf16x8 load_frag(const unsigned short *fr, int lane) {
  return as_bf16x8(*reinterpret_cast<const uint4 *>(fr + lane * 8));
}
__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// GELU in inference: linear interpolation in a table of GELU at 1024 points
// on [-8, 8) (|error| < 3e-5, below BF16 resolution); beyond the ends the
// first and last segments extrapolate linearly (GELU ~ 0 and ~ x there).
constexpr int SG_GTAB = 1024;
constexpr float SG_GLO = -8.0f, SG_GSCALE = (float)SG_GTAB / 16.0f;
typedef float f32x8 __attribute__((ext_vector_type(8)));

// GELU(a + b) from the table position u = a S + (b - lo) S; ub holds (b - lo) S
__device__ __forceinline__ float gelu_tab(const float2 *tab, float a, float ub) {
  const float u = __builtin_fmaf(a, SG_GSCALE, ub);
  const int i = min(max((int)u, 0), SG_GTAB - 1);  // one v_med3_i32
  const float2 e = tab[i];
  return __builtin_fmaf(e.y, u - (float)i, e.x);
}

// bias + GELU of 8 accumulator tiles -> the 4 B-operand chunks of the next layer
// (ub: the layer's biases as table offsets)
__device__ __forceinline__ void act_to_chunks(const f32x4 *acc, const float *ub, const float2 *tab, int g,
                                              bf16x8 *h) {
#pragma unroll
  for (int c = 0; c < 4; c++) {
    f32x8 v;
    const float4 b0 = *reinterpret_cast<const float4 *>(ub + 16 * (2 * c) + 4 * g);
    const float4 b1 = *reinterpret_cast<const float4 *>(ub + 16 * (2 * c + 1) + 4 * g);
    v[0] = gelu_tab(tab, acc[2 * c][0], b0.x);
    v[1] = gelu_tab(tab, acc[2 * c][1], b0.y);
    v[2] = gelu_tab(tab, acc[2 * c][2], b0.z);
    v[3] = gelu_tab(tab, acc[2 * c][3], b0.w);
    v[4] = gelu_tab(tab, acc[2 * c + 1][0], b1.x);
    v[5] = gelu_tab(tab, acc[2 * c + 1][1], b1.y);
    v[6] = gelu_tab(tab, acc[2 * c + 1][2], b1.z);
    v[7] = gelu_tab(tab, acc[2 * c + 1][3], b1.w);
    h[c] = __builtin_convertvector(v, bf16x8);  // round to nearest even
  }
}

// layer-0 inputs of one sample row: features 8g..8g+7 (chunk 0), 32+8g..+7 (chunk 1);
// returns nonzero when a feature lies outside the training range
__device__ __forceinline__ int load_inputs(const SgMlpArgs &A, const SgNorm &N, int slot, bool live, int g,
                                           bf16x8 &x0, bf16x8 &x1) {
  f32x8 v0, v1;
  int ood = 0;
  const float *fr = A.F + (long)(live ? slot : 0) * SG_FSTRIDE;
#pragma unroll
  for (int jj = 0; jj < 8; jj++) {
    const int f = 8 * g + jj;
    const float v = fr[f];
    ood |= (v < N.lo[f]) | (v > N.hi[f]);
    v0[jj] = (v - N.mean[f]) * N.inv_std[f];
    const int f1 = 32 + 8 * g + jj;
    float w1 = 0.0f;
    if (f1 < SG_NIN) {
      const float r = fr[f1];
      ood |= (r < N.lo[f1]) | (r > N.hi[f1]);
      w1 = (r - N.mean[f1]) * N.inv_std[f1];
    }
    v1[jj] = w1;
  }
  x0 = __builtin_convertvector(v0, bf16x8);
  x1 = __builtin_convertvector(v1, bf16x8);
  ood |= __shfl_xor(ood, 16, 64);
  ood |= __shfl_xor(ood, 32, 64);
  return live ? ood : 0;
}

// Blocks of 16 wavefronts (4 per SIMD) share one copy of the weights in LDS;
// each wavefront evaluates 16-sample tiles.
constexpr int SG_MLP_WAVES = 16;

// Block bid of nblk persistent blocks walking the list (N: the normalization,
// A.norm or a copy of it). LDS: wfrag (dynamic): [SG_NFRAG][64][8] weights,
// then the GELU table; bias: the hidden biases as GELU-table offsets
// (b - lo) S, then the raw output biases; wmax, wfall: one entry per wavefront.
__device__ __forceinline__ void sg_mlp_block(const DevParams &P, const SgMlpArgs &A, const SgNorm &N, int bid,
                                             int nblk, unsigned short *wfrag, float *bias, double *wmax,
                                             int *wfall) {
  float2 *gtab = reinterpret_cast<float2 *>(wfrag + SG_NFRAG * 512);
  {
    const uint4 *src = reinterpret_cast<const uint4 *>(A.frag);
    uint4 *dst = reinterpret_cast<uint4 *>(wfrag);
    for (int e = threadIdx.x; e < SG_NFRAG * 64; e += blockDim.x) dst[e] = src[e];
    for (int e = threadIdx.x; e < SG_WIDTH; e += blockDim.x) {
      bias[e] = (A.bias[SG_B1 + e] - SG_GLO) * SG_GSCALE;
      bias[SG_WIDTH + e] = (A.bias[SG_B2 + e] - SG_GLO) * SG_GSCALE;
      bias[2 * SG_WIDTH + e] = (A.bias[SG_B3 + e] - SG_GLO) * SG_GSCALE;
      bias[3 * SG_WIDTH + e] = (A.bias[SG_B4 + e] - SG_GLO) * SG_GSCALE;
    }
    for (int e = threadIdx.x; e < SG_NOUT; e += blockDim.x) bias[4 * SG_WIDTH + e] = A.bias[SG_B5 + e];
    for (int e = threadIdx.x; e < SG_GTAB; e += blockDim.x) {
      const float x0 = SG_GLO + (float)e / SG_GSCALE, x1 = SG_GLO + (float)(e + 1) / SG_GSCALE;
      const float g0 = gelu_f(x0), g1 = gelu_f(x1);
      gtab[e] = make_float2(g0, g1 - g0);
    }
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int smp = lane & 15, g = lane >> 4;
  const int n = *A.n_list;
  const int ntile = (n + 15) / 16;
  double dmax = 0.0;
  int nfall = 0;
  for (int tile = bid * SG_MLP_WAVES + wv; tile < ntile; tile += nblk * SG_MLP_WAVES) {
    const int slot = tile * 16 + smp;
    const bool live = slot < n;
    bf16x8 x0, x1;
    const int ood = load_inputs(A, N, slot, live, g, x0, x1);
    f32x4 acc[8];
    bf16x8 h[4];
#pragma unroll
    for (int t = 0; t < 8; t++) {
      f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
      c = mfma(load_frag(wfrag + frag_index(0, t, 0) * 512, lane), x0, c);
      acc[t] = mfma(load_frag(wfrag + frag_index(0, t, 1) * 512, lane), x1, c);
    }
    act_to_chunks(acc, bias, gtab, g, h);
    for (int layer = 1; layer < 4; layer++) {
#pragma unroll
      for (int t = 0; t < 8; t++) {
        f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int cc = 0; cc < 4; cc++) c = mfma(load_frag(wfrag + frag_index(layer, t, cc) * 512, lane), h[cc], c);
        acc[t] = c;
        // keep each tile's weight reads next to its MFMAs (hoisting them all spills)
        __builtin_amdgcn_sched_barrier(0);
      }
      act_to_chunks(acc, bias + layer * SG_WIDTH, gtab, g, h);
    }
    f32x4 o0 = {0.0f, 0.0f, 0.0f, 0.0f}, o1 = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int cc = 0; cc < 4; cc++) {
      o0 = mfma(load_frag(wfrag + frag_index(4, 0, cc) * 512, lane), h[cc], o0);
      o1 = mfma(load_frag(wfrag + frag_index(4, 1, cc) * 512, lane), h[cc], o1);
    }
    // D = C S_t + sum_m t_m with t_m = C Qb_m exp(r_m); this lane holds cells
    // 4g..4g+3, 16+4g..16+4g+3. The fallback score sum_m t_m |r_m| / D is the
    // predicted-share-weighted size of the log-correction.
    const float *bo = bias + 4 * SG_WIDTH;
    float part = 0.0f, spart = 0.0f;
    if (live) {
      const float4 q0 = *reinterpret_cast<const float4 *>(A.CQb + (long)slot * SG_NCELL + 4 * g);
      const float4 q1 = *reinterpret_cast<const float4 *>(A.CQb + (long)slot * SG_NCELL + 16 + 4 * g);
      const float rs = N.rscale;
      const float q[8] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w};
      const float o[8] = {o0[0], o0[1], o0[2], o0[3], o1[0], o1[1], o1[2], o1[3]};
#pragma unroll
      for (int c = 0; c < 8; c++) {
        const float r = (o[c] + bo[(c < 4 ? 0 : 16) + 4 * g + (c & 3)]) * rs;
        const float t = q[c] * __expf(r);
        part = (c == 0) ? t : part + t;
        spart = __builtin_fmaf(t, __builtin_fabsf(r), spart);
      }
    }
    part += __shfl_xor(part, 16, 64);
    part += __shfl_xor(part, 32, 64);
    spart += __shfl_xor(spart, 16, 64);
    spart += __shfl_xor(spart, 32, 64);
    if (g == 0 && live) {
      const int s = A.list[slot];
      const double Dp = A.CSt[slot] + (double)part;
      const float score = ood ? -1.0f : (Dp > 0.0 ? (float)((double)spart / Dp) : 0.0f);
      if (A.score) A.score[slot] = score;
      if (ood || score > A.tau) {
        A.status[s] = SG_FALLBACK;
        A.fb_list[atomicAdd(A.n_fb, 1)] = s;
        nfall++;
      } else {
        const double D = Dp;
        const int q = s >> 1, oo = s & 1;
        const int i = q / P.My, j = q - i * P.My;
        const double slope = (oo == 0) ? A.hx[s] : A.hy[s];
        const double Dout = (i >= P.Mx - 1 || j >= P.My - 1) ? 0.0 : D;  // SIAFD.cc:694-697
        A.Dstag[s] = D;
        A.flux[s] = -Dout * slope;
        dmax = fmax(dmax, Dout);
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    dmax = fmax(dmax, __shfl_xor(dmax, off, 64));
    nfall += __shfl_xor(nfall, off, 64);
  }
  if (lane == 0) {
    wmax[wv] = dmax;
    wfall[wv] = nfall;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double mx = wmax[0];
    int nf = wfall[0];
    for (int w = 1; w < SG_MLP_WAVES; w++) {
      mx = fmax(mx, wmax[w]);
      nf += wfall[w];
    }
    A.dmax_part[bid] = mx;
    if (nf) atomicAdd(&A.counts[2], nf);
    if (nf) atomicSub(&A.counts[1], nf);
  }
}

__global__ __launch_bounds__(64 * SG_MLP_WAVES) void k_sg_mlp(DevParams P, SgMlpArgs A) {
  extern __shared__ unsigned short wfrag[];
  __shared__ __align__(16) float bias[4 * SG_WIDTH + SG_NOUT];
  __shared__ double wmax[SG_MLP_WAVES];
  __shared__ int wfall[SG_MLP_WAVES];
  sg_mlp_block(P, A, A.norm, blockIdx.x, gridDim.x, wfrag, bias, wmax, wfall);
}

static size_t sg_mlp_lds() { return (size_t)SG_NFRAG * 512 * sizeof(unsigned short) + sizeof(float2) * SG_GTAB; }

void launch_sg_mlp(const DevParams &P, const SgMlpArgs &A, hipStream_t st) {
  const size_t lds = sg_mlp_lds();
  hipLaunchKernelGGL(k_sg_mlp, dim3(SG_MLP_BLOCKS), dim3(64 * SG_MLP_WAVES), lds, st, P, A);
}

// ---------------------------------------------------------------- ensembles
// Stage grids of all members bound to one surrogate: block b runs block
// b - start[m] of member m's single-member kernel body (the member's table
// entry is at a wave-uniform address: scalar loads). The feature pass and the
// MLP give each member a persistent share of blocks (start[m + 1] - start[m])
// that walks its own list, so each member's rows are those of its own
// single-member update, whatever the share.
__global__ __launch_bounds__(SG_CLS_BLOCK) void k_ens_sg_classify(const EnsSgMember *__restrict__ T,
                                                                   const int *__restrict__ start, int nm) {
  const int m = ens_member(start, nm, blockIdx.x);
  const SgColumnArgs C = global_view(T[m].C);
  const SgListArgs L = global_view(T[m].L);
  sg_classify_block(T[m].P, C, L, blockIdx.x - start[m]);
}

// one block per member
__global__ __launch_bounds__(1024) void k_ens_sg_scan(const EnsSgMember *__restrict__ T,
                                                      const int *__restrict__ start, int nm) {
  __shared__ long long part[1024], partf[1024];
  __shared__ long long tot[3];
  const int m = blockIdx.x;
  if (m >= nm) return;
  const SgListArgs L = global_view(T[m].L);
  sg_scan_block(start[m + 1] - start[m], L, part, partf, tot);
}

__global__ __launch_bounds__(SG_CLS_BLOCK) void k_ens_sg_scatter(const EnsSgMember *__restrict__ T,
                                                                  const int *__restrict__ start, int nm) {
  __shared__ int wsum[2 * SG_CLS_BLOCK / 64];
  const int m = ens_member(start, nm, blockIdx.x);
  const SgListArgs L = global_view(T[m].L);
  sg_scatter_block(T[m].P, L, blockIdx.x - start[m], wsum);
}

template <int NP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_ens_sg_features(const EnsSgMember *__restrict__ T,
                                                         const int *__restrict__ start, int nm) {
  extern __shared__ double2 zsk[];
  __shared__ int bucket[SG_NBUCKET];
  const int m = ens_member(start, nm, blockIdx.x);
  const SgColumnArgs C = global_view(T[m].C);
  const SgFeatureArgs F = global_view(T[m].F);
  sg_infer_features_block<NP>(T[m].P, C, F, blockIdx.x - start[m], start[m + 1] - start[m], zsk, bucket);
}

__global__ __launch_bounds__(64 * SG_MLP_WAVES) void k_ens_sg_mlp(const EnsSgMember *__restrict__ T,
                                                                  const int *__restrict__ start, int nm) {
  extern __shared__ unsigned short wfrag[];
  __shared__ __align__(16) float bias[4 * SG_WIDTH + SG_NOUT];
  __shared__ double wmax[SG_MLP_WAVES];
  __shared__ int wfall[SG_MLP_WAVES];
  const int m = ens_member(start, nm, blockIdx.x);
  const SgMlpArgs M = global_view(T[m].M);
  sg_mlp_block(T[m].P, M, *glb(&T[m].M.norm), blockIdx.x - start[m], start[m + 1] - start[m], wfrag, bias, wmax,
               wfall);
}

int sg_features_waves(int Mz) { return sg_features_lds(Mz, 4, false) <= 64 * 1024 ? 4 : 1; }

void launch_ens_sg(const EnsSgLaunch &L, hipStream_t st) {
  hipLaunchKernelGGL(k_ens_sg_classify, dim3(L.cls_blocks), dim3(SG_CLS_BLOCK), 0, st, L.table, L.cls_start, L.nm);
  hipLaunchKernelGGL(k_ens_sg_scan, dim3(L.nm), dim3(1024), 0, st, L.table, L.cls_start, L.nm);
  hipLaunchKernelGGL(k_ens_sg_scatter, dim3(L.cls_blocks), dim3(SG_CLS_BLOCK), 0, st, L.table, L.cls_start, L.nm);
  const size_t lds = sg_features_lds(L.max_mz, L.feat_waves, false);
  const dim3 fg(L.feat_blocks), fb(64 * L.feat_waves);
  if (L.nm1 == 1) {
    hipLaunchKernelGGL(k_ens_sg_features<2>, fg, fb, lds, st, L.table, L.feat_start, L.nm);
  } else if (L.nm1 == 2) {
    hipLaunchKernelGGL(k_ens_sg_features<3>, fg, fb, lds, st, L.table, L.feat_start, L.nm);
  } else {
    hipLaunchKernelGGL(k_ens_sg_features<4>, fg, fb, lds, st, L.table, L.feat_start, L.nm);
  }
  hipLaunchKernelGGL(k_ens_sg_mlp, dim3(L.mlp_blocks), dim3(64 * SG_MLP_WAVES), sg_mlp_lds(), st, L.table,
                     L.mlp_start, L.nm);
}

// ---------------------------------------------------------------- training
__device__ inline unsigned long long sg_mix(unsigned long long x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// random minibatch: normalized bf16 inputs, targets, shares, masks
__global__ void k_sg_gather(const SgSample *__restrict__ samples, long long n, SgNorm norm, unsigned long long seed,
                            long long step, SgTrainBuffers T) {
  const int b = blockIdx.x;
  const int e = threadIdx.x;  // 64 threads
  const long long idx = (long long)(sg_mix(seed ^ sg_mix((unsigned long long)(step * 1000003ll + b))) % (unsigned long long)n);
  const SgSample &S = samples[idx];
  const float v = (e < SG_NIN) ? (S.f[e] - norm.mean[e]) * norm.inv_std[e] : 0.0f;
  T.X0[(long)b * SG_KIN + e] = f2bf(v);
  if (e < SG_NCELL) {
    T.T[b * SG_NCELL + e] = S.r[e] / norm.rscale;
    T.S[b * SG_NCELL + e] = S.share[e];
    T.Hm[b * SG_NCELL + e] = ((S.has >> e) & 1u) ? 1.0f : 0.0f;
  }
}

void launch_sg_gather(const SgSample *samples, long long n, const SgNorm &norm, unsigned long long seed, long long step,
                      const SgTrainBuffers &T, hipStream_t st) {
  hipLaunchKernelGGL(k_sg_gather, dim3(T.B), dim3(64), 0, st, samples, n, norm, seed, step, T);
}

// explicit inputs (tests): Xn [B][64] f32 already normalized; r in output units
__global__ void k_sg_set_inputs(const float *Xn, const float *r, const float *share, const float *has, int B,
                                SgTrainBuffers T) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < B * SG_KIN) T.X0[e] = f2bf(Xn[e]);
  if (e < B * SG_NCELL) {
    T.T[e] = r[e];
    T.S[e] = share[e];
    T.Hm[e] = has[e];
  }
}

void launch_sg_set_inputs(const float *Xn, const float *r, const float *share, const float *has, int B,
                          const SgTrainBuffers &T, hipStream_t st) {
  hipLaunchKernelGGL(k_sg_set_inputs, dim3((B * SG_KIN + 255) / 256), dim3(256), 0, st, Xn, r, share, has, B, T);
}

// bf16 copies of the weights, W1 padded to 64 inputs: [W1 128x64][W2][W3][W4][W5 32x128]
constexpr int SG_WBF_W1 = 0, SG_WBF_W2 = SG_WIDTH * SG_KIN, SG_WBF_W3 = SG_WBF_W2 + SG_WIDTH * SG_WIDTH,
              SG_WBF_W4 = SG_WBF_W3 + SG_WIDTH * SG_WIDTH, SG_WBF_W5 = SG_WBF_W4 + SG_WIDTH * SG_WIDTH,
              SG_WBF_N = SG_WBF_W5 + SG_NOUT * SG_WIDTH;

__global__ void k_sg_weights_bf16(const float *__restrict__ w, unsigned short *__restrict__ wb) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= SG_WBF_N) return;
  float v;
  if (e < SG_WBF_W2) {
    const int r = e / SG_KIN, c = e % SG_KIN;
    v = (c < SG_NIN) ? w[SG_W1 + r * SG_NIN + c] : 0.0f;
  } else if (e < SG_WBF_W3) v = w[SG_W2 + (e - SG_WBF_W2)];
  else if (e < SG_WBF_W4) v = w[SG_W3 + (e - SG_WBF_W3)];
  else if (e < SG_WBF_W5) v = w[SG_W4 + (e - SG_WBF_W4)];
  else v = w[SG_W5 + (e - SG_WBF_W5)];
  wb[e] = f2bf(v);
}

void launch_sg_weights_bf16(const float *master, unsigned short *Wbf, hipStream_t st) {
  hipLaunchKernelGGL(k_sg_weights_bf16, dim3((SG_WBF_N + 255) / 256), dim3(256), 0, st, master, Wbf);
}

// C[M][N] (f32, row-major, ldc = N) = sum_k A(m,k) B(k,n), with bf16
// A(m,k) = Ap[m sam + k sak], B(k,n) = Bp[k sbk + n sbn]. M, N multiples of 16,
// K a multiple of 32. One wavefront per 16x16 tile.
__global__ __launch_bounds__(256) void k_sg_gemm(int M, int N, int K, const unsigned short *__restrict__ Ap, long sam,
                                                 long sak, const unsigned short *__restrict__ Bp, long sbk, long sbn,
                                                 float *__restrict__ Cm) {
  const int lane = threadIdx.x & 63;
  const int tile = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int tn = N / 16;
  if (tile >= (M / 16) * tn) return;
  const int m0 = (tile / tn) * 16, n0 = (tile % tn) * 16;
  const int r = lane & 15, g = lane >> 4;
  f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int k0 = 0; k0 < K; k0 += 32) {
    unsigned short ua[8], ub[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const long k = k0 + 8 * g + j;
      ua[j] = Ap[(long)(m0 + r) * sam + k * sak];
      ub[j] = Bp[k * sbk + (long)(n0 + r) * sbn];
    }
    bf16x8 a, b;
    __builtin_memcpy(&a, ua, 16);
    __builtin_memcpy(&b, ub, 16);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
  }
#pragma unroll
  for (int i = 0; i < 4; i++) Cm[(long)(m0 + 4 * g + i) * N + n0 + r] = acc[i];
}

static void gemm(int M, int N, int K, const unsigned short *A, long sam, long sak, const unsigned short *B, long sbk,
                 long sbn, float *C, hipStream_t st) {
  const int tiles = (M / 16) * (N / 16);
  hipLaunchKernelGGL(k_sg_gemm, dim3((tiles + 3) / 4), dim3(256), 0, st, M, N, K, A, sam, sak, B, sbk, sbn, C);
}

// Z += bias; A = bf16(gelu(Z)) (hidden) or nothing (output layer)
__global__ void k_sg_bias_act(float *Z, const float *bias, int B, int N, unsigned short *Abf) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= B * N) return;
  const float z = Z[e] + bias[e % N];
  Z[e] = z;
  if (Abf) Abf[e] = f2bf(gelu_f(z));
}

// loss = mean_b [ (sum_m s (y - t))^2 + lambda sum_m h (y - t)^2 ] rscale^2, y = Z5 (output units),
// t = target in output units; dZ5 and its bf16 copy
__global__ void k_sg_loss(const float *Z5, SgTrainBuffers T, float rscale, float lambda, double *loss_acc) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= T.B) return;
  float e = 0.0f, l2 = 0.0f;
  for (int m = 0; m < SG_NCELL; m++) {
    const float d = (Z5[b * SG_NOUT + m] - T.T[b * SG_NCELL + m]) * rscale;
    e += T.S[b * SG_NCELL + m] * d;
    l2 += T.Hm[b * SG_NCELL + m] * d * d;
  }
  const float inv = 1.0f / (float)T.B;
  for (int m = 0; m < SG_NCELL; m++) {
    const float d = (Z5[b * SG_NOUT + m] - T.T[b * SG_NCELL + m]) * rscale;
    const float gr = 2.0f * inv * rscale * (e * T.S[b * SG_NCELL + m] + lambda * T.Hm[b * SG_NCELL + m] * d);
    T.dZ[b * SG_NOUT + m] = gr;
    T.dZbf[b * SG_NOUT + m] = f2bf(gr);
  }
  atomicAdd(loss_acc, (double)((e * e + lambda * l2) * inv));
}

// dZ = dA * gelu'(Z) (f32 and bf16)
__global__ void k_sg_act_bwd(const float *dA, const float *Z, int n, float *dZ, unsigned short *dZbf) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const float v = dA[e] * gelu_grad_f(Z[e]);
  dZ[e] = v;
  dZbf[e] = f2bf(v);
}

// db[n] = sum_b dZ[b][n]: one block per n, fixed-order tree
__global__ __launch_bounds__(256) void k_sg_colsum(const float *dZ, int B, int N, float *db) {
  __shared__ float red[256];
  const int nn = blockIdx.x;
  float s = 0.0f;
  for (int b = threadIdx.x; b < B; b += blockDim.x) s += dZ[(long)b * N + nn];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[nn] = red[0];
}

// copies a [128][64] padded gradient of W1 into the [128][34] master layout
__global__ void k_sg_unpad_w1(const float *dWpad, float *g) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= SG_WIDTH * SG_NIN) return;
  g[SG_W1 + e] = dWpad[(e / SG_NIN) * SG_KIN + (e % SG_NIN)];
}

static inline unsigned nblk(long n) { return (unsigned)((n + 255) / 256); }

void launch_sg_forward_backward(const float *master, const SgTrainBuffers &T, float rscale, float lambda,
                                hipStream_t st) {
  const int B = T.B;
  const int Kin[SG_NLAYER] = {SG_KIN, SG_WIDTH, SG_WIDTH, SG_WIDTH, SG_WIDTH};
  const int Nout[SG_NLAYER] = {SG_WIDTH, SG_WIDTH, SG_WIDTH, SG_WIDTH, SG_NOUT};
  const int woff[SG_NLAYER] = {SG_WBF_W1, SG_WBF_W2, SG_WBF_W3, SG_WBF_W4, SG_WBF_W5};
  const int boff[SG_NLAYER] = {SG_B1, SG_B2, SG_B3, SG_B4, SG_B5};
  const int gw[SG_NLAYER] = {SG_W1, SG_W2, SG_W3, SG_W4, SG_W5};
  launch_sg_weights_bf16(master, T.Wbf, st);
  // forward: Z_l = A_l W_l^T + b_l
  for (int l = 0; l < SG_NLAYER; l++) {
    const unsigned short *Ain = T.Abf[l];
    gemm(B, Nout[l], Kin[l], Ain, Kin[l], 1, T.Wbf + woff[l], 1, Kin[l], T.Z[l], st);
    hipLaunchKernelGGL(k_sg_bias_act, dim3(nblk((long)B * Nout[l])), dim3(256), 0, st, T.Z[l], master + boff[l], B,
                       Nout[l], l + 1 < SG_NLAYER ? T.Abf[l + 1] : nullptr);
  }
  hipLaunchKernelGGL(k_sg_loss, dim3(nblk(B)), dim3(256), 0, st, T.Z[4], T, rscale, lambda, T.loss);
  // backward
  for (int l = SG_NLAYER - 1; l >= 0; l--) {
    const int K = Kin[l], N = Nout[l];
    // dW_l[N][K] = dZ^T A_l  (k-dimension = batch)
    float *dW = (l == 0) ? T.dWpad : T.grad + gw[l];
    gemm(N, K, B, T.dZbf, 1, N, T.Abf[l], K, 1, dW, st);
    if (l == 0) hipLaunchKernelGGL(k_sg_unpad_w1, dim3(nblk(SG_WIDTH * SG_NIN)), dim3(256), 0, st, T.dWpad, T.grad);
    hipLaunchKernelGGL(k_sg_colsum, dim3(N), dim3(256), 0, st, T.dZ, B, N, T.grad + boff[l]);
    if (l > 0) {
      // dA_{l}[B][K] = dZ W_l ; dZ_{l-1} = dA * gelu'(Z_{l-1})
      gemm(B, K, N, T.dZbf, N, 1, T.Wbf + woff[l], K, 1, T.dA, st);
      hipLaunchKernelGGL(k_sg_act_bwd, dim3(nblk((long)B * K)), dim3(256), 0, st, T.dA, T.Z[l - 1], B * K, T.dZ,
                         T.dZbf);
    }
  }
}

__global__ void k_sg_adam(float *w, float *m, float *v, const float *g, float lr, float b1, float b2, float eps,
                          float c1, float c2) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= SG_NPARAM) return;
  const float gg = g[e];
  const float mm = b1 * m[e] + (1.0f - b1) * gg;
  const float vv = b2 * v[e] + (1.0f - b2) * gg * gg;
  m[e] = mm;
  v[e] = vv;
  w[e] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

void launch_sg_adam(float *master, float *m, float *v, const float *grad, float lr, float b1, float b2, float eps,
                    long long t, hipStream_t st) {
  const float c1 = 1.0f - powf(b1, (float)t), c2 = 1.0f - powf(b2, (float)t);
  hipLaunchKernelGGL(k_sg_adam, dim3(nblk(SG_NPARAM)), dim3(256), 0, st, master, m, v, grad, lr, b1, b2, eps, c1, c2);
}

// per-wavefront partial statistics of the features (lane f < SG_NIN takes
// feature f of every sample of the wave's share) and of max |r| (lanes 32..63)
__global__ void k_sg_norm_stats(const SgSample *__restrict__ samples, long long n, double *part) {
  const int lane = threadIdx.x & 63;
  const int gw = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int nw = gridDim.x * (blockDim.x >> 6);
  double s = 0.0, s2 = 0.0, lo = 1e300, hi = -1e300, rmax = 0.0, r2 = 0.0, rn = 0.0;
  for (long long i = gw; i < n; i += nw) {
    const SgSample &S = samples[i];
    if (lane < SG_NIN) {
      const double v = S.f[lane];
      s += v;
      s2 += v * v;
      lo = fmin(lo, v);
      hi = fmax(hi, v);
    }
    if (lane >= 32 && ((S.has >> (lane - 32)) & 1u)) {
      const double r = S.r[lane - 32];
      rmax = fmax(rmax, fabs(r));
      r2 += r * r;
      rn += 1.0;
    }
  }
  if (lane < SG_NIN) {
    double *o = part + ((long)gw * SG_NIN + lane) * 4;
    o[0] = s;
    o[1] = s2;
    o[2] = lo;
    o[3] = hi;
  }
  for (int off = 16; off > 0; off >>= 1) {
    rmax = fmax(rmax, __shfl_xor(rmax, off, 64));
    r2 += __shfl_xor(r2, off, 64);
    rn += __shfl_xor(rn, off, 64);
  }
  if (lane == 32) {
    double *o = part + (long)nw * SG_NIN * 4 + (long)gw * 3;
    o[0] = rmax;
    o[1] = r2;
    o[2] = rn;
  }
}

void launch_sg_norm_stats(const SgSample *samples, long long n, double *partial, hipStream_t st) {
  hipLaunchKernelGGL(k_sg_norm_stats, dim3(SG_STAT_WAVES / 4), dim3(256), 0, st, samples, n, partial);
}

}  // namespace pism_sia
