// sia_kernels.hip -- HIP kernels of the SIA hot path for gfx950 (MI355X).
//
// Field layout is PISM's (IceGrid.cc:183-195): 2-D a[i*My+j], staggered
// a[(i*My+j)*2+o], 3-D a[(i*My+j)*Mz+k]; ghosts are periodic index wraps.
//
// Kernels:
//   k_smooth_mean / k_smooth_coeffs  BedSmoother (PISMBedSmoother.cc:162-269)
//   k_gradient                       surface gradient, all 3 methods (SIAFD.cc:210-503),
//                                    get_theta + get_smoothed_thk (PISMBedSmoother.cc:286-420)
//   k_column<LAW, FULL, CPW>         diffusive flux, D, D_max, and (FULL) the column
//                                    integral I (SIAFD.cc:546-722, 819-879); a 64-lane
//                                    wavefront walks CPW (32 or 64) staggered columns,
//                                    one lane per level, wave-tree reductions and scans
//   k_velocity_tile<LAW, FASTLN>     u, v (SIAFD.cc:918-958), w (PISMStressBalance.cc:209-286)
//                                    and (LAW >= 0) the volumetric strain heating
//                                    (PISMStressBalance.cc:370-500) in one pass over
//                                    4x4 tiles of map columns
//   k_mass_continuity, k_mask_surface, k_dt, accounting reductions
//                                    iMgeometry.cc:56-149, 497-822; iMadaptive.cc:118-171
#include "sia_kernels.hpp"

namespace pism_sia {

// ---------------------------------------------------------------- wave helpers
// Cross-lane moves of doubles by DPP (one 32-bit move per half). With
// BOUND_ZERO a lane whose source lies outside its row reads 0; rows outside
// ROW_MASK keep the old value 0. Callers keep all 64 lanes active.
template <int CTRL, int ROW_MASK, bool BOUND_ZERO>
__device__ __forceinline__ double dpp_f64(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROW_MASK, 0xf, BOUND_ZERO);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROW_MASK, 0xf, BOUND_ZERO);
  return __hiloint2double(hi, lo);
}

// value of lane l-1, 0 in lane 0 (wave_shr:1)
__device__ __forceinline__ double wave_shift_up1(double v) { return dpp_f64<0x138, 0xf, true>(v); }
// value of lane l-1, `first` in lane 0 (wave_shr:1; without bound_ctrl a lane
// whose source lies outside the wave keeps the old value)
__device__ __forceinline__ double wave_shift_up1_from(double v, double first) {
  const int lo = __builtin_amdgcn_update_dpp(__double2loint(first), __double2loint(v), 0x138, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(__double2hiint(first), __double2hiint(v), 0x138, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}
// value of lane l+1, 0 in lane 63 (wave_shl:1)
__device__ __forceinline__ double wave_shift_down1(double v) { return dpp_f64<0x130, 0xf, true>(v); }

// inclusive prefix sum over the 64 lanes: Hillis-Steele with row_shr 1, 2,
// 4, 8 inside each row of 16 lanes, then row_bcast:15 adds the total of row
// 0 (2) to row 1 (3), and row_bcast:31 the total of lanes 0-31 to rows 2, 3.
__device__ __forceinline__ double wave_inclusive_scan(double v) {
  v += dpp_f64<0x111, 0xf, true>(v);
  v += dpp_f64<0x112, 0xf, true>(v);
  v += dpp_f64<0x114, 0xf, true>(v);
  v += dpp_f64<0x118, 0xf, true>(v);
  v += dpp_f64<0x142, 0xa, false>(v);
  v += dpp_f64<0x143, 0xc, false>(v);
  return v;
}

// value held by lane l (l wave-uniform), as a scalar
__device__ __forceinline__ double lane_value(double v, int l) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ int lane_value(int v, int l) { return __builtin_amdgcn_readlane(v, l); }

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

constexpr int WAVES_PER_BLOCK = 4;

// ---------------------------------------------------------------- work ordering
// MI355X dispatches workgroups round-robin over its 8 XCDs, each with its own
// 4 MB L2; kernels whose neighbouring work items share data map blocks to
// XCDs explicitly (blockIdx.x % N_XCD is the XCD).
constexpr int N_XCD = 8;

__global__ void k_reset(unsigned long long *dmax_bits, int *err, int *err_acc, unsigned long long *cfl_bits,
                        double max_dt) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    *dmax_bits = 0ull;  // +0.0
    *err_acc |= *err;
    *err = 0;
    *cfl_bits = (unsigned long long)__double_as_longlong(max_dt);
  }
}

// clears the current and the accumulated error flags after they were read
__global__ void k_clear_flags(int *err, int *err_acc) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    *err = 0;
    *err_acc = 0;
  }
}

void launch_clear_flags(int *err, int *err_acc, hipStream_t st) {
  hipLaunchKernelGGL(k_clear_flags, dim3(1), dim3(64), 0, st, err, err_acc);
}

__global__ void k_fill(double *__restrict__ a, long n, double v) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) a[q] = v;
}

// ---------------------------------------------------------------- bed smoother
// smooth_the_bed_on_proc0 (PISMBedSmoother.cc:162-193): box mean over in-grid
// points, same summation order as the reference.
__global__ void k_smooth_mean(DevParams P, const double *__restrict__ b0, double *__restrict__ bs) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int Mx = P.Mx, My = P.My;
  if (q >= (long)Mx * My) return;
  int i, j;
  map_ij(q, My, i, j);
  double sum = 0.0, count = 0.0;
  for (int r = -P.Nx; r <= P.Nx; r++) {
    for (int s = -P.Ny; s <= P.Ny; s++) {
      if ((i + r >= 0) && (i + r < Mx) && (j + s >= 0) && (j + s < My)) {
        sum += b0[(long)(i + r) * My + (j + s)];
        count += 1.0;
      }
    }
  }
  bs[q] = sum / count;
}

// compute_coefficients_on_proc0 (PISMBedSmoother.cc:196-269)
__global__ void k_smooth_coeffs(DevParams P, double s2, double s3, double s4,
                                const double *__restrict__ b0, const double *__restrict__ bs,
                                double *__restrict__ mt, double *__restrict__ c2,
                                double *__restrict__ c3, double *__restrict__ c4) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int Mx = P.Mx, My = P.My;
  if (q >= (long)Mx * My) return;
  int i, j;
  map_ij(q, My, i, j);
  const double topgs = bs[q];
  double maxtltemp = 0.0, sum2 = 0.0, sum3 = 0.0, sum4 = 0.0, count = 0.0;
  for (int r = -P.Nx; r <= P.Nx; r++) {
    for (int s = -P.Ny; s <= P.Ny; s++) {
      if ((i + r >= 0) && (i + r < Mx) && (j + s >= 0) && (j + s < My)) {
        const double tl = b0[(long)(i + r) * My + (j + s)] - topgs;
        maxtltemp = (maxtltemp < tl) ? tl : maxtltemp;
        const double tl2 = tl * tl;
        sum2 += tl2;
        sum3 += tl2 * tl;
        sum4 += tl2 * tl2;
        count += 1.0;
      }
    }
  }
  mt[q] = maxtltemp;
  c2[q] = (sum2 / count) * s2;
  c3[q] = (sum3 / count) * s3;
  c4[q] = (sum4 / count) * s4;
}

// ---------------------------------------------------------------- gradient

__device__ __forceinline__ long idx2(const DevParams &P, int i, int j) {
  return (long)wrapi(i, P.Mx) * P.My + wrapi(j, P.My);
}

// haseloff first pass (SIAFD.cc:402-438): x-derivative on the i-offset grid and its weight
template <class Acc>
__device__ __forceinline__ void haseloff_x(const DevParams &P, const Acc &F, int i, int j, double &hx0,
                                           double &wi) {
  const int m0 = F.mask(i, j), m1 = F.mask(i + 1, j);
  const double h0 = F.h(i, j), h1 = F.h(i + 1, j);
  if ((m_floating_ice(m0) && m_ice_free_ocean(m1)) || (m_ice_free_ocean(m0) && m_floating_ice(m1))) {
    hx0 = 0.0; wi = 0.0;
  } else if ((m_icy(m0) && m_ice_free(m1) && F.b(i + 1, j) > h0) || (m_ice_free(m0) && m_icy(m1) && F.b(i, j) > h1)) {
    hx0 = 0.0; wi = 0.0;
  } else {
    hx0 = (h1 - h0) / P.dx; wi = 1.0;
  }
}
template <class Acc>
__device__ __forceinline__ void haseloff_y(const DevParams &P, const Acc &F, int i, int j, double &hy1,
                                           double &wj) {
  const int m0 = F.mask(i, j), m1 = F.mask(i, j + 1);
  const double h0 = F.h(i, j), h1 = F.h(i, j + 1);
  if ((m_floating_ice(m0) && m_ice_free_ocean(m1)) || (m_ice_free_ocean(m0) && m_floating_ice(m1))) {
    hy1 = 0.0; wj = 0.0;
  } else if ((m_icy(m0) && m_ice_free(m1) && F.b(i, j + 1) > h0) || (m_ice_free(m0) && m_icy(m1) && F.b(i, j) > h1)) {
    hy1 = 0.0; wj = 0.0;
  } else {
    hy1 = (h1 - h0) / P.dy; wj = 1.0;
  }
}

// surface gradient at map point q = (i, j) on both staggered grids; F reads
// H, b, h and the mask at (i', j') for |i' - i|, |j' - j| <= 1
template <class Acc>
__device__ __forceinline__ void gradient_at(const DevParams &P, const Acc &F, double *__restrict__ hx,
                                            double *__restrict__ hy, long q, int i, int j) {
  const double dx = P.dx, dy = P.dy;
  double hx0, hy0, hx1, hy1;
  if (P.grad == 1) {  // mahaffy, SIAFD.cc:305-316
#define HH(a, c) F.h((a), (c))
    hx0 = (HH(i + 1, j) - HH(i, j)) / dx;
    hy0 = (+HH(i + 1, j + 1) + HH(i, j + 1) - HH(i + 1, j - 1) - HH(i, j - 1)) / (4.0 * dy);
    hy1 = (HH(i, j + 1) - HH(i, j)) / dy;
    hx1 = (+HH(i + 1, j + 1) + HH(i + 1, j) - HH(i - 1, j + 1) - HH(i - 1, j)) / (4.0 * dx);
#undef HH
  } else if (P.grad == 2) {  // eta, SIAFD.cc:210-284
    const double n = P.n, etapow = (2.0 * n + 2.0) / n, invpow = 1.0 / etapow,
                 dinvpow = (-n - 2.0) / (2.0 * n + 2.0);
#define ETA(a, c) pow(F.H((a), (c)), etapow)
#define BB(a, c) F.b((a), (c))
    {
      const double e00 = ETA(i, j), e10 = ETA(i + 1, j);
      const double mean_eta = 0.5 * (e10 + e00);
      if (mean_eta > 0.0) {
        const double factor = invpow * pow(mean_eta, dinvpow);
        hx0 = factor * (e10 - e00) / dx;
        hy0 = factor * (+ETA(i + 1, j + 1) + ETA(i, j + 1) - ETA(i + 1, j - 1) - ETA(i, j - 1)) / (4.0 * dy);
      } else {
        hx0 = 0.0; hy0 = 0.0;
      }
      hx0 += (BB(i + 1, j) - BB(i, j)) / dx;
      hy0 += (BB(i + 1, j + 1) + BB(i, j + 1) - BB(i + 1, j - 1) - BB(i, j - 1)) / (4 * dy);
    }
    {
      const double e00 = ETA(i, j), e01 = ETA(i, j + 1);
      const double mean_eta = 0.5 * (e01 + e00);
      if (mean_eta > 0.0) {
        const double factor = invpow * pow(mean_eta, dinvpow);
        hy1 = factor * (e01 - e00) / dy;
        hx1 = factor * (+ETA(i + 1, j + 1) + ETA(i + 1, j) - ETA(i - 1, j + 1) - ETA(i - 1, j)) / (4.0 * dx);
      } else {
        hy1 = 0.0; hx1 = 0.0;
      }
      hy1 += (BB(i, j + 1) - BB(i, j)) / dy;
      hx1 += (BB(i + 1, j + 1) + BB(i + 1, j) - BB(i - 1, j + 1) - BB(i - 1, j)) / (4 * dx);
    }
#undef ETA
#undef BB
  } else {  // haseloff, SIAFD.cc:366-503
    double wi_c, wj_c;
    haseloff_x(P, F, i, j, hx0, wi_c);
    haseloff_y(P, F, i, j, hy1, wj_c);
    const bool icy = m_icy(F.mask(i, j));
    // x-derivative, j-offset (SIAFD.cc:444-470)
    {
      double a1, w1, a2, w2, a3, w3;  // (i-1,j), (i-1,j+1), (i,j+1)
      haseloff_x(P, F, i - 1, j, a1, w1);
      haseloff_x(P, F, i - 1, j + 1, a2, w2);
      haseloff_x(P, F, i, j + 1, a3, w3);
      if (wj_c > 0) {
        const double W = wi_c + w1 + w2 + w3;
        hx1 = (W > 0) ? 1.0 / W * (hx0 + a1 + a2 + a3) : 0.0;
      } else if (icy) {
        const double W = wi_c + w1;
        hx1 = (W > 0) ? 1.0 / W * (hx0 + a1) : 0.0;
      } else {
        const double W = w3 + w2;
        hx1 = (W > 0) ? 1.0 / W * (a2 + a3) : 0.0;
      }
    }
    // y-derivative, i-offset (SIAFD.cc:472-498)
    {
      double a1, w1, a2, w2, a3, w3;  // (i,j-1), (i+1,j-1), (i+1,j)
      haseloff_y(P, F, i, j - 1, a1, w1);
      haseloff_y(P, F, i + 1, j - 1, a2, w2);
      haseloff_y(P, F, i + 1, j, a3, w3);
      if (wi_c > 0) {
        const double W = wj_c + w1 + w2 + w3;
        hy0 = (W > 0) ? 1.0 / W * (hy1 + a1 + a2 + a3) : 0.0;
      } else if (icy) {
        const double W = wj_c + w1;
        hy0 = (W > 0) ? 1.0 / W * (hy1 + a1) : 0.0;
      } else {
        const double W = w2 + w3;
        hy0 = (W > 0) ? 1.0 / W * (a2 + a3) : 0.0;
      }
    }
  }
  hx[2 * q] = hx0;
  hx[2 * q + 1] = hx1;
  hy[2 * q] = hy0;
  hy[2 * q + 1] = hy1;
}

// field reads through global memory, with the periodic index wraps
struct GlobalFields {
  const DevParams &P;
  const Fields2D &F;
  __device__ double H(int i, int j) const { return F.H[idx2(P, i, j)]; }
  __device__ double b(int i, int j) const { return F.b[idx2(P, i, j)]; }
  __device__ double h(int i, int j) const { return F.h[idx2(P, i, j)]; }
  __device__ int mask(int i, int j) const { return as_int(F.mask[idx2(P, i, j)]); }
};

__device__ __forceinline__ void gradient_point(const DevParams &P, const Fields2D &F, double *__restrict__ hx,
                                               double *__restrict__ hy, long q) {
  int i, j;
  map_ij(q, P.My, i, j);
  gradient_at(P, GlobalFields{P, F}, hx, hy, q, i, j);
}

// k_gradient: a block stages a tile of GT_I x GT_J map points and its ring of
// one point (periodic wraps) in LDS, then each thread evaluates one point from
// there with the same arithmetic as gradient_point, and theta and the smoothed
// thickness of that point (theta_thk_point, which reads no neighbour)
constexpr int GT_I = 8, GT_J = 32, GT_SJ = GT_J + 2, GT_N = (GT_I + 2) * GT_SJ;

__device__ __forceinline__ void theta_thk_point(const DevParams &P, const Fields2D &F, const double *__restrict__ topgs,
                                                const double *__restrict__ maxtl, const double *__restrict__ C2,
                                                const double *__restrict__ C3, const double *__restrict__ C4,
                                                double *__restrict__ theta, double *__restrict__ thks, int *err,
                                                long q);

struct TileFields {
  const double *sH, *sb, *sh;
  const int *sm;
  int i0, j0;  // global index of the tile's first point; the ring starts at i0 - 1, j0 - 1
  __device__ int at(int i, int j) const { return (i - i0 + 1) * GT_SJ + (j - j0 + 1); }
  __device__ double H(int i, int j) const { return sH[at(i, j)]; }
  __device__ double b(int i, int j) const { return sb[at(i, j)]; }
  __device__ double h(int i, int j) const { return sh[at(i, j)]; }
  __device__ int mask(int i, int j) const { return sm[at(i, j)]; }
};

__global__ __launch_bounds__(GT_I * GT_J) void k_gradient(DevParams P, Fields2D F, double *__restrict__ hx,
                                                          double *__restrict__ hy, ThetaArgs T) {
  __shared__ double sH[GT_N], sb[GT_N], sh[GT_N];
  __shared__ int sm[GT_N];
  const int tj = (P.My + GT_J - 1) / GT_J;
  const int i0 = (blockIdx.x / tj) * GT_I, j0 = (blockIdx.x % tj) * GT_J;
  const bool eta = P.grad == 2, has_b = P.grad != 1;
  for (int e = threadIdx.x; e < GT_N; e += blockDim.x) {
    const int a = e / GT_SJ, c = e - a * GT_SJ;
    const long g = idx2(P, i0 - 1 + a, j0 - 1 + c);
    sh[e] = F.h[g];
    sm[e] = as_int(F.mask[g]);
    if (has_b) sb[e] = F.b[g];
    if (eta) sH[e] = F.H[g];
  }
  __syncthreads();
  const int i = i0 + (int)threadIdx.x / GT_J, j = j0 + (int)threadIdx.x % GT_J;
  if (i >= P.Mx || j >= P.My) return;
  const long q = (long)i * P.My + j;
  gradient_at(P, TileFields{sH, sb, sh, sm, i0, j0}, hx, hy, q, i, j);
  theta_thk_point(P, F, T.topgs, T.maxtl, T.C2, T.C3, T.C4, T.theta, T.thks, T.err, q);
}

// SIAFD_Regional::compute_surface_gradient's override (regional.cc:45-109): where
// the no-model mask touches a staggered gradient's stencil, the gradient of the
// stored surface usurfstore (zero where the stencil leaves the domain)
__global__ void k_regional_gradient(DevParams P, const double *__restrict__ nmm, const double *__restrict__ hst,
                                    double *__restrict__ hx, double *__restrict__ hy) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int Mx = P.Mx, My = P.My;
  if (q >= (long)Mx * My) return;
  int i, j;
  map_ij(q, My, i, j);
#define NM(a, c) nmm[idx2(P, (a), (c))]
#define HS(a, c) hst[idx2(P, (a), (c))]
  if (NM(i, j) > 0.5 || NM(i + 1, j) > 0.5) {
    hx[2 * q] = (i + 1 > Mx - 1) ? 0.0 : (HS(i + 1, j) - HS(i, j)) / P.dx;
  }
  if (NM(i - 1, j + 1) > 0.5 || NM(i + 1, j + 1) > 0.5 || NM(i - 1, j) > 0.5 || NM(i + 1, j) > 0.5) {
    hx[2 * q + 1] = (i - 1 < 0 || j + 1 > My - 1 || i + 1 > Mx - 1)
                        ? 0.0
                        : (+HS(i + 1, j + 1) + HS(i + 1, j) - HS(i - 1, j + 1) - HS(i - 1, j)) / (4.0 * P.dx);
  }
  if (NM(i, j + 1) > 0.5 || NM(i + 1, j + 1) > 0.5 || NM(i, j - 1) > 0.5 || NM(i + 1, j - 1) > 0.5) {
    hy[2 * q] = (j + 1 > My - 1 || i + 1 > Mx - 1 || j - 1 < 0)
                    ? 0.0
                    : (+HS(i + 1, j + 1) + HS(i, j + 1) - HS(i + 1, j - 1) - HS(i, j - 1)) / (4.0 * P.dy);
  }
  if (NM(i, j) > 0.5 || NM(i, j + 1) > 0.5) {
    hy[2 * q + 1] = (j + 1 > My - 1) ? 0.0 : (HS(i, j + 1) - HS(i, j)) / P.dy;
  }
#undef NM
#undef HS
}

// ---------------------------------------------------------------- theta, thksmooth
// get_theta (PISMBedSmoother.cc:360-420) and get_smoothed_thk (286-336)
__device__ __forceinline__ void theta_thk_point(const DevParams &P, const Fields2D &F, const double *__restrict__ topgs,
                                                const double *__restrict__ maxtl, const double *__restrict__ C2,
                                                const double *__restrict__ C3, const double *__restrict__ C4,
                                                double *__restrict__ theta, double *__restrict__ thks, int *err,
                                                long q) {
  const double usurf = F.h[q], thk = F.H[q], mt = maxtl[q], bs = topgs[q];
  // theta
  double th;
  if (P.Nx < 0 || P.Ny < 0) {
    th = 1.0;
  } else {
    const double H = usurf - bs;
    if (H > mt) {
      const double Hinv = 1.0 / ((H < 1.0) ? 1.0 : H);
      double omega = 1.0 + Hinv * Hinv * (C2[q] + Hinv * (C3[q] + Hinv * C4[q]));
      if (omega <= 0) atomicOr(err, ERR_OMEGA);
      if (omega < 0.001) omega = 0.001;
      th = pow(omega, -P.n);
      if (th > 1.0) th = 1.0;
      if (th < 0.0) th = 0.0;
    } else {
      th = 0.00;
    }
  }
  theta[q] = th;
  // smoothed thickness
  double r;
  if (thk < 0.0) {
    atomicOr(err, ERR_NEG_THK);
    r = 0.0;
  } else if (thk == 0.0) {
    r = 0.0;
  } else if (mt >= thk) {
    r = thk;
  } else if (m_grounded(as_int(F.mask[q]))) {
    const double thks_try = usurf - bs;
    r = (thks_try > 0.0) ? thks_try : 0.0;
  } else {
    r = thk;
  }
  thks[q] = r;
}

// ---------------------------------------------------------------- column kernel
// compute_diffusive_flux (SIAFD.cc:620-721) and, when FULL, compute_I
// (SIAFD.cc:845-878). Each wavefront owns COLS_PER_WAVE consecutive
// staggered columns s = 2*(i*My+j)+o. Lane c first prepares column c (2-D
// inputs, kBelowHeight), so the 2-D loads of all its columns are in flight
// together; the wave then walks the columns in chunks of 64 levels, one lane
// per level, and the enthalpy of the next chunk (of this column or of the
// next one) is loaded while the current chunk is evaluated. D, flux and ks
// leave as one coalesced store per wave. D_max is reduced per block into
// dmax_part[block]; a one-block kernel takes the max of the partials.
// GPBLD with the default enthalpy converter and an integer Glen exponent
// takes flow_gpbld_fast, every other configuration the reference-form flow<LAW>
// (the choice depends on the parameters only, so a column gets the same
// arithmetic from every kernel that computes it).
// dynamic LDS of the column kernels: (z_k, dz_k) and z_k per level
static inline size_t column_lds(int Mz) { return 3 * sizeof(double) * (size_t)Mz; }

// block bid of the column grid with COLS_PER_WAVE = CPW; zs: column_lds(Mz)
// bytes of LDS (16-byte aligned), wmax: WAVES_PER_BLOCK doubles of LDS. With a
// list, the block's columns are list entries bid * WAVES_PER_BLOCK * CPW .. (n_list
// of them in all) instead of staggered columns. Returns the block's D_max in
// thread 0.

template <int LAW, bool FULL, int CPW, int FASTNM1>
__device__ __forceinline__ double column_block_impl(const DevParams &P, const ColumnArgs &A, int bid, double *zs,
                                                    double *wmax, const int *list, int n_list, bool strips) {
  constexpr int COLS_PER_WAVE = CPW;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int Mx = P.Mx, My = P.My, Mz = P.Mz;
  const int nstag = 2 * Mx * My;
  // LDS (column_lds bytes): per level (z_k, z_k - z_{k-1}) (0 at k = 0), one
  // 16-byte read a level; then z_k alone for kBelowHeight
  double2 *zh = reinterpret_cast<double2 *>(zs);
  double *zl = zs + 2 * Mz;
  for (int k = threadIdx.x; k < Mz; k += blockDim.x) {
    const double zk = A.z[k];
    zl[k] = zk;
    zh[k] = make_double2(zk, (k > 0) ? zk - A.z[k - 1] : 0.0);
  }
  __syncthreads();
  const int s0 = (bid * WAVES_PER_BLOCK + wv) * COLS_PER_WAVE;
  // list mode: lane c < COLS_PER_WAVE takes list entry s0 + c (nstag: none).
  // strips (32 or 64 columns a wavefront): the wavefront takes the staggered
  // columns of one row of a strip of COLS_PER_WAVE / 2 map columns along j;
  // strip k goes to XCD k % 8, and each XCD walks its strips row by row, so
  // the rows i and i + 1 that both read the enthalpy of column (i + 1, j) run
  // close in time on one L2 (in staggered order they are a row of the grid
  // apart, mostly on different XCDs). Otherwise consecutive staggered columns.
  int s;
  if (list) {
    s = (lane < COLS_PER_WAVE && s0 + lane < n_list) ? list[s0 + lane] : nstag;
  } else if ((COLS_PER_WAVE == 32 || COLS_PER_WAVE == 64) && strips) {
    const long seg = (long)(bid >> 3) * WAVES_PER_BLOCK + wv;  // row segment of this XCD
    const int sl = (int)(seg / Mx), row = (int)(seg - (long)sl * Mx);
    const int strip = sl * N_XCD + (bid & (N_XCD - 1));
    const int j = strip * (COLS_PER_WAVE / 2) + (lane >> 1);
    s = (lane < COLS_PER_WAVE && j < My) ? 2 * (row * My + j) + (lane & 1) : nstag;
  } else {
    s = s0 + lane;
  }

  // lane c < COLS_PER_WAVE: column s0 + c. c_ks: -3 no column, -2 zero thickness (delta = 0, SIAFD.cc:635-641), -1
  // thickness above Lz, else kBelowHeight
  int c_ks = -3, c_qa = 0, c_qb = 0;
  double c_thk = 0.0, c_alpha = 0.0, c_c0 = 0.0, c_slope = 0.0;
  bool c_edge = false;
  if (lane < COLS_PER_WAVE && s < nstag) {
    const int q = s >> 1, o = s & 1;
    const int i = q / My, j = q - i * My;
    const int q2 = (o == 0) ? wrapi(i + 1, Mx) * My + j : i * My + wrapi(j + 1, My);
    c_qa = q;
    c_qb = q2;
    c_thk = 0.5 * (A.thks[q] + A.thks[q2]);
    c_edge = (i >= Mx - 1 || j >= My - 1);  // SIAFD.cc:694-697
    if (c_thk == 0.0) {
      c_ks = -2;
    } else {
      c_ks = kbelow(zl, Mz, P.Lz, c_thk);
      if (c_ks < 0) {
        atomicOr(A.err, ERR_THK_LZ);
        c_ks = -1;
      }
    }
    if (c_ks >= 0) {
      const double hxs = A.hx[s], hys = A.hy[s];
      c_slope = (o == 0) ? hxs : hys;
      c_alpha = sqrt(hxs * hxs + hys * hys);
      const double theta_local = 0.5 * (A.theta[q] + A.theta[q2]);
      c_c0 = P.e * theta_local * 2.0;
    }
  }

  auto Icol_of = [&](int c) -> double * {
    const int s_c = (list || strips) ? lane_value(s, c) : s0 + c;
    return ((s_c & 1) ? A.I1 : A.I0) + (long)(s_c >> 1) * i_stride(Mz);
  };
  // first column >= c with a level to evaluate. I is 0 at level 0 of every
  // column (the integral from the base), and the columns passed over have ks
  // stored as 0: readers take I = 0 at index 0 and never read it, so it is
  // not written.
  auto next_valid = [&](int c) -> int {
    for (; c < COLS_PER_WAVE; c++) {
      const int kk = lane_value(c_ks, c);
      if (kk == -3) return COLS_PER_WAVE;
      if (kk >= 0) return c;
    }
    return COLS_PER_WAVE;
  };
  constexpr bool FASTGP = FASTNM1 > 0;
  const bool with_age = !FASTGP && P.use_age;
  // the chunk after (c, k0): the rest of column c, else the next column's first
  auto advance = [&](int &c, int &k0) {
    if (c >= COLS_PER_WAVE) return;
    k0 += 64;
    if (k0 > lane_value(c_ks, c)) {
      c = next_valid(c + 1);
      k0 = 0;
    }
  };
  // enthalpy (and age) of the chunk (c, k0): level k0 + lane of both cells;
  // lanes above ks load level ks (a valid address, the value is not used).
  // The column bases are wave-uniform (scalar registers); a lane adds its
  // 32-bit byte offset.
  int lc = -1, l_ks = 0;
  long l_qa = 0, l_qb = 0;
  auto at = [](const double *base, long col, unsigned off) {
    return *reinterpret_cast<const double *>(reinterpret_cast<const char *>(base + col) + off);
  };
  auto load = [&](int c, int k0, double &ea, double &eb, double &ga, double &gb) {
    if (c >= COLS_PER_WAVE) return;
    if (c != lc) {  // wave-uniform: the column's values are read once
      lc = c;
      l_ks = lane_value(c_ks, c);
      l_qa = (long)lane_value(c_qa, c) * Mz;
      l_qb = (long)lane_value(c_qb, c) * Mz;
    }
    const unsigned off = (unsigned)min(k0 + lane, l_ks) * 8u;
    ea = at(A.E, l_qa, off);
    eb = at(A.E, l_qb, off);
    if (with_age) {
      ga = at(A.age, l_qa, off);
      gb = at(A.age, l_qb, off);
    }
  };

  double myD = 0.0;  // lane c: D of column c before the edge override
  // three chunks in flight: the one evaluated (B0) and the next two (B1, B2),
  // whose loads are issued two chunks ahead of their use; the buffers move up
  // after each chunk (unrolling the ring instead was measured slower)
  int c = next_valid(0), k0 = 0;
  int c1 = c, k1 = 0;
  advance(c1, k1);
  int c2 = c1, k2 = k1;
  advance(c2, k2);
  double B0[4] = {0.0, 0.0, 0.0, 0.0}, B1[4] = {0.0, 0.0, 0.0, 0.0}, B2[4] = {0.0, 0.0, 0.0, 0.0};
  load(c, 0, B0[0], B0[1], B0[2], B0[3]);
  load(c1, k1, B1[0], B1[1], B1[2], B1[3]);
  double Dpart = 0.0, dprev_carry = 0.0, Icarry = 0.0;
  // values of the column being evaluated, read once per column
  int vc = -1, ks = 0;
  double thk = 0.0, alpha = 0.0, c0 = 0.0, top = 0.0;
  // evaluates chunk (c, k0) from Ecur, issues chunk (c2, k2) into Enext
  auto step = [&](const double (&Ecur)[4], double (&Enext)[4]) -> bool {
    if (c >= COLS_PER_WAVE) return false;
    if (c != vc) {
      vc = c;
      ks = lane_value(c_ks, c);
      thk = lane_value(c_thk, c);
      alpha = lane_value(c_alpha, c);
      c0 = lane_value(c_c0, c);
      const double dzt = thk - zl[ks];  // SIAFD.cc:682-684
      top = 0.5 * dzt * dzt;
    }
    load(c2, k2, Enext[0], Enext[1], Enext[2], Enext[3]);

    const int k = k0 + lane;
    const bool act = k <= ks;
    const int kz = min(k, Mz - 1);
    const double2 zk2 = zh[kz];  // z_kz, z_kz - z_{kz-1} (0 at kz = 0)
    // depth >= 0 on the active lanes, where this is EnthalpyConverter::pressure
    // (enthalpyConverter.cc:92-98: p_air + rho g 0 = p_air at depth 0)
    const double depth = thk - zk2.x;
    const double pressure = P.p_air + P.rho_i * P.g * depth;
    const double Ek = 0.5 * (Ecur[0] + Ecur[1]);
    double fl;
    if (FASTGP) {
      fl = flow_gpbld_fast<(FASTNM1 > 0 ? FASTNM1 : 1)>(P, alpha * pressure, Ek, pressure);
    } else {
      const double gs = with_age ? grain_size_vostok(P, 0.5 * (Ecur[2] + Ecur[3])) : P.gs;
      fl = flow<LAW>(P, alpha * pressure, Ek, pressure, gs);
    }
    // lanes above ks evaluate a value no sum takes: every use below is masked
    // to the active lanes, and lane k > 0 reads lane k - 1, active when k is
    const double delta = c0 * pressure * fl;
    const double dz = zk2.y;
    const double dprev = wave_shift_up1_from(delta, dprev_carry);
    const double trap = 0.5 * dz * ((depth + dz) * dprev + depth * delta);  // SIAFD.cc:677-680
    if (act && k > 0) Dpart += trap;
    if (FULL) {
      const double t = (act && k > 0) ? 0.5 * dz * (dprev + delta) : 0.0;  // SIAFD.cc:863-867
      const double Ik = Icarry + wave_inclusive_scan(t);
      if (act && k > 0) Icol_of(c)[k] = Ik;
      Icarry = lane_value(Ik, 63);
    }
    dprev_carry = lane_value(delta, 63);
    if (c1 != c) {
      // the chunk holds level ks: D = the trapezoids + d_ks^2 / 2 delta_ks
      // (SIAFD.cc:677-684)
      const double D = lane_value(wave_inclusive_scan(Dpart), 63) + top * lane_value(delta, ks - k0);
      if (lane == c) myD = D;
      Dpart = 0.0;
      dprev_carry = 0.0;
      Icarry = 0.0;
    }
    c = c1;
    k0 = k1;
    c1 = c2;
    k1 = k2;
    advance(c2, k2);
    return true;
  };
  while (step(B0, B2)) {
#pragma unroll
    for (int t = 0; t < 4; t++) {
      B0[t] = B1[t];
      B1[t] = B2[t];
    }
  }

  double dmax = 0.0;
  if (lane < COLS_PER_WAVE && c_ks > -3) {
    const double Dout = c_edge ? 0.0 : myD;
    A.Dstag[s] = myD;
    A.flux[s] = (c_ks < 0) ? 0.0 : -Dout * c_slope;
    if (FULL) A.ks[s] = (c_ks < 0) ? 0 : c_ks;
    dmax = Dout;
  }
  dmax = wave_max(dmax);  // D >= 0: +0.0 is the identity (std::max, SIAFD.cc:699)
  if (lane == 0) wmax[wv] = dmax;
  __syncthreads();
  double m = 0.0;
  if (threadIdx.x == 0) {
    m = wmax[0];
    for (int w = 1; w < WAVES_PER_BLOCK; w++) m = (m < wmax[w]) ? wmax[w] : m;
  }
  return m;
}

// flow_gpbld_fast applies (see column_block_impl)
__device__ __forceinline__ bool use_gpbld_fast(const DevParams &P) {
  return P.ec == EC_DEFAULT && P.nm1_int > 0 && !P.use_age;
}

template <int LAW, bool FULL, int CPW>
__device__ __forceinline__ double column_block(const DevParams &P, const ColumnArgs &A, int bid, double *zs,
                                               double *wmax, const int *list = nullptr, int n_list = 0,
                                               bool strips = false) {
  if (LAW == LAW_GPBLD && use_gpbld_fast(P)) {
    if (P.nm1_int == 1) return column_block_impl<LAW, FULL, CPW, 1>(P, A, bid, zs, wmax, list, n_list, strips);
    if (P.nm1_int == 2) return column_block_impl<LAW, FULL, CPW, 2>(P, A, bid, zs, wmax, list, n_list, strips);
    return column_block_impl<LAW, FULL, CPW, 3>(P, A, bid, zs, wmax, list, n_list, strips);
  }
  return column_block_impl<LAW, FULL, CPW, 0>(P, A, bid, zs, wmax, list, n_list, strips);
}

// columns a wavefront of the plain pass. On the 1501x2801x401 sheet 64 (in
// strips of 32 map columns) is faster than 32 (in staggered order the two
// measured the same; 8, 16 and 48 slower); on grids of 801x801 and below,
// stepped one at a time, 64 is slower: half as many wavefronts leave more of
// the chip idle in the last round. 64 is taken where the grid gives each of
// the chip's 4096 wavefront slots (256 CUs x 4 SIMDs x 4 waves) at least 16
// wavefronts.
constexpr int COLS_PER_WAVE_WIDE = 64, COLS_PER_WAVE_NARROW = 32;
static int column_cpw(const DevParams &P) {
  return 2L * P.Mx * P.My >= 16L * 4096 * COLS_PER_WAVE_WIDE ? COLS_PER_WAVE_WIDE : COLS_PER_WAVE_NARROW;
}
// the widest of the ensemble's tile widths (8, 16, 32 columns a wavefront),
// which takes strips
constexpr int ENS_STRIP_CPW = 32;
// the FP64 pass over a list (the surrogate's fallback columns, a small share of
// deep columns) takes 4 columns per wavefront, so that more wavefronts share it
constexpr int LIST_COLS_PER_WAVE = 4;
constexpr int LIST_COLS_PER_BLOCK = WAVES_PER_BLOCK * LIST_COLS_PER_WAVE;

// wavefronts per SIMD the column kernels are compiled for (at most 512 / n
// VGPRs each, the rest spilled). The level walk waits on its enthalpy loads
// more than on issue, so occupancy pays for the spills. On the 1501x2801x401
// sheet, against 4 (107 VGPRs, none spilled), fast updates: 5 waves -3 %, 6
// -4 to -5 %, 7 -0 to -1 %, 8 +14 %; full updates: 6 -10 %, 7 -12 %, 8 -16 %.
// The 64-member ensemble's column stage (101-201 levels): 5 -6 %, 6 -10 %, 7
// -11 %, 8 +10 %.
constexpr int COLUMN_WAVES_FAST = 6, COLUMN_WAVES_FULL = 8, COLUMN_WAVES_ENS = 7;

template <int LAW, bool FULL, int CPW>
__global__ __launch_bounds__(256, FULL ? COLUMN_WAVES_FULL : COLUMN_WAVES_FAST) void k_column(DevParams P,
                                                                                             ColumnArgs A,
                                                                                             int strips) {
  extern __shared__ __align__(16) double zs[];  // column_lds(Mz) bytes
  __shared__ double wmax[WAVES_PER_BLOCK];
  const double m = column_block<LAW, FULL, CPW>(P, A, blockIdx.x, zs, wmax, nullptr, 0, strips != 0);
  if (threadIdx.x == 0) A.dmax_part[blockIdx.x] = m;
}

// fast-update column pass over a list of staggered columns (the surrogate's
// FP64 fallback): gridDim.x blocks walk the list's chunks of COLS_PER_BLOCK;
// each block's D_max goes to dmax_part[blockIdx.x]
template <int LAW>
__global__ __launch_bounds__(256) void k_column_list(DevParams P, ColumnArgs A, const int *__restrict__ list,
                                                     const int *__restrict__ n_list) {
  extern __shared__ __align__(16) double zs[];  // column_lds(Mz) bytes
  __shared__ double wmax[WAVES_PER_BLOCK];
  const int n = *n_list;
  double m = 0.0;
  for (int vb = blockIdx.x; vb * LIST_COLS_PER_BLOCK < n; vb += gridDim.x) {
    const double mb = column_block<LAW, false, LIST_COLS_PER_WAVE>(P, A, vb, zs, wmax, list, n);
    m = (m < mb) ? mb : m;
  }
  if (threadIdx.x == 0) A.dmax_part[blockIdx.x] = m;
}

// max of the per-block partials (one block); D >= 0 so +0.0 is the identity
__global__ void k_reduce_max(const double *__restrict__ part, int n, unsigned long long *dmax_bits) {
  __shared__ double red[1024];
  double m = 0.0;
  for (int b = threadIdx.x; b < n; b += blockDim.x) m = (m < part[b]) ? part[b] : m;
  red[threadIdx.x] = m;
  __syncthreads();
  for (int off = blockDim.x >> 1; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) red[threadIdx.x] = (red[threadIdx.x] < red[threadIdx.x + off]) ? red[threadIdx.x + off] : red[threadIdx.x];
    __syncthreads();
  }
  if (threadIdx.x == 0) *dmax_bits = (unsigned long long)__double_as_longlong(red[0]);
}

// ---------------------------------------------------------------- u, v, w
// -0.25 * (I_e h_e + I_w h_w + I_n h_n + I_s h_s) + basal (SIAFD.cc:949-956)
__device__ __forceinline__ double uv_formula(double Ie, double he, double Iw, double hw, double In, double hn,
                                             double Is, double hs, double basal) {
  return -0.25 * (Ie * he + Iw * hw + In * hn + Is * hs) + basal;
}

__device__ __forceinline__ double strain_D2(double u_x, double u_y, double u_z, double v_x, double v_y,
                                            double v_z) {  // PISMStressBalance.cc:309-311
  return 0.5 * ((u_x + v_y) * (u_x + v_y) + u_x * u_x + v_y * v_y +
                0.5 * ((u_y + v_x) * (u_y + v_x) + u_z * u_z + v_z * v_z));
}

// ---------------------------------------------------------------- u, v, w by tiles
// compute_3d_horizontal_velocity (SIAFD.cc:918-958) fused with
// StressBalance::compute_vertical_velocity (PISMStressBalance.cc:209-286) and,
// for LAW >= 0, the volumetric strain heating (PISMStressBalance.cc:370-500),
// computed by tiles of TI x TJ map columns:
// per chunk of KC levels, u and v of the tile and of a ring of one column
// around it (without the ring's corners, which no difference reads) go to LDS
// once (each from the four I columns of its faces, with the reference's
// formula, so a ring value is bitwise the one its own tile stores); the w
// integrand, w (a prefix scan over the wavefront's lanes, one level each) and
// the strain heating read them there.
// A chunk holds the levels k0-1 .. k0+KC of u and v (u_z at k0 and k0+KC-1);
// its slot KC (level k0+KC-1) becomes the next chunk's slot 0. Sigma =
// 2 e^(-1/n) B D^(1+1/n) is evaluated as 2 e^(-1/n) exp(-ln A / n + (1/n + 1)/2
// ln D^2) with ln A from the flow law, 0 above kBelowHeight(H). Above the tile's
// largest kBelowHeight (of its faces and of H) u, v and the integrand are
// constant and the strain heating is 0, so those levels are written directly,
// w by the trapezoid of a constant integrand (S + int (z_k - z_top)), on
// 128-byte lines; a tile whose faces and columns all lie below the second
// level has no chunk at all and is written as contiguous rows. Each column's
// KC levels of a chunk leave as runs of 512 bytes (shorter runs of this
// layout, whose columns are not aligned to cache lines, write HBM at a
// fraction of the rate).
// workgroup barrier that orders LDS only: the global stores in flight are
// not waited for (__syncthreads would wait for them at every chunk)
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

namespace vtile {
constexpr int TI = 4, TJ = 4, KC = 64;
// u, v columns: the tile and the ring around it without its corners, which
// no difference reads: row a = -1 (TJ), rows 0 .. TI-1 (TJ + 2 each), row TI (TJ)
constexpr int RJ = TJ + 2, NR = 2 * TJ + TI * RJ;
constexpr int FJ0 = TJ + 2, NF0 = (TI + 3) * FJ0;                    // I0 faces (a, b): a in [-2, TI], b in [-1, TJ]
constexpr int FJ1 = TJ + 3, NF1 = (TI + 2) * FJ1;                    // I1 faces (a, b): a in [-1, TI], b in [-2, TJ]
constexpr int NF = NF0 + NF1, NC = TI * TJ, NS = KC + 2;
constexpr int VW = NC / 4;  // wavefronts of a tile block: 4 tile columns each in phase C
__device__ __forceinline__ int f0(int a, int b) { return (a + 2) * FJ0 + (b + 1); }
__device__ __forceinline__ int f1(int a, int b) { return NF0 + (a + 1) * FJ1 + (b + 2); }
__device__ __forceinline__ int rr(int a, int b) {
  return (a < 0) ? b : ((a < TI) ? TJ + a * RJ + (b + 1) : TJ + TI * RJ + b);
}
// (a, b) of ring slot r
__device__ __forceinline__ void rr_inv(int r, int &a, int &b) {
  if (r < TJ) {
    a = -1;
    b = r;
  } else if (r < TJ + TI * RJ) {
    a = (r - TJ) / RJ;
    b = (r - TJ) - a * RJ - 1;
  } else {
    a = TI;
    b = r - TJ - TI * RJ;
  }
}
__device__ __forceinline__ int wrapm(int i, int M) { const int r = i % M; return r < 0 ? r + M : r; }
}  // namespace vtile

// Tiles go to the XCDs in bands of TILE_BAND tile columns: XCD x takes bands
// x, x + 8, ..., each walked row by row (j fastest), so a tile's neighbours
// along both i and j run on the same XCD at about the same time and the I
// columns of the faces they share come from one L2; the bands interleave so
// that every XCD gets a share of the ice sheet's middle. Bands of 4 measured
// fastest on the 1501x2801x401 sheet (1, 8 and 16 slower).
constexpr int TILE_BAND = 4;
struct TileMap {
  int ti, tj, per_xcd;  // tiles along i and j; blocks reserved per XCD
};

__host__ __device__ inline TileMap tile_map(int Mx, int My) {
  TileMap m;
  m.ti = (Mx + vtile::TI - 1) / vtile::TI;
  m.tj = (My + vtile::TJ - 1) / vtile::TJ;
  const int bands = (m.tj + TILE_BAND - 1) / TILE_BAND;
  m.per_xcd = ((bands + N_XCD - 1) / N_XCD) * m.ti * TILE_BAND;
  return m;
}

// tile (ti, tj) of block b; false for a block past the grid
__device__ __forceinline__ bool tile_of_block(const TileMap &tm, int b, int &ti, int &tj) {
  const int x = b % N_XCD, L = b / N_XCD;
  const int per_band = tm.ti * TILE_BAND;
  const int bi = L / per_band, r = L - bi * per_band;
  ti = r / TILE_BAND;
  tj = (x + N_XCD * bi) * TILE_BAND + (r - ti * TILE_BAND);
  return tj < tm.tj;
}

template <int LAW, bool FASTLN>
__device__ __forceinline__ double log_softness_sel(const DevParams &P, double E, double p) {
  if (FASTLN) return log_softness_gpbld_fast(P, E, p);
  return log_softness<LAW>(P, E, p);
}

template <int LAW, bool FASTLN>
__global__ __launch_bounds__(64 * vtile::VW, 4) void k_velocity_tile(DevParams P, VelArgs A, TileMap tm) {
  using namespace vtile;
  constexpr bool HEAT = LAW >= 0;
  // 37.3 KB of static LDS and z (8 Mz bytes): up to Mz = 459, 4 blocks (16
  // wavefronts) fit on a CU. z is read from LDS so that the store loops (the
  // fill, the rows of tiles below the second level) issue no vector-memory
  // load: one would wait for every store before it on the in-order
  // vector-memory counter.
  extern __shared__ double zs[];
  __shared__ int fcq[NF];  // map column of the face's I column
  __shared__ int fks[NF];
  __shared__ double fhx[NF], fhy[NF];
  __shared__ double rub[NR], rvb[NR];
  __shared__ double sU[NR][NS], sV[NR][NS];
  __shared__ double cDx[NC], cDy[NC], cw0[NC], cH[NC], cS[NC], cI[NC];
  __shared__ int cks[NC], cq[NC], cfac[NC];  // cfac: bit 0 west, 1 east, 2 south, 3 north
  __shared__ int s_kend;
  const int tid = threadIdx.x;
  const int Mx = P.Mx, My = P.My, Mz = P.Mz;
  const double *zg = zs;
  int ti, tj;
  if (!tile_of_block(tm, blockIdx.x, ti, tj)) return;
  const int i0 = ti * TI, j0 = tj * TJ;
  for (int k = tid; k < Mz; k += blockDim.x) zs[k] = A.z[k];
  if (tid == 0) s_kend = 0;
  __syncthreads();
  int kend = 0;
  // faces: I column, kBelowHeight, h_x, h_y
  for (int f = tid; f < NF; f += blockDim.x) {
    int a, b, o;
    if (f < NF0) { a = f / FJ0 - 2; b = f % FJ0 - 1; o = 0; }
    else { a = (f - NF0) / FJ1 - 1; b = (f - NF0) % FJ1 - 2; o = 1; }
    const int gi = wrapm(i0 + a, Mx), gj = wrapm(j0 + b, My);
    const int sid = 2 * (gi * My + gj) + o;
    fcq[f] = gi * My + gj;
    const int k = A.ks[sid];
    fks[f] = k;
    fhx[f] = A.hx[sid];
    fhy[f] = A.hy[sid];
    kend = max(kend, k);
  }
  // ring columns: basal velocity
  for (int r = tid; r < NR; r += blockDim.x) {
    int a, b;
    rr_inv(r, a, b);
    const long q = (long)wrapm(i0 + a, Mx) * My + wrapm(j0 + b, My);
    rub[r] = A.ub ? A.ub[q] : 0.0;
    rvb[r] = A.vb ? A.vb[q] : 0.0;
  }
  // tile columns: margin-aware differences (PISMStressBalance.cc:223-260), w_0, H and its kBelowHeight
  if (tid < NC) {
    const int a = tid / TJ, b = tid % TJ;
    const int gi = i0 + a, gj = j0 + b;
    int ksH = -1;
    if (gi < Mx && gj < My) {
      const int q = gi * My + gj;
      cq[tid] = q;
      const int im = wrapm(gi - 1, Mx), ip = wrapm(gi + 1, Mx), jm = wrapm(gj - 1, My), jp = wrapm(gj + 1, My);
      const int m = as_int(A.mask[q]);
      const int me = as_int(A.mask[ip * My + gj]), mw = as_int(A.mask[im * My + gj]);
      const int mn = as_int(A.mask[gi * My + jp]), ms = as_int(A.mask[gi * My + jm]);
      double west = 1, east = 1, south = 1, north = 1;
      if ((m_icy(m) && m_ice_free(me)) || (m_ice_free(m) && m_icy(me))) east = 0;
      if ((m_icy(m) && m_ice_free(mw)) || (m_ice_free(m) && m_icy(mw))) west = 0;
      if ((m_icy(m) && m_ice_free(mn)) || (m_ice_free(m) && m_icy(mn))) north = 0;
      if ((m_icy(m) && m_ice_free(ms)) || (m_ice_free(m) && m_icy(ms))) south = 0;
      cDx[tid] = (east + west > 0) ? 1.0 / (P.dx * (east + west)) : 0.0;
      cDy[tid] = (north + south > 0) ? 1.0 / (P.dy * (north + south)) : 0.0;
      cfac[tid] = (west != 0 ? 1 : 0) | (east != 0 ? 2 : 0) | (south != 0 ? 4 : 0) | (north != 0 ? 8 : 0);
      cw0[tid] = A.bmr ? -A.bmr[q] : 0.0;
      if (HEAT) {
        const double Hq = A.H[q];
        cH[tid] = Hq;
        ksH = kbelow(zg, Mz, P.Lz, Hq);  // throws above Lz (IceGrid.cc:381-389)
        if (ksH < 0) atomicOr(A.err, ERR_THK_LZ);
        kend = max(kend, ksH);
      }
    } else {
      cq[tid] = -1;
    }
    cks[tid] = ksH;
    cS[tid] = 0.0;
    cI[tid] = 0.0;
  }
  kend = max(kend, __shfl_xor(kend, 1));  // cheap partial reduction before the LDS atomic
  atomicMax(&s_kend, kend);
  __syncthreads();
  kend = min(s_kend, Mz - 1);

  int k0 = 0;
  // u and v of ring slot r at level k0 - 1 + sl (slot sl) from the I of its four faces
  auto uv_at = [&](int r, int sl) {
    const int kk = min(k0 - 1 + sl, Mz - 1);
    int a, b;
    rr_inv(r, a, b);
    const int fe = f0(a, b), fw = f0(a - 1, b), fn = f1(a, b), fs = f1(a, b - 1);
    // I is 0 at level 0 (and never stored there)
    const long Is_ = i_stride(Mz);
    const int le = min(kk, fks[fe]), lw = min(kk, fks[fw]), ln = min(kk, fks[fn]), ls = min(kk, fks[fs]);
    const double Ie = le > 0 ? A.I0[(long)fcq[fe] * Is_ + le] : 0.0;
    const double Iw = lw > 0 ? A.I0[(long)fcq[fw] * Is_ + lw] : 0.0;
    const double In = ln > 0 ? A.I1[(long)fcq[fn] * Is_ + ln] : 0.0;
    const double Is = ls > 0 ? A.I1[(long)fcq[fs] * Is_ + ls] : 0.0;
    sU[r][sl] = uv_formula(Ie, fhx[fe], Iw, fhx[fw], In, fhx[fn], Is, fhx[fs], rub[r]);
    sV[r][sl] = uv_formula(Ie, fhy[fe], Iw, fhy[fw], In, fhy[fn], Is, fhy[fs], rvb[r]);
  };
  if (kend == 0) {
    // Every face of the tile has kBelowHeight 0, so its I is 0 at every level,
    // and every column's H lies below the second level: u and v of every column
    // are their base values at every level, the integrand is constant (w is
    // its trapezoid, w_0 - int z_k) and the strain heating is 0 above the
    // base. Each tile row (TJ map columns, consecutive in memory) is written
    // as one contiguous run.
    if (tid < NR) uv_at(tid, 1);
    lds_barrier();
    if (tid < NC && cq[tid] >= 0) {
      const int c = tid, a = c / TJ, b = c - a * TJ;
      const int r = rr(a, b), rw = rr(a - 1, b), re = rr(a + 1, b), rs = rr(a, b - 1), rn = rr(a, b + 1);
      const double uc = sU[r][1], vc = sV[r][1];
      const double Dx = cDx[c], Dy = cDy[c];
      const int fac = cfac[c];
      const double west = (fac & 1) ? 1.0 : 0.0, east = (fac & 2) ? 1.0 : 0.0;
      const double south = (fac & 4) ? 1.0 : 0.0, north = (fac & 8) ? 1.0 : 0.0;
      const double u_x = Dx * (west * (uc - sU[rw][1]) + east * (sU[re][1] - uc));
      const double v_y = Dy * (south * (vc - sV[rs][1]) + north * (sV[rn][1] - vc));
      double sig = 0.0;
      if (HEAT && cks[c] >= 0) {
        const double pressure = ec_pressure(P, cH[c] - zg[0]);
        const double lnA = log_softness_sel<LAW, FASTLN>(P, glb(A.E)[(long)cq[c] * Mz], pressure);
        const double hu_y = Dy * (south * (uc - sU[rs][1]) + north * (sU[rn][1] - uc));
        const double hv_x = Dx * (west * (vc - sV[rw][1]) + east * (sV[re][1] - vc));
        const double d2 = strain_D2(u_x, hu_y, 0.0, hv_x, v_y, 0.0);  // u, v constant in z: u_z = v_z = 0
        sig = (d2 > 0.0) ? 2.0 * P.e_pow * exp_fast(P.hard_pow * lnA + P.heat_exp * log_fast(d2, *(const LogCoef *)P.log_c), *(const ExpCoef *)P.exp_c) : 0.0;
      }
      cI[c] = u_x + v_y;
      cS[c] = sig;  // the strain heating at the base
    }
    lds_barrier();
    const float inv_Mz = 1.0f / (float)Mz;
#pragma unroll 1
    for (int a = 0; a < TI; a++) {
      if (i0 + a >= Mx) break;
      const int n = min(TJ, My - j0) * Mz;
      const long base = ((long)(i0 + a) * My + j0) * Mz;
      // lanes on 128-byte lines: a wavefront's store covers 4 whole lines
      const int sh = (int)(base & 15);
      for (int e = tid - sh; e < n; e += blockDim.x) {
        if (e < 0) continue;
        int b = (int)((float)e * inv_Mz), k = e - b * Mz;  // e = b Mz + k
        if (k < 0) { b--; k += Mz; } else if (k >= Mz) { b++; k -= Mz; }
        const int c = a * TJ + b, r = rr(a, b);
        A.u[base + e] = sU[r][1];
        A.v[base + e] = sV[r][1];
        A.w[base + e] = cw0[c] - cI[c] * zg[k];
        if (HEAT) A.Sigma[base + e] = (k == 0) ? cS[c] : 0.0;
      }
    }
    return;
  }
  for (; k0 <= kend; k0 += KC) {
    // enthalpy of this chunk's levels of the wavefront's tile columns (phase
    // C), loaded now so that phase B hides its latency
    double Ec[NC / VW];
    {
      const int lane = tid & 63;
#pragma unroll
      for (int t = 0; t < NC / VW; t++) {
        const int c = (tid >> 6) + t * VW;
        const int k = k0 + lane;
        Ec[t] = (HEAT && cq[c] >= 0 && k <= cks[c]) ? glb(A.E)[(long)cq[c] * Mz + k] : 0.0;
      }
    }
    // level coordinates of this chunk's lanes: z_k, z_{k-1} and the centred
    // difference's 1 / (z_{k+1} - z_{k-1}) (1 / (z_1 - z_0) at k = 0)
    const int kl = min(k0 + (tid & 63), Mz - 1);
    const double z_k = zg[kl], z_km1 = zg[max(kl - 1, 0)];
    const double idz_k = HEAT ? A.idz[kl] : 0.0;
    // phase B: u, v at levels k0 .. k0+KC-1 (slots 1 .. KC), one wavefront
    // per ring column, one lane per level; then level k0+KC (slot KC+1) of
    // every ring column. Slot 0 (level k0-1) came from the last chunk.
    {
      const int lane = tid & 63;
#pragma unroll 4
      for (int r = tid >> 6; r < NR; r += VW) uv_at(r, lane + 1);
      if (tid < NR) uv_at(tid, KC + 1);
    }
    lds_barrier();
    // phase C: integrand, w and the strain heating of levels k0 .. k0+KC-1;
    // one wavefront per tile column, one lane per level
#pragma unroll
    for (int tc = 0; tc < NC / VW; tc++) {
      const int c = (tid >> 6) + tc * VW;
      const int lane = tid & 63;
      const int a = c / TJ, b = c - a * TJ;
      const int r = rr(a, b);
      const int q = cq[c];
      const int k = k0 + lane, sl = lane + 1;
      const bool act = q >= 0 && k < Mz;
      const double uc = sU[r][sl], vc = sV[r][sl];
      const int rw = rr(a - 1, b), re = rr(a + 1, b), rs = rr(a, b - 1), rn = rr(a, b + 1);
      const double uw = sU[rw][sl], ue = sU[re][sl], vs = sV[rs][sl], vn = sV[rn][sl];
      const double Dx = cDx[c], Dy = cDy[c];
      const int fac = cfac[c];
      const double west = (fac & 1) ? 1.0 : 0.0, east = (fac & 2) ? 1.0 : 0.0;
      const double south = (fac & 4) ? 1.0 : 0.0, north = (fac & 8) ? 1.0 : 0.0;
      const double u_x = Dx * (west * (uc - uw) + east * (ue - uc));
      const double v_y = Dy * (south * (vc - vs) + north * (vn - vc));
      double integrand = act ? u_x + v_y : 0.0;
      double iprev = wave_shift_up1(integrand);
      if (lane == 0) iprev = cI[c];
      double dz = 0.0;
      if (act && k > 0) dz = z_k - z_km1;
      const double t = (act && k > 0) ? 0.5 * (integrand + iprev) * dz : 0.0;  // PISMStressBalance.cc:280-282
      const double S = cS[c] + wave_inclusive_scan(t);
      double sig = 0.0;
      if (HEAT && act && k <= cks[c]) {
        const double us = sU[rs][sl], un = sU[rn][sl], vw = sV[rw][sl], ve = sV[re][sl];
        const double pressure = ec_pressure(P, cH[c] - z_k);
        const double lnA = log_softness_sel<LAW, FASTLN>(P, Ec[tc], pressure);
        const double hu_x = u_x, hu_y = Dy * (south * (uc - us) + north * (un - uc)),
                     hv_x = Dx * (west * (vc - vw) + east * (ve - vc)), hv_y = v_y;
        // one-sided at the base (PISMStressBalance.cc:464-476)
        const double u_z = (k > 0) ? (sU[r][sl + 1] - sU[r][sl - 1]) * idz_k : (sU[r][sl + 1] - uc) * idz_k;
        const double v_z = (k > 0) ? (sV[r][sl + 1] - sV[r][sl - 1]) * idz_k : (sV[r][sl + 1] - vc) * idz_k;
        const double d2 = strain_D2(hu_x, hu_y, u_z, hv_x, hv_y, v_z);
        sig = (d2 > 0.0) ? 2.0 * P.e_pow * exp_fast(P.hard_pow * lnA + P.heat_exp * log_fast(d2, *(const LogCoef *)P.log_c), *(const ExpCoef *)P.exp_c) : 0.0;
      }
      if (act) {
        const long o = (long)q * Mz + k;
        A.u[o] = uc;
        A.v[o] = vc;
        A.w[o] = cw0[c] - S;
        if (HEAT) A.Sigma[o] = sig;
      }
      // carries of the wavefront's last lane (the last level of the chunk)
      if (lane == 63) {
        cS[c] = S;
        cI[c] = integrand;
      }
    }
    lds_barrier();
    // slot KC (level k0+KC-1) becomes the next chunk's slot 0
    if (tid < NR) {
      sU[tid][0] = sU[tid][KC];
      sV[tid][0] = sV[tid][KC];
    }
    lds_barrier();
  }
  // levels k0 .. Mz-1: constant u, v, integrand, Sigma = 0; w continues the trapezoid
  if (k0 < Mz) {
    const double ztop = zg[k0 - 1];
#pragma unroll 1
    for (int c = 0; c < NC; c++) {
      const int q = cq[c];
      if (q < 0) continue;
      const int a = c / TJ, b = c - a * TJ;
      const int r = rr(a, b);
      const double ut = sU[r][0], vt = sV[r][0];  // level k0 - 1 (moved to slot 0)
      const double it_ = cI[c], St = cS[c], w0 = cw0[c];
      const long base = (long)q * Mz;
      // lanes on 128-byte lines: a wavefront's store covers 4 whole lines
      const int sh = (int)((base + k0) & 15);
      for (int k = k0 + tid - sh; k < Mz; k += blockDim.x) {
        if (k < k0) continue;
        A.u[base + k] = ut;
        A.v[base + k] = vt;
        A.w[base + k] = w0 - (St + it_ * (zg[k] - ztop));
        if (HEAT) A.Sigma[base + k] = 0.0;
      }
    }
  }
}

// ---------------------------------------------------------------- diagnostics
// IceModelVec3::getSurfaceValues (iceModelVec3.cc:206-222) = getValZ at z = H
// (iceModelVec3.cc:131-155): linear interpolation between the levels that
// bracket z (GSL bsearch: largest m <= Mz - 2 with z_m <= z), the end levels
// beyond the column.
__global__ void k_surface_values(DevParams P, const double *__restrict__ z, const double *__restrict__ f,
                                 const double *__restrict__ H, double *__restrict__ out) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (long)P.Mx * P.My) return;
  const int Mz = P.Mz;
  const double *col = f + q * Mz;
  const double zq = H[q];
  double r;
  if (zq >= z[Mz - 1]) {
    r = col[Mz - 1];
  } else if (zq <= z[0]) {
    r = col[0];
  } else {
    int lo = 0, hi = Mz - 1;
    while (hi > lo + 1) {
      const int mid = (lo + hi) >> 1;
      if (z[mid] > zq) hi = mid; else lo = mid;
    }
    const double incr = (zq - z[lo]) / (z[lo + 1] - z[lo]);
    const double valm = col[lo];
    r = valm + incr * (col[lo + 1] - valm);
  }
  out[q] = r;
}

// SIAFD::compute_diffusivity (SIAFD.cc:736-744): the staggered D averaged to
// the regular grid by IceModelVec2Stag::staggered_to_regular (iceModelVec2.cc:679-689)
__global__ void k_diffusivity_regular(DevParams P, const double *__restrict__ D, double *__restrict__ out) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (long)P.Mx * P.My) return;
  int i, j;
  map_ij(q, P.My, i, j);
  out[q] = 0.25 * (D[2 * q] + D[2 * q + 1] + D[2 * idx2(P, i, j - 1) + 1] + D[2 * idx2(P, i - 1, j)]);
}

// ---------------------------------------------------------------- SIA_Sliding
// SIA_Sliding::surface_gradient_haseloff (SIA_Sliding.cc:374-498): thickness
// threshold H <= 0 marks ice-free points, whose pseudo-surface is the
// neighbour's where the bed stands above it. Its mahaffy and eta variants are
// SIAFD's (SIA_Sliding.cc:292-372, 501-531).
__global__ void k_sliding_gradient(DevParams P, SlideArgs A) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (long)P.Mx * P.My) return;
  if (P.grad != 0) {
    gradient_point(P, A.F, A.hx, A.hy, q);
    return;
  }
  int i, j;
  map_ij(q, P.My, i, j);
#define HH(a, c) A.F.H[idx2(P, (a), (c))]
#define BB(a, c) A.F.b[idx2(P, (a), (c))]
#define SS(a, c) A.F.h[idx2(P, (a), (c))]
  {  // i-offset
    const bool icefreeP = HH(i, j) <= 0.0, icefreeE = HH(i + 1, j) <= 0.0, icefreeN = HH(i, j + 1) <= 0.0,
               icefreeS = HH(i, j - 1) <= 0.0, icefreeNE = HH(i + 1, j + 1) <= 0.0,
               icefreeSE = HH(i + 1, j - 1) <= 0.0;
    double hhE = SS(i + 1, j);
    if (icefreeE && (BB(i + 1, j) > SS(i, j))) hhE = SS(i, j);
    if (icefreeP && (BB(i, j) > SS(i + 1, j))) hhE = SS(i, j);
    A.hx[2 * q] = (hhE - SS(i, j)) / P.dx;
    double hhN = SS(i, j + 1);
    if (icefreeN && (BB(i, j + 1) > SS(i, j))) hhN = SS(i, j);
    if (icefreeP && (BB(i, j) > SS(i, j + 1))) hhN = SS(i, j);
    double hhS = SS(i, j - 1);
    if (icefreeS && (BB(i, j - 1) > SS(i, j))) hhS = SS(i, j);
    if (icefreeP && (BB(i, j) > SS(i, j - 1))) hhS = SS(i, j);
    double hhNE = SS(i + 1, j + 1);
    if (icefreeNE && (BB(i + 1, j + 1) > SS(i + 1, j))) hhNE = SS(i + 1, j);
    if (icefreeE && (BB(i + 1, j) > SS(i + 1, j + 1))) hhNE = SS(i + 1, j);
    double hhSE = SS(i + 1, j - 1);
    if (icefreeSE && (BB(i + 1, j - 1) > SS(i + 1, j))) hhSE = SS(i + 1, j);
    if (icefreeE && (BB(i + 1, j) > SS(i + 1, j - 1))) hhSE = SS(i + 1, j);
    A.hy[2 * q] = (hhNE + hhN - hhSE - hhS) / (4.0 * P.dy);
  }
  {  // j-offset
    const bool icefreeP = HH(i, j) <= 0.0, icefreeN = HH(i, j + 1) <= 0.0, icefreeE = HH(i + 1, j) <= 0.0,
               icefreeW = HH(i - 1, j) <= 0.0, icefreeNE = HH(i + 1, j + 1) <= 0.0,
               icefreeNW = HH(i - 1, j + 1) <= 0.0;
    double hhN = SS(i, j + 1);
    if (icefreeN && (BB(i, j + 1) > SS(i, j))) hhN = SS(i, j);
    if (icefreeP && (BB(i, j) > SS(i, j + 1))) hhN = SS(i, j);
    A.hy[2 * q + 1] = (hhN - SS(i, j)) / P.dy;
    double hhE = SS(i + 1, j);
    if (icefreeE && (BB(i + 1, j) > SS(i, j))) hhE = SS(i, j);
    if (icefreeP && (BB(i, j) > SS(i + 1, j))) hhE = SS(i, j);
    double hhW = SS(i - 1, j);
    if (icefreeW && (BB(i - 1, j) > SS(i, j))) hhW = SS(i, j);
    if (icefreeP && (BB(i, j) > SS(i - 1, j))) hhW = SS(i, j);
    double hhNE = SS(i + 1, j + 1);
    if (icefreeNE && (BB(i + 1, j + 1) > SS(i, j + 1))) hhNE = SS(i, j + 1);
    if (icefreeN && (BB(i, j + 1) > SS(i + 1, j + 1))) hhNE = SS(i, j + 1);
    double hhNW = SS(i - 1, j + 1);
    if (icefreeNW && (BB(i - 1, j + 1) > SS(i, j + 1))) hhNW = SS(i, j + 1);
    if (icefreeN && (BB(i, j + 1) > SS(i - 1, j + 1))) hhNW = SS(i, j + 1);
    A.hx[2 * q + 1] = (hhNE + hhE - hhNW - hhW) / (4.0 * P.dx);
  }
#undef HH
#undef BB
#undef SS
}

// SIA_Sliding::basalVelocitySIA (SIA_Sliding.cc:209-267): C in U_b = -C grad h
__device__ __forceinline__ double basal_velocity_sia(const DevParams &P, const SlideArgs &A, double x_in,
                                                     double y_in, double H, double T) {
  const double ice_rho = P.rho_i, beta_CC_grad = P.beta * ice_rho * P.g;
  if (A.mode == 1) {  // verification test E
    const double r1 = 200e3, r2 = 700e3, theta1 = 10 * (M_PI / 180), theta2 = 40 * (M_PI / 180);
    const double x = fabs(x_in), y = fabs(y_in);
    const double r = sqrt(x * x + y * y);
    const double theta = (x < 1.0) ? M_PI / 2.0 : atan(y / x);
    if ((r > r1) && (r < r2) && (theta > theta1) && (theta < theta2)) {
      const double rbot = (r2 - r1) * (r2 - r1), thetabot = (theta2 - theta1) * (theta2 - theta1);
      const double mu_max = 2.5e-11;
      const double muE = mu_max * (4.0 * (r - r1) * (r2 - r) / rbot) *
                         (4.0 * (theta - theta1) * (theta2 - theta) / thetabot);
      return muE * ice_rho * P.g * H;
    }
    return 0.0;
  }
  if (A.mode == 2 || A.mode == 3) {  // EISMINT II experiments G, H
    const double Bfactor = 1e-3 / A.secpera;
    const double pressure = ec_pressure(P, H);
    const double E = ec_enthalpy_permissive(P, T, 0.0, pressure);
    if (A.mode == 2) return Bfactor * ice_rho * P.g * H;
    return ec_is_temperate(P, E, pressure) ? Bfactor * ice_rho * P.g * H : 0.0;
  }
  if (T + beta_CC_grad * H > A.min_T) return A.mu * (ice_rho * P.g * H);
  return 0.0;
}

// SIA_Sliding::update (SIA_Sliding.cc:106-206)
__global__ void k_sia_sliding(DevParams P, SlideArgs A) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (long)P.Mx * P.My) return;
  if (m_ocean(as_int(A.F.mask[q]))) {
    A.ub[q] = 0.0;
    A.vb[q] = 0.0;
    A.heat[q] = 0.0;
    return;
  }
  int i, j;
  map_ij(q, P.My, i, j);
  const long qw = idx2(P, i - 1, j), qs = idx2(P, i, j - 1);
  const double myhx = 0.25 * (A.hx[2 * q] + A.hx[2 * qw] + A.hx[2 * q + 1] + A.hx[2 * qs + 1]),
               myhy = 0.25 * (A.hy[2 * q] + A.hy[2 * qw] + A.hy[2 * q + 1] + A.hy[2 * qs + 1]);
  const double H = A.F.h[q] - A.F.b[q];
  const double T = ec_temperature(P, A.E[q * P.Mz], ec_pressure(P, H));
  const double basalC = basal_velocity_sia(P, A, A.x ? A.x[i] : 0.0, A.y ? A.y[j] : 0.0, H, T);
  const double u = -basalC * myhx, v = -basalC * myhy;
  A.ub[q] = u;
  A.vb[q] = v;
  const double Pov = P.rho_i * P.g * H;
  A.heat[q] = -(Pov * myhx) * u - (Pov * myhy) * v;
}

// ---------------------------------------------------------------- time step
// 2-D CFL over icy cells (iMadaptive.cc:118-144); the min is taken on the bit
// patterns of positive doubles, which is order independent.
__global__ void k_cfl2d(DevParams P, const double *__restrict__ mask, const double *__restrict__ ub,
                        const double *__restrict__ vb, unsigned long long *cfl_bits) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (long)P.Mx * P.My) return;
  if (!m_icy(as_int(mask[q]))) return;
  const double denom = fabs(ub[q]) / P.dx + fabs(vb[q]) / P.dy;
  if (denom > 0.0) {
    const double c = 1.0 / denom;
    atomicMin(cfl_bits, (unsigned long long)__double_as_longlong(c));
  }
}

// dt = min(max, end of run, 2-D CFL, diffusivity) (iMadaptive.cc:157-171, 214-327);
// accumulates t += dt when t is given.
__device__ __forceinline__ void dt_single(const DevParams &P, const unsigned long long *dmax_bits,
                                          const unsigned long long *cfl_bits, double t_end, double *t,
                                          double *dt_out) {
  const double D_max = __longlong_as_double((long long)*dmax_bits);
  double dt = P.max_dt;
  if (t != nullptr && t_end > 0.0) {
    const double tte = t_end - *t;
    if (tte > 0.0 && tte < dt) dt = tte;
  }
  const double cfl = __longlong_as_double((long long)*cfl_bits);
  if (cfl < dt) dt = cfl;
  if (D_max > 0.0) {
    const double grid_factor = 1.0 / (P.dx * P.dx) + 1.0 / (P.dy * P.dy);
    const double dtD = P.adapt_ratio * 2.0 / (D_max * grid_factor);
    if (dtD < dt) dt = dtD;
  }
  *dt_out = dt;
  if (t != nullptr) *t = *t + dt;
}

__global__ void k_dt(DevParams P, const unsigned long long *dmax_bits,
                     const unsigned long long *cfl_bits, double t_end, double *t, double *dt_out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  dt_single(P, dmax_bits, cfl_bits, t_end, t, dt_out);
}

// ---------------------------------------------------------------- mass continuity
// IceModel::massContExplicitStep with cell_interface_fluxes and adjust_flow
// (iMgeometry.cc:172-415, 599-776), no part_grid, no Dirichlet BC.
// which of a face's fluxes adjust_flow zeroes: bit 0 the SIA flux, bit 1 the
// SSA velocity (values, not references: stores through references to the four
// faces' values were merged into one store to a selected address, which kept
// them in scratch memory)
constexpr int ZERO_SIA = 1, ZERO_SSA = 2;
__device__ __forceinline__ int adjust_flow_dir(int c, int nb) {
  if (m_grounded_ice(c) && m_grounded_ice(nb)) return 0;
  if ((m_grounded_ice(c) && m_floating_ice(nb)) || (m_floating_ice(c) && m_grounded_ice(nb))) return 0;
  if ((m_grounded_ice(c) && m_ice_free_land(nb)) || (m_ice_free_land(c) && m_grounded_ice(nb))) return 0;
  if ((m_grounded_ice(c) && m_ice_free_ocean(nb)) || (m_ice_free_ocean(c) && m_grounded_ice(nb))) return 0;
  if (m_floating_ice(c) && m_floating_ice(nb)) return ZERO_SIA;
  if ((m_floating_ice(c) && m_ice_free_land(nb)) || (m_ice_free_land(c) && m_floating_ice(nb))) {
    return ZERO_SIA | ZERO_SSA;
  }
  if ((m_floating_ice(c) && m_ice_free_ocean(nb)) || (m_ice_free_ocean(c) && m_floating_ice(nb))) return ZERO_SIA;
  if (m_ice_free_land(c) && m_ice_free_land(nb)) return ZERO_SIA | ZERO_SSA;
  if ((m_ice_free_land(c) && m_ice_free_ocean(nb)) || (m_ice_free_ocean(c) && m_ice_free_land(nb))) {
    return ZERO_SIA | ZERO_SSA;
  }
  if (m_ice_free_ocean(c) && m_ice_free_ocean(nb)) return ZERO_SIA | ZERO_SSA;
  return 0;
}

// block bid (256 map points) of the mass-continuity grid; red: LDS [8][4]
__device__ __forceinline__ void mass_block(const DevParams &P, const MassArgs &A, long bid, double (*red)[4]) {
  const long q = bid * 256 + threadIdx.x;
  const int Mx = P.Mx, My = P.My;
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (q < (long)Mx * My) {
    int i, j;
    map_ij(q, My, i, j);
    const double dt = *A.dt;
    const double dx = P.dx, dy = P.dy, ice_density = P.rho_i;
    const double cell_area = dx * dy;
    const double meter_to_kg = cell_area * ice_density, meter_per_s_to_kg = meter_to_kg * dt;
    double surface_mass_balance = (A.smb ? A.smb[q] : 0.0) / ice_density;
    double my_basal_melt_rate = 0.0;
    if (P.include_bmr) my_basal_melt_rate = A.bmr ? A.bmr[q] : 0.0;
    // neighbours N, E, S, W
    const long qn = (long)i * My + wrapi(j + 1, My), qe = (long)wrapi(i + 1, Mx) * My + j;
    const long qs = (long)i * My + wrapi(j - 1, My), qw = (long)wrapi(i - 1, Mx) * My + j;
    const int mc = as_int(A.mask[q]);
    // faces N, E, S, W: the SIA flux and the advective velocity after adjust_flow
    const long nb[4] = {qn, qe, qs, qw};
    double Q[4] = {A.flux[2 * q + 1], A.flux[2 * q], A.flux[2 * qs + 1], A.flux[2 * qw]};
    double V[4];
#pragma unroll
    for (int d = 0; d < 4; d++) {
      const int mn = as_int(A.mask[nb[d]]);
      const bool ew = (d == 1 || d == 3);
      const double *vel = ew ? A.ub : A.vb;
      const double vij = vel ? vel[q] : 0.0, vnb = vel ? vel[nb[d]] : 0.0;
      double v;
      if (m_icy(mc) && m_icy(mn)) v = 0.5 * (vij + vnb);
      else if (m_icy(mc) && m_ice_free(mn)) v = vij;
      else if (m_ice_free(mc) && m_icy(mn)) v = vnb;
      else v = 0.0;
      const int z = adjust_flow_dir(mc, mn);
      V[d] = (z & ZERO_SSA) ? 0.0 : v;
      Q[d] = (z & ZERO_SIA) ? 0.0 : Q[d];
    }
    const double Hc = A.H[q];
    const double divQ_SIA = (Q[1] - Q[3]) / dx + (Q[0] - Q[2]) / dy;
    double divQ_SSA = 0.0;
    divQ_SSA += (V[1] * (V[1] > 0 ? Hc : A.H[qe]) - V[3] * (V[3] > 0 ? A.H[qw] : Hc)) / dx;
    divQ_SSA += (V[0] * (V[0] > 0 ? Hc : A.H[qn]) - V[2] * (V[2] > 0 ? A.H[qs] : Hc)) / dy;
    if (m_ice_free_ocean(mc)) {
      surface_mass_balance = 0.0;
      my_basal_melt_rate = 0.0;
    }
    double Hn = Hc;
    Hn += (dt * (surface_mass_balance - my_basal_melt_rate - (divQ_SIA + divQ_SSA)) + 0.0);
    double nonneg_rule_flux = 0.0;
    if (Hn < 0.0) {
      nonneg_rule_flux += -Hn;
      Hn = 0.0;
    }
    A.H_new[q] = Hn;
    if (m_grounded(mc)) acc[0] = -my_basal_melt_rate * meter_per_s_to_kg;
    else acc[2] = -my_basal_melt_rate * meter_per_s_to_kg;
    acc[1] = nonneg_rule_flux * meter_to_kg;
    acc[3] = surface_mass_balance * meter_per_s_to_kg;
    acc[4] = -divQ_SIA * meter_per_s_to_kg;
    acc[5] = -divQ_SSA * meter_per_s_to_kg;
  }
  // fixed-order block reduction of the accounting terms: a DPP scan in each
  // wavefront, then the four wavefront totals in order
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < 8; c++) {
    const double t = wave_inclusive_scan(acc[c]);
    if (lane == 63) red[c][wv] = t;
  }
  __syncthreads();
  if (threadIdx.x < 8 && A.partials) {
    const int c = threadIdx.x;
    A.partials[bid * 8 + c] = ((red[c][0] + red[c][1]) + red[c][2]) + red[c][3];
  }
}

__global__ __launch_bounds__(256) void k_mass_continuity(DevParams P, MassArgs A) {
  __shared__ double red[8][4];
  mass_block(P, A, blockIdx.x, red);
}

// sums the per-block partials in a fixed order (one block)
__global__ void k_reduce_partials(const double *__restrict__ partials, long nblocks, double *out) {
  __shared__ double red[8][256];
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  // unrolled: the loads of four rounds are in flight together (same order)
#pragma unroll 4
  for (long b = threadIdx.x; b < nblocks; b += blockDim.x) {
#pragma unroll
    for (int c = 0; c < 8; c++) acc[c] += partials[b * 8 + c];
  }
#pragma unroll
  for (int c = 0; c < 8; c++) red[c][threadIdx.x] = acc[c];
  __syncthreads();
  for (int off = blockDim.x >> 1; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) {
#pragma unroll
      for (int c = 0; c < 8; c++) red[c][threadIdx.x] += red[c][threadIdx.x + off];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
#pragma unroll
    for (int c = 0; c < 8; c++) out[c] = red[c][0];
  }
}

// GeometryCalculator::compute over the grid (Mask.hh:82-118; iMgeometry.cc:80-149);
// copies H_new into H.
__device__ __forceinline__ void mask_point(const DevParams &P, const double *__restrict__ H_new,
                                           const double *__restrict__ bed, double *__restrict__ H,
                                           double *__restrict__ mask, double *__restrict__ h, int *err, long q) {
  const double thickness = H_new[q];
  if (thickness < 0) atomicOr(err, ERR_NEG_THK);
  const double hgrounded = bed[q] + thickness;
  const double hfloating = P.sea_level + P.alpha_float * thickness;
  const bool is_floating = hfloating > hgrounded + P.floating_thk;
  const bool ice_free = thickness < P.icefree_thk;
  double m, s;
  if (is_floating && !P.is_dry) {
    s = hfloating;
    m = ice_free ? 4.0 : 3.0;
  } else {
    s = hgrounded;
    m = ice_free ? 0.0 : 2.0;
  }
  H[q] = thickness;
  mask[q] = m;
  h[q] = s;
}

__global__ void k_mask_surface(DevParams P, const double *__restrict__ H_new, const double *__restrict__ bed,
                               double *__restrict__ H, double *__restrict__ mask, double *__restrict__ h,
                               int *err) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (long)P.Mx * P.My) return;
  mask_point(P, H_new, bed, H, mask, h, err, q);
}

// ---------------------------------------------------------------- launchers
template <int LAW>
static void launch_column_law(const DevParams &P, const ColumnArgs &A, bool full, hipStream_t st, int extra,
                              bool allow_strips) {
  const bool strips = allow_strips && column_strips_ok(P);
  const int nb = strips ? column_strip_blocks(P) : column_blocks(P);
  const size_t lds = column_lds(P.Mz);
  const bool wide = column_cpw(P) == COLS_PER_WAVE_WIDE;
  if (full && wide) {
    hipLaunchKernelGGL((k_column<LAW, true, COLS_PER_WAVE_WIDE>), dim3(nb), dim3(256), lds, st, P, A, (int)strips);
  } else if (full) {
    hipLaunchKernelGGL((k_column<LAW, true, COLS_PER_WAVE_NARROW>), dim3(nb), dim3(256), lds, st, P, A,
                       (int)strips);
  } else if (wide) {
    hipLaunchKernelGGL((k_column<LAW, false, COLS_PER_WAVE_WIDE>), dim3(nb), dim3(256), lds, st, P, A,
                       (int)strips);
  } else {
    hipLaunchKernelGGL((k_column<LAW, false, COLS_PER_WAVE_NARROW>), dim3(nb), dim3(256), lds, st, P, A,
                       (int)strips);
  }
  hipLaunchKernelGGL(k_reduce_max, dim3(1), dim3(1024), 0, st, A.dmax_part, nb + extra, A.dmax_bits);
}

template <int LAW>
static void launch_column_list_law(const DevParams &P, const ColumnArgs &A, const int *list, const int *n_list,
                                   hipStream_t st, int extra) {
  const int nb = column_list_blocks(P);
  hipLaunchKernelGGL((k_column_list<LAW>), dim3(nb), dim3(256), column_lds(P.Mz), st, P, A, list,
                     n_list);
  hipLaunchKernelGGL(k_reduce_max, dim3(1), dim3(1024), 0, st, A.dmax_part, nb + extra, A.dmax_bits);
}

void launch_column_list(const DevParams &P, const ColumnArgs &A, const int *list, const int *n_list,
                        hipStream_t st, int extra) {
  switch (P.law) {
  case LAW_ISOGLEN: launch_column_list_law<LAW_ISOGLEN>(P, A, list, n_list, st, extra); break;
  case LAW_PB: launch_column_list_law<LAW_PB>(P, A, list, n_list, st, extra); break;
  case LAW_GPBLD: launch_column_list_law<LAW_GPBLD>(P, A, list, n_list, st, extra); break;
  case LAW_HOOKE: launch_column_list_law<LAW_HOOKE>(P, A, list, n_list, st, extra); break;
  case LAW_ARR: launch_column_list_law<LAW_ARR>(P, A, list, n_list, st, extra); break;
  case LAW_ARRWARM: launch_column_list_law<LAW_ARRWARM>(P, A, list, n_list, st, extra); break;
  default: launch_column_list_law<LAW_GK>(P, A, list, n_list, st, extra); break;
  }
}

void launch_column(const DevParams &P, const ColumnArgs &A, bool full, hipStream_t st, int extra, bool allow_strips) {
  switch (P.law) {
  case LAW_ISOGLEN: launch_column_law<LAW_ISOGLEN>(P, A, full, st, extra, allow_strips); break;
  case LAW_PB: launch_column_law<LAW_PB>(P, A, full, st, extra, allow_strips); break;
  case LAW_GPBLD: launch_column_law<LAW_GPBLD>(P, A, full, st, extra, allow_strips); break;
  case LAW_HOOKE: launch_column_law<LAW_HOOKE>(P, A, full, st, extra, allow_strips); break;
  case LAW_ARR: launch_column_law<LAW_ARR>(P, A, full, st, extra, allow_strips); break;
  case LAW_ARRWARM: launch_column_law<LAW_ARRWARM>(P, A, full, st, extra, allow_strips); break;
  default: launch_column_law<LAW_GK>(P, A, full, st, extra, allow_strips); break;
  }
}

static inline unsigned blocks_for(long n, int per_block) { return (unsigned)((n + per_block - 1) / per_block); }

int column_blocks(const DevParams &P) {
  return (int)blocks_for(2L * P.Mx * P.My, WAVES_PER_BLOCK * column_cpw(P));
}

// strips of cpw / 2 map columns (one row of a strip a wavefront of cpw
// columns), ceil(strips / 8) of them on each of the 8 XCDs; blocks of a strip
// slot past the last strip exit at once
static int strip_blocks(const DevParams &P, int cpw) {
  const int w = cpw / 2;
  const long per_xcd_strips = ((P.My + w - 1) / w + N_XCD - 1) / N_XCD;
  return (int)(N_XCD * blocks_for(per_xcd_strips * P.Mx, WAVES_PER_BLOCK));
}
int column_strip_blocks(const DevParams &P) { return strip_blocks(P, column_cpw(P)); }


int column_list_blocks(const DevParams &P) { return column_blocks(P) < 1024 ? column_blocks(P) : 1024; }

void launch_reset(unsigned long long *dmax_bits, int *err, int *err_acc, unsigned long long *cfl_bits,
                  double max_dt, hipStream_t st) {
  hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, st, dmax_bits, err, err_acc, cfl_bits, max_dt);
}

void launch_fill(double *a, long n, double v, hipStream_t st) {
  hipLaunchKernelGGL(k_fill, dim3(blocks_for(n, 256)), dim3(256), 0, st, a, n, v);
}

void launch_bed_smoother(const DevParams &P, double s2, double s3, double s4, const double *b0, double *bs,
                         double *mt, double *c2, double *c3, double *c4, hipStream_t st) {
  const long n = (long)P.Mx * P.My;
  hipLaunchKernelGGL(k_smooth_mean, dim3(blocks_for(n, 256)), dim3(256), 0, st, P, b0, bs);
  hipLaunchKernelGGL(k_smooth_coeffs, dim3(blocks_for(n, 256)), dim3(256), 0, st, P, s2, s3, s4, b0, bs,
                     mt, c2, c3, c4);
}

void launch_gradient(const DevParams &P, const double *H, const double *b, const double *h,
                     const double *mask, double *hx, double *hy, const ThetaArgs &T, hipStream_t st) {
  Fields2D F{H, b, h, mask};
  const long tiles = (long)((P.Mx + GT_I - 1) / GT_I) * ((P.My + GT_J - 1) / GT_J);
  hipLaunchKernelGGL(k_gradient, dim3((unsigned)tiles), dim3(GT_I * GT_J), 0, st, P, F, hx, hy, T);
}

void launch_regional_gradient(const DevParams &P, const double *nmm, const double *hst, double *hx, double *hy,
                              hipStream_t st) {
  const long n = (long)P.Mx * P.My;
  hipLaunchKernelGGL(k_regional_gradient, dim3(blocks_for(n, 256)), dim3(256), 0, st, P, nmm, hst, hx, hy);
}


template <int LAW>
static void launch_vel_law(const DevParams &P, const VelArgs &A, hipStream_t st) {
  const TileMap tm = tile_map(P.Mx, P.My);
  const size_t lds = sizeof(double) * (size_t)P.Mz;
  if (LAW == LAW_GPBLD && P.ec == EC_DEFAULT) {
    hipLaunchKernelGGL((k_velocity_tile<LAW, true>), dim3(tm.per_xcd * N_XCD), dim3(64 * vtile::VW), lds, st, P, A,
                       tm);
  } else {
    hipLaunchKernelGGL((k_velocity_tile<LAW, false>), dim3(tm.per_xcd * N_XCD), dim3(64 * vtile::VW), lds, st, P, A,
                       tm);
  }
}

void launch_velocity_uvw(const DevParams &P, const VelArgs &A, hipStream_t st) {
  if (!A.Sigma) {
    launch_vel_law<-1>(P, A, st);
    return;
  }
  switch (P.law) {
  case LAW_ISOGLEN: launch_vel_law<LAW_ISOGLEN>(P, A, st); break;
  case LAW_PB: launch_vel_law<LAW_PB>(P, A, st); break;
  case LAW_GPBLD: launch_vel_law<LAW_GPBLD>(P, A, st); break;
  case LAW_HOOKE: launch_vel_law<LAW_HOOKE>(P, A, st); break;
  case LAW_ARR: launch_vel_law<LAW_ARR>(P, A, st); break;
  case LAW_ARRWARM: launch_vel_law<LAW_ARRWARM>(P, A, st); break;
  default: launch_vel_law<LAW_GK>(P, A, st); break;
  }
}

void launch_surface_values(const DevParams &P, const double *z, const double *f, const double *H, double *out,
                           hipStream_t st) {
  const long n = (long)P.Mx * P.My;
  hipLaunchKernelGGL(k_surface_values, dim3(blocks_for(n, 256)), dim3(256), 0, st, P, z, f, H, out);
}

void launch_diffusivity_regular(const DevParams &P, const double *D, double *out, hipStream_t st) {
  const long n = (long)P.Mx * P.My;
  hipLaunchKernelGGL(k_diffusivity_regular, dim3(blocks_for(n, 256)), dim3(256), 0, st, P, D, out);
}

void launch_sia_sliding(const DevParams &P, const SlideArgs &A, hipStream_t st) {
  const long n = (long)P.Mx * P.My;
  hipLaunchKernelGGL(k_sliding_gradient, dim3(blocks_for(n, 256)), dim3(256), 0, st, P, A);
  hipLaunchKernelGGL(k_sia_sliding, dim3(blocks_for(n, 256)), dim3(256), 0, st, P, A);
}

void launch_cfl2d(const DevParams &P, const double *mask, const double *ub, const double *vb,
                  unsigned long long *cfl_bits, hipStream_t st) {
  const long n = (long)P.Mx * P.My;
  hipLaunchKernelGGL(k_cfl2d, dim3(blocks_for(n, 256)), dim3(256), 0, st, P, mask, ub, vb, cfl_bits);
}

void launch_dt(const DevParams &P, const unsigned long long *dmax_bits, const unsigned long long *cfl_bits,
               double t_end, double *t, double *dt_out, hipStream_t st) {
  hipLaunchKernelGGL(k_dt, dim3(1), dim3(64), 0, st, P, dmax_bits, cfl_bits, t_end, t, dt_out);
}

long mass_partial_blocks(const DevParams &P) { return blocks_for((long)P.Mx * P.My, 256); }

void launch_mass_continuity(const DevParams &P, const MassArgs &A, double *acc_out, hipStream_t st) {
  const long nb = mass_partial_blocks(P);
  hipLaunchKernelGGL(k_mass_continuity, dim3((unsigned)nb), dim3(256), 0, st, P, A);
  if (acc_out && A.partials) {
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(256), 0, st, A.partials, nb, acc_out);
  }
}

void launch_mask_surface(const DevParams &P, const double *H_new, const double *bed, double *H,
                         double *mask, double *h, int *err, hipStream_t st) {
  const long n = (long)P.Mx * P.My;
  hipLaunchKernelGGL(k_mask_surface, dim3(blocks_for(n, 256)), dim3(256), 0, st, P, H_new, bed, H, mask, h,
                     err);
}

// ---------------------------------------------------------------- ensembles
// One launch per stage for all members of an ensemble: block b of a stage grid
// belongs to the member m with start[m] <= b < start[m + 1] and runs that
// member's block b - start[m] of the single-member kernel, with the member's
// parameters and fields read from the device table (a wave-uniform address:
// scalar loads). The arithmetic is the single-member kernels', so every member's
// results are bitwise those of stepping it alone.
// (ens_member: sia_kernels.hpp)

// per member: D_max accumulators, error flags, 2-D CFL bound
__global__ void k_ens_reset(const EnsMember *__restrict__ T, int nm) {
  const int m = blockIdx.x;
  if (m >= nm || threadIdx.x != 0) return;
  const EnsMember &E = T[m];
  *E.A.dmax_bits = 0ull;
  *E.err_acc |= *E.A.err;
  *E.A.err = 0;
  *E.cfl_bits = (unsigned long long)__double_as_longlong(E.P.max_dt);
}

// surface gradient, theta and the smoothed thickness (256 map points a block)
__global__ __launch_bounds__(256) void k_ens_gradient(const EnsMember *__restrict__ T, const int *__restrict__ start,
                                                      int nm) {
  const int m = ens_member(start, nm, blockIdx.x);
  const EnsMember &E = T[m];
  const long q = (long)(blockIdx.x - start[m]) * 256 + threadIdx.x;
  if (q >= (long)E.P.Mx * E.P.My) return;
  const Fields2D F = global_view(E.F);
  gradient_point(E.P, F, glb(E.hx), glb(E.hy), q);
  theta_thk_point(E.P, F, glb(E.topgs), glb(E.maxtl), glb(E.C2), glb(E.C3), glb(E.C4), glb(E.theta),
                  glb(E.thks), glb(E.A.err), q);
}

template <int LAW, bool FULL, int CPW>
__global__ __launch_bounds__(256, COLUMN_WAVES_ENS) void k_ens_column(const EnsMember *__restrict__ T,
                                                                       const int *__restrict__ start, int nm) {
  extern __shared__ __align__(16) double zs[];  // column_lds(Mz) bytes
  __shared__ double wmax[WAVES_PER_BLOCK];
  const int m = ens_member(start, nm, blockIdx.x);
  const ColumnArgs A = global_view(T[m].A);
  const bool strips = CPW == ENS_STRIP_CPW && column_strips_ok(T[m].P);  // as ens_column_blocks
  const double mb = column_block<LAW, FULL, CPW>(T[m].P, A, blockIdx.x - start[m], zs, wmax, nullptr, 0, strips);
  if (threadIdx.x == 0) A.dmax_part[blockIdx.x - start[m]] = mb;
}

// per member (one block each): D_max from the column partials, then dt and t += dt
__global__ __launch_bounds__(1024) void k_ens_dt(const EnsMember *__restrict__ T, int nm) {
  __shared__ double red[1024];
  const int mb = blockIdx.x;
  if (mb >= nm) return;
  const EnsMember &E = T[mb];
  double m = 0.0;
  const double *part = glb(E.A.dmax_part);
  for (int b = threadIdx.x; b < E.n_part; b += blockDim.x) m = (m < part[b]) ? part[b] : m;
  red[threadIdx.x] = m;
  __syncthreads();
  for (int off = blockDim.x >> 1; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) red[threadIdx.x] = (red[threadIdx.x] < red[threadIdx.x + off]) ? red[threadIdx.x + off] : red[threadIdx.x];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    *E.A.dmax_bits = (unsigned long long)__double_as_longlong(red[0]);
    dt_single(E.P, E.A.dmax_bits, E.cfl_bits, E.t_end, E.t, E.dt);
  }
}

__global__ __launch_bounds__(256) void k_ens_mass(const EnsMember *__restrict__ T, const int *__restrict__ start,
                                                  int nm) {
  __shared__ double red[8][4];
  const int m = ens_member(start, nm, blockIdx.x);
  const MassArgs M = global_view(T[m].M);
  mass_block(T[m].P, M, blockIdx.x - start[m], red);
}

__global__ __launch_bounds__(256) void k_ens_mask(const EnsMember *__restrict__ T, const int *__restrict__ start,
                                                  int nm) {
  const int m = ens_member(start, nm, blockIdx.x);
  const EnsMember &E = T[m];
  const long q = (long)(blockIdx.x - start[m]) * 256 + threadIdx.x;
  if (q >= (long)E.P.Mx * E.P.My) return;
  mask_point(E.P, glb(E.M.H_new), glb(E.F.b), glb(E.H), glb(E.mask), glb(E.h), glb(E.A.err), q);
}

template <int LAW>
__global__ __launch_bounds__(256) void k_ens_column_list(const EnsListMember *__restrict__ T,
                                                         const int *__restrict__ start, int nm) {
  extern __shared__ __align__(16) double zs[];  // column_lds(Mz) bytes
  __shared__ double wmax[WAVES_PER_BLOCK];
  const int m = ens_member(start, nm, blockIdx.x);
  const ColumnArgs A = global_view(T[m].A);
  const int *list = glb(T[m].list);
  const int n = *glb(T[m].n_list);
  const int bid = blockIdx.x - start[m], nblk = start[m + 1] - start[m];
  double mx = 0.0;
  for (int vb = bid; vb * LIST_COLS_PER_BLOCK < n; vb += nblk) {
    const double mb = column_block<LAW, false, LIST_COLS_PER_WAVE>(T[m].P, A, vb, zs, wmax, list, n);
    mx = (mx < mb) ? mb : mx;
  }
  if (threadIdx.x == 0) A.dmax_part[bid] = mx;
}

void launch_ens_column_list(const EnsListMember *T, const int *start, int nm, int nblocks, int max_mz, int law,
                            hipStream_t st) {
  const size_t lds = column_lds(max_mz);
  switch (law) {
  case LAW_ISOGLEN: hipLaunchKernelGGL(k_ens_column_list<LAW_ISOGLEN>, dim3(nblocks), dim3(256), lds, st, T, start, nm); break;
  case LAW_PB: hipLaunchKernelGGL(k_ens_column_list<LAW_PB>, dim3(nblocks), dim3(256), lds, st, T, start, nm); break;
  case LAW_GPBLD: hipLaunchKernelGGL(k_ens_column_list<LAW_GPBLD>, dim3(nblocks), dim3(256), lds, st, T, start, nm); break;
  case LAW_HOOKE: hipLaunchKernelGGL(k_ens_column_list<LAW_HOOKE>, dim3(nblocks), dim3(256), lds, st, T, start, nm); break;
  case LAW_ARR: hipLaunchKernelGGL(k_ens_column_list<LAW_ARR>, dim3(nblocks), dim3(256), lds, st, T, start, nm); break;
  case LAW_ARRWARM: hipLaunchKernelGGL(k_ens_column_list<LAW_ARRWARM>, dim3(nblocks), dim3(256), lds, st, T, start, nm); break;
  default: hipLaunchKernelGGL(k_ens_column_list<LAW_GK>, dim3(nblocks), dim3(256), lds, st, T, start, nm); break;
  }
}

int ens_point_blocks(const DevParams &P) { return (int)(((long)P.Mx * P.My + 255) / 256); }

// a member's column blocks: in strips (16 map columns, as the plain pass at
// 32 columns a wavefront) at ENS_STRIP_CPW columns a wavefront where
// column_strips_ok holds, else in staggered order; a multiple of the XCD
// count, so that every member's block range starts on XCD 0
int ens_column_blocks(const DevParams &P, int cpw) {
  const long nb = (cpw == ENS_STRIP_CPW && column_strips_ok(P))
                      ? strip_blocks(P, ENS_STRIP_CPW)
                      : (2L * P.Mx * P.My + WAVES_PER_BLOCK * cpw - 1) / (WAVES_PER_BLOCK * cpw);
  return (int)((nb + N_XCD - 1) / N_XCD * N_XCD);
}

template <int LAW, int CPW>
static void ens_column_law(const EnsMember *T, const int *start, int nm, int nblocks, int max_mz, hipStream_t st) {
  hipLaunchKernelGGL((k_ens_column<LAW, false, CPW>), dim3(nblocks), dim3(256), column_lds(max_mz),
                     st, T, start, nm);
}

template <int CPW>
static void ens_column_cpw(int law, const EnsMember *T, const int *start, int nm, int nblocks, int max_mz,
                           hipStream_t st) {
  switch (law) {
  case LAW_ISOGLEN: ens_column_law<LAW_ISOGLEN, CPW>(T, start, nm, nblocks, max_mz, st); break;
  case LAW_PB: ens_column_law<LAW_PB, CPW>(T, start, nm, nblocks, max_mz, st); break;
  case LAW_GPBLD: ens_column_law<LAW_GPBLD, CPW>(T, start, nm, nblocks, max_mz, st); break;
  case LAW_HOOKE: ens_column_law<LAW_HOOKE, CPW>(T, start, nm, nblocks, max_mz, st); break;
  case LAW_ARR: ens_column_law<LAW_ARR, CPW>(T, start, nm, nblocks, max_mz, st); break;
  case LAW_ARRWARM: ens_column_law<LAW_ARRWARM, CPW>(T, start, nm, nblocks, max_mz, st); break;
  default: ens_column_law<LAW_GK, CPW>(T, start, nm, nblocks, max_mz, st); break;
  }
}

void launch_ens_front(const EnsLaunch &L, hipStream_t st) {
  hipLaunchKernelGGL(k_ens_reset, dim3(L.nm), dim3(64), 0, st, L.table, L.nm);
  hipLaunchKernelGGL(k_ens_gradient, dim3(L.pt_blocks), dim3(256), 0, st, L.table, L.pt_start, L.nm);
}

void launch_ens_back(const EnsLaunch &L, hipStream_t st) {
  // (no column blocks when every member runs through a surrogate)
  switch (L.col_blocks > 0 ? L.cpw : 0) {
  case 0: break;
  case 8: ens_column_cpw<8>(L.law, L.table, L.col_start, L.nm, L.col_blocks, L.max_mz, st); break;
  case 32: ens_column_cpw<32>(L.law, L.table, L.col_start, L.nm, L.col_blocks, L.max_mz, st); break;
  default: ens_column_cpw<16>(L.law, L.table, L.col_start, L.nm, L.col_blocks, L.max_mz, st); break;
  }
  hipLaunchKernelGGL(k_ens_dt, dim3(L.nm), dim3(1024), 0, st, L.table, L.nm);
  hipLaunchKernelGGL(k_ens_mass, dim3(L.pt_blocks), dim3(256), 0, st, L.table, L.pt_start, L.nm);
  hipLaunchKernelGGL(k_ens_mask, dim3(L.pt_blocks), dim3(256), 0, st, L.table, L.pt_start, L.nm);
}

void launch_ens_step(const EnsLaunch &L, hipStream_t st) {
  launch_ens_front(L, st);
  launch_ens_back(L, st);
}

}  // namespace pism_sia
