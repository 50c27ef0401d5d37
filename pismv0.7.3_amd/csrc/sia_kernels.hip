// sia_kernels.hip -- HIP kernels of the SIA hot path for gfx950 (MI355X).
//
// Field layout is PISM's (IceGrid.cc:183-195): 2-D a[i*My+j], staggered
// a[(i*My+j)*2+o], 3-D a[(i*My+j)*Mz+k]; ghosts are periodic index wraps.
//
// Kernels:
//   k_smooth_mean / k_smooth_coeffs  BedSmoother (PISMBedSmoother.cc:162-269)
//   k_gradient                       surface gradient, all 3 methods (SIAFD.cc:210-503)
//   k_theta_thk                      get_theta + get_smoothed_thk (PISMBedSmoother.cc:286-420)
//   k_column<LAW, FULL>              diffusive flux, D, D_max, and (FULL) the column
//                                    integral I (SIAFD.cc:546-722, 819-879); one
//                                    64-lane wavefront per staggered column, one lane
//                                    per level, wave-tree reductions and scans
//   k_velocity_uvw                   u, v (SIAFD.cc:918-958) and w
//                                    (PISMStressBalance.cc:209-286) in one pass
//   k_mass_continuity, k_mask_surface, k_dt, accounting reductions
//                                    iMgeometry.cc:56-149, 497-822; iMadaptive.cc:118-171
#include "sia_kernels.hpp"

namespace pism_sia {

// ---------------------------------------------------------------- wave helpers
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ __forceinline__ double wave_inclusive_scan(double v, int lane) {
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const double t = __shfl_up(v, off, 64);
    if (lane >= off) v += t;
  }
  return v;
}

// ---------------------------------------------------------------- resets
__global__ void k_reset(unsigned long long *dmax_bits, int *err, unsigned long long *cfl_bits,
                        double max_dt) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    *dmax_bits = 0ull;  // +0.0
    *err = 0;
    *cfl_bits = (unsigned long long)__double_as_longlong(max_dt);
  }
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
  const int i = (int)(q / My), j = (int)(q % My);
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
  const int i = (int)(q / My), j = (int)(q % My);
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
struct Fields2D {
  const double *H, *b, *h, *mask;
};

__device__ __forceinline__ long idx2(const DevParams &P, int i, int j) {
  return (long)wrapi(i, P.Mx) * P.My + wrapi(j, P.My);
}

// haseloff first pass (SIAFD.cc:402-438): x-derivative on the i-offset grid and its weight
__device__ __forceinline__ void haseloff_x(const DevParams &P, const Fields2D &F, int i, int j,
                                           double &hx0, double &wi) {
  const long a = idx2(P, i, j), c = idx2(P, i + 1, j);
  const int m0 = as_int(F.mask[a]), m1 = as_int(F.mask[c]);
  const double h0 = F.h[a], h1 = F.h[c];
  if ((m_floating_ice(m0) && m_ice_free_ocean(m1)) || (m_ice_free_ocean(m0) && m_floating_ice(m1))) {
    hx0 = 0.0; wi = 0.0;
  } else if ((m_icy(m0) && m_ice_free(m1) && F.b[c] > h0) || (m_ice_free(m0) && m_icy(m1) && F.b[a] > h1)) {
    hx0 = 0.0; wi = 0.0;
  } else {
    hx0 = (h1 - h0) / P.dx; wi = 1.0;
  }
}
__device__ __forceinline__ void haseloff_y(const DevParams &P, const Fields2D &F, int i, int j,
                                           double &hy1, double &wj) {
  const long a = idx2(P, i, j), c = idx2(P, i, j + 1);
  const int m0 = as_int(F.mask[a]), m1 = as_int(F.mask[c]);
  const double h0 = F.h[a], h1 = F.h[c];
  if ((m_floating_ice(m0) && m_ice_free_ocean(m1)) || (m_ice_free_ocean(m0) && m_floating_ice(m1))) {
    hy1 = 0.0; wj = 0.0;
  } else if ((m_icy(m0) && m_ice_free(m1) && F.b[c] > h0) || (m_ice_free(m0) && m_icy(m1) && F.b[a] > h1)) {
    hy1 = 0.0; wj = 0.0;
  } else {
    hy1 = (h1 - h0) / P.dy; wj = 1.0;
  }
}

__global__ void k_gradient(DevParams P, Fields2D F, double *__restrict__ hx, double *__restrict__ hy) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (long)P.Mx * P.My) return;
  const int i = (int)(q / P.My), j = (int)(q % P.My);
  const double dx = P.dx, dy = P.dy;
  double hx0, hy0, hx1, hy1;
  if (P.grad == 1) {  // mahaffy, SIAFD.cc:305-316
#define HH(a, c) F.h[idx2(P, (a), (c))]
    hx0 = (HH(i + 1, j) - HH(i, j)) / dx;
    hy0 = (+HH(i + 1, j + 1) + HH(i, j + 1) - HH(i + 1, j - 1) - HH(i, j - 1)) / (4.0 * dy);
    hy1 = (HH(i, j + 1) - HH(i, j)) / dy;
    hx1 = (+HH(i + 1, j + 1) + HH(i + 1, j) - HH(i - 1, j + 1) - HH(i - 1, j)) / (4.0 * dx);
#undef HH
  } else if (P.grad == 2) {  // eta, SIAFD.cc:210-284
    const double n = P.n, etapow = (2.0 * n + 2.0) / n, invpow = 1.0 / etapow,
                 dinvpow = (-n - 2.0) / (2.0 * n + 2.0);
#define ETA(a, c) pow(F.H[idx2(P, (a), (c))], etapow)
#define BB(a, c) F.b[idx2(P, (a), (c))]
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
    const bool icy = m_icy(as_int(F.mask[q]));
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

// ---------------------------------------------------------------- theta, thksmooth
// get_theta (PISMBedSmoother.cc:360-420) and get_smoothed_thk (286-336)
__global__ void k_theta_thk(DevParams P, Fields2D F, const double *__restrict__ topgs,
                            const double *__restrict__ maxtl, const double *__restrict__ C2,
                            const double *__restrict__ C3, const double *__restrict__ C4,
                            double *__restrict__ theta, double *__restrict__ thks, int *err) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (long)P.Mx * P.My) return;
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
// (SIAFD.cc:845-878). One wavefront per staggered column s = 2*(i*My+j)+o.
template <int LAW, bool FULL>
__global__ __launch_bounds__(256) void k_column(DevParams P, ColumnArgs A) {
  const int lane = threadIdx.x & 63;
  const long s = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int Mx = P.Mx, My = P.My, Mz = P.Mz;
  if (s >= 2L * Mx * My) return;
  const int o = (int)(s & 1);
  const long q = s >> 1;
  const int i = (int)(q / My), j = (int)(q % My);
  const long q2 = (o == 0) ? (long)wrapi(i + 1, Mx) * My + j : (long)i * My + wrapi(j + 1, My);
  const double *__restrict__ z = A.z;

  const double thk = 0.5 * (A.thks[q] + A.thks[q2]);
  double *Icol = FULL ? ((o == 0) ? A.I0 : A.I1) + q * Mz : nullptr;

  // zero thickness (SIAFD.cc:635-641): delta = 0, so I = 0 at every level.
  // Above ks, I is constant (SIAFD.cc:870-872); only I[0..ks] is stored and
  // readers use I[min(k, ks)].
  if (thk == 0.0) {
    if (lane == 0) {
      A.flux[s] = 0.0;
      A.Dstag[s] = 0.0;
      if (FULL) {
        Icol[0] = 0.0;
        A.ks[s] = 0;
      }
    }
    return;
  }
  const int ks = kbelow(z, Mz, P.Lz, thk);
  if (ks < 0) {
    if (lane == 0) {
      atomicOr(A.err, ERR_THK_LZ);
      A.flux[s] = 0.0;
      A.Dstag[s] = 0.0;
      if (FULL) {
        Icol[0] = 0.0;
        A.ks[s] = 0;
      }
    }
    return;
  }
  const double hxs = A.hx[s], hys = A.hy[s];
  const double slope = (o == 0) ? hxs : hys;
  const double alpha = sqrt(hxs * hxs + hys * hys);
  const double theta_local = 0.5 * (A.theta[q] + A.theta[q2]);
  const double *__restrict__ Ea = A.E + q * Mz;
  const double *__restrict__ Eb = A.E + q2 * Mz;
  const double *__restrict__ aa = P.use_age ? A.age + q * Mz : nullptr;
  const double *__restrict__ ab = P.use_age ? A.age + q2 * Mz : nullptr;
  const double c0 = P.e * theta_local * 2.0;

  double Dpart = 0.0, dprev_carry = 0.0, Icarry = 0.0;
  for (int k0 = 0; k0 <= ks; k0 += 64) {
    const int k = k0 + lane;
    const bool act = k <= ks;
    double delta = 0.0, depth = 0.0, dz = 0.0;
    if (act) {
      depth = thk - z[k];
      const double pressure = ec_pressure(P, depth);
      const double gs = P.use_age ? grain_size_vostok(P, 0.5 * (aa[k] + ab[k])) : P.gs;
      const double Ek = 0.5 * (Ea[k] + Eb[k]);
      const double fl = flow<LAW>(P, alpha * pressure, Ek, pressure, gs);
      delta = c0 * pressure * fl;
      if (k > 0) dz = z[k] - z[k - 1];
    }
    double dprev = __shfl_up(delta, 1, 64);
    if (lane == 0) dprev = dprev_carry;
    if (act && k > 0) {
      Dpart += 0.5 * dz * ((depth + dz) * dprev + depth * delta);
    }
    if (k == ks) {
      const double dzt = thk - z[ks];
      Dpart += 0.5 * dzt * dzt * delta;
    }
    if (FULL) {
      const double t = (act && k > 0) ? 0.5 * dz * (dprev + delta) : 0.0;
      const double Ik = Icarry + wave_inclusive_scan(t, lane);
      if (act) Icol[k] = Ik;
      Icarry = __shfl(Ik, 63, 64);
    }
    dprev_carry = __shfl(delta, 63, 64);
  }
  if (FULL && lane == 0) A.ks[s] = ks;
  double D = wave_sum(Dpart);
  if (lane == 0) {
    A.Dstag[s] = D;
    if (i >= Mx - 1 || j >= My - 1) D = 0.0;  // SIAFD.cc:694-697
    atomicMax(A.dmax_bits, (unsigned long long)__double_as_longlong(D));
    A.flux[s] = -D * slope;
  }
}

// ---------------------------------------------------------------- u, v, w
// compute_3d_horizontal_velocity (SIAFD.cc:918-958) fused with
// StressBalance::compute_vertical_velocity (PISMStressBalance.cc:209-286).
// One wavefront per column (i,j), one lane per level. The wave evaluates u at
// (i-1,j), (i,j), (i+1,j) and v at (i,j-1), (i,j), (i,j+1) with the
// reference's formula from the I columns of the 16 staggered neighbours (so
// each neighbour value is bitwise the one that neighbour writes), then w by a
// trapezoid prefix scan over the levels. u, v, w are written once; u, v are
// never read back from HBM.
struct UV4 {
  long e, w, n, s;   // staggered column indices (2 q + o) of I_e, I_w, I_n, I_s
  double hx_e, hx_w, hx_n, hx_s, hy_e, hy_w, hy_n, hy_s;
  double ub, vb;
};

__device__ __forceinline__ UV4 uv_setup(const DevParams &P, const double *__restrict__ hx,
                                        const double *__restrict__ hy, const double *__restrict__ ub,
                                        const double *__restrict__ vb, int i, int j) {
  const int Mx = P.Mx, My = P.My;
  const long q = (long)i * My + j;
  const long qw = (long)wrapi(i - 1, Mx) * My + j, qs = (long)i * My + wrapi(j - 1, My);
  UV4 c;
  c.e = 2 * q; c.w = 2 * qw; c.n = 2 * q + 1; c.s = 2 * qs + 1;
  c.hx_w = hx[c.w]; c.hx_e = hx[c.e]; c.hx_n = hx[c.n]; c.hx_s = hx[c.s];
  c.hy_w = hy[c.w]; c.hy_e = hy[c.e]; c.hy_n = hy[c.n]; c.hy_s = hy[c.s];
  c.ub = ub ? ub[q] : 0.0;
  c.vb = vb ? vb[q] : 0.0;
  return c;
}

__device__ __forceinline__ double Iat(const double *__restrict__ I, const int *__restrict__ ks, long sidx,
                                      int Mz, int k) {
  const int kk = min(k, ks[sidx]);
  return I[(sidx >> 1) * (long)Mz + kk];
}

__global__ __launch_bounds__(256) void k_velocity_uvw(DevParams P, VelArgs A) {
  const int lane = threadIdx.x & 63;
  const long q = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int Mx = P.Mx, My = P.My, Mz = P.Mz;
  if (q >= (long)Mx * My) return;
  const int i = (int)(q / My), j = (int)(q % My);
  const int im = wrapi(i - 1, Mx), ip = wrapi(i + 1, Mx), jm = wrapi(j - 1, My), jp = wrapi(j + 1, My);
  const UV4 cc = uv_setup(P, A.hx, A.hy, A.ub, A.vb, i, j);
  const UV4 cw = uv_setup(P, A.hx, A.hy, A.ub, A.vb, im, j);
  const UV4 ce = uv_setup(P, A.hx, A.hy, A.ub, A.vb, ip, j);
  const UV4 cs = uv_setup(P, A.hx, A.hy, A.ub, A.vb, i, jm);
  const UV4 cn = uv_setup(P, A.hx, A.hy, A.ub, A.vb, i, jp);
  // margin-aware differences (PISMStressBalance.cc:223-260)
  const double *mask = A.mask;
  const int m = as_int(mask[q]);
  const int me = as_int(mask[(long)ip * My + j]), mw = as_int(mask[(long)im * My + j]);
  const int mn = as_int(mask[(long)i * My + jp]), ms = as_int(mask[(long)i * My + jm]);
  double west = 1, east = 1, south = 1, north = 1;
  if ((m_icy(m) && m_ice_free(me)) || (m_ice_free(m) && m_icy(me))) east = 0;
  if ((m_icy(m) && m_ice_free(mw)) || (m_ice_free(m) && m_icy(mw))) west = 0;
  const double D_x = (east + west > 0) ? 1.0 / (P.dx * (east + west)) : 0.0;
  if ((m_icy(m) && m_ice_free(mn)) || (m_ice_free(m) && m_icy(mn))) north = 0;
  if ((m_icy(m) && m_ice_free(ms)) || (m_ice_free(m) && m_icy(ms))) south = 0;
  const double D_y = (north + south > 0) ? 1.0 / (P.dy * (north + south)) : 0.0;
  const double w0 = A.bmr ? -A.bmr[q] : 0.0;
  const double *I0 = A.I0, *I1 = A.I1;
  const int *ks = A.ks;
  const double *__restrict__ z = A.z;
  double carry_int = 0.0, carry_sum = 0.0;
  for (int k0 = 0; k0 < Mz; k0 += 64) {
    const int k = k0 + lane;
    const bool act = k < Mz;
    const int kc = act ? k : Mz - 1;
    double uc, vc, integrand = 0.0, dz = 0.0;
    {
      // u, v at (i,j): SIAFD.cc:949-956
      const double ie = Iat(I0, ks, cc.e, Mz, kc), iw = Iat(I0, ks, cc.w, Mz, kc);
      const double in = Iat(I1, ks, cc.n, Mz, kc), is = Iat(I1, ks, cc.s, Mz, kc);
      uc = -0.25 * (ie * cc.hx_e + iw * cc.hx_w + in * cc.hx_n + is * cc.hx_s) + cc.ub;
      vc = -0.25 * (ie * cc.hy_e + iw * cc.hy_w + in * cc.hy_n + is * cc.hy_s) + cc.vb;
    }
    double ue = 0.0, uw = 0.0, vn = 0.0, vs = 0.0;
    if (D_x != 0.0) {
      if (east != 0.0) {
        const double ie = Iat(I0, ks, ce.e, Mz, kc), iw = Iat(I0, ks, ce.w, Mz, kc);
        const double in = Iat(I1, ks, ce.n, Mz, kc), is = Iat(I1, ks, ce.s, Mz, kc);
        ue = -0.25 * (ie * ce.hx_e + iw * ce.hx_w + in * ce.hx_n + is * ce.hx_s) + ce.ub;
      }
      if (west != 0.0) {
        const double ie = Iat(I0, ks, cw.e, Mz, kc), iw = Iat(I0, ks, cw.w, Mz, kc);
        const double in = Iat(I1, ks, cw.n, Mz, kc), is = Iat(I1, ks, cw.s, Mz, kc);
        uw = -0.25 * (ie * cw.hx_e + iw * cw.hx_w + in * cw.hx_n + is * cw.hx_s) + cw.ub;
      }
    }
    if (D_y != 0.0) {
      if (north != 0.0) {
        const double ie = Iat(I0, ks, cn.e, Mz, kc), iw = Iat(I0, ks, cn.w, Mz, kc);
        const double in = Iat(I1, ks, cn.n, Mz, kc), is = Iat(I1, ks, cn.s, Mz, kc);
        vn = -0.25 * (ie * cn.hy_e + iw * cn.hy_w + in * cn.hy_n + is * cn.hy_s) + cn.vb;
      }
      if (south != 0.0) {
        const double ie = Iat(I0, ks, cs.e, Mz, kc), iw = Iat(I0, ks, cs.w, Mz, kc);
        const double in = Iat(I1, ks, cs.n, Mz, kc), is = Iat(I1, ks, cs.s, Mz, kc);
        vs = -0.25 * (ie * cs.hy_e + iw * cs.hy_w + in * cs.hy_n + is * cs.hy_s) + cs.vb;
      }
    }
    if (act) {
      A.u[q * Mz + k] = uc;
      A.v[q * Mz + k] = vc;
      // PISMStressBalance.cc:276-278; a zero weight multiplies a zero here,
      // as in the reference where west*(...) is 0 * finite
      const double u_x = D_x * (west * (uc - uw) + east * (ue - uc));
      const double v_y = D_y * (south * (vc - vs) + north * (vn - vc));
      integrand = u_x + v_y;
      if (k > 0) dz = z[k] - z[k - 1];
    }
    double iprev = __shfl_up(integrand, 1, 64);
    if (lane == 0) iprev = carry_int;
    const double t = (act && k > 0) ? 0.5 * (integrand + iprev) * dz : 0.0;
    const double S = carry_sum + wave_inclusive_scan(t, lane);
    if (act) A.w[q * Mz + k] = w0 - S;
    carry_sum = __shfl(S, 63, 64);
    carry_int = __shfl(integrand, 63, 64);
  }
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
__global__ void k_dt(DevParams P, const unsigned long long *dmax_bits,
                     const unsigned long long *cfl_bits, double t_end, double *t, double *dt_out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
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

// ---------------------------------------------------------------- mass continuity
// IceModel::massContExplicitStep with cell_interface_fluxes and adjust_flow
// (iMgeometry.cc:172-415, 599-776), no part_grid, no Dirichlet BC.
__device__ __forceinline__ void adjust_flow_dir(int c, int nb, double &ssa, double &sia) {
  if (m_grounded_ice(c) && m_grounded_ice(nb)) return;
  if ((m_grounded_ice(c) && m_floating_ice(nb)) || (m_floating_ice(c) && m_grounded_ice(nb))) return;
  if ((m_grounded_ice(c) && m_ice_free_land(nb)) || (m_ice_free_land(c) && m_grounded_ice(nb))) return;
  if ((m_grounded_ice(c) && m_ice_free_ocean(nb)) || (m_ice_free_ocean(c) && m_grounded_ice(nb))) return;
  if (m_floating_ice(c) && m_floating_ice(nb)) { sia = 0.0; return; }
  if ((m_floating_ice(c) && m_ice_free_land(nb)) || (m_ice_free_land(c) && m_floating_ice(nb))) {
    sia = 0.0; ssa = 0.0; return;
  }
  if ((m_floating_ice(c) && m_ice_free_ocean(nb)) || (m_ice_free_ocean(c) && m_floating_ice(nb))) {
    sia = 0.0; return;
  }
  if (m_ice_free_land(c) && m_ice_free_land(nb)) { sia = 0.0; ssa = 0.0; return; }
  if ((m_ice_free_land(c) && m_ice_free_ocean(nb)) || (m_ice_free_ocean(c) && m_ice_free_land(nb))) {
    sia = 0.0; ssa = 0.0; return;
  }
  if (m_ice_free_ocean(c) && m_ice_free_ocean(nb)) { sia = 0.0; ssa = 0.0; return; }
}

__global__ __launch_bounds__(256) void k_mass_continuity(DevParams P, MassArgs A) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int Mx = P.Mx, My = P.My;
  __shared__ double red[8][256];
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (q < (long)Mx * My) {
    const int i = (int)(q / My), j = (int)(q % My);
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
    const long nb[4] = {qn, qe, qs, qw};
    const int mc = as_int(A.mask[q]);
    double Q[4] = {A.flux[2 * q + 1], A.flux[2 * q], A.flux[2 * qs + 1], A.flux[2 * qw]};
    double V[4];
#pragma unroll
    for (int d = 0; d < 4; d++) {
      const int mn = as_int(A.mask[nb[d]]);
      const bool ew = (d == 1 || d == 3);
      const double *vel = ew ? A.ub : A.vb;
      const double vij = vel ? vel[q] : 0.0, vnb = vel ? vel[nb[d]] : 0.0;
      if (m_icy(mc) && m_icy(mn)) V[d] = 0.5 * (vij + vnb);
      else if (m_icy(mc) && m_ice_free(mn)) V[d] = vij;
      else if (m_ice_free(mc) && m_icy(mn)) V[d] = vnb;
      else V[d] = 0.0;
      adjust_flow_dir(mc, mn, V[d], Q[d]);
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
  // fixed-order block reduction of the accounting terms
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
  if (threadIdx.x == 0 && A.partials) {
#pragma unroll
    for (int c = 0; c < 8; c++) A.partials[(long)blockIdx.x * 8 + c] = red[c][0];
  }
}

// sums the per-block partials in a fixed order (one block)
__global__ void k_reduce_partials(const double *__restrict__ partials, long nblocks, double *out) {
  __shared__ double red[8][256];
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
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
__global__ void k_mask_surface(DevParams P, const double *__restrict__ H_new, const double *__restrict__ bed,
                               double *__restrict__ H, double *__restrict__ mask, double *__restrict__ h,
                               int *err) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (long)P.Mx * P.My) return;
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

// ---------------------------------------------------------------- launchers
template <int LAW>
static void launch_column_law(const DevParams &P, const ColumnArgs &A, bool full, hipStream_t st) {
  const long nstag = 2L * P.Mx * P.My;
  const dim3 block(256);
  const dim3 grid((unsigned)((nstag + 3) / 4));
  if (full) {
    hipLaunchKernelGGL((k_column<LAW, true>), grid, block, 0, st, P, A);
  } else {
    hipLaunchKernelGGL((k_column<LAW, false>), grid, block, 0, st, P, A);
  }
}

void launch_column(const DevParams &P, const ColumnArgs &A, bool full, hipStream_t st) {
  switch (P.law) {
  case LAW_ISOGLEN: launch_column_law<LAW_ISOGLEN>(P, A, full, st); break;
  case LAW_PB: launch_column_law<LAW_PB>(P, A, full, st); break;
  case LAW_GPBLD: launch_column_law<LAW_GPBLD>(P, A, full, st); break;
  case LAW_HOOKE: launch_column_law<LAW_HOOKE>(P, A, full, st); break;
  case LAW_ARR: launch_column_law<LAW_ARR>(P, A, full, st); break;
  case LAW_ARRWARM: launch_column_law<LAW_ARRWARM>(P, A, full, st); break;
  default: launch_column_law<LAW_GK>(P, A, full, st); break;
  }
}

static inline unsigned blocks_for(long n, int per_block) { return (unsigned)((n + per_block - 1) / per_block); }

void launch_reset(unsigned long long *dmax_bits, int *err, unsigned long long *cfl_bits, double max_dt,
                  hipStream_t st) {
  hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, st, dmax_bits, err, cfl_bits, max_dt);
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
                     const double *mask, double *hx, double *hy, hipStream_t st) {
  const long n = (long)P.Mx * P.My;
  Fields2D F{H, b, h, mask};
  hipLaunchKernelGGL(k_gradient, dim3(blocks_for(n, 256)), dim3(256), 0, st, P, F, hx, hy);
}

void launch_theta_thk(const DevParams &P, const double *H, const double *b, const double *h,
                      const double *mask, const double *topgs, const double *maxtl, const double *C2,
                      const double *C3, const double *C4, double *theta, double *thks, int *err,
                      hipStream_t st) {
  const long n = (long)P.Mx * P.My;
  Fields2D F{H, b, h, mask};
  hipLaunchKernelGGL(k_theta_thk, dim3(blocks_for(n, 256)), dim3(256), 0, st, P, F, topgs, maxtl, C2, C3,
                     C4, theta, thks, err);
}

void launch_velocity_uvw(const DevParams &P, const VelArgs &A, hipStream_t st) {
  const long n = (long)P.Mx * P.My;
  hipLaunchKernelGGL(k_velocity_uvw, dim3(blocks_for(n, 4)), dim3(256), 0, st, P, A);
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

}  // namespace pism_sia
