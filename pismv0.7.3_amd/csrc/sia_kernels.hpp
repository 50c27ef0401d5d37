// sia_kernels.hpp -- launch interface of the SIA kernels (host side).
#pragma once

#include "sia_device.hpp"

namespace pism_sia {

struct Fields2D {
  const double *H, *b, *h, *mask;
};

struct ColumnArgs {
  const double *z;           // [Mz]
  const double *E;           // [Mx][My][Mz]
  const double *age;         // [Mx][My][Mz] or nullptr
  const double *thks;        // [Mx][My]
  const double *theta;       // [Mx][My]
  const double *hx, *hy;     // [Mx][My][2]
  double *flux;              // [Mx][My][2]
  double *Dstag;             // [Mx][My][2]
  double *I0, *I1;           // [Mx][My][Mz] (full updates; levels 0..ks stored)
  int *ks;                   // [Mx][My][2] kBelowHeight of each staggered column (full updates)
  unsigned long long *dmax_bits;
  double *dmax_part;         // per-block maxima of D [max(column_blocks(P), column_strip_blocks(P))]
  int *err;
};

// levels allocated per staggered column of I (the column kernel's output on
// full updates, read by the velocity tiles): Mz rounded up to 16, so that every
// column starts on a 128-byte line and a wavefront's 64 levels cover 4 lines
__host__ __device__ inline int i_stride(int Mz) { return (Mz + 15) & ~15; }

struct VelArgs {
  const double *z, *mask, *I0, *I1, *hx, *hy, *ub, *vb, *bmr;
  const int *ks;
  double *u, *v, *w;
  // the volumetric strain heating in the same pass when Sigma is given:
  // thickness, enthalpy, Sigma [Mx][My][Mz], error flags
  const double *H, *E;
  double *Sigma;
  int *err;
  // [Mz] 1 / (z_{k+1} - z_{k-1}), 1 / (z_1 - z_0) at k = 0, 0 at k = Mz - 1 (the
  // centred vertical differences of the strain heating)
  const double *idz;
};

// SIA_Sliding (SIA_Sliding.cc:106-267): mode 0 pressure-melting-activated
// mu_sliding, 1 verification test E, 2 EISMINT II G, 3 EISMINT II H
struct SlideArgs {
  Fields2D F;                // H, b, h, mask
  const double *E;           // [Mx][My][Mz] (the basal level is read)
  const double *x, *y;       // [Mx], [My] grid coordinates (test E)
  double *hx, *hy;           // [Mx][My][2] the sliding law's own surface gradient
  double *ub, *vb, *heat;    // [Mx][My]
  int mode;
  double mu, min_T, secpera;
};

struct MassArgs {
  const double *dt;          // device scalar
  const double *H, *flux, *mask, *ub, *vb, *smb, *bmr;
  double *H_new;
  double *partials;          // [nblocks][8] or nullptr
};

// clears D_max and the per-update error flags (which are first ORed into
// err_acc) and sets the CFL bound to max_dt
void launch_reset(unsigned long long *dmax_bits, int *err, int *err_acc, unsigned long long *cfl_bits,
                  double max_dt, hipStream_t st);
void launch_clear_flags(int *err, int *err_acc, hipStream_t st);
void launch_fill(double *a, long n, double v, hipStream_t st);
void launch_bed_smoother(const DevParams &P, double s2, double s3, double s4, const double *b0, double *bs,
                         double *mt, double *c2, double *c3, double *c4, hipStream_t st);
// theta (get_theta) and the smoothed thickness (get_smoothed_thk) of every map
// point, computed by the gradient kernel beside the surface gradient
struct ThetaArgs {
  const double *topgs, *maxtl, *C2, *C3, *C4;
  double *theta, *thks;
  int *err;
};
void launch_gradient(const DevParams &P, const double *H, const double *b, const double *h,
                     const double *mask, double *hx, double *hy, const ThetaArgs &T, hipStream_t st);
void launch_regional_gradient(const DevParams &P, const double *nmm, const double *hst, double *hx, double *hy,
                              hipStream_t st);
// blocks (= per-block D_max partials) of the column pass in staggered order,
// and in XCD-interleaved strips (used where column_strips_ok holds and the
// caller allows it); A.dmax_part must hold the larger of the two
int column_blocks(const DevParams &P);
int column_strip_blocks(const DevParams &P);
// strips pay where a row of the grid holds several of them
__host__ __device__ inline bool column_strips_ok(const DevParams &P) { return P.My >= 64; }
// extra_partials: further per-block D_max partials that other kernels wrote
// after the column pass's own; all are reduced into A.dmax_bits
void launch_column(const DevParams &P, const ColumnArgs &A, bool full, hipStream_t st, int extra_partials = 0,
                   bool allow_strips = true);
// fast-update FP64 pass over the n_list (device count) staggered columns of
// list: column_list_blocks(P) <= column_blocks(P) per-block D_max partials,
// then extra_partials more, reduced into A.dmax_bits
int column_list_blocks(const DevParams &P);
void launch_column_list(const DevParams &P, const ColumnArgs &A, const int *list, const int *n_list,
                        hipStream_t st, int extra_partials = 0);
void launch_velocity_uvw(const DevParams &P, const VelArgs &A, hipStream_t st);
void launch_sia_sliding(const DevParams &P, const SlideArgs &A, hipStream_t st);
void launch_surface_values(const DevParams &P, const double *z, const double *f, const double *H, double *out,
                           hipStream_t st);
void launch_diffusivity_regular(const DevParams &P, const double *D, double *out, hipStream_t st);
void launch_cfl2d(const DevParams &P, const double *mask, const double *ub, const double *vb,
                  unsigned long long *cfl_bits, hipStream_t st);
void launch_dt(const DevParams &P, const unsigned long long *dmax_bits, const unsigned long long *cfl_bits,
               double t_end, double *t, double *dt_out, hipStream_t st);
long mass_partial_blocks(const DevParams &P);
void launch_mass_continuity(const DevParams &P, const MassArgs &A, double *acc_out, hipStream_t st);
void launch_mask_surface(const DevParams &P, const double *H_new, const double *bed, double *H,
                         double *mask, double *h, int *err, hipStream_t st);

// Pointers read from a device table are generic to the compiler, and its
// accesses through them become flat instructions, which also wait on the LDS
// counter. Everything the tables point to is device memory: rebuilding the
// pointer from its integer value as a global-address-space pointer lets the
// compiler issue global loads and stores (null stays null).
template <typename T>
__device__ __forceinline__ T *glb(T *p) {
  return (T *)(__attribute__((address_space(1))) T *)(uintptr_t)p;
}

__device__ __forceinline__ Fields2D global_view(Fields2D f) {
  f.H = glb(f.H);
  f.b = glb(f.b);
  f.h = glb(f.h);
  f.mask = glb(f.mask);
  return f;
}

__device__ __forceinline__ ColumnArgs global_view(ColumnArgs a) {
  a.z = glb(a.z);
  a.E = glb(a.E);
  a.age = glb(a.age);
  a.thks = glb(a.thks);
  a.theta = glb(a.theta);
  a.hx = glb(a.hx);
  a.hy = glb(a.hy);
  a.flux = glb(a.flux);
  a.Dstag = glb(a.Dstag);
  a.I0 = glb(a.I0);
  a.I1 = glb(a.I1);
  a.ks = glb(a.ks);
  a.dmax_bits = glb(a.dmax_bits);
  a.dmax_part = glb(a.dmax_part);
  a.err = glb(a.err);
  return a;
}

__device__ __forceinline__ MassArgs global_view(MassArgs a) {
  a.dt = glb(a.dt);
  a.H = glb(a.H);
  a.flux = glb(a.flux);
  a.mask = glb(a.mask);
  a.ub = glb(a.ub);
  a.vb = glb(a.vb);
  a.smb = glb(a.smb);
  a.bmr = glb(a.bmr);
  a.H_new = glb(a.H_new);
  a.partials = glb(a.partials);
  return a;
}

// ---- ensembles: one fast SIA step + dt + mass continuity of every member, one
// launch per stage (device table of members; the member's fields are its own)
struct EnsMember {
  DevParams P;
  Fields2D F;                       // H, b, h, mask read by the update
  const double *topgs, *maxtl, *C2, *C3, *C4;
  double *hx, *hy, *theta, *thks;
  ColumnArgs A;                     // flux, D, D_max partials (n_part of them), errors
  MassArgs M;                       // dt, H, flux, mask -> H_new (no accounting partials)
  double *H, *mask, *h;             // geometry written back after mass continuity
  unsigned long long *cfl_bits;
  double *t, *dt;                   // device time and time step of the member
  int *err_acc;                     // error flags of all earlier steps (A.err: this step)
  double t_end;
  int n_part;
};

// member of block b of a stage grid: start[m] <= b < start[m + 1] (start: device [nm + 1])
__device__ __forceinline__ int ens_member(const int *__restrict__ start, int nm, int b) {
  int lo = 0, hi = nm;  // start[lo] <= b < start[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (start[mid] <= b) lo = mid; else hi = mid;
  }
  return lo;
}

struct EnsLaunch {
  const EnsMember *table;           // device [nm]
  const int *pt_start, *col_start;  // device [nm + 1]: first block of each member
  int nm, pt_blocks, col_blocks, cpw, law, max_mz;
};

// list-driven FP64 fast column pass of many members (the surrogate's
// fallback lists): block b of the grid belongs to member m with start[m] <= b <
// start[m + 1] and walks that member's list with its share of blocks; its D_max
// goes to A.dmax_part[b - start[m]]
struct EnsListMember {
  DevParams P;
  ColumnArgs A;
  const int *list, *n_list;         // staggered columns, device count
};
void launch_ens_column_list(const EnsListMember *T, const int *start, int nm, int nblocks, int max_mz, int law,
                            hipStream_t st);

int ens_point_blocks(const DevParams &P);
int ens_column_blocks(const DevParams &P, int cpw);
void launch_ens_step(const EnsLaunch &L, hipStream_t st);
// the step in two halves: reset + gradient, then column + dt + mass + mask
// (the surrogate stages of surrogate-bound members go in between)
void launch_ens_front(const EnsLaunch &L, hipStream_t st);
void launch_ens_back(const EnsLaunch &L, hipStream_t st);

}  // namespace pism_sia
