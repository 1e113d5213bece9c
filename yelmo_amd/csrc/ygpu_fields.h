// ygpu_fields.h -- per-Picard-iteration field kernels of the SSA solve.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ygpu_common.h"

// BND_* codes (yelmo_tools.f90:14-20)
#define YG_BC_ZEROS 0
#define YG_BC_INFINITE 1
#define YG_BC_MISMIP3D 2
#define YG_BC_TROUGH 3
#define YG_BC_PERIODIC 4
#define YG_BC_PERIODIC_X 5
#define YG_BC_PERIODIC_Y 6

// boundaries strings, numbered as YGPU_BND_* in include/ygpu.h
#define YG_B_ZEROS 0
#define YG_B_INFINITE 1
#define YG_B_MASK 2
#define YG_B_MISMIP3D 3
#define YG_B_TROUGH 4
#define YG_B_PERIODIC 5
#define YG_B_PERIODIC_X 6
#define YG_B_PERIODIC_Y 7

struct YgFieldParams {
  int bnd;             // YG_B_*
  int BC;              // YG_BC_*
  int visc_method, beta_method, beta_gl_scale, beta_gl_stag, vel_lim_method;
  int diva;            // 1: DIVA viscosity / beta_eff
  int no_slip;         // DIVA no-slip variant
  int energy;          // 1: energy-formulation assembly (ssa_solver "energy")
  double visc_const, beta_const, beta_q, beta_u0, beta_gl_f, H_grnd_lim, beta_min, eps_0,
      vel_lim_tau, vel_max, iter_rel, iter_conv, n_glen, rho_ice, rho_sw;
};

// Device arrays of one batch, natural (Fortran) cell order, type R.
template <typename R>
struct YgFields {
  // inputs (2D unless noted)
  const R *c_bed, *f_slide, *taud_acx, *taud_acy, *taul_acx, *taul_acy, *H_ice, *f_ice, *H_grnd,
      *f_grnd, *z_bed, *z_sl;
  const R *f_grnd_acx, *f_grnd_acy;   // grounded fraction on the faces (subgrid staggering)
  const R* ATT;      // 3D, per member at zoff
  // ATT^(-1/n) at the 4 quadrature nodes of every cell and level (node q,
  // level k at (q nz + k) ncell + cell, per member at 4 zoff), filled once
  // per solve by k_init for visc_method 1 (ATT is an input of the solve);
  // null: computed in every Picard iteration
  R* attp;
  const R* zeta;     // nz per member at member index * zmax
  int zstride;
  const int32_t *mx, *my;
  // state
  R *ux, *uy, *ux_nm1, *uy_nm1;
  R *visc;   // 3D
  R *visc_int;
  R *beta, *beta_acx, *beta_acy, *bss_x, *bss_y;  // bss: friction used by the matrix
  R *lkx, *lky, *lrx, *lry;                       // limit drag
  R *dudx, *dudy, *dvdx, *dvdy;                   // strain rates
  R *err_x, *err_y, *taub_x, *taub_y;
  // DIVA
  R *duxdz, *duydz;           // 3D
  R *F2, *F2_acx, *F2_acy;
  R *beff_acx, *beff_acy, *beta_eff;
  R *ubar_x, *ubar_y;         // DIVA depth-averaged velocity (the solved field)
  // energy formulation: per unknown (2 per cell, global row order) the
  // member-local unknown whose solution a tied unknown takes, or -1; and a
  // per-member flag set when a row falls outside the stencil
  int32_t* cfrom;
  int* asm_err;
};

struct YgBatchGeom {
  const YgMember* members;
  const long long* coffs;     // member cell offsets (device)
  const long long* zoffs;     // member 3D offsets
  const int* soff;            // step tables
  int nmem;
  long long total_cells;      // per-cell kernels cover cells [0, total_cells)
  // Occupied cell ranges in ascending order of first cell: seg[k] is the
  // first cell of member segm[k] (segm null: member k).  A batch packs its
  // members back to back (seg = coffs); a server's slots are placed by its
  // cell allocator, so ranges can be out of member order and gaps can exist.
  const long long* seg;
  const int* segm;
  int nseg;
  // For a batch (seg = coffs): the member of the first cell of every chunk
  // of YG_CHUNK cells, and nmem - 1 after the last chunk (null: search).
  const int* chunk_seg;
};
#define YG_CHUNK 256

// member owning cell gc, or -1 for a cell outside every member's domain
YG_DEV int yg_member_at(const YgBatchGeom& G, long long gc) {
  int lo = 0;
  if (G.chunk_seg) {
    // from the chunk's first member forward (a chunk rarely holds a member start)
    lo = G.chunk_seg[gc / YG_CHUNK];
    while (lo + 1 < G.nseg && G.seg[lo + 1] <= gc) ++lo;
  } else {
    int hi = G.nseg - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (G.seg[mid] <= gc) lo = mid; else hi = mid - 1;
    }
  }
  if (G.nseg <= 0 || gc < G.seg[lo]) return -1;
  const int m = G.segm ? G.segm[lo] : lo;
  const YgMember& mb = G.members[m];
  return gc - G.seg[lo] < (long long)mb.nx * mb.ny ? m : -1;
}

template <typename R>
void yg_launch_fields_pre(const YgBatchGeom& G, const YgFields<R>& f, const YgFieldParams& p,
                          const int* skip, hipStream_t st);

template <typename R>
void yg_launch_assemble(const YgBatchGeom& G, const YgFields<R>& f, const YgFieldParams& p,
                        const int* skip, double* A64, float* A32, double* b, double* x0,
                        uint32_t* mask, hipStream_t st);

struct YgPicardState {
  int* iter_now;      // per member
  int* done;          // per member: 0 running, 1 converged, 2 iteration limit, -1 non-finite
  double* resid;      // per member Picard residual of the last iteration
  int* lin_iter;      // per member sums
  int* lin_fail;
  int* lim_n;
  // the epilogue's blocks per member (eparts > 1 needs both arrays): partial
  // sums [member][part][4] and per-member arrival counters, zero between launches
  double* epart = nullptr;
  int* eticket = nullptr;
  int eparts = 1;
};

#define YG_EPI_BLOCKS 8   // epilogue blocks per member

// Cells of the per-cell field kernels in XCD-local ranges.  The dispatcher
// deals workgroups round-robin over the 8 XCDs (workgroup b on XCD b % 8), so
// workgroup b takes logical block (b % 8) * (gridDim / 8) + b / 8: each XCD
// walks one contiguous eighth of the cells, and the neighbours of a cell one
// grid line away are mostly read through that XCD's L2.  A grid that is not a
// multiple of 8 blocks keeps the plain order.
YG_DEV long long yg_xcd_cell() {
  unsigned lb = blockIdx.x;
  if ((gridDim.x & 7u) == 0) lb = (blockIdx.x & 7u) * (gridDim.x >> 3) + (blockIdx.x >> 3);
  return (long long)lb * blockDim.x + threadIdx.x;
}
// blocks of b threads for n cells, rounded up to a multiple of 8
static inline unsigned yg_xcd_grid(long long n, int b) {
  const unsigned nb = (unsigned)((n + b - 1) / b);
  return (nb + 7u) & ~7u;
}

// One Picard epilogue per member not yet done; each member's iteration count
// (iter_now) advances by one.
template <typename R>
void yg_launch_epilogue(const YgBatchGeom& G, const YgFields<R>& f, const YgFieldParams& p,
                        const double* x, const int* lin_iters, const int* lin_status,
                        int iter_max, YgPicardState ps, hipStream_t st);

// skip (may be null): members left untouched
template <typename R>
void yg_launch_finish(const YgBatchGeom& G, const YgFields<R>& f, const YgFieldParams& p,
                      YgPicardState ps, const int* skip, hipStream_t st);

template <typename R>
void yg_launch_init(const YgBatchGeom& G, const YgFields<R>& f, const YgFieldParams& p,
                    const int* skip, hipStream_t st);

