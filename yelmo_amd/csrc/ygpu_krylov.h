// ygpu_krylov.h -- host/device interface of the persistent Krylov kernel.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ygpu_common.h"

#define YG_MODE_FP64 0
#define YG_MODE_MIXED 1
#define YG_MAX_STEPS 3072   // ny + 2*(nx-1) limit of the LDS step table
#define YG_MAX_LINES 1024   // one lane per grid line, one block per member
// Most cells of one wavefront step (min(nx, ceil(ny/2))): one step of the
// FP64 factorization stage (19 doubles per cell, two buffers) must fit the
// kernel's dynamic LDS.
#define YG_MAX_STEP_CELLS 400

// BF16 line-block factors, 16-bit values allocated per cell: 20 are used (L
// 8, scaled U 8, the two FP32 inverse diagonals 4); 24 keep every member's
// factors 16-byte aligned
#define YG_NF16 24

// Krylov workspace vectors per member (2 values per cell each)
#define YG_NW64 9
#define YG_NW32 11

struct YgKrylovOpts {
  int solver;       // 1 cg, 4 bicgstab (LIS numbering)
  int precon;       // 0 none, 1 jacobi, 2 ilu
  int maxiter;
  int initx_zeros;
  int conv_cond;    // 0 ||r||/||r0||, 1 ||r||/||b||
  int ir_max;       // mixed mode: refinement sweeps
  int ilu_cols;     // block-Jacobi ILU(0): column blocks per line block (tiles of
                    // lines x columns; 1: whole lines)
  int ilu_blocks;   // 0 exact ILU(0); n > 0 block-Jacobi ILU(0) over n line blocks
                    // (at least one per 64 lines); -1 automatic: exact in the
                    // FP64 mode, one block per 64 lines in the mixed mode
  double tol;
  double inner_tol; // mixed mode: relative tolerance of each FP32 inner solve
  double ir_target; // mixed mode: residual the refinement stops at (0: tol / 10)
  int factor16;     // mixed mode: 1 = ILU(0) factors stored in BF16 (U rows scaled
                    // by their FP32 inverse diagonals); 2 = the U part in BF16 and
                    // the L part in FP32
  int graph_chunk;  // host side: Picard iterations per captured graph (0: no graph)
  int ilu_rev;      // block-Jacobi ILU(0): 1 = columns of each line block numbered in
                    // reverse (wavefront skew 1); used with FP32/FP64 factors when the
                    // factorization's stage holds a step of the reversed wavefront
  double ir_loose;  // mixed mode: a member whose last Picard residual (picard_resid)
                    // is above ir_loose solves to the tolerance itself with inner
                    // solves to loose_inner; the others (and every member when 0) to
                    // ir_target with inner solves to inner_tol
  double loose_inner;
  int keep_factors; // mixed mode, ILU(0): 1 = a member flagged in YgKrylovArgs::warm uses
                    // the line-block factors left by its previous solve in the first
                    // Picard iteration of the next one instead of factorizing anew
  double ir_delta;  // mixed mode, BiCGSTAB: > 0 (1e-3 by default) one FP32 BiCGSTAB
                    // whose residual is replaced by the FP64 one (b - A x in FP64, x
                    // accumulated in FP64) whenever the iterated residual fell below
                    // ir_delta times its value at the last replacement, the search
                    // direction kept (reliable updates); 0 restarts a new inner solve
                    // at every FP64 residual instead (inner_tol)
};

struct YgKrylovArgs {
  const YgMember* members;
  const int* soff;          // packed step offsets, T+1 per member at members[m].toff
  const uint32_t* mask;     // per natural cell (n = i*ny + j) presence bits
  const double* A64;        // 18 per cell: [slot][natural cell][comp]
  const float* A32;
  double* F64;              // ILU factors (skewed wavefront order) / jacobi 1/diag
  float* F32;
  uint16_t* F16;            // mixed mode, BF16 factors: YG_NF16 values per cell (or null)
  const double* b;          // 2 per cell, the reference's row order 2n, 2n+1
  double* x;                // 2 per cell, x0 in, solution out
  double* W64;              // YG_NW64 * 2 per cell
  float* W32;               // YG_NW32 * 2 per cell
  long long total_cells;
  int wrapx, wrapy;
  const int* skip;          // members to skip (may be null)
  int* out_iters;
  int* out_status;
  double* out_resid;
  int* out_sweeps;          // optional (mixed mode): FP64 residual passes of the refinement
  const double* picard_resid;   // optional (mixed mode): per member Picard residual of the
                                // last iteration (opts.ir_loose)
  const int* warm;          // optional (mixed mode, opts.keep_factors): per member, nonzero
                            // when its solve continues a sequence, whose previous solve's
                            // factors then stand in while iter_now of the member is 0
  const int* iter_now;      // per member Picard iterations finished in the current solve
  long long* timers;        // optional: per member 8 phase totals (wall clock ticks)
  YgKrylovOpts opts;
  // member groups (yg_launch_krylov_group): workgroups per member, members,
  // per-group synchronization words (YG_GSYNC_BYTES each, cleared by the
  // launcher) and a flag set when a group wait timed out (cleared by the host)
  int group;
  int nmembers;
  int* gsync;
  int* group_fail;
  // member groups of different sizes (uneven batches; null gtab: every
  // group has a.group workgroups, workgroup b is rank (b/8) % G of group
  // (b/8/G)*8 + b%8, and group g takes members g, g + ngroups, ...):
  // gtab[3 b + {0, 1, 2}] = group, rank, group size of workgroup b (group
  // -1: the workgroup has no part and returns); group g takes the members
  // gmem[gmoff[g] .. gmoff[g+1]) in turn; nwg workgroups and ngrp groups
  const int* gtab;
  const int* gmoff;
  const int* gmem;
  int nwg, ngrp;
  // member groups in the plain FP64 mode (Jacobi or no preconditioner; W64's
  // nine vectors hold the Krylov vectors); 0: the mixed mode
  int plain64;
};

void yg_launch_krylov(const YgKrylovArgs& a, int mode, int nmembers, int block, hipStream_t st);
// The member-group kernel: ngroups groups of a.group workgroups (threads 512,
// one workgroup per CU, or 256, two per CU; ngroups a multiple of 8, the
// workgroups at most what the CUs hold at once) take the a.nmembers members
// in turn: the mixed mode (BiCGSTAB with line-block ILU(0), Jacobi or none;
// CG with Jacobi or none) or, with a.plain64, the FP64 mode with Jacobi or
// no preconditioner.  Clears the groups' synchronization words first.
void yg_launch_krylov_group(const YgKrylovArgs& a, int ngroups, int threads, hipStream_t st);
#define YG_GSYNC_BYTES 2048
// threads per workgroup of the member groups that batches and servers pick
// when all their groups fit the CUs (two workgroups per CU at most)
#define YG_GROUP_THREADS 256
// sets the kernels' launch attributes; call before capturing a launch in a graph
void yg_krylov_prepare();
