// ygpu_common.h -- device-side shared definitions of the SSA velocity solver.
//
// Cell ordering.  The reference numbers unknowns i-outer, j-inner
// (solver_linear.F90:134-143): cell n = i*ny + j, ux -> row 2n, uy -> 2n+1.
// A grid line is one i (2*ny consecutive rows).  Every lower-triangular
// coupling of the C-grid stencil reaches at most line i-1, cells j and j+1,
// so the ILU(0) recurrences advance as a wavefront: cell (i,j) can be done at
// step t = j + 2*i.  All solver vectors, the stencil matrix and the ILU
// factors are stored in that "skewed" order: the cells of step t are
// contiguous, in increasing i.  A lane per line then reads and writes
// contiguous memory at every step of a triangular sweep, and the SpMV, dots
// and updates stream the same arrays.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define YG_DEV __device__ __forceinline__

// Stencil slots of one row (residual assembly, solver_ssa_ac.f90:473-545,
// 765-835).  ux row (i,j):
//   0 ux(i,j) 1 ux(i+1,j) 2 ux(i-1,j) 3 ux(i,j+1) 4 ux(i,j-1)
//   5 uy(i,j) 6 uy(i+1,j) 7 uy(i+1,j-1) 8 uy(i,j-1)
// uy row (i,j):
//   0 uy(i,j) 1 uy(i,j+1) 2 uy(i,j-1) 3 uy(i+1,j) 4 uy(i-1,j)
//   5 ux(i,j) 6 ux(i,j+1) 7 ux(i-1,j+1) 8 ux(i-1,j)
// Lateral-front and free-slip rows use subsets of these slots.
#define YG_NSLOT 9

// The stencil matrix on the device: one record of YG_NA values per cell,
// value 2*s + comp = slot s of the cell's ux (comp 0) or uy (comp 1) row,
// then two zero pad values.  Absent slots hold exact zeros.  One record is
// five (FP32) or ten (FP64) 16-byte chunks, so a cell's row pair arrives in
// a few wide loads from one address.
#define YG_NA 20

// (di, dj, comp) of each slot; comp 0 = ux, 1 = uy
__constant__ static const int8_t kSlotUx[9][3] = {
    {0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0},
    {0, 0, 1}, {1, 0, 1}, {1, -1, 1}, {0, -1, 1}};
__constant__ static const int8_t kSlotUy[9][3] = {
    {0, 0, 1}, {0, 1, 1}, {0, -1, 1}, {1, 0, 1}, {-1, 0, 1},
    {0, 0, 0}, {0, 1, 0}, {-1, 1, 0}, {-1, 0, 0}};

// ILU(0) factor values of a cell:
//   L ux row : slots 2,4,8            -> F0..F2
//   L uy row : slots 8,4,7,2,5        -> F3..F7
//   U ux row : slots 5,3,7,1,6 + 1/d  -> F8..F13
//   U uy row : slots 6,1,3 + 1/d      -> F14..F17
// Two storage layouts, YG_NF values per cell:
//   coupled (exact ILU(0) over more than 64 lines): SoA, F[v*ncell + q] over
//     the member's skewed cell index q;
//   streamed (line blocks of at most 64 lines, one per wave): AoS in the
//     rotated tile order of the sweeps (cell (k, j) of a tile of L lines and
//     M columns at slot ((j + 2k) mod M) L + k), L part [q][8] (F0..F7) then
//     U part [q][12] (F8..F17 and two pad values), so that one lane's cell is
//     2 (3) 16-byte chunks and one wavefront step of a tile is a contiguous
//     range.
#define YG_NF 20

// One member (one independent system) of a batch.
struct YgMember {
  int nx, ny, nz;
  int T;             // wavefront steps = ny + 2*(nx-1)
  int maxcnt;        // most cells in one step
  long long coff;    // first cell of this member in packed per-cell arrays
  long long toff;    // first entry of this member in the packed step tables
  long long woff;    // first value of this member's work vectors (multiple of 4)
  int vlen;          // length of one work vector: 2*ncell rounded up to a multiple of 4
  int nb;            // ILU(0) line blocks of this member (0: the solve options decide)
  int ncol;          // column blocks per line block (tiles; 0: the solve options decide)
  double dx, dy;
};

// first line active at step t
YG_DEV int yg_glo(int t, int ny) {
  int a = t - (ny - 1);
  return a <= 0 ? 0 : (a + 1) >> 1;
}

// skewed cell index of (i,j) within a member
YG_DEV int yg_q(const int* __restrict__ soff, int i, int j, int ny) {
  int t = j + 2 * i;
  return soff[t] + i - yg_glo(t, ny);
}

// ---- reductions --------------------------------------------------------
template <typename T>
YG_DEV T yg_wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Block-wide sum of NV values; every thread gets the totals.  `red` is
// __shared__ scratch of at least 16*NV doubles.
template <int NV>
YG_DEV void yg_block_sum(double (&v)[NV], double* red) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = yg_wave_sum(v[k]);
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) red[w * NV + k] = v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double s = 0.0;
    for (int q = 0; q < nw; ++q) s += red[q * NV + k];
    v[k] = s;
  }
  __syncthreads();
}
