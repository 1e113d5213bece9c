// ygpu_krylov_impl.h -- persistent per-system Krylov solve on the stencil matrix.
//
// Included once per block size by ygpu_krylov.hip (512 threads) and
// ygpu_krylov_1024.hip (1024 threads, grids of more than 512 lines): each
// translation unit holds its own copies of the device functions, so their
// register budget follows the one kernel that calls them (256 VGPRs per
// lane at 512 threads, 128 at 1024).
//
// One workgroup owns one system of a batch for the whole linear solve:
// ILU(0) factorization, preconditioned BiCGSTAB (or CG) and, in the mixed
// precision mode, the FP64 iterative-refinement loop around an FP32 inner
// solve.  Nothing is launched per Krylov step and no host synchronisation
// happens inside a solve; every member of a batch runs concurrently on its
// own CU.  The semantics follow the LIS call behind
// linear_solver_matrix_solve (solver_linear.F90:151-400): b == 0 gives x = 0,
// status 0 converged / 2 breakdown / 4 iteration limit, the reported residual
// is ||b - A x|| relative to ||r0|| (LIS conv_cond 0) or ||b|| (conv_cond 1).
#pragma once
#include "ygpu_common.h"
#include "ygpu_krylov.h"

// Wavefront phases are separate device functions: each gets its own register
// allocation instead of sharing the pressure of the whole inlined solver.
#define YG_NOINLINE __device__ __attribute__((noinline))

namespace {

// --------------------------------------------------------------------------
// per-block context
//
// Grid lines are mapped onto the waves of the block for the wavefront
// recurrences (ILU(0) factorization and the triangular sweeps).  Wave w owns
// lines [w*lpw, w*lpw + lpw), lane l line w*lpw + l.  Two schemes:
//   coupled  (exact ILU(0)): lpw = 64; the first lane of a wave reads the
//            last lane of the wave before it through an LDS ring, one block
//            barrier per wavefront step;
//   blocked  (block-Jacobi ILU(0)): the lines are split into nb line blocks
//            of lpw <= 64 lines, one per wave; couplings between blocks are
//            dropped from the factorization, so waves never wait for each
//            other inside a sweep.  A line block may further be cut into
//            tiles of consecutive columns (one per wave in the sweeps, which
//            then need mcol + 2(lpw-1) steps instead of ny + 2(lpw-1)).
struct Grp {
  int* cnt;       // arrivals at group barriers (monotonic over the launch)
  int* err;       // set when a wait timed out
  double* part;   // partial sums: [2 (barrier parity)][G][4]
  int G, rank;
};
#define YG_GMAX 16   // largest member group
struct Ctx {
  const YgMember* mb;
  int nx, ny, T;
  int ncell;
  int vlen;                   // stride of the work vectors (values, multiple of 4)
  int wrapx, wrapy;
  const uint32_t* mask;       // per natural cell: presence bits of its two rows
  int maxcnt;                 // most cells in one wavefront step
  int line;                   // grid line of this thread in the recurrences, -1 if none
  int coupled;                // waves exchange boundary values every step
  int nwl;                    // waves that own lines
  int lpw;                    // lines per wave (the line block of a wave)
  // blocked scheme: each line block is split into ncol tiles of mcol
  // consecutive columns (the last one narrower); couplings between tiles
  // are dropped like those between line blocks, and each wave of a sweep
  // runs one tile (ncol = 1: whole lines)
  int ncol, mcol;
  // blocked scheme: the columns of every line block (tile) are numbered in
  // reverse (line i, column j before column j-1), which makes each line's
  // lower couplings to the line before it reach only columns j and j+1, so
  // line i can take column index t - i at step t (skew 1: ny + L - 1 steps
  // instead of ny + 2(L-1)).  The natural order (rev = 0) has skew 2.
  int rev;
  long long* tm;              // phase timers (thread 0), may be null
  // Offsets (values) from a factor pointer to the U part of its line-block
  // factors: 8 ncell for a whole member (FP32/FP64 and BF16 alike), and, for
  // BF16 factors, to the inverse diagonals (16 ncell); a part of a member
  // (the lines of one block of a member group) addresses the member's U
  // part and diagonals from its own L part.
  long long fu, fu16, fd16;
  // The cells [cb, ce) of the member whose rows this block computes in the
  // SpMV and vector passes (the whole member, or this block's share in a
  // member group), and the group (null for one block per member) with a copy
  // of it (gv), read without a pointer
  int cb, ce;
  const Grp* grp;
  Grp gv;
  // the context of the preconditioner's part of the member (itself for one
  // block per member)
  const Ctx* pc;
  int lds;   // bytes of the kernel's dynamic LDS (the factorization's stage)
};

// optional phase timing: 0 factor, 1 sweep L, 2 sweep U, 3 spmv, 4 vectors
struct PhaseTimer {
  const Ctx& c;
  int k;
  long long t0;
  YG_DEV PhaseTimer(const Ctx& cc, int kk) : c(cc), k(kk), t0(0) {
    if (c.tm && threadIdx.x == 0) t0 = wall_clock64();
  }
  YG_DEV ~PhaseTimer() {
    if (c.tm && threadIdx.x == 0) c.tm[k] += wall_clock64() - t0;
  }
};

// The stage buffers live in the kernel's dynamic LDS.  They are always
// addressed from this symbol (never through a stored pointer) so that the
// compiler emits LDS instructions for them.
extern __shared__ __align__(16) unsigned char yg_dyn_lds[];

// Element k of the dynamic LDS viewed as T.  Stage buffers are addressed by
// integer offsets through this accessor (never through a pointer held in a
// variable), which keeps every access an LDS instruction: a generic pointer
// would make the compiler emit flat accesses, whose completion waits also
// cover every outstanding global load.
template <typename T>
YG_DEV T& lds(int k) { return reinterpret_cast<T*>(yg_dyn_lds)[k]; }

// The block's copy of its member's step table and the ring through which
// neighbouring waves pass boundary values, addressed by symbol for the same
// reason.
__shared__ int yg_soff[YG_MAX_STEPS + 1];
__shared__ double yg_ring[16][3][10];
__shared__ double yg_red[16 * 4];   // block reductions

// Column tiles of the blocked scheme: want tiles per line block, at most
// `fit` (the waves per line block), at least 8 columns each.
YG_DEV void yg_tiles(int want, int fit, int ny, Ctx& c) {
  int n = max(1, min(want, min(fit, ny / 8)));
  c.mcol = (ny + n - 1) / n;
  c.ncol = (ny + c.mcol - 1) / c.mcol;
}

// ---- member groups ------------------------------------------------------
// G workgroups (one per CU) can share one member's solve.  Workgroup b runs
// on XCD b % 8 (the dispatcher deals workgroups round-robin over the 8
// XCDs), so a group is made of workgroups with equal b % 8 and shares one
// L2: stores reach that L2 when they complete (the L1 writes through), and a
// load that skips the L1 (device scope) sees them.  A group barrier is then
//   every wave waits for its own stores, block barrier, one lane adds 1 to
//   the group's counter and waits for it to reach G x (barriers so far);
// and data written by another block of the group (the halo lines of x in the
// SpMV, the partial sums of the reductions) is read only with device-scope
// loads.  No cache is written back or invalidated.  Every wait is bounded in
// time: a block that waits longer than YG_GRP_SPIN sets the group's error
// flag, after which no block of the group waits again (the solve runs on to
// its iteration limit on meaningless data and is reported failed).
// The placement is checked, not assumed: every block reads its XCD from the
// XCC_ID hardware register and the group compares them behind a first
// barrier with device-scope fences; a group that finds itself spread over
// several XCDs keeps those fences on every barrier (L2 write-back before the
// arrival, invalidation after the wait: correct across L2s, slower).
#define YG_GRP_SPIN 200000000ll   // wall-clock ticks (100 MHz): 2 s
__shared__ int yg_gbar;    // group barriers this block has passed
__shared__ int yg_gfail;   // a wait of this group timed out
__shared__ int yg_gfence;  // the group spans several XCDs: barriers fence
__shared__ long long yg_gwait;   // wall-clock ticks this block spent waiting at group barriers

YG_DEV int grp_load(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
YG_DEV double grp_loadd(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Barrier k of the group (k = yg_gbar + 1 on every block).
// The wait for the wave's stores is bracketed by compiler fences: the
// s_waitcnt intrinsic touches no memory as far as the compiler knows, so
// without them a store issued before the barrier could be scheduled after the
// wait and reach the L2 after this block's arrival is counted.
YG_DEV void grp_barrier(const Grp& g) {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);   // this wave's stores have reached the L2
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __syncthreads();
  if (threadIdx.x == 0) {
    const int k = yg_gbar + 1;
    if (yg_gfence) __threadfence();   // release at device scope
    if (!yg_gfail) {
      __hip_atomic_fetch_add(g.cnt, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const long long t0 = wall_clock64();
      while (grp_load(g.cnt) < g.G * k) {
        if (grp_load(g.err) || wall_clock64() - t0 > YG_GRP_SPIN) {
          yg_gfail = 1;
          __hip_atomic_store(g.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
      yg_gwait += wall_clock64() - t0;
    }
    if (yg_gfence) __threadfence();   // acquire at device scope
    yg_gbar = k;
  }
  __syncthreads();
}

// XCD of the calling workgroup (hardware register XCC_ID, bits 3:0)
YG_DEV int xcc_id() { return __builtin_amdgcn_s_getreg((20) | (0 << 6) | ((4 - 1) << 11)) & 0xf; }

// The group's first barrier, with device-scope fences, also compares the
// XCDs of its workgroups (published in the synchronization words); from
// then on the barriers fence only if they differ.
YG_DEV void grp_start(const Grp& g, int* xcc_words) {
  if (threadIdx.x == 0) {
    yg_gbar = 0;
    yg_gfail = 0;
    yg_gfence = 1;
    yg_gwait = 0;
    xcc_words[g.rank] = xcc_id();
  }
  __syncthreads();
  grp_barrier(g);
  if (threadIdx.x == 0) {
    const int x0 = grp_load(xcc_words);
    int spread = 0;
    for (int r = 1; r < g.G; ++r) spread |= grp_load(xcc_words + r) != x0;
    yg_gfence = spread;
  }
  __syncthreads();
}

// Sums of K values over the block (one block per member) or over the group:
// every block publishes its block sums, passes a group barrier and adds the
// G partial sums in rank order, so every block of the group gets the same
// bits.  Partial sums alternate between two buffers by barrier parity: a
// block can reach the next sum only after every block has passed this
// barrier, and the sum after that only after every block has read this one.
// The partial sums are loaded together (one round trip to the L2) and then
// added in rank order.  The loads cover YG_GMAX ranks whatever G is: the
// partial-sum buffer lies inside the group's synchronization block
// (YG_GSYNC_BYTES from byte 128), which holds [2][YG_GMAX][4] doubles, so
// every load stays inside it; values of ranks past G are not used.
static_assert(128 + 2 * YG_GMAX * 4 * sizeof(double) <= YG_GSYNC_BYTES, "partial sums");
template <int K>
YG_DEV void csum(const Ctx& c, double (&v)[K]) {
  static_assert(K <= 4, "four partial sums per rank");
  yg_block_sum<K>(v, yg_red);
  if (!c.grp) return;
  const Grp g = c.gv;
  double* slot = g.part + (size_t)((yg_gbar + 1) & 1) * g.G * 4;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int e = 0; e < K; ++e) slot[g.rank * 4 + e] = v[e];
  }
  grp_barrier(g);
  double pv[YG_GMAX][K];
#pragma unroll
  for (int r = 0; r < YG_GMAX; ++r) {
#pragma unroll
    for (int e = 0; e < K; ++e) pv[r][e] = grp_loadd(slot + r * 4 + e);
  }
#pragma unroll
  for (int e = 0; e < K; ++e) {
    double t = 0.0;
#pragma unroll
    for (int r = 0; r < YG_GMAX; ++r)
      if (r < g.G) t += pv[r][e];
    v[e] = t;
  }
}

// Global-memory views of pointers handed to the (non-inlined) wavefront
// phases: the compiler then emits global (not flat) accesses, whose
// completion counter it can track per instruction, so loads issued ahead
// stay in flight.
#define YG_GLOBAL __attribute__((address_space(1)))
template <typename T>
YG_DEV const YG_GLOBAL T* gc(const T* p) { return (const YG_GLOBAL T*)p; }
template <typename T>
YG_DEV YG_GLOBAL T* gm(T* p) { return (YG_GLOBAL T*)p; }

template <typename T> struct Vec2;
template <> struct Vec2<float> { using type = float2; };
template <> struct Vec2<double> { using type = double2; };

// first line of step t that holds a cell: max(0, ceil((t - ny + 1) / s))
// for the wavefront of skew s (2 natural, 1 reversed columns)
YG_DEV int glo_of(const Ctx& c, int t) {
  if (c.rev) { const int a = t - c.ny + 1; return a > 0 ? a : 0; }
  const int a = t - c.ny + 2;
  return a > 0 ? a >> 1 : 0;
}

// sum over s < t of floor(s / 2)
YG_DEV int half_sum(int t) {
  if (t <= 0) return 0;
  const int m = t >> 1;
  return (t & 1) ? m * m : m * (m - 1);
}

// Cells before step t of the wavefront over L lines of ny cells: the sum
// over s < t of min(L-1, floor(s/2)) - max(0, ceil((s-ny+1)/2)) + 1, in
// closed form (the step table of an L x ny grid, without a table).
YG_DEV int blk_soff(int t, int L, int ny) {
  const int a = t <= 2 * L ? half_sum(t) : half_sum(2 * L) + (t - 2 * L) * (L - 1);
  const int b = t > ny ? half_sum(t - ny) + (t - ny) : 0;
  return t + a - b;
}
// the same for skew 1 (line i takes column index t - i at step t): the sum
// over s < t of min(L-1, s) - max(0, s-ny+1) + 1
YG_DEV int blk_soff1(int t, int L, int ny) {
  const int a = t <= L ? t * (t - 1) / 2 : L * (L - 1) / 2 + (t - L) * (L - 1);
  const int b = t > ny ? (t - ny) * (t - ny + 1) / 2 : 0;
  return t + a - b;
}
// column of line i at step t
YG_DEV int col_of(const Ctx& c, int t, int i) { return c.rev ? c.ny - 1 - (t - i) : t - 2 * i; }

// Cell of a wavefront chunk in stage slot tid: the chunk's steps t0..t0+C-1
// hold at most C*maxcnt cells; slot tid takes step offset tid / maxcnt and
// line offset tid % maxcnt (thread g serves slots g and g + blockDim).  Returns false if that slot is
// past the step's last cell.  Gives the skewed index q, the natural index
// nat and the stage slot l = dt*maxcnt + k of the cell.
YG_DEV bool chunk_cell(const Ctx& c, int t0, int t1, int tid, int& q, int& nat, int& l) {
  const int dt = tid / c.maxcnt, k = tid - dt * c.maxcnt;
  const int t = t0 + dt;
  if (t >= t1) return false;
  const int cnt = yg_soff[t + 1] - yg_soff[t];
  if (k >= cnt) return false;
  const int i = glo_of(c, t) + k, j = col_of(c, t, i);
  q = yg_soff[t] + k;
  nat = i * c.ny + j;
  l = tid;
  return true;
}

// stage slot of this thread's line at step t of the chunk starting at t0,
// or -1 if its line has no cell at step t
YG_DEV int step_slot(const Ctx& c, int t, int t0) {
  const int jj = t - (c.rev ? 1 : 2) * c.line;   // the column's index in the line's order
  if (c.line < 0 || jj < 0 || jj >= c.ny) return -1;
  return (t - t0) * c.maxcnt + c.line - glo_of(c, t);
}

// ---- LDS staging of the wavefront sweeps ------------------------------------
// A wavefront sweep touches, at step t, the contiguous cells [soff[t],
// soff[t+1]) of the skewed order.  The steps are walked in chunks whose
// inputs fit one LDS buffer; while the lanes run the recurrence of one chunk
// out of LDS, every thread of the block already has the next chunk's global
// loads in flight (in registers), so the per-step cost is the recurrence and
// not a memory round trip.  Stage slot of a cell: (step - t0)*maxcnt + its
// line's offset in the step, computed without a table lookup.

#define YG_DYN_LDS (120 * 1024)

// Barrier of one wavefront step: waits for this wave's LDS traffic (the ring
// of boundary values, the staged outputs) and nothing else, so global loads
// and stores stay in flight across it.  The memory clobber keeps the
// compiler from moving LDS accesses across it.
YG_DEV void step_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }


template <typename T> YG_DEV T mask_to_stage(uint32_t m);
template <> YG_DEV float mask_to_stage<float>(uint32_t m) { return __uint_as_float(m); }
template <> YG_DEV double mask_to_stage<double>(uint32_t m) { return (double)m; }
template <typename T> YG_DEV uint32_t stage_to_mask(T v);
template <> YG_DEV uint32_t stage_to_mask<float>(float v) { return __float_as_uint(v); }
template <> YG_DEV uint32_t stage_to_mask<double>(double v) { return (uint32_t)v; }

// Neighbour lanes exchange values by DPP wave shifts (a VALU operation).
// bound_ctrl set: the lane without a source lane reads 0.
YG_DEV float dpp_shr1(float v) {   // lane i <- lane i-1, lane 0 <- 0
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xf, 0xf, true));
}
YG_DEV float dpp_shl1(float v) {   // lane i <- lane i+1, lane 63 <- 0
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xf, 0xf, true));
}
YG_DEV double dpp_shr1(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), 0x138, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x138, 0xf, 0xf, true);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
YG_DEV double dpp_shl1(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), 0x130, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x130, 0xf, 0xf, true);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// ---- ILU(0) factorization as a wavefront over grid lines -----------------
// The line of a thread handles cell (i, j = t - 2 i) at step t (reversed
// columns: j = ny - 1 - (t - i)).  The U rows of the two most recent cells of
// each lane stay in registers; a lane reads the line before it through wave
// shifts, and across waves (coupled scheme) through an LDS ring.
//
// Stencil slots of a cell's two rows (Ctx::mask bit s of each row):
//   ux row: 0 diag, 1 ux(i+1,j), 2 ux(i-1,j), 3 ux(i,j+1), 4 ux(i,j-1),
//           5 uy(i,j), 6 uy(i+1,j), 7 uy(i+1,j-1), 8 uy(i,j-1)
//   uy row: 0 diag, 1 uy(i,j+1), 2 uy(i,j-1), 3 uy(i+1,j), 4 uy(i-1,j),
//           5 ux(i,j), 6 ux(i,j+1), 7 ux(i-1,j+1), 8 ux(i-1,j)
template <typename CT>
struct URow {   // U part of the two rows of one cell (natural column order)
  CT x5, x3, x7, x1, x6, xd;  // ux row: slots 5,3,7,1,6, 1/diag
  CT y6, y1, y3, yd;          // uy row: slots 6,1,3, 1/diag
};
template <typename CT>
struct URowR {  // the same in the reversed column order
  CT x5, x4, x8, x1, x6, x7, xd;  // ux row: slots 5,4,8,1,6,7, 1/diag
  CT y2, y3, yd;                  // uy row: slots 2,3, 1/diag
};
template <typename TA, bool REV> struct URowOf { using type = URow<TA>; };
template <typename TA> struct URowOf<TA, true> { using type = URowR<TA>; };

template <typename CT>
YG_DEV CT shup(CT v) { return dpp_shr1(v); }

// cells of one stage buffer of the factorization (two buffers of 19 values
// per cell in the dynamic LDS, at most two cells per thread)
YG_DEV int factor_cap_of(int lds_bytes, int value_bytes) {
  return min(2 * (int)blockDim.x, (lds_bytes / value_bytes) / (2 * 19));
}
template <typename TA>
YG_DEV int factor_cap(const Ctx& c) { return factor_cap_of(c.lds, (int)sizeof(TA)); }

// Pivot reciprocal: exact division in FP64 (the reference's ILU(0)); the
// hardware reciprocal (1 ulp) for the FP32 factors of the mixed path's
// inner preconditioner, where it shortens each wavefront step.
YG_DEV double pivot_rcp(double v) { return 1.0 / v; }
YG_DEV float pivot_rcp(float v) { return __builtin_amdgcn_rcpf(v); }

// TA is the matrix and the factor type, and the arithmetic type of the
// elimination (FP64 path: double; the FP32 inner solve of the mixed path:
// float, whose factors are FP32 in any case).  REV: the reversed column order
// of the blocked scheme (Ctx::rev).  CPL: the coupled scheme (Ctx::coupled).
template <typename TA, bool REV, bool CPL>
YG_NOINLINE void ilu0_factor_t(const Ctx& cin, const TA* __restrict__ A, TA* __restrict__ F,
                               int* bad) {
  static_assert(!(REV && CPL), "the reversed order is a blocked-scheme option");
  const Ctx c = cin;   // fields in registers, not reloaded after stores
  using UR = typename URowOf<TA, REV>::type;
  PhaseTimer ptm(c, 0);
  const YG_GLOBAL TA* __restrict__ Ag = gc(A);
  YG_GLOBAL TA* __restrict__ Fg = gm(F);
  const int g = threadIdx.x, lane = g & 63, w = g >> 6;
  const bool ring_in = CPL && w > 0 && lane == 0;
  const bool ring_out = CPL && w + 1 < c.nwl && lane == 63;
  const size_t nc = c.ncell;
  constexpr int NV = 19;  // 18 slot values + the presence mask
  const int capc = factor_cap<TA>(c);
  const int C = max(1, capc / c.maxcnt);
  const int stride = C * c.maxcnt;
  // The chunk starting at t0 -> stage at offset sb: the threads s0, s0 + ds,
  // ... take its slots, two at a time, both slots' loads issued before the
  // first LDS store waits for them.
  constexpr int NL = sizeof(TA) == 4 ? 5 : 9;   // 16-byte chunks of a record's 18 values
  typedef TA V16 __attribute__((ext_vector_type(16 / sizeof(TA))));
  auto stage = [&](int sb, int t0, int s0, int ds) {
    const int t1s = min(c.T, t0 + C);
    for (int base = s0; base < stride; base += 2 * ds) {
      int q[2], nat[2], l[2];
      bool ok[2];
      V16 v[2][NL];
      uint32_t mk[2];
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        ok[k] = chunk_cell(c, t0, t1s, base + k * ds, q[k], nat[k], l[k]);
        if (ok[k]) {
          // the record's 18 slot values in 16-byte loads (records are 16-byte aligned)
          const YG_GLOBAL V16* rp = reinterpret_cast<const YG_GLOBAL V16*>(Ag + (size_t)nat[k] * YG_NA);
#pragma unroll
          for (int e = 0; e < NL; ++e) v[k][e] = rp[e];
          mk[k] = gc(c.mask)[nat[k]];
        }
      }
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        if (!ok[k]) continue;
        constexpr int E = 16 / sizeof(TA);
#pragma unroll
        for (int e = 0; e < NL; ++e)
#pragma unroll
          for (int u = 0; u < E; ++u)
            if (E * e + u < 18) lds<TA>(sb + (E * e + u) * stride + l[k]) = v[k][e][u];
        lds<TA>(sb + 18 * stride + l[k]) = mask_to_stage<TA>(mk[k]);
      }
    }
  };
  UR o1{}, o2{};  // own cells at steps t-1 and t-2
  const int nch = (c.T + C - 1) / C;
  // blocked scheme with waves that own no lines: those waves stage the next
  // chunk while the others eliminate this one (no barrier inside a chunk)
  const int nw = (int)(blockDim.x >> 6);
  const bool split = !CPL && c.nwl < nw;
  const bool stager = split && w >= c.nwl;
  stage(0, 0, g, (int)blockDim.x);
  __syncthreads();
  for (int ch = 0; ch < nch; ++ch) {
    const int t0 = ch * C, t1 = min(c.T, t0 + C);
    const int sb = (ch & 1) * NV * stride;
    if (stager) {
      if (ch + 1 < nch) stage(((ch + 1) & 1) * NV * stride, t1, g - c.nwl * 64, (nw - c.nwl) * 64);
      __syncthreads();
      continue;
    }
    for (int t = t0; t < t1; ++t) {
      // natural order: P = line i-1 cell j (its step t-2), Q = line i-1 cell
      // j+1 (its step t-1); reversed: P = line i-1 cell j (its step t-1), Q =
      // line i-1 cell j+1 (its step t-2)
      UR P{}, Q{};
      if constexpr (REV) {
        P.x5 = shup(o1.x5); P.x1 = shup(o1.x1); P.x6 = shup(o1.x6); P.x7 = shup(o1.x7);
        P.xd = shup(o1.xd); P.y3 = shup(o1.y3); P.yd = shup(o1.yd);
        Q.x4 = shup(o2.x4); Q.x8 = shup(o2.x8); Q.x1 = shup(o2.x1); Q.x6 = shup(o2.x6);
        Q.x7 = shup(o2.x7); Q.xd = shup(o2.xd);
      } else {
        P.x5 = shup(o2.x5); P.x3 = shup(o2.x3); P.x7 = shup(o2.x7); P.x1 = shup(o2.x1);
        P.x6 = shup(o2.x6); P.xd = shup(o2.xd); P.y6 = shup(o2.y6); P.y3 = shup(o2.y3);
        P.yd = shup(o2.yd);
        Q.x7 = shup(o1.x7); Q.x1 = shup(o1.x1); Q.x6 = shup(o1.x6); Q.xd = shup(o1.xd);
        if (CPL && ring_in) {
          const double* rp = yg_ring[w - 1][(t + 1) % 3];  // step t-2
          const double* rq = yg_ring[w - 1][(t + 2) % 3];  // step t-1
          P.x5 = rp[0]; P.x3 = rp[1]; P.x7 = rp[2]; P.x1 = rp[3]; P.x6 = rp[4]; P.xd = rp[5];
          P.y6 = rp[6]; P.y1 = rp[7]; P.y3 = rp[8]; P.yd = rp[9];
          Q.x5 = rq[0]; Q.x3 = rq[1]; Q.x7 = rq[2]; Q.x1 = rq[3]; Q.x6 = rq[4]; Q.xd = rq[5];
          Q.y6 = rq[6]; Q.y1 = rq[7]; Q.y3 = rq[8]; Q.yd = rq[9];
        }
      }
      const int l = step_slot(c, t, t0);
      UR X{};
      if (l >= 0) {
        const int j = col_of(c, t, c.line);
        const uint32_t m = stage_to_mask<TA>(lds<TA>(sb + 18 * stride + l));
        TA a[9], b[9];
#pragma unroll
        for (int s = 0; s < 9; ++s) {
          a[s] = lds<TA>(sb + (2 * s) * stride + l);
          b[s] = lds<TA>(sb + (2 * s + 1) * stride + l);
        }
        uint32_t mx = m & 0x1ffu, my = (m >> 16) & 0x1ffu;
        if (!CPL && c.ncol > 1) {
          // tile edges: couplings to the column before the tile's first
          // (ux row: ux(i,j-1), uy(i+1,j-1), uy(i,j-1); uy row: uy(i,j-1))
          // and after its last (ux row: ux(i,j+1); uy row: uy(i,j+1),
          // ux(i,j+1), ux(i-1,j+1)) are dropped
          const int jt = j % c.mcol;
          if (jt == 0 && j > 0) {
            a[4] = a[7] = a[8] = 0.0; b[2] = 0.0;
            mx &= ~((1u << 4) | (1u << 7) | (1u << 8)); my &= ~(1u << 2);
          }
          if (jt == c.mcol - 1 && j + 1 < c.ny) {
            a[3] = 0.0; b[1] = b[6] = b[7] = 0.0;
            mx &= ~(1u << 3); my &= ~((1u << 1) | (1u << 6) | (1u << 7));
          }
        }
        // an update of slot s only where the row holds slot s (ILU(0) keeps
        // the pattern): a select of the operand (or zero) rather than a branch
#define PX(s) ((mx >> (s)) & 1u)
#define PY(s) ((my >> (s)) & 1u)
        const UR& O = o1;
        TA lv;
        if constexpr (REV) {
          // ux row, lower entries in column order: ux(i-1,j), ux(i,j+1)
          lv = a[2] * P.xd; a[2] = lv;
          a[0] -= lv * (PX(0) ? P.x1 : TA(0));
          a[5] -= lv * (PX(5) ? P.x6 : TA(0));
          a[8] -= lv * (PX(8) ? P.x7 : TA(0));
          lv = a[3] * O.xd; a[3] = lv;
          a[0] -= lv * (PX(0) ? O.x4 : TA(0));
          a[5] -= lv * (PX(5) ? O.x8 : TA(0));
          a[6] -= lv * (PX(6) ? O.x7 : TA(0));
          X.xd = pivot_rcp(a[0]);
          X.x5 = a[5]; X.x4 = a[4]; X.x8 = a[8]; X.x1 = a[1]; X.x6 = a[6]; X.x7 = a[7];
          // uy row, lower entries: ux(i-1,j+1), ux(i-1,j), uy(i-1,j), ux(i,j+1),
          // uy(i,j+1), ux(i,j)
          lv = b[7] * Q.xd; b[7] = lv;
          b[8] -= lv * (PY(8) ? Q.x4 : TA(0));
          b[4] -= lv * (PY(4) ? Q.x8 : TA(0));
          b[6] -= lv * (PY(6) ? Q.x1 : TA(0));
          b[1] -= lv * (PY(1) ? Q.x6 : TA(0));
          b[0] -= lv * (PY(0) ? Q.x7 : TA(0));
          lv = b[8] * P.xd; b[8] = lv;
          b[4] -= lv * (PY(4) ? P.x5 : TA(0));
          b[5] -= lv * (PY(5) ? P.x1 : TA(0));
          b[0] -= lv * (PY(0) ? P.x6 : TA(0));
          b[2] -= lv * (PY(2) ? P.x7 : TA(0));
          lv = b[4] * P.yd; b[4] = lv;
          b[0] -= lv * (PY(0) ? P.y3 : TA(0));
          lv = b[6] * O.xd; b[6] = lv;
          b[1] -= lv * (PY(1) ? O.x5 : TA(0));
          b[5] -= lv * (PY(5) ? O.x4 : TA(0));
          b[0] -= lv * (PY(0) ? O.x8 : TA(0));
          b[3] -= lv * (PY(3) ? O.x7 : TA(0));
          lv = b[1] * O.yd; b[1] = lv;
          b[0] -= lv * (PY(0) ? O.y2 : TA(0));
          lv = b[5] * X.xd; b[5] = lv;
          b[0] -= lv * (PY(0) ? X.x5 : TA(0));
          b[2] -= lv * (PY(2) ? X.x8 : TA(0));
          b[3] -= lv * (PY(3) ? X.x6 : TA(0));
          X.yd = pivot_rcp(b[0]);
          X.y2 = b[2]; X.y3 = b[3];
        } else {
          // ux row, lower entries in column order: ux(i-1,j), ux(i,j-1), uy(i,j-1)
          lv = a[2] * P.xd; a[2] = lv;
          a[8] -= lv * (PX(8) ? P.x7 : TA(0));
          a[0] -= lv * (PX(0) ? P.x1 : TA(0));
          a[5] -= lv * (PX(5) ? P.x6 : TA(0));
          lv = a[4] * O.xd; a[4] = lv;
          a[8] -= lv * (PX(8) ? O.x5 : TA(0));
          a[0] -= lv * (PX(0) ? O.x3 : TA(0));
          a[7] -= lv * (PX(7) ? O.x6 : TA(0));
          lv = a[8] * O.yd; a[8] = lv;
          a[0] -= lv * (PX(0) ? O.y6 : TA(0));
          a[5] -= lv * (PX(5) ? O.y1 : TA(0));
          a[7] -= lv * (PX(7) ? O.y3 : TA(0));
          X.xd = pivot_rcp(a[0]);
          X.x5 = a[5]; X.x3 = a[3]; X.x7 = a[7]; X.x1 = a[1]; X.x6 = a[6];
          // uy row, lower entries: ux(i-1,j), uy(i-1,j), ux(i-1,j+1), uy(i,j-1), ux(i,j)
          lv = b[8] * P.xd; b[8] = lv;
          b[4] -= lv * (PY(4) ? P.x5 : TA(0));
          b[7] -= lv * (PY(7) ? P.x3 : TA(0));
          b[2] -= lv * (PY(2) ? P.x7 : TA(0));
          b[5] -= lv * (PY(5) ? P.x1 : TA(0));
          b[0] -= lv * (PY(0) ? P.x6 : TA(0));
          lv = b[4] * P.yd; b[4] = lv;
          b[7] -= lv * (PY(7) ? P.y6 : TA(0));
          b[0] -= lv * (PY(0) ? P.y3 : TA(0));
          lv = b[7] * Q.xd; b[7] = lv;
          b[0] -= lv * (PY(0) ? Q.x7 : TA(0));
          b[6] -= lv * (PY(6) ? Q.x1 : TA(0));
          b[1] -= lv * (PY(1) ? Q.x6 : TA(0));
          lv = b[2] * O.yd; b[2] = lv;
          b[5] -= lv * (PY(5) ? O.y6 : TA(0));
          b[0] -= lv * (PY(0) ? O.y1 : TA(0));
          lv = b[5] * X.xd; b[5] = lv;
          b[0] -= lv * (PY(0) ? X.x5 : TA(0));
          b[6] -= lv * (PY(6) ? X.x3 : TA(0));
          b[3] -= lv * (PY(3) ? X.x6 : TA(0));
          X.yd = pivot_rcp(b[0]);
          X.y6 = b[6]; X.y1 = b[1]; X.y3 = b[3];
        }
#undef PX
#undef PY
        if (!(a[0] != 0.0) || !(b[0] != 0.0)) *bad = 1;
        if constexpr (CPL) {
          const int q = yg_soff[t] + c.line - glo_of(c, t);
          Fg[0 * nc + q] = (TA)a[2]; Fg[1 * nc + q] = (TA)a[4]; Fg[2 * nc + q] = (TA)a[8];
          Fg[3 * nc + q] = (TA)b[8]; Fg[4 * nc + q] = (TA)b[4]; Fg[5 * nc + q] = (TA)b[7];
          Fg[6 * nc + q] = (TA)b[2]; Fg[7 * nc + q] = (TA)b[5];
          Fg[8 * nc + q] = (TA)X.x5; Fg[9 * nc + q] = (TA)X.x3; Fg[10 * nc + q] = (TA)X.x7;
          Fg[11 * nc + q] = (TA)X.x1; Fg[12 * nc + q] = (TA)X.x6; Fg[13 * nc + q] = (TA)X.xd;
          Fg[14 * nc + q] = (TA)X.y6; Fg[15 * nc + q] = (TA)X.y1; Fg[16 * nc + q] = (TA)X.y3;
          Fg[17 * nc + q] = (TA)X.yd;
        } else {
          // rotated tile order (the streamed sweeps): the tiles of a line
          // block follow each other; cell (k, jr) of a tile, jr its column's
          // index in the tile's order (jt, or M - 1 - jt reversed), at slot
          // ((jr + s k) mod M) L + k for skew s
          const int i0 = w * c.lpw, k = c.line - i0, Lb = min(c.lpw, c.nx - i0);
          const int jt = c.ncol > 1 ? j % c.mcol : j;
          const int j0 = j - jt, M = min(c.mcol, c.ny - j0);
          // tb mod M with tb = jr + s k < M + 2(Lb - 1): a few subtractions
          int rq = REV ? (M - 1 - jt) + k : jt + 2 * k;
          while (rq >= M) rq -= M;
          const size_t qb = (size_t)i0 * c.ny + (size_t)Lb * j0 + (size_t)rq * Lb + k;
          // 16-byte stores (the L and U records are 16-byte aligned)
          TA lv8[8], uv12[12];
          if constexpr (REV) {
            // L: ux row 2,3; uy row 7,8,4,6,1,5 (column order)
            // U: ux row 5,4,8,1,6, 1/diag, 7; uy row 2,3, 1/diag
            const TA l8[8] = {a[2], a[3], b[7], b[8], b[4], b[6], b[1], b[5]};
            const TA u12[12] = {X.x5, X.x4, X.x8, X.x1, X.x6, X.xd,
                                X.x7, X.y2, X.y3, X.yd, TA(0), TA(0)};
#pragma unroll
            for (int e = 0; e < 8; ++e) lv8[e] = l8[e];
#pragma unroll
            for (int e = 0; e < 12; ++e) uv12[e] = u12[e];
          } else {
            const TA l8[8] = {a[2], a[4], a[8], b[8], b[4], b[7], b[2], b[5]};
            const TA u12[12] = {X.x5, X.x3, X.x7, X.x1, X.x6, X.xd,
                                X.y6, X.y1, X.y3, X.yd, TA(0), TA(0)};
#pragma unroll
            for (int e = 0; e < 8; ++e) lv8[e] = l8[e];
#pragma unroll
            for (int e = 0; e < 12; ++e) uv12[e] = u12[e];
          }
          constexpr int E = 16 / sizeof(TA);
          typedef TA V __attribute__((ext_vector_type(E)));
          YG_GLOBAL V* fl = reinterpret_cast<YG_GLOBAL V*>(Fg + 8 * qb);
          YG_GLOBAL V* fu = reinterpret_cast<YG_GLOBAL V*>(Fg + c.fu + 12 * qb);
#pragma unroll
          for (int e = 0; e < 8 / E; ++e) {
            V v;
#pragma unroll
            for (int u = 0; u < E; ++u) v[u] = lv8[E * e + u];
            fl[e] = v;
          }
#pragma unroll
          for (int e = 0; e < 12 / E; ++e) {
            V v;
#pragma unroll
            for (int u = 0; u < E; ++u) v[u] = uv12[E * e + u];
            fu[e] = v;
          }
        }
      }
      o2 = o1;
      o1 = X;
      if constexpr (CPL) {
        if (ring_out) {
          double* rr = yg_ring[w][t % 3];
          rr[0] = X.x5; rr[1] = X.x3; rr[2] = X.x7; rr[3] = X.x1; rr[4] = X.x6; rr[5] = X.xd;
          rr[6] = X.y6; rr[7] = X.y1; rr[8] = X.y3; rr[9] = X.yd;
        }
        step_barrier();
      }
    }
    if (split) {   // the stagers have loaded the next chunk meanwhile
      __syncthreads();
      continue;
    }
    // every wave owns lines: the next chunk is loaded after the steps, which
    // keeps its 19 values per cell out of the registers the recurrence needs
    __syncthreads();
    if (ch + 1 < nch) stage(((ch + 1) & 1) * NV * stride, t1, g, (int)blockDim.x);
    __syncthreads();
  }
}

template <typename TA>
YG_DEV void ilu0_factor(const Ctx& c, const TA* __restrict__ A, TA* __restrict__ F, int* bad) {
  if (c.coupled) ilu0_factor_t<TA, false, true>(c, A, F, bad);
  else if (c.rev) ilu0_factor_t<TA, true, false>(c, A, F, bad);
  else ilu0_factor_t<TA, false, false>(c, A, F, bad);
}

// ---- triangular sweeps ---------------------------------------------------
// Forward: y = L^{-1} rhs; backward: z = U^{-1} y.  T is the arithmetic type
// of the sweep (the factor type).  The steps are walked in chunks of C
// steps.  Per chunk: every thread loads one cell's factor slots and right-
// hand side into registers for the next chunk while the lanes run the
// recurrence of the current chunk entirely in LDS (inputs staged, outputs
// collected); the chunk's outputs then leave in one coalesced store.  Inside
// a chunk each lane reads the next step's slots from LDS before it runs the
// current step, so the LDS latency overlaps the recurrence.  No global
// memory operation sits on the per-step critical path.

template <typename T, int NV>
YG_DEV int sweep_cap_cells() {
  const int c = (int)(YG_DYN_LDS / sizeof(T)) / (2 * NV + 2);
  return c < 2 * (int)blockDim.x ? c : 2 * (int)blockDim.x;
}

template <typename T>
YG_NOINLINE void sweep_L(const Ctx& cin, const T* __restrict__ F, const T* __restrict__ rhs,
                    T* __restrict__ y) {
  const Ctx c = cin;   // fields in registers, not reloaded after stores
  PhaseTimer ptm(c, 1);
  const YG_GLOBAL T* __restrict__ Fg = gc(F);
  const YG_GLOBAL T* __restrict__ rg = gc(rhs);
  YG_GLOBAL T* __restrict__ yg = gm(y);
  constexpr int NV = 10;  // L slots of both rows + the right-hand side
  const int g = threadIdx.x, lane = g & 63, w = g >> 6;
  const bool ring_in = c.coupled && w > 0 && lane == 0;
  const bool ring_out = c.coupled && w + 1 < c.nwl && lane == 63;
  const size_t nc = c.ncell;
  const int capc = sweep_cap_cells<T, NV>();
  const int C = max(1, capc / c.maxcnt);
  const int stride = C * c.maxcnt;
  const int ob = 2 * stride * NV;   // output slots follow the two stage buffers
  T reg[2][NV];
  int lreg[2];
  auto load = [&](int t0) {   // this thread's cells of the chunk at step t0 -> registers
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      int q, nat, l;
      lreg[r] = -1;
      if (chunk_cell(c, t0, min(c.T, t0 + C), g + r * (int)blockDim.x, q, nat, l)) {
#pragma unroll
        for (int v = 0; v < 8; ++v) reg[r][v] = Fg[(size_t)v * nc + q];
        reg[r][8] = rg[2 * nat];
        reg[r][9] = rg[2 * nat + 1];
        lreg[r] = l;
      }
    }
  };
  auto stash = [&](int sb) {
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      if (lreg[r] >= 0) {
#pragma unroll
        for (int v = 0; v < NV; ++v) lds<T>(sb + v * stride + lreg[r]) = reg[r][v];
      }
    }
  };
  auto scatter = [&](YG_GLOBAL T* out, int t0, int t1) {   // staged outputs -> natural order
    for (int r = 0; r < 2; ++r) {
      int q, nat, l;
      if (chunk_cell(c, t0, t1, g + r * (int)blockDim.x, q, nat, l)) {
        out[2 * nat] = lds<T>(ob + 2 * l);
        out[2 * nat + 1] = lds<T>(ob + 2 * l + 1);
      }
    }
  };
  T o1x = 0, o1y = 0, o2x = 0, o2y = 0;
  load(0);
  stash(0);
  __syncthreads();
  const int nch = (c.T + C - 1) / C;
  for (int ch = 0; ch < nch; ++ch) {
    const int t0 = ch * C, t1 = min(c.T, t0 + C);
    if (ch + 1 < nch) load(t1);
    const int sb = (ch & 1) * NV * stride;
    T bc[NV];
    int lc = step_slot(c, t0, t0);
#pragma unroll
    for (int v = 0; v < NV; ++v) bc[v] = lc >= 0 ? lds<T>(sb + v * stride + lc) : (T)0;
    for (int t = t0; t < t1; ++t) {
      const int ln = t + 1 < t1 ? step_slot(c, t + 1, t0) : -1;
      T bn[NV];
#pragma unroll
      for (int v = 0; v < NV; ++v) bn[v] = ln >= 0 ? lds<T>(sb + v * stride + ln) : (T)0;
      T n1x = dpp_shr1(o1x), n2x = dpp_shr1(o2x), n2y = dpp_shr1(o2y);
      if (ring_in) {
        n1x = (T)yg_ring[w - 1][(t + 2) % 3][0];
        n2x = (T)yg_ring[w - 1][(t + 1) % 3][0];
        n2y = (T)yg_ring[w - 1][(t + 1) % 3][1];
      }
      T yx = 0, yy = 0;
      if (lc >= 0) {
        yx = bc[8] - bc[0] * n2x - bc[1] * o1x - bc[2] * o1y;
        yy = bc[9] - bc[3] * n2x - bc[4] * n2y - bc[5] * n1x - bc[6] * o1y - bc[7] * yx;
        lds<T>(ob + 2 * lc) = yx;
        lds<T>(ob + 2 * lc + 1) = yy;
      }
      o2x = o1x; o2y = o1y; o1x = yx; o1y = yy;
      if (c.coupled) {
        if (ring_out) {
          yg_ring[w][t % 3][0] = (double)yx;
          yg_ring[w][t % 3][1] = (double)yy;
        }
        step_barrier();
      }
      lc = ln;
#pragma unroll
      for (int v = 0; v < NV; ++v) bc[v] = bn[v];
    }
    __syncthreads();
    if (ch + 1 < nch) stash(((ch + 1) & 1) * NV * stride);
    scatter(yg, t0, t1);
    __syncthreads();
  }
}

template <typename T>
YG_NOINLINE void sweep_U(const Ctx& cin, const T* __restrict__ F, const T* __restrict__ y,
                    T* __restrict__ z) {
  const Ctx c = cin;   // fields in registers, not reloaded after stores
  PhaseTimer ptm(c, 2);
  const YG_GLOBAL T* __restrict__ Fg = gc(F);
  const YG_GLOBAL T* __restrict__ yg = gc(y);
  YG_GLOBAL T* __restrict__ zg = gm(z);
  constexpr int NV = 12;  // U slots and 1/diag of both rows + y
  const int g = threadIdx.x, lane = g & 63, w = g >> 6;
  const bool ring_in = c.coupled && w + 1 < c.nwl && lane == 63;
  const bool ring_out = c.coupled && w > 0 && lane == 0;
  const size_t nc = c.ncell;
  const int capc = sweep_cap_cells<T, NV>();
  const int C = max(1, capc / c.maxcnt);
  const int stride = C * c.maxcnt;
  const int ob = 2 * stride * NV;   // output slots follow the two stage buffers
  T reg[2][NV];
  int lreg[2];
  const int nch = (c.T + C - 1) / C;
  auto load = [&](int ch) {   // this thread's cells of chunk ch -> registers
    const int t0 = ch * C;
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      int q, nat, l;
      lreg[r] = -1;
      if (chunk_cell(c, t0, min(c.T, t0 + C), g + r * (int)blockDim.x, q, nat, l)) {
#pragma unroll
        for (int v = 0; v < 10; ++v) reg[r][v] = Fg[(size_t)(8 + v) * nc + q];
        reg[r][10] = yg[2 * nat];
        reg[r][11] = yg[2 * nat + 1];
        lreg[r] = l;
      }
    }
  };
  auto stash = [&](int sb) {
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      if (lreg[r] >= 0) {
#pragma unroll
        for (int v = 0; v < NV; ++v) lds<T>(sb + v * stride + lreg[r]) = reg[r][v];
      }
    }
  };
  auto scatter = [&](YG_GLOBAL T* out, int t0, int t1) {   // staged outputs -> natural order
    for (int r = 0; r < 2; ++r) {
      int q, nat, l;
      if (chunk_cell(c, t0, t1, g + r * (int)blockDim.x, q, nat, l)) {
        out[2 * nat] = lds<T>(ob + 2 * l);
        out[2 * nat + 1] = lds<T>(ob + 2 * l + 1);
      }
    }
  };
  T o1x = 0, o1y = 0, o2x = 0, o2y = 0;  // own cells j+1 (step u-1), j+2 (u-2)
  load(nch - 1);
  stash(0);
  __syncthreads();
  for (int k = 0; k < nch; ++k) {
    const int ch = nch - 1 - k;
    const int t0 = ch * C, t1 = min(c.T, t0 + C);
    if (ch > 0) load(ch - 1);
    const int sb = (k & 1) * NV * stride;
    T bc[NV];
    int lc = step_slot(c, t1 - 1, t0);
#pragma unroll
    for (int v = 0; v < NV; ++v) bc[v] = lc >= 0 ? lds<T>(sb + v * stride + lc) : (T)0;
    for (int t = t1 - 1; t >= t0; --t) {
      const int u = c.T - 1 - t;
      const int ln = t - 1 >= t0 ? step_slot(c, t - 1, t0) : -1;
      T bn[NV];
#pragma unroll
      for (int v = 0; v < NV; ++v) bn[v] = ln >= 0 ? lds<T>(sb + v * stride + ln) : (T)0;
      // line i+1: cell j-1 (its step u-1) and cell j (its step u-2)
      T n1y = dpp_shl1(o1y), n2x = dpp_shl1(o2x), n2y = dpp_shl1(o2y);
      if (ring_in) {
        n1y = (T)yg_ring[w + 1][(u + 2) % 3][1];
        n2x = (T)yg_ring[w + 1][(u + 1) % 3][0];
        n2y = (T)yg_ring[w + 1][(u + 1) % 3][1];
      }
      T zx = 0, zy = 0;
      if (lc >= 0) {
        zy = (bc[11] - bc[6] * o1x - bc[7] * o1y - bc[8] * n2y) * bc[9];
        zx = (bc[10] - bc[0] * zy - bc[1] * o1x - bc[2] * n1y - bc[3] * n2x - bc[4] * n2y) *
             bc[5];
        lds<T>(ob + 2 * lc) = zx;
        lds<T>(ob + 2 * lc + 1) = zy;
      }
      o2x = o1x; o2y = o1y; o1x = zx; o1y = zy;
      if (c.coupled) {
        if (ring_out) {
          yg_ring[w][u % 3][0] = (double)zx;
          yg_ring[w][u % 3][1] = (double)zy;
        }
        step_barrier();
      }
      lc = ln;
#pragma unroll
      for (int v = 0; v < NV; ++v) bc[v] = bn[v];
    }
    __syncthreads();
    if (ch > 0) stash(((k + 1) & 1) * NV * stride);
    scatter(zg, t0, t1);
    __syncthreads();
  }
}

// ---- streamed triangular sweeps (line blocks of at most 64 lines) ---------
// Each wave sweeps its own line block over the block's own wavefront
// (T_b = ny + 2(L-1) steps) and never waits for another wave.  At each
// step a lane loads its cell's factor values as 16-byte chunks of the
// rotated AoS layout (below; one step of a tile is one contiguous range, so
// the loads of a wave are coalesced) into a ring of D register slots filled
// D steps ahead of use, so the memory latency overlaps D steps of the
// recurrence.  The forward sweep writes y in the same rotated order (read
// back coalesced by the backward sweep); the right-hand side and the result
// are read and written in the natural row order of the Krylov vectors.

// ---- bounds-checked buffer access ------------------------------------------
// The streamed sweeps read and write through buffer resources (base, byte
// range): a lane without a cell at a step passes an offset past the range,
// which the hardware turns into a zero load or a dropped store.  Every step
// then issues the same memory instructions on every lane with no branch
// around them, so the compiler counts the loads in flight exactly and the
// loads issued D steps ahead are waited for only when their step comes.
#define YG_RSRC_WORD3 0x00020000   // raw buffer, 32-bit data format (gfx9)
#define YG_OOB 0x80000000u

// A buffer resource must sit in scalar registers: the base address and the
// range are made wave-uniform first (pointer arguments of a called device
// function arrive in vector registers).
template <typename T>
YG_DEV __amdgpu_buffer_rsrc_t yg_rsrc(const T* base, size_t count) {
  const unsigned long long a = (unsigned long long)base;
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)a);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(a >> 32));
  void* u = (void*)(((unsigned long long)hi << 32) | lo);
  const int bytes = __builtin_amdgcn_readfirstlane((int)(count * sizeof(T)));
  return __builtin_amdgcn_make_buffer_rsrc(u, (short)0, bytes, YG_RSRC_WORD3);
}

template <typename T> struct Pair16;   // T pair (8 or 16 bytes) through buffer loads
// load<true>: a device-scope load (cache policy SC1), which skips the L1:
// data another workgroup of a member group wrote.
template <> struct Pair16<float> {
  template <bool COH = false>
  static YG_DEV float2 load(__amdgpu_buffer_rsrc_t r, unsigned off) {
    return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, COH ? 16 : 0));
  }
  static YG_DEV void store(float2 v, __amdgpu_buffer_rsrc_t r, unsigned off) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(__attribute__((ext_vector_type(2))) unsigned, v), r, off, 0, 0);
  }
};
template <> struct Pair16<double> {
  template <bool COH = false>
  static YG_DEV double2 load(__amdgpu_buffer_rsrc_t r, unsigned off) {
    return __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, COH ? 16 : 0));
  }
  static YG_DEV void store(double2 v, __amdgpu_buffer_rsrc_t r, unsigned off) {
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(__attribute__((ext_vector_type(4))) unsigned, v), r, off, 0, 0);
  }
};

// N values of T from consecutive 16-byte chunks at byte offset off
template <int N>
YG_DEV void load_vals(__amdgpu_buffer_rsrc_t r, unsigned off, float (&v)[N]) {
#pragma unroll
  for (int e = 0; e < N / 4; ++e) {
    const float4 x = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, off + 16 * e, 0, 0));
    v[4 * e] = x.x; v[4 * e + 1] = x.y; v[4 * e + 2] = x.z; v[4 * e + 3] = x.w;
  }
}
template <int N>
YG_DEV void load_vals(__amdgpu_buffer_rsrc_t r, unsigned off, double (&v)[N]) {
#pragma unroll
  for (int e = 0; e < N / 2; ++e) {
    const double2 x = __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, off + 16 * e, 0, 0));
    v[2 * e] = x.x; v[2 * e + 1] = x.y;
  }
}

// BF16 storage of the line-block factors (mixed mode, option factor16).
// Three arrays over the cells, in the same rotated tile order as the FP32
// factors, so that each sweep load is one contiguous range across the lanes:
//   L part: 8 BF16 values (F0..F7), 16 bytes at FB + 16 q bytes;
//   U part: 8 BF16 values of the off-diagonal U entries already multiplied by
//     their row's inverse diagonal (ux row: slots 5,3,7,1,6; uy row: 6,1,3),
//     16 bytes at FB + 16 ncell + 16 q bytes;
//   the two inverse diagonals in FP32, 8 bytes at FB + 32 ncell + 8 q bytes.
// The inverse diagonal is the row scale, so a U row is
// x = inv_d y - sum (inv_d U) x_j.
YG_DEV uint16_t to_bf16(float v) {
  const uint32_t b = __builtin_bit_cast(uint32_t, v);
  return (uint16_t)((b + 0x7FFFu + ((b >> 16) & 1u)) >> 16);   // nearest, ties to even
}
YG_DEV float from_bf16(uint32_t h16) { return __builtin_bit_cast(float, h16 << 16); }

// The sweeps keep prefetched BF16 words packed in registers and widen them
// only at their step.  Widened at load time, each value would be an
// arithmetic result of a load still in flight; the compiler then moves the
// widening of every prefetched step to the top of the loop (an operation on
// all incoming values of a loop-carried value becomes one operation after
// it), where it waits for all of them, and the prefetch depth collapses.
YG_DEV uint4 load_raw16(__amdgpu_buffer_rsrc_t r, unsigned off) {
  return __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
}
YG_DEV float bf_lo(uint32_t w) { return __builtin_bit_cast(float, w << 16); }
YG_DEV float bf_hi(uint32_t w) { return __builtin_bit_cast(float, w & 0xFFFF0000u); }

// Rotated tile order of the line-block factors.  A tile has L lines (one
// lane each) and M columns and is swept in Tb = M + s(L-1) wavefront steps
// (skew s = 2, or 1 for reversed columns): line k handles the column of
// index jr = t - s k in the line's order at step t (jr = j natural, M - 1 - j
// reversed).  Cell (k, jr) is stored at slot ((jr + s k) mod M) L + k of the
// tile's range.  For each line, jr -> (jr + s k) mod M is a rotation, so the
// tile's M x L slots hold its L x M cells once each, and the cells of step t
// are the contiguous slots (t mod M) L + [0, L): the loads of a wave are
// coalesced, and the next step's slots are one add and a wrap away.  Tiles
// follow each other by line block, then by column.
//
// A lane without a cell at step t (j outside [0, M), or a lane past the
// tile's lines) passes an out-of-range offset, which loads zero factors and a
// zero right-hand side, so its result is an exact zero (the value its
// neighbours must see for a cell that does not exist) without a select, and
// its store is dropped.
YG_DEV int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

struct SweepTile {
  int L, M, Tb, LM;   // lines, columns, steps, cells (wave-uniform)
  int base;           // the tile's first slot in the member's factor order
  int tl;             // -s lane, or far out of range for a lane past the lines
  int rev;            // reversed columns (skew 1)
  // the natural cell of this lane's line at step t: lane_n + t (natural
  // order), or lane_n - 64 + (Tb - 1 - t) (reversed order, whose lines run
  // backwards; see nat_rsrc)
  int lane_n;
};
// the wave's tile: line block wb, columns [j0, j0 + M); false for a wave without one
YG_DEV bool sweep_tile(const Ctx& c, SweepTile& g) {
  const int lane = threadIdx.x & 63, w = uni(threadIdx.x >> 6);
  const int nwl = uni(c.nwl), ny = uni(c.ny), nx = uni(c.nx), lpw = uni(c.lpw);
  const int ncol = uni(c.ncol), mcol = uni(c.mcol);
  if (w >= nwl * ncol) return false;
  const int wb = w / ncol, j0 = (w - wb * ncol) * mcol, i0 = wb * lpw;
  const int rev = uni(c.rev), sk = rev ? 1 : 2;
  g.L = min(lpw, nx - i0);
  g.M = min(mcol, ny - j0);
  g.Tb = g.M + sk * (g.L - 1);
  g.LM = g.L * g.M;
  g.base = i0 * ny + g.L * j0;
  g.tl = lane < g.L ? -sk * lane : -(1 << 30);
  g.rev = rev;
  // reversed: column j0 + M - 1 - (t - k) of line i0 + k; at least 65 - L >= 1
  g.lane_n = rev ? (i0 + lane) * ny + j0 + g.M - 1 + lane - (g.Tb - 1) + 64
                 : (i0 + lane) * ny + j0 - 2 * lane;
  return true;
}
// does this lane have a cell at step t
YG_DEV bool on_step(const SweepTile& g, int t) { return (unsigned)(t + g.tl) < (unsigned)g.M; }
// byte offset of this lane's cell at step t in the tile's rotated order
// (sq = (t mod M) L), or out of range.  A slot index is below 2^24, so a
// cell size that is not a power of two takes a full-rate 24-bit multiply.
YG_DEV unsigned rot_off(const SweepTile& g, int t, int sq, unsigned bytes) {
  const unsigned q = (unsigned)(sq + (int)(threadIdx.x & 63));
  const unsigned o = (bytes & (bytes - 1)) ? __umul24(q, bytes) : q * bytes;
  return on_step(g, t) ? o : YG_OOB;
}
// The step's part of a natural-order offset (cells; the scalar offset of the
// buffer instructions): t, or the steps left in the reversed order.
template <bool REV>
YG_DEV int nat_step(const SweepTile& g, int t) { return REV ? g.Tb - 1 - t : t; }
// A natural-order vector (2 values per cell, nc cells) as a buffer resource
// for nat_step offsets.  In the reversed order a lane's fixed offset is taken
// from 64 cells before the vector (the lines run backwards, and the scalar
// offset can only add), so the resource starts there; a lane with a cell at
// a step always lands inside the vector.
template <typename T>
YG_DEV __amdgpu_buffer_rsrc_t nat_rsrc(const SweepTile& g, const T* v, int nc) {
  const size_t pad = g.rev ? 2 * 64 : 0;
  const T* b = reinterpret_cast<const T*>(reinterpret_cast<uintptr_t>(v) - pad * sizeof(T));
  return yg_rsrc(b, 2 * (size_t)nc + pad);
}
// (t mod M) L of the next and of the previous step
YG_DEV int rot_next(const SweepTile& g, int sq) { sq += g.L; return sq == g.LM ? 0 : sq; }
YG_DEV int rot_prev(const SweepTile& g, int sq) { return (sq == 0 ? g.LM : sq) - g.L; }

// Every sweep keeps D steps in flight in a ring of register slots: slot s is
// refilled right after its step is computed, and a sched_barrier after each
// step keeps the steps in program order, so the wait before a step covers
// exactly that step's loads.  The store offset of a step is computed with its
// loads and kept in the slot.
//
// Addressing: a lane's offset inside a step is fixed (lane x cell bytes in
// the rotated order, its line's natural cell in the natural order), and the
// step's own offset is wave-uniform ((t mod M) L or t cells), so it goes in
// the scalar offset of the buffer instruction; per step a lane only selects
// its fixed offset or the out-of-range one.  A lane on no cell keeps the
// out-of-range vector offset whatever the scalar offset adds (below 2^31), so
// its loads read zeros and its store is dropped.
//
// Wave shifts: the line before (forward) or after (backward) is read at the
// previous step and at the one before.  The shift of a value two steps back
// is the shift taken one step back, so each step shifts only its two newest
// values and keeps them for the next step.
//
// One wave runs a step's recurrence alone, in order, so the chain of
// dependent operations from the previous step's results sets the step's
// time: the terms of values two steps back are summed first, and the
// previous step's results enter last.

template <int N>
YG_DEV void load_vals_so(__amdgpu_buffer_rsrc_t r, unsigned off, int soff, float (&v)[N]) {
#pragma unroll
  for (int e = 0; e < N / 4; ++e) {
    const float4 x = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, off + 16 * e, soff, 0));
    v[4 * e] = x.x; v[4 * e + 1] = x.y; v[4 * e + 2] = x.z; v[4 * e + 3] = x.w;
  }
}
template <int N>
YG_DEV void load_vals_so(__amdgpu_buffer_rsrc_t r, unsigned off, int soff, double (&v)[N]) {
#pragma unroll
  for (int e = 0; e < N / 2; ++e) {
    const double2 x = __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, off + 16 * e, soff, 0));
    v[2 * e] = x.x; v[2 * e + 1] = x.y;
  }
}
YG_DEV float2 pair_load_so(__amdgpu_buffer_rsrc_t r, unsigned off, int soff, float) {
  return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, off, soff, 0));
}
YG_DEV double2 pair_load_so(__amdgpu_buffer_rsrc_t r, unsigned off, int soff, double) {
  return __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, off, soff, 0));
}
YG_DEV void pair_store_so(float2 v, __amdgpu_buffer_rsrc_t r, unsigned off, int soff) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(__attribute__((ext_vector_type(2))) unsigned, v), r, off, soff, 0);
}
YG_DEV void pair_store_so(double2 v, __amdgpu_buffer_rsrc_t r, unsigned off, int soff) {
  __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(__attribute__((ext_vector_type(4))) unsigned, v), r, off, soff, 0);
}
YG_DEV uint4 load_raw16_so(__amdgpu_buffer_rsrc_t r, unsigned off, int soff) {
  return __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(r, off, soff, 0));
}

// REV: the reversed column order (factor values in its own slot order)
template <typename T, bool REV>
YG_NOINLINE void sweep_L_str(const Ctx& cin, const T* __restrict__ F, const T* __restrict__ rhs,
                             T* __restrict__ ys) {
  const Ctx c = cin;   // fields in registers, not reloaded after stores
  PhaseTimer ptm(c, 1);
  constexpr int D = sizeof(T) == 4 ? 12 : 6;   // register slots (steps in flight)
  constexpr int SF = 8 * sizeof(T), SV = 2 * sizeof(T);   // bytes per cell
  using V2 = typename Vec2<T>::type;
  SweepTile g;
  if (!sweep_tile(c, g)) return;
  const int nc = uni(c.ncell);
  const __amdgpu_buffer_rsrc_t rF = yg_rsrc(F + (size_t)8 * g.base, (size_t)8 * g.LM);
  const __amdgpu_buffer_rsrc_t rR = nat_rsrc(g, rhs, nc);
  const __amdgpu_buffer_rsrc_t rY = yg_rsrc(ys + (size_t)2 * g.base, (size_t)2 * g.LM);
  const unsigned lane = threadIdx.x & 63;
  const unsigned voF = lane * SF, voY = lane * SV, voN = (unsigned)g.lane_n * SV;
  T f[D][8];
  V2 r[D];
  unsigned so[D];
  int ss[D];
  int sq = 0;   // (t mod M) L of the step loaded next
  auto issue = [&](int s, int t) {
    const bool v = on_step(g, t);
    load_vals_so<8>(rF, v ? voF : YG_OOB, sq * SF, f[s]);
    r[s] = pair_load_so(rR, v ? voN : YG_OOB, nat_step<REV>(g, t) * SV, T());
    so[s] = v ? voY : YG_OOB;
    ss[s] = sq * SV;
    sq = rot_next(g, sq);
  };
#pragma unroll
  for (int s = 0; s < D; ++s) {
    issue(s, s);
    __builtin_amdgcn_sched_barrier(0);
  }
  // o1: this lane's results one step back; n1, p1: the line before, one and
  // two steps back
  T o1x = 0, o1y = 0, p1x = 0, p1y = 0;
  for (int t0 = 0; t0 < g.Tb; t0 += D) {
#pragma unroll
    for (int s = 0; s < D; ++s) {
      const T n1x = dpp_shr1(o1x), n1y = dpp_shr1(o1y);
      T yx, yy;
      if constexpr (REV) {
        // ux row: ux(i-1,j) = n1, ux(i,j+1) = o1; uy row: ux(i-1,j+1) = p1,
        // ux(i-1,j), uy(i-1,j) = n1, ux(i,j+1), uy(i,j+1) = o1, ux(i,j)
        const T ax = r[s].x - f[s][1] * o1x;
        const T ay = r[s].y - f[s][2] * p1x - f[s][5] * o1x - f[s][6] * o1y;
        yx = ax - f[s][0] * n1x;
        yy = ay - f[s][3] * n1x - f[s][4] * n1y - f[s][7] * yx;
      } else {
        const T ax = r[s].x - f[s][0] * p1x, ay = r[s].y - f[s][3] * p1x - f[s][4] * p1y;
        yx = ax - f[s][1] * o1x - f[s][2] * o1y;
        yy = ay - f[s][6] * o1y - f[s][5] * n1x - f[s][7] * yx;
      }
      V2 o;
      o.x = yx;
      o.y = yy;
      pair_store_so(o, rY, so[s], ss[s]);
      p1x = n1x; p1y = n1y; o1x = yx; o1y = yy;
      issue(s, t0 + s + D);
      __builtin_amdgcn_sched_barrier(0);
    }
  }
}

// The backward sweeps write z in the natural order, either straight to z
// (ZL false: a step's cells lie on different lines, one cache line each) or
// (ZL true) into the dynamic LDS, cell q of the block's part at value pair q,
// which the block then copies to z with coalesced stores (zl_copy_out).  A
// lane without a cell at a step writes a pair of its own past the part.
// Staging needs (cells + 64) value pairs of LDS (zl_fits).
template <typename T>
YG_DEV bool zl_fits(const Ctx& c) { return (size_t)(c.ncell + 64) * 2 * sizeof(T) <= (size_t)c.lds; }
template <typename T>
YG_DEV void zl_copy_out(const Ctx& c, T* __restrict__ z) {
  using V2 = typename Vec2<T>::type;
  V2* __restrict__ zv = reinterpret_cast<V2*>(z);
  for (int q = threadIdx.x; q < c.ncell; q += blockDim.x) zv[q] = lds<V2>(q);
}

template <typename T, bool REV, bool ZL>
YG_NOINLINE void sweep_U_str(const Ctx& cin, const T* __restrict__ F, const T* __restrict__ ys,
                             T* __restrict__ z) {
  const Ctx c = cin;   // fields in registers, not reloaded after stores
  PhaseTimer ptm(c, 2);
  constexpr int D = sizeof(T) == 4 ? 10 : 6;
  constexpr int SF = 12 * sizeof(T), SV = 2 * sizeof(T);
  using V2 = typename Vec2<T>::type;
  SweepTile g;
  if (!sweep_tile(c, g)) return;
  const int nc = uni(c.ncell);
  const __amdgpu_buffer_rsrc_t rF = yg_rsrc(F + c.fu + (size_t)12 * g.base, (size_t)12 * g.LM);
  const __amdgpu_buffer_rsrc_t rY = yg_rsrc(ys + (size_t)2 * g.base, (size_t)2 * g.LM);
  const __amdgpu_buffer_rsrc_t rZ = nat_rsrc(g, z, nc);
  const unsigned lane = threadIdx.x & 63;
  const unsigned voF = lane * SF, voY = lane * SV, voN = (unsigned)g.lane_n * SV;
  // LDS byte address of this lane's cell: zl0 + t SV (natural) or zl0 - t SV
  // (reversed); its own pair past the part at a step without a cell
  const unsigned zl0 = REV ? (unsigned)(g.lane_n - 64 + g.Tb - 1) * SV : voN;
  const unsigned zlx = (unsigned)(nc + (int)lane) * SV;
  T f[D][12];
  V2 yv[D];
  unsigned so[D];
  int ss[D];
  const int tz = g.Tb - 1;   // the last step; the sweep runs t = tz - u
  int sq = (tz % g.M) * g.L;
  auto issue = [&](int s, int t) {
    const bool v = on_step(g, t);
    load_vals_so<12>(rF, v ? voF : YG_OOB, sq * SF, f[s]);
    yv[s] = pair_load_so(rY, v ? voY : YG_OOB, sq * SV, T());
    if constexpr (ZL) {
      // a bit mask rather than a select: the compiler turns the select into
      // a branch around the step's loads
      const unsigned za = REV ? zl0 - (unsigned)t * SV : zl0 + (unsigned)t * SV;
      const unsigned m = 0u - (unsigned)v;
      so[s] = (za & m) | (zlx & ~m);
    } else {
      so[s] = v ? voN : YG_OOB;
      ss[s] = nat_step<REV>(g, t) * SV;
    }
    sq = rot_prev(g, sq);
  };
#pragma unroll
  for (int s = 0; s < D; ++s) {
    issue(s, tz - s);
    __builtin_amdgcn_sched_barrier(0);
  }
  // o1: own results one step back; n1, p1: the line after, one and two steps back
  T o1x = 0, o1y = 0, p1x = 0, p1y = 0;
  for (int u0 = 0; u0 < g.Tb; u0 += D) {
#pragma unroll
    for (int s = 0; s < D; ++s) {
      const T n1y = dpp_shl1(o1y), n1x = dpp_shl1(o1x);
      T zx, zy;
      if constexpr (REV) {
        // ux row: uy(i,j), ux(i,j-1), uy(i,j-1) = o1, ux(i+1,j), uy(i+1,j) = n1,
        // uy(i+1,j-1) = p1; uy row: uy(i,j-1) = o1, uy(i+1,j) = n1
        const T by = yv[s].y - f[s][7] * o1y;
        const T bx = yv[s].x - f[s][6] * p1y - f[s][1] * o1x - f[s][2] * o1y;
        zy = (by - f[s][8] * n1y) * f[s][9];
        zx = (bx - f[s][3] * n1x - f[s][4] * n1y - f[s][0] * zy) * f[s][5];
      } else {
        // line i+1: cell j-1 (one step back) and cell j (two steps back)
        const T by = yv[s].y - f[s][8] * p1y, bx = yv[s].x - f[s][3] * p1x - f[s][4] * p1y;
        zy = (by - f[s][6] * o1x - f[s][7] * o1y) * f[s][9];
        zx = (bx - f[s][1] * o1x - f[s][2] * n1y - f[s][0] * zy) * f[s][5];
      }
      V2 o;
      o.x = zx;
      o.y = zy;
      if constexpr (ZL) *reinterpret_cast<V2*>(yg_dyn_lds + so[s]) = o;
      else pair_store_so(o, rZ, so[s], ss[s]);
      p1x = n1x; p1y = n1y; o1x = zx; o1y = zy;
      issue(s, tz - (u0 + s + D));
      __builtin_amdgcn_sched_barrier(0);
    }
  }
}

// The streamed sweeps over BF16 factors (FP32 arithmetic): the same
// recurrences as sweep_L_str / sweep_U_str<float, false>, with a factor cell
// of 16 (L) or 24 (U) bytes instead of 32 and 48.  Natural column order only
// (the kernels never set Ctx::rev with BF16 factors).
YG_NOINLINE void sweep_L_str16(const Ctx& cin, const uint16_t* __restrict__ FB,
                               const float* __restrict__ rhs, float* __restrict__ ys) {
  const Ctx c = cin;   // fields in registers, not reloaded after stores
  PhaseTimer ptm(c, 1);
  constexpr int D = 20;   // steps in flight: 7 registers per step (4 packed factor words, 2 rhs, 1 offset)
  constexpr int SF = 16, SV = 8;
  using T = float;
  using V2 = typename Vec2<T>::type;
  SweepTile g;
  if (!sweep_tile(c, g)) return;
  const int nc = uni(c.ncell);
  const __amdgpu_buffer_rsrc_t rF = yg_rsrc(FB + (size_t)8 * g.base, (size_t)8 * g.LM);
  const __amdgpu_buffer_rsrc_t rR = yg_rsrc(rhs, (size_t)2 * nc);
  const __amdgpu_buffer_rsrc_t rY = yg_rsrc(ys + (size_t)2 * g.base, (size_t)2 * g.LM);
  const unsigned lane = threadIdx.x & 63;
  const unsigned voF = lane * SF, voY = lane * SV, voN = (unsigned)g.lane_n * SV;
  uint4 fw[D];
  V2 r[D];
  unsigned so[D];
  int ss[D];
  int sq = 0;
  auto issue = [&](int s, int t) {
    const bool v = on_step(g, t);
    fw[s] = load_raw16_so(rF, v ? voF : YG_OOB, sq * SF);
    r[s] = pair_load_so(rR, v ? voN : YG_OOB, t * SV, T());
    so[s] = v ? voY : YG_OOB;
    ss[s] = sq * SV;
    sq = rot_next(g, sq);
  };
#pragma unroll
  for (int s = 0; s < D; ++s) {
    issue(s, s);
    __builtin_amdgcn_sched_barrier(0);
  }
  T o1x = 0, o1y = 0, p1x = 0, p1y = 0;
  for (int t0 = 0; t0 < g.Tb; t0 += D) {
#pragma unroll
    for (int s = 0; s < D; ++s) {
      const T f0 = bf_lo(fw[s].x), f1 = bf_hi(fw[s].x), f2 = bf_lo(fw[s].y), f3 = bf_hi(fw[s].y);
      const T f4 = bf_lo(fw[s].z), f5 = bf_hi(fw[s].z), f6 = bf_lo(fw[s].w), f7 = bf_hi(fw[s].w);
      const T n1x = dpp_shr1(o1x), n1y = dpp_shr1(o1y);
      const T ax = r[s].x - f0 * p1x, ay = r[s].y - f3 * p1x - f4 * p1y;
      T yx = ax - f1 * o1x - f2 * o1y;
      T yy = ay - f6 * o1y - f5 * n1x - f7 * yx;
      V2 o;
      o.x = yx;
      o.y = yy;
      pair_store_so(o, rY, so[s], ss[s]);
      p1x = n1x; p1y = n1y; o1x = yx; o1y = yy;
      issue(s, t0 + s + D);
      __builtin_amdgcn_sched_barrier(0);
    }
  }
}

YG_NOINLINE void sweep_U_str16(const Ctx& cin, const uint16_t* __restrict__ FB,
                               const float* __restrict__ ys, float* __restrict__ z) {
  const Ctx c = cin;   // fields in registers, not reloaded after stores
  PhaseTimer ptm(c, 2);
  constexpr int D = 18;   // 9 registers per step: 4 packed words, the 2 inverse diagonals, y, offset
  constexpr int SF = 16, SD = 8, SV = 8;
  using T = float;
  using V2 = typename Vec2<T>::type;
  SweepTile g;
  if (!sweep_tile(c, g)) return;
  const int nc = uni(c.ncell);
  const __amdgpu_buffer_rsrc_t rF = yg_rsrc(FB + c.fu16 + (size_t)8 * g.base, (size_t)8 * g.LM);
  const __amdgpu_buffer_rsrc_t rD = yg_rsrc(FB + c.fd16 + (size_t)4 * g.base, (size_t)4 * g.LM);
  const __amdgpu_buffer_rsrc_t rY = yg_rsrc(ys + (size_t)2 * g.base, (size_t)2 * g.LM);
  const __amdgpu_buffer_rsrc_t rZ = yg_rsrc(z, (size_t)2 * nc);
  const unsigned lane = threadIdx.x & 63;
  const unsigned voF = lane * SF, voD = lane * SD, voY = lane * SV, voN = (unsigned)g.lane_n * SV;
  uint4 gw[D];   // the 8 scaled off-diagonal U entries
  V2 dg[D];      // ux-row and uy-row inverse diagonals
  V2 yv[D];
  unsigned so[D];
  int ss[D];
  const int tz = g.Tb - 1;   // the last step
  int sq = (tz % g.M) * g.L;
  auto issue = [&](int s, int t) {
    const bool v = on_step(g, t);
    gw[s] = load_raw16_so(rF, v ? voF : YG_OOB, sq * SF);
    dg[s] = pair_load_so(rD, v ? voD : YG_OOB, sq * SD, T());
    yv[s] = pair_load_so(rY, v ? voY : YG_OOB, sq * SV, T());
    so[s] = v ? voN : YG_OOB;
    ss[s] = t * SV;
    sq = rot_prev(g, sq);
  };
#pragma unroll
  for (int s = 0; s < D; ++s) {
    issue(s, tz - s);
    __builtin_amdgcn_sched_barrier(0);
  }
  T o1x = 0, o1y = 0, p1x = 0, p1y = 0;
  for (int u0 = 0; u0 < g.Tb; u0 += D) {
#pragma unroll
    for (int s = 0; s < D; ++s) {
      // 0..4 scaled ux-row U entries (slots 5,3,7,1,6), 5..7 scaled uy-row
      // entries (slots 6,1,3)
      const T g0 = bf_lo(gw[s].x), g1 = bf_hi(gw[s].x), g2 = bf_lo(gw[s].y), g3 = bf_hi(gw[s].y);
      const T g4 = bf_lo(gw[s].z), g5 = bf_hi(gw[s].z), g6 = bf_lo(gw[s].w), g7 = bf_hi(gw[s].w);
      const T n1y = dpp_shl1(o1y), n1x = dpp_shl1(o1x);
      const T by = dg[s].y * yv[s].y - g7 * p1y, bx = dg[s].x * yv[s].x - g3 * p1x - g4 * p1y;
      T zy = by - g5 * o1x - g6 * o1y;
      T zx = bx - g1 * o1x - g2 * n1y - g0 * zy;
      V2 o;
      o.x = zx;
      o.y = zy;
      pair_store_so(o, rZ, so[s], ss[s]);
      p1x = n1x; p1y = n1y; o1x = zx; o1y = zy;
      issue(s, tz - (u0 + s + D));
      __builtin_amdgcn_sched_barrier(0);
    }
  }
}

// FP32 line-block factors -> BF16 storage (layout above)
YG_DEV void factors_to_bf16(const Ctx& c, const float* __restrict__ F, uint16_t* __restrict__ FB) {
  const int nc = c.ncell;
  const float* FU = F + c.fu;
  uint16_t* BU = FB + c.fu16;
  uint16_t* BD = FB + c.fd16;
  for (int q = threadIdx.x; q < nc; q += blockDim.x) {
    uint32_t lw[4], uw[4];
#pragma unroll
    for (int e = 0; e < 4; ++e)
      lw[e] = (uint32_t)to_bf16(F[8 * q + 2 * e]) | ((uint32_t)to_bf16(F[8 * q + 2 * e + 1]) << 16);
    const float* u = FU + 12 * q;
    const float g[8] = {u[0] * u[5], u[1] * u[5], u[2] * u[5], u[3] * u[5], u[4] * u[5],
                        u[6] * u[9], u[7] * u[9], u[8] * u[9]};
#pragma unroll
    for (int e = 0; e < 4; ++e)
      uw[e] = (uint32_t)to_bf16(g[2 * e]) | ((uint32_t)to_bf16(g[2 * e + 1]) << 16);
    uint4* dl = reinterpret_cast<uint4*>(FB + 8 * q);
    *dl = make_uint4(lw[0], lw[1], lw[2], lw[3]);
    *reinterpret_cast<uint4*>(BU + 8 * q) = make_uint4(uw[0], uw[1], uw[2], uw[3]);
    *reinterpret_cast<float2*>(BD + 4 * q) = make_float2(u[5], u[9]);
  }
  __syncthreads();
}

// ---- preconditioner --------------------------------------------------------
template <typename T>
struct Precon {
  int kind;        // 0 none, 1 jacobi, 2 ilu
  const T* F;      // ILU factors, or 1/diag pairs for jacobi
  const uint16_t* FB = nullptr;   // BF16 copy of line-block factors (mixed mode)
  bool fb_l = true;                // the forward sweep also reads BF16 (else F's FP32 L part)
};

// out = M^-1 in; tmp is scratch of the same size.  The sweeps run on the
// part of the member this block owns (c.pc: the whole member, or the lines
// of this block of a member group, whose line blocks do not couple to other
// blocks' lines); in, tmp, out are member vectors.
template <typename T>
YG_DEV void precon_apply(const Ctx& c, const Precon<T>& P, const T* in, T* tmp, T* out) {
  const Ctx& pc = *c.pc;
  const size_t o = 2 * (size_t)c.cb;
  if (P.kind == 2 && pc.coupled) {
    sweep_L<T>(pc, P.F, in, tmp);
    sweep_U<T>(pc, P.F, tmp, out);
  } else if (P.kind == 2 && P.FB) {
    if constexpr (sizeof(T) == 4) {
      if (pc.ncell > 0) {
        if (P.fb_l) sweep_L_str16(pc, P.FB, in + o, tmp + o);
        else sweep_L_str<T, false>(pc, P.F, in + o, tmp + o);
      }
      __syncthreads();
      if (pc.ncell > 0) sweep_U_str16(pc, P.FB, tmp + o, out + o);
      __syncthreads();
    }
  } else if (P.kind == 2) {
    if (pc.ncell > 0) {
      if (pc.rev) sweep_L_str<T, true>(pc, P.F, in + o, tmp + o);
      else sweep_L_str<T, false>(pc, P.F, in + o, tmp + o);
    }
    __syncthreads();
    // z staged in LDS when the part fits (coalesced stores afterwards)
    const bool zl = zl_fits<T>(pc);
    if (pc.ncell > 0) {
      if (zl) {
        if (pc.rev) sweep_U_str<T, true, true>(pc, P.F, tmp + o, out + o);
        else sweep_U_str<T, false, true>(pc, P.F, tmp + o, out + o);
      } else {
        if (pc.rev) sweep_U_str<T, true, false>(pc, P.F, tmp + o, out + o);
        else sweep_U_str<T, false, false>(pc, P.F, tmp + o, out + o);
      }
    }
    __syncthreads();
    if (zl) {
      zl_copy_out<T>(pc, out + o);
      __syncthreads();
    }
  } else {
    for (int q = 2 * c.cb + threadIdx.x; q < 2 * c.ce; q += blockDim.x)
      out[q] = P.kind == 1 ? in[q] * P.F[q] : in[q];
    __syncthreads();
  }
}

// ---- BiCGSTAB / CG ------------------------------------------------------
struct SolveOut {
  int iters, status;
  double resid;
};

// ---- stencil SpMV ------------------------------------------------------
// y = A x for the two rows of natural cell n = i*ny + j (the reference's
// unknown numbering: rows 2n, 2n+1).  The cell's matrix record arrives as
// 16-byte chunks from one offset; x is read as the (ux,uy) pairs of the 7
// cells of the stencil through a bounds-checked buffer: a neighbour index
// past either end of the member reads 0, and one that steps across a
// non-periodic border into another line meets a zero coefficient (absent
// slots hold exact zeros).  Periodic dimensions wrap explicitly.
// wrap: bit 0 periodic in x, bit 1 periodic in y (wave-uniform)
template <typename TA, typename TV, bool COH>
YG_DEV void spmv_cell(int nx, int ny, int wrap, __amdgpu_buffer_rsrc_t rA,
                      __amdgpu_buffer_rsrc_t rX, int n, double& yx, double& yy) {
  using V2 = typename Vec2<TV>::type;
  constexpr unsigned SV = 2 * sizeof(TV), SA = YG_NA * sizeof(TA);
  int nip = n + ny, nim = n - ny, njp = n + 1, njm = n - 1, nipjm = n + ny - 1, nimjp = n - ny + 1;
  if (wrap) {
    const int i = n / ny, j = n - i * ny;
    const int ip = (wrap & 1) ? (i + 1 == nx ? 0 : i + 1) : i + 1;
    const int im = (wrap & 1) ? (i == 0 ? nx - 1 : i - 1) : i - 1;
    const int jp = (wrap & 2) ? (j + 1 == ny ? 0 : j + 1) : j + 1;
    const int jm = (wrap & 2) ? (j == 0 ? ny - 1 : j - 1) : j - 1;
    nip = ip * ny + j; nim = im * ny + j; njp = i * ny + jp; njm = i * ny + jm;
    nipjm = ip * ny + jm; nimjp = im * ny + jp;
  }
  auto ldx = [&](int k) { return Pair16<TV>::template load<COH>(rX, (unsigned)k * SV); };
  const V2 x0 = ldx(n), xip = ldx(nip), xim = ldx(nim), xjp = ldx(njp), xjm = ldx(njm),
           xipjm = ldx(nipjm), ximjp = ldx(nimjp);
  TA a[sizeof(TA) == 4 ? 20 : 18];   // the record (FP64: its 18 values, not the pad)
  load_vals<sizeof(TA) == 4 ? 20 : 18>(rA, (unsigned)n * SA, a);
  // products summed in the vector type (FP32 in the inner solve of the
  // mixed mode, FP64 otherwise)
  // ux row: slots ux(i,j) ux(i+1,j) ux(i-1,j) ux(i,j+1) ux(i,j-1)
  //              uy(i,j) uy(i+1,j) uy(i+1,j-1) uy(i,j-1)
  const TV sx = (TV)a[0] * x0.x + (TV)a[2] * xip.x + (TV)a[4] * xim.x + (TV)a[6] * xjp.x +
                (TV)a[8] * xjm.x + (TV)a[10] * x0.y + (TV)a[12] * xip.y + (TV)a[14] * xipjm.y +
                (TV)a[16] * xjm.y;
  // uy row: slots uy(i,j) uy(i,j+1) uy(i,j-1) uy(i+1,j) uy(i-1,j)
  //              ux(i,j) ux(i,j+1) ux(i-1,j+1) ux(i-1,j)
  const TV sy = (TV)a[1] * x0.y + (TV)a[3] * xjp.y + (TV)a[5] * xjm.y + (TV)a[7] * xip.y +
                (TV)a[9] * xim.y + (TV)a[11] * x0.x + (TV)a[13] * xjp.x + (TV)a[15] * ximjp.x +
                (TV)a[17] * xim.x;
  yx = (double)sx;
  yy = (double)sy;
}

// What a SpMV computes beside y = A x, with its block-reduced sums:
//   YG_SPMV_RES : r = u - y (u = b), w2 = r when given;  sums (r,r), (u,u)
//   YG_SPMV_DOT : sum (u, y)
//   YG_SPMV_DOT2: sums (y, y), (y, u)
#define YG_SPMV_RES 0
#define YG_SPMV_DOT 1
#define YG_SPMV_DOT2 2
struct Sums2 {
  double s0, s1;
};

// One SpMV over the rows of cells [c.cb, c.ce) of the member, four cells per
// thread per trip (the loads of all four in flight together), with the fused
// dots (summed over the member: over the group when the member has one).
// COH: x is read with device-scope loads (a member group: the halo lines
// belong to other workgroups).  A device function of its own (not inlined)
// so that its addresses and resources get their own register allocation.
template <int MODE, typename TA, typename TV, bool COH = false>
YG_NOINLINE Sums2 spmv_dots(const Ctx& cin, const TA* A, const TV* x, TV* y, const TV* u, TV* w1,
                            TV* w2) {
  const Ctx c = cin;   // fields in registers, not reloaded after stores
  PhaseTimer ptm(c, 3);
  using V2 = typename Vec2<TV>::type;
  const int nc = uni(c.ncell), B = blockDim.x, cb = uni(c.cb), ce = uni(c.ce);
  const int nx = uni(c.nx), ny = uni(c.ny), wrap = uni((c.wrapx ? 1 : 0) | (c.wrapy ? 2 : 0));
  const __amdgpu_buffer_rsrc_t rA = yg_rsrc(A, (size_t)YG_NA * nc);
  const __amdgpu_buffer_rsrc_t rX = yg_rsrc(x, (size_t)2 * nc);
  constexpr unsigned SV = 2 * sizeof(TV);
  const __amdgpu_buffer_rsrc_t rY = yg_rsrc(y, (size_t)2 * nc);
  const __amdgpu_buffer_rsrc_t rU = yg_rsrc(u, (size_t)2 * nc);
  const __amdgpu_buffer_rsrc_t rW1 = yg_rsrc(w1, w1 ? (size_t)2 * nc : 0);
  const __amdgpu_buffer_rsrc_t rW2 = yg_rsrc(w2, w2 ? (size_t)2 * nc : 0);   // range 0: stores dropped
  double s0 = 0.0, s1 = 0.0;
  // Four cells per thread per trip, with no branch around the memory
  // operations: a cell index past the member reads zeros (matrix record and
  // u out of range) and its stores are dropped, so every load of the trip
  // (u's too, ahead of the stores, which the compiler cannot move loads
  // across) is issued before the first is waited for.
  for (int q0 = cb + threadIdx.x; q0 < ce; q0 += 4 * B) {
    double ax[4], ay[4];
    V2 uqs[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      // a cell past the range reads as the zero record past the member
      const int q = q0 + k * B < ce ? q0 + k * B : nc;
      uqs[k] = Pair16<TV>::load(rU, q0 + k * B < ce ? (unsigned)q * SV : YG_OOB);
      // only the first and last line of the range read lines that other
      // workgroups wrote; the other cells' x loads may go through the L1
      if (COH && (q - cb < ny || ce - q <= ny))
        spmv_cell<TA, TV, true>(nx, ny, wrap, rA, rX, q, ax[k], ay[k]);
      else
        spmv_cell<TA, TV, false>(nx, ny, wrap, rA, rX, q, ax[k], ay[k]);
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int q = q0 + k * B;
      const unsigned off = q < ce ? (unsigned)q * SV : YG_OOB;
      V2 o;
      o.x = (TV)ax[k];
      o.y = (TV)ay[k];
      Pair16<TV>::store(o, rY, off);
      const double yx = (double)o.x, yy = (double)o.y;
      const V2 uq = uqs[k];
      if (MODE == YG_SPMV_RES) {
        V2 r;
        r.x = (TV)((double)uq.x - yx);
        r.y = (TV)((double)uq.y - yy);
        Pair16<TV>::store(r, rW1, off);
        Pair16<TV>::store(r, rW2, off);
        s0 += (double)r.x * (double)r.x + (double)r.y * (double)r.y;
        s1 += (double)uq.x * (double)uq.x + (double)uq.y * (double)uq.y;
      } else if (MODE == YG_SPMV_DOT) {
        s0 += (double)uq.x * yx + (double)uq.y * yy;
      } else {
        s0 += yx * yx + yy * yy;
        s1 += yx * (double)uq.x + yy * (double)uq.y;
      }
    }
  }
  double sums[2] = {s0, s1};
  csum<2>(c, sums);
  return Sums2{sums[0], sums[1]};
}

// The FP64 residual of the reliable-update solve over the rows of cells
// [c.cb, c.ce): with e = b - A x (FP64 matrix records, x and b), stores
// r = (float)(e sc) and returns the member sums (e, e) (FP64, unscaled) and
// (rt, r).  The same trip structure as spmv_dots (four cells per thread, every
// load of a trip in flight together; cells past the range read zeros and
// their stores are dropped).  COH: a member group (x's halo lines were
// written by other workgroups).
template <bool COH>
YG_NOINLINE Sums2 spmv_res32(const Ctx& cin, const double* A, const double* x, const double* b,
                             const float* rt, float* r, double sc) {
  const Ctx c = cin;
  PhaseTimer ptm(c, 3);
  const int nc = uni(c.ncell), B = blockDim.x, cb = uni(c.cb), ce = uni(c.ce);
  const int nx = uni(c.nx), ny = uni(c.ny), wrap = uni((c.wrapx ? 1 : 0) | (c.wrapy ? 2 : 0));
  const __amdgpu_buffer_rsrc_t rA = yg_rsrc(A, (size_t)YG_NA * nc);
  const __amdgpu_buffer_rsrc_t rX = yg_rsrc(x, (size_t)2 * nc);
  const __amdgpu_buffer_rsrc_t rB = yg_rsrc(b, (size_t)2 * nc);
  const __amdgpu_buffer_rsrc_t rT = yg_rsrc(rt, (size_t)2 * nc);
  const __amdgpu_buffer_rsrc_t rR = yg_rsrc(r, (size_t)2 * nc);
  double s0 = 0.0, s1 = 0.0;
  for (int q0 = cb + threadIdx.x; q0 < ce; q0 += 4 * B) {
    double ax[4], ay[4];
    double2 bq[4];
    float2 tq[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const bool in = q0 + k * B < ce;
      const int q = in ? q0 + k * B : nc;
      bq[k] = Pair16<double>::load(rB, in ? (unsigned)q * 16u : YG_OOB);
      tq[k] = Pair16<float>::load(rT, in ? (unsigned)q * 8u : YG_OOB);
      if (COH && (q - cb < ny || ce - q <= ny))
        spmv_cell<double, double, true>(nx, ny, wrap, rA, rX, q, ax[k], ay[k]);
      else
        spmv_cell<double, double, false>(nx, ny, wrap, rA, rX, q, ax[k], ay[k]);
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int q = q0 + k * B;
      const double ex = bq[k].x - ax[k], ey = bq[k].y - ay[k];
      float2 o;
      o.x = (float)(ex * sc);
      o.y = (float)(ey * sc);
      Pair16<float>::store(o, rR, q < ce ? (unsigned)q * 8u : YG_OOB);
      s0 += ex * ex + ey * ey;
      s1 += (double)tq[k].x * (double)o.x + (double)tq[k].y * (double)o.y;
    }
  }
  double sums[2] = {s0, s1};
  csum<2>(c, sums);
  return Sums2{sums[0], sums[1]};
}

// A 16-byte chunk of a vector (4 floats or 2 doubles).
template <typename T>
struct alignas(16) C16 {
  T v[16 / sizeof(T)];
};

// Element-wise pass over the values [lo, hi): out[o][k] = f(in[0][k], ...,
// in[NI-1][k]).  The vectors are 16-byte aligned; thread g takes the aligned
// chunks g, g + B, ... inside the range, four per trip (two for passes of
// four or more inputs) with every load of a trip issued before its
// arithmetic; the values before the first and after the last whole chunk go
// one by one.  Arithmetic in double: f(const double (&xi)[NI], double
// (&xo)[NO]); dots accumulate in f's captures.
template <typename T, int NI, int NO, typename F>
YG_DEV void vmap(int lo, int hi, const T* const (&in)[NI], T* const (&out)[NO], F&& f) {
  constexpr int E = 16 / sizeof(T), U = NI >= 4 ? 2 : 4;
  const int B = blockDim.x;
  const int c0 = (lo + E - 1) / E, c1 = max(c0, hi / E);   // whole chunks [c0, c1)
  for (int k0 = c0 + threadIdx.x; k0 < c1; k0 += U * B) {
    C16<T> ci[U][NI];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = k0 + u * B;
      if (k < c1) {
#pragma unroll
        for (int i = 0; i < NI; ++i) ci[u][i] = reinterpret_cast<const C16<T>*>(in[i])[k];
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = k0 + u * B;
      if (k < c1) {
        C16<T> co[NO > 0 ? NO : 1];
#pragma unroll
        for (int e = 0; e < E; ++e) {
          double xi[NI], xo[NO > 0 ? NO : 1];
#pragma unroll
          for (int i = 0; i < NI; ++i) xi[i] = (double)ci[u][i].v[e];
          f(xi, xo);
#pragma unroll
          for (int o = 0; o < NO; ++o) co[o].v[e] = (T)xo[o];
        }
#pragma unroll
        for (int o = 0; o < NO; ++o) reinterpret_cast<C16<T>*>(out[o])[k] = co[o];
      }
    }
  }
  // the values outside the whole chunks: [lo, c0 E) and [c1 E, hi)
  const int h0 = min(hi, c0 * E), t0 = max(h0, c1 * E);
  const int nh = h0 - lo, nt = max(0, hi - t0);
  for (int g = threadIdx.x; g < nh + nt; g += B) {
    const int q = g < nh ? lo + g : t0 + (g - nh);
    double xi[NI], xo[NO > 0 ? NO : 1];
#pragma unroll
    for (int i = 0; i < NI; ++i) xi[i] = (double)in[i][q];
    f(xi, xo);
#pragma unroll
    for (int o = 0; o < NO; ++o) out[o][q] = (T)xo[o];
  }
}

// The refinement's passes over value pairs [lo, hi) (lo, hi even; the
// vectors start 16-byte aligned): eight pairs per thread per trip, every
// load of a trip issued before the first is used (a plain loop waits for
// each load in turn, about a memory latency per element of a thread).
constexpr int YG_RU = 8;

// rin = (float)(r64 sc), d = 0
YG_DEV void refine_scale(const double* __restrict__ r64, float* __restrict__ rin,
                         float* __restrict__ d, double sc, int lo, int hi) {
  const int p0 = lo >> 1, p1 = hi >> 1, B = blockDim.x;
  for (int k0 = p0 + (int)threadIdx.x; k0 < p1; k0 += YG_RU * B) {
    double2 v[YG_RU];
#pragma unroll
    for (int u = 0; u < YG_RU; ++u) {
      const int k = k0 + u * B;
      if (k < p1) v[u] = reinterpret_cast<const double2*>(r64)[k];
    }
#pragma unroll
    for (int u = 0; u < YG_RU; ++u) {
      const int k = k0 + u * B;
      if (k < p1) {
        float2 o;
        o.x = (float)(v[u].x * sc);
        o.y = (float)(v[u].y * sc);
        reinterpret_cast<float2*>(rin)[k] = o;
        reinterpret_cast<float2*>(d)[k] = make_float2(0.0f, 0.0f);
      }
    }
  }
}

// x64 += rn d
YG_DEV void refine_update(double* __restrict__ x64, const float* __restrict__ d, double rn,
                          int lo, int hi) {
  const int p0 = lo >> 1, p1 = hi >> 1, B = blockDim.x;
  for (int k0 = p0 + (int)threadIdx.x; k0 < p1; k0 += YG_RU * B) {
    double2 x[YG_RU];
    float2 dv[YG_RU];
#pragma unroll
    for (int u = 0; u < YG_RU; ++u) {
      const int k = k0 + u * B;
      if (k < p1) {
        x[u] = reinterpret_cast<const double2*>(x64)[k];
        dv[u] = reinterpret_cast<const float2*>(d)[k];
      }
    }
#pragma unroll
    for (int u = 0; u < YG_RU; ++u) {
      const int k = k0 + u * B;
      if (k < p1) {
        double2 o;
        o.x = x[u].x + rn * (double)dv[u].x;
        o.y = x[u].y + rn * (double)dv[u].y;
        reinterpret_cast<double2*>(x64)[k] = o;
      }
    }
  }
}

// number of nonzero values of b64 in [lo, hi) (this thread's share)
YG_DEV double count_nonzero(const double* __restrict__ b64, int lo, int hi) {
  const int p0 = lo >> 1, p1 = hi >> 1, B = blockDim.x;
  double n = 0.0;
  for (int k0 = p0 + (int)threadIdx.x; k0 < p1; k0 += YG_RU * B) {
    double2 v[YG_RU];
#pragma unroll
    for (int u = 0; u < YG_RU; ++u) {
      const int k = k0 + u * B;
      v[u] = k < p1 ? reinterpret_cast<const double2*>(b64)[k] : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int u = 0; u < YG_RU; ++u) n += (v[u].x != 0.0 ? 1.0 : 0.0) + (v[u].y != 0.0 ? 1.0 : 0.0);
  }
  return n;
}

// BiCGSTAB over the rows [c.cb, c.ce) of the member.  GRP: a member group
// (the rows are this block's share; every sum is a group sum and a group
// barrier precedes each SpMV whose x other blocks wrote).  With x0zero
// (GRP only: the refinement's inner solves) x = 0 on entry and r0 = b needs
// no SpMV.
template <typename TA, typename TV, bool GRP = false>
YG_DEV SolveOut bicgstab(const Ctx& c, const TA* A, const Precon<TV>& P, const TV* b, TV* x,
                         TV* W, const YgKrylovOpts& o, bool x0zero = GRP) {
  const int lo = 2 * c.cb, hi = 2 * c.ce;
  const size_t vl = c.vlen;   // vector stride in W (16-byte aligned)
  TV* __restrict__ r = W;
  TV* __restrict__ rt = W + 1 * vl;
  TV* __restrict__ p = W + 2 * vl;
  TV* __restrict__ v = W + 3 * vl;
  TV* __restrict__ ph = W + 4 * vl;
  TV* __restrict__ s = W + 5 * vl;
  TV* __restrict__ sh = W + 6 * vl;
  TV* __restrict__ tt = W + 7 * vl;
  TV* __restrict__ ytmp = W + 8 * vl;
  SolveOut out{0, 4, 0.0};
  double red[2];
  // r = b - A x, rt = r, (r,r), (b,b)
  if (GRP && x0zero) {
    red[0] = 0.0;
    vmap<TV, 1, 2>(lo, hi, {b}, {r, rt}, [&](const double (&xi)[1], double (&xo)[2]) {
      xo[0] = xo[1] = xi[0];
      red[0] += xi[0] * xi[0];
    });
    csum<1>(c, *reinterpret_cast<double(*)[1]>(red));
    red[1] = red[0];
  } else {
    if (GRP) grp_barrier(c.gv);   // the halo lines of x
    const Sums2 z = spmv_dots<YG_SPMV_RES, TA, TV, GRP>(c, A, x, tt, b, r, rt);
    red[0] = z.s0;
    red[1] = z.s1;
  }
  const double nrm0 = o.conv_cond == 1 ? sqrt(red[1]) : sqrt(red[0]);
  if (nrm0 == 0.0) { out.status = 0; out.resid = 0.0; return out; }
  const double bnrm = 1.0 / nrm0;
  out.resid = sqrt(red[0]) * bnrm;
  if (out.resid <= o.tol) { out.status = 0; return out; }
  double rho = red[0], rho_old = 1.0, alpha = 1.0, omega = 1.0;  // (rt, r) = (r, r) at start
  for (int it = 1; it <= o.maxiter; ++it) {
    if (rho == 0.0) { out.status = 2; out.iters = it; return out; }
    // p = r + beta (p - omega v); phat = M^-1 p
    if (it == 1) {
      PhaseTimer ptm(c, 4);
      vmap<TV, 1, 1>(lo, hi, {r}, {p}, [&](const double (&xi)[1], double (&xo)[1]) { xo[0] = xi[0]; });
    } else {
      PhaseTimer ptm(c, 4);
      const double beta = (rho / rho_old) * (alpha / omega);
      vmap<TV, 3, 1>(lo, hi, {r, p, v}, {p}, [&](const double (&xi)[3], double (&xo)[1]) {
        xo[0] = xi[0] + beta * (xi[1] - omega * xi[2]);
      });
    }
    __syncthreads();
    precon_apply<TV>(c, P, p, ytmp, ph);
    if (GRP) grp_barrier(c.gv);   // the halo lines of phat
    // v = A phat, (rt, v)
    red[0] = spmv_dots<YG_SPMV_DOT, TA, TV, GRP>(c, A, ph, v, rt, nullptr, nullptr).s0;
    if (red[0] == 0.0) { out.status = 2; out.iters = it; return out; }
    alpha = rho / red[0];
    // s = r - alpha v and ||s||: converged on s ends the iteration early
    red[0] = 0.0;
    {
      PhaseTimer ptm(c, 4);
      vmap<TV, 2, 1>(lo, hi, {r, v}, {s}, [&](const double (&xi)[2], double (&xo)[1]) {
        const double sq = (double)(TV)(xi[0] - alpha * xi[1]);
        xo[0] = sq;
        red[0] += sq * sq;
      });
    }
    if (GRP) {
      // shat first: the sum of ||s|| then doubles as the barrier before the
      // SpMV on shat (on the last iteration shat goes unused)
      __syncthreads();
      precon_apply<TV>(c, P, s, ytmp, sh);
    }
    csum<1>(c, *reinterpret_cast<double(*)[1]>(red));
    const double sn = sqrt(red[0]) * bnrm;
    if (sn <= o.tol) {
      vmap<TV, 2, 1>(lo, hi, {x, ph}, {x}, [&](const double (&xi)[2], double (&xo)[1]) {
        xo[0] = xi[0] + alpha * xi[1];
      });
      __syncthreads();
      out.resid = sn; out.iters = it; out.status = 0;
      return out;
    }
    if (!GRP) precon_apply<TV>(c, P, s, ytmp, sh);
    // t = A shat, (t,t), (t,s)
    {
      const Sums2 z = spmv_dots<YG_SPMV_DOT2, TA, TV, GRP>(c, A, sh, tt, s, nullptr, nullptr);
      red[0] = z.s0;
      red[1] = z.s1;
    }
    if (red[0] == 0.0) { out.status = 2; out.iters = it; return out; }
    omega = red[1] / red[0];
    if (omega == 0.0) { out.status = 2; out.iters = it; return out; }
    // x += alpha phat + omega shat; r = s - omega t; (r,r), (rt,r)
    red[0] = red[1] = 0.0;
    {
      PhaseTimer ptm(c, 4);
      vmap<TV, 6, 2>(lo, hi, {x, ph, sh, s, tt, rt}, {x, r},
                     [&](const double (&xi)[6], double (&xo)[2]) {
        xo[0] = xi[0] + alpha * xi[1] + omega * xi[2];
        const double rq = (double)(TV)(xi[3] - omega * xi[4]);
        xo[1] = rq;
        red[0] += rq * rq;
        red[1] += xi[5] * rq;
      });
    }
    csum<2>(c, red);
    out.resid = sqrt(red[0]) * bnrm;
    out.iters = it;
    if (out.resid <= o.tol) { out.status = 0; return out; }
    rho_old = rho;
    rho = red[1];
  }
  out.status = 4;
  return out;
}

// ---- mixed precision with reliable updates ----------------------------------
// r = rt = (float)(r64 sc), d = 0 over the value pairs [lo, hi); returns this
// thread's share of (r, r)
YG_DEV double ir_start(const double* __restrict__ r64, float* __restrict__ r,
                       float* __restrict__ rt, float* __restrict__ d, double sc, int lo, int hi) {
  const int p0 = lo >> 1, p1 = hi >> 1, B = blockDim.x;
  double rr = 0.0;
  for (int k0 = p0 + (int)threadIdx.x; k0 < p1; k0 += YG_RU * B) {
    double2 v[YG_RU];
#pragma unroll
    for (int u = 0; u < YG_RU; ++u) {
      const int k = k0 + u * B;
      if (k < p1) v[u] = reinterpret_cast<const double2*>(r64)[k];
    }
#pragma unroll
    for (int u = 0; u < YG_RU; ++u) {
      const int k = k0 + u * B;
      if (k < p1) {
        float2 o;
        o.x = (float)(v[u].x * sc);
        o.y = (float)(v[u].y * sc);
        reinterpret_cast<float2*>(r)[k] = o;
        reinterpret_cast<float2*>(rt)[k] = o;
        reinterpret_cast<float2*>(d)[k] = make_float2(0.0f, 0.0f);
        rr += (double)o.x * (double)o.x + (double)o.y * (double)o.y;
      }
    }
  }
  return rr;
}

// x64 += sb d, d = 0 over the value pairs [lo, hi)
YG_DEV void ir_fold(double* __restrict__ x64, float* __restrict__ d, double sb, int lo, int hi) {
  const int p0 = lo >> 1, p1 = hi >> 1, B = blockDim.x;
  for (int k0 = p0 + (int)threadIdx.x; k0 < p1; k0 += YG_RU * B) {
    double2 x[YG_RU];
    float2 dv[YG_RU];
#pragma unroll
    for (int u = 0; u < YG_RU; ++u) {
      const int k = k0 + u * B;
      if (k < p1) {
        x[u] = reinterpret_cast<const double2*>(x64)[k];
        dv[u] = reinterpret_cast<const float2*>(d)[k];
      }
    }
#pragma unroll
    for (int u = 0; u < YG_RU; ++u) {
      const int k = k0 + u * B;
      if (k < p1) {
        double2 o;
        o.x = x[u].x + sb * (double)dv[u].x;
        o.y = x[u].y + sb * (double)dv[u].y;
        reinterpret_cast<double2*>(x64)[k] = o;
        reinterpret_cast<float2*>(d)[k] = make_float2(0.0f, 0.0f);
      }
    }
  }
}

struct IrOut {
  int iters, status, nres;
  double res;
};

// The mixed-precision solve of A x = b as one FP32 BiCGSTAB with reliable
// updates (residual replacement with group-wise updates of x, van der Vorst
// and Ye, SIAM J. Sci. Comput. 22 (2000) 835-852): the recurrences run in
// FP32 on the FP32 matrix records and factors, in units of the initial FP64
// residual norm rn0; the correction d accumulates in FP32 and is folded into
// the FP64 x (x += rn0 d) whenever the iterated residual has fallen below
// ir_delta times its value at the last replacement (or below the target),
// after which the iterated residual is replaced by the FP64 residual
// b - A x (FP64 records) and the iteration continues with its search
// direction.  A breakdown (rho, (rt, v), (t, t) or omega zero) or
// convergence on s also folds d and replaces r, and then restarts the
// recurrences from the true residual (rt = r, p = r); a breakdown right
// after a (re)start ends the solve with status 2.  The solve stops when the
// FP64 residual reaches the refinement target ||r|| <= ir_target ||r0||
// (ir_target 0: a tenth of tol), or at maxiter iterations or ir_max FP64
// residuals; status 0 when ||r|| <= tol ||r0||, else 4.  Every decision
// rests on group sums, so every block of a member group takes the same
// branches.  GRP: a member group (x's halo lines go through group barriers).
// W holds the 9 FP32 Krylov vectors, d the FP32 correction; r64, y64 are
// FP64 scratch vectors.
template <bool GRP>
YG_DEV IrOut ir_bicgstab(const Ctx& c, const double* A64, const float* A32,
                         const Precon<float>& P, const double* b64, double* x64, double* r64,
                         double* y64, float* W, float* d, const YgKrylovOpts& o) {
  const int lo = 2 * c.cb, hi = 2 * c.ce;
  const size_t vl = c.vlen;
  float* __restrict__ r = W;
  float* __restrict__ rt = W + 1 * vl;
  float* __restrict__ p = W + 2 * vl;
  float* __restrict__ v = W + 3 * vl;
  float* __restrict__ ph = W + 4 * vl;
  float* __restrict__ s = W + 5 * vl;
  float* __restrict__ sh = W + 6 * vl;
  float* __restrict__ tt = W + 7 * vl;
  float* __restrict__ ytmp = W + 8 * vl;
  IrOut out{0, 4, 0, 1.0};
  const double target = o.ir_target > 0.0 ? o.ir_target : 1e-1 * o.tol;
  // the FP64 residual of x0
  if (GRP) grp_barrier(c.gv);
  double red[2];
  {
    const Sums2 z = spmv_dots<YG_SPMV_RES, double, double, GRP>(c, A64, x64, y64, b64, r64, nullptr);
    red[0] = z.s0;
    red[1] = z.s1;
  }
  out.nres = 1;
  const double rn0 = sqrt(red[0]);
  const double nrm0 = o.conv_cond == 1 ? sqrt(red[1]) : rn0;
  if (nrm0 == 0.0 || rn0 == 0.0) { out.res = 0.0; out.status = 0; return out; }
  out.res = rn0 / nrm0;
  out.status = out.res <= o.tol ? 0 : 4;
  if (out.res <= target || o.maxiter <= 0) return out;
  const double sc = 1.0 / rn0;             // the FP32 vectors are residuals / rn0
  const double tgt = target * nrm0 * sc;   // the iterated residual's target in those units
  red[0] = ir_start(r64, r, rt, d, sc, lo, hi);
  csum<1>(c, *reinterpret_cast<double(*)[1]>(red));
  __syncthreads();
  double rho = red[0], rho_old = 1.0, alpha = 1.0, omega = 1.0;
  double rupd = sqrt(red[0]);   // the iterated residual at the last replacement
  bool fresh = true;            // p starts from r (first iteration after a (re)start)
  int since = 0;                // iterations since the last (re)start
  for (;;) {
    ++out.iters;
    ++since;
    bool brk = false, conv_s = false, upd = false;
    if (rho == 0.0) {
      brk = true;
    } else {
      // p = r + beta (p - omega v); phat = M^-1 p
      {
        PhaseTimer ptm(c, 4);
        if (fresh) {
          vmap<float, 1, 1>(lo, hi, {r}, {p}, [&](const double (&xi)[1], double (&xo)[1]) { xo[0] = xi[0]; });
        } else {
          const double beta = (rho / rho_old) * (alpha / omega);
          vmap<float, 3, 1>(lo, hi, {r, p, v}, {p}, [&](const double (&xi)[3], double (&xo)[1]) {
            xo[0] = xi[0] + beta * (xi[1] - omega * xi[2]);
          });
        }
      }
      fresh = false;
      __syncthreads();
      precon_apply<float>(c, P, p, ytmp, ph);
      if (GRP) grp_barrier(c.gv);   // the halo lines of phat
      // v = A phat, (rt, v)
      const double rtv = spmv_dots<YG_SPMV_DOT, float, float, GRP>(c, A32, ph, v, rt, nullptr, nullptr).s0;
      if (rtv == 0.0) {
        brk = true;
      } else {
        alpha = rho / rtv;
        // s = r - alpha v, ||s||
        red[0] = 0.0;
        {
          PhaseTimer ptm(c, 4);
          vmap<float, 2, 1>(lo, hi, {r, v}, {s}, [&](const double (&xi)[2], double (&xo)[1]) {
            const double sq = (double)(float)(xi[0] - alpha * xi[1]);
            xo[0] = sq;
            red[0] += sq * sq;
          });
        }
        if (GRP) {
          // shat first: the sum of ||s|| doubles as the barrier before the
          // SpMV on shat
          __syncthreads();
          precon_apply<float>(c, P, s, ytmp, sh);
        }
        csum<1>(c, *reinterpret_cast<double(*)[1]>(red));
        if (sqrt(red[0]) <= tgt) {
          // converged on s: d += alpha phat
          PhaseTimer ptm(c, 4);
          vmap<float, 2, 1>(lo, hi, {d, ph}, {d}, [&](const double (&xi)[2], double (&xo)[1]) {
            xo[0] = xi[0] + alpha * xi[1];
          });
          conv_s = true;
        } else {
          if (!GRP) precon_apply<float>(c, P, s, ytmp, sh);
          // t = A shat, (t, t), (t, s)
          const Sums2 z = spmv_dots<YG_SPMV_DOT2, float, float, GRP>(c, A32, sh, tt, s, nullptr, nullptr);
          omega = z.s0 != 0.0 ? z.s1 / z.s0 : 0.0;
          if (omega == 0.0) {
            brk = true;
          } else {
            // d += alpha phat + omega shat; r = s - omega t; (r, r), (rt, r)
            red[0] = red[1] = 0.0;
            {
              PhaseTimer ptm(c, 4);
              vmap<float, 6, 2>(lo, hi, {d, ph, sh, s, tt, rt}, {d, r},
                                [&](const double (&xi)[6], double (&xo)[2]) {
                xo[0] = xi[0] + alpha * xi[1] + omega * xi[2];
                const double rq = (double)(float)(xi[3] - omega * xi[4]);
                xo[1] = rq;
                red[0] += rq * rq;
                red[1] += xi[5] * rq;
              });
            }
            csum<2>(c, red);
            rho_old = rho;
            rho = red[1];
            const double rit = sqrt(red[0]);
            upd = rit <= tgt || rit <= o.ir_delta * rupd;
          }
        }
      }
    }
    const bool last = out.iters >= o.maxiter || out.nres >= o.ir_max;
    if (!(brk || conv_s || upd || last)) continue;
    // fold the correction into x and replace r by the FP64 residual
    __syncthreads();
    {
      PhaseTimer ptm(c, 4);
      ir_fold(x64, d, rn0, lo, hi);
    }
    if (GRP) grp_barrier(c.gv);   // x's halo lines
    else __syncthreads();
    const Sums2 z = spmv_res32<GRP>(c, A64, x64, b64, rt, r, sc);
    ++out.nres;
    out.res = sqrt(z.s0) / nrm0;
    out.status = out.res <= o.tol ? 0 : 4;
    if (out.res <= target) return out;
    if (brk && since <= 1) { out.status = 2; return out; }
    if (last) return out;
    rupd = sqrt(z.s0) * sc;
    if (brk || conv_s) {
      // restart the recurrences from the true residual: rt = r, rho = (r, r)
      red[0] = 0.0;
      __syncthreads();
      {
        PhaseTimer ptm(c, 4);
        vmap<float, 1, 1>(lo, hi, {r}, {rt}, [&](const double (&xi)[1], double (&xo)[1]) {
          xo[0] = xi[0];
          red[0] += xi[0] * xi[0];
        });
      }
      csum<1>(c, *reinterpret_cast<double(*)[1]>(red));
      __syncthreads();
      rho = red[0];
      fresh = true;
      since = 0;
    } else {
      rho = z.s1;   // (rt, r) of the replaced residual
    }
  }
}

// CG over the rows [c.cb, c.ce) of a member group's member, with Jacobi
// (P.kind 1: z = r * 1/diag, P.F the inverse diagonal per value) or no
// preconditioner, fused into the update pass.  With x0zero (the
// refinement's inner solves) x = 0 on entry and r0 = b; else r0 = b - A x0
// (conv_cond 0: relative to ||r0||, 1: to ||b||).  Per iteration a group
// barrier (the halo lines of p), the sum (p, A p) and the sums (r, r),
// (r, z): every block takes the same branches.
template <typename TA, typename TV>
YG_DEV SolveOut cg_grp(const Ctx& c, const TA* A, const Precon<TV>& P, const TV* b, TV* x, TV* W,
                       const YgKrylovOpts& o, bool x0zero = true) {
  const int lo = 2 * c.cb, hi = 2 * c.ce;
  const size_t vl = c.vlen;
  TV* __restrict__ r = W;
  TV* __restrict__ z = W + 1 * vl;
  TV* __restrict__ p = W + 2 * vl;
  TV* __restrict__ q = W + 3 * vl;
  const bool jac = P.kind == 1;
  const TV* dinv = P.F;
  SolveOut out{0, 4, 0.0};
  double red[2] = {0.0, 0.0};
  double bb = 0.0;   // (b, b) when x0 is not zero
  if (!x0zero) {
    // r = b - A x0 in place of b below
    grp_barrier(c.gv);   // the halo lines of x
    const Sums2 zs = spmv_dots<YG_SPMV_RES, TA, TV, true>(c, A, x, q, b, r, nullptr);
    bb = zs.s1;
    b = r;
  }
  // r = b, z = M^-1 r, p = z; (r, r), (r, z)
  {
    PhaseTimer ptm(c, 4);
    if (jac) {
      vmap<TV, 2, 3>(lo, hi, {b, dinv}, {r, z, p}, [&](const double (&xi)[2], double (&xo)[3]) {
        const double zz = (double)(TV)(xi[0] * xi[1]);
        xo[0] = xi[0]; xo[1] = zz; xo[2] = zz;
        red[0] += xi[0] * xi[0];
        red[1] += xi[0] * zz;
      });
    } else {
      vmap<TV, 1, 3>(lo, hi, {b}, {r, z, p}, [&](const double (&xi)[1], double (&xo)[3]) {
        xo[0] = xo[1] = xo[2] = xi[0];
        red[0] += xi[0] * xi[0];
        red[1] += xi[0] * xi[0];
      });
    }
  }
  csum<2>(c, red);
  const double nrm0 = !x0zero && o.conv_cond == 1 ? sqrt(bb) : sqrt(red[0]);
  if (nrm0 == 0.0) { out.status = 0; return out; }
  const double bnrm = 1.0 / nrm0;
  out.resid = sqrt(red[0]) * bnrm;
  if (out.resid <= o.tol) { out.status = 0; return out; }
  double rho = red[1];
  for (int it = 1; it <= o.maxiter; ++it) {
    if (rho == 0.0) { out.status = 2; out.iters = it; return out; }
    grp_barrier(c.gv);   // the halo lines of p
    const double pq = spmv_dots<YG_SPMV_DOT, TA, TV, true>(c, A, p, q, p, nullptr, nullptr).s0;
    if (pq == 0.0) { out.status = 2; out.iters = it; return out; }
    const double alpha = rho / pq;
    // x += alpha p, r -= alpha q, z = M^-1 r; (r, r), (r, z)
    red[0] = red[1] = 0.0;
    {
      PhaseTimer ptm(c, 4);
      if (jac) {
        vmap<TV, 5, 3>(lo, hi, {x, p, r, q, dinv}, {x, r, z},
                       [&](const double (&xi)[5], double (&xo)[3]) {
          xo[0] = xi[0] + alpha * xi[1];
          const double rq = (double)(TV)(xi[2] - alpha * xi[3]);
          const double zq = (double)(TV)(rq * xi[4]);
          xo[1] = rq; xo[2] = zq;
          red[0] += rq * rq;
          red[1] += rq * zq;
        });
      } else {
        vmap<TV, 4, 3>(lo, hi, {x, p, r, q}, {x, r, z}, [&](const double (&xi)[4], double (&xo)[3]) {
          xo[0] = xi[0] + alpha * xi[1];
          const double rq = (double)(TV)(xi[2] - alpha * xi[3]);
          xo[1] = xo[2] = rq;
          red[0] += rq * rq;
          red[1] += rq * rq;
        });
      }
    }
    csum<2>(c, red);
    out.resid = sqrt(red[0]) * bnrm;
    out.iters = it;
    if (out.resid <= o.tol) { out.status = 0; return out; }
    const double beta = red[1] / rho;
    rho = red[1];
    {
      PhaseTimer ptm(c, 4);
      vmap<TV, 2, 1>(lo, hi, {z, p}, {p}, [&](const double (&xi)[2], double (&xo)[1]) {
        xo[0] = xi[0] + beta * xi[1];
      });
    }
  }
  out.status = 4;
  return out;
}

template <typename TA, typename TV>
YG_DEV SolveOut cg(const Ctx& c, const TA* A, const Precon<TV>& P, const TV* b, TV* x, TV* W,
                   const YgKrylovOpts& o) {
  const int n2 = 2 * c.ncell;
  const size_t vl = c.vlen;
  TV* r = W;
  TV* z = W + 1 * vl;
  TV* p = W + 2 * vl;
  TV* qv = W + 3 * vl;
  TV* ytmp = W + 8 * vl;
  SolveOut out{0, 4, 0.0};
  double red[2];
  {
    const Sums2 z = spmv_dots<YG_SPMV_RES, TA, TV>(c, A, x, qv, b, r, nullptr);
    red[0] = z.s0;
    red[1] = z.s1;
  }
  const double nrm0 = o.conv_cond == 1 ? sqrt(red[1]) : sqrt(red[0]);
  if (nrm0 == 0.0) { out.status = 0; return out; }
  const double bnrm = 1.0 / nrm0;
  out.resid = sqrt(red[0]) * bnrm;
  if (out.resid <= o.tol) { out.status = 0; return out; }
  double rho_old = 0.0;
  for (int it = 1; it <= o.maxiter; ++it) {
    precon_apply<TV>(c, P, r, ytmp, z);
    double rr[1] = {0.0};
    for (int q = threadIdx.x; q < n2; q += blockDim.x) rr[0] += (double)r[q] * (double)z[q];
    yg_block_sum<1>(rr, yg_red);
    const double rho = rr[0];
    if (rho == 0.0) { out.status = 2; out.iters = it; return out; }
    const double beta = it == 1 ? 0.0 : rho / rho_old;
    for (int q = threadIdx.x; q < n2; q += blockDim.x)
      p[q] = it == 1 ? z[q] : (TV)((double)z[q] + beta * (double)p[q]);
    __syncthreads();
    double pq[1];
    pq[0] = spmv_dots<YG_SPMV_DOT, TA, TV>(c, A, p, qv, p, nullptr, nullptr).s0;
    if (pq[0] == 0.0) { out.status = 2; out.iters = it; return out; }
    const double alpha = rho / pq[0];
    double rn[1] = {0.0};
    for (int q = threadIdx.x; q < n2; q += blockDim.x) {
      x[q] = (TV)((double)x[q] + alpha * (double)p[q]);
      r[q] = (TV)((double)r[q] - alpha * (double)qv[q]);
      rn[0] += (double)r[q] * (double)r[q];
    }
    yg_block_sum<1>(rn, yg_red);
    out.resid = sqrt(rn[0]) * bnrm;
    out.iters = it;
    if (out.resid <= o.tol) { out.status = 0; return out; }
    rho_old = rho;
  }
  out.status = 4;
  return out;
}

// Preconditioner residency: a member whose previous solve of the same
// sequence left its ILU(0) factors in place (warm flag) uses them in the
// first Picard iteration of the next solve, whose matrix differs from the
// last one factorized by one thickness update; later iterations factorize
// their own matrices.  Mixed mode only (the FP64 path factorizes every call,
// as the reference's solver does).
YG_DEV bool yg_keep_factors(const YgKrylovArgs& a, int m) {
  return a.opts.keep_factors && a.opts.precon == 2 && a.warm && a.iter_now && a.warm[m] != 0 &&
         a.iter_now[m] == 0;
}

// Loose refinement (opts.ir_loose): a member whose last Picard residual is
// above ir_loose refines only to the requested tolerance.
YG_DEV bool yg_loose(const YgKrylovArgs& a, int m) {
  return a.opts.ir_loose > 0.0 && a.picard_resid && a.picard_resid[m] > a.opts.ir_loose;
}

template <typename T>
YG_DEV void build_precon(const Ctx& c, const T* A, T* F, Precon<T>& P, int kind, int* bad) {
  P.kind = kind;
  P.F = F;
  if (kind == 2) {
    ilu0_factor<T>(c, A, F, bad);
  } else if (kind == 1) {
    for (int q = threadIdx.x; q < c.ncell; q += blockDim.x) {
      F[2 * q] = (T)(1.0 / (double)A[(size_t)q * YG_NA + 0]);
      F[2 * q + 1] = (T)(1.0 / (double)A[(size_t)q * YG_NA + 1]);
    }
    __syncthreads();
  }
}

template <typename T>
YG_DEV SolveOut krylov(const Ctx& c, const T* A, const Precon<T>& P, const T* b, T* x, T* W,
                       const YgKrylovOpts& o) {
  return o.solver == 1 ? cg<T, T>(c, A, P, b, x, W, o)
                       : bicgstab<T, T>(c, A, P, b, x, W, o);
}

}  // namespace

// ---------------------------------------------------------------------------
// The kernel: one block per member.  Modes:
//   YG_MODE_FP64  : FP64 matrix, factors, vectors (the plain FP64 path)
//   YG_MODE_MIXED : FP64 residual b - A x and correction accumulation,
//                   FP32 matrix copy, factors and inner Krylov solve
template <int MODE, int NT>
__global__ __launch_bounds__(NT) void yg_krylov_kernel(YgKrylovArgs a) {
  const int m = blockIdx.x;
  if (a.skip && a.skip[m]) return;
  const YgMember mb = a.members[m];
  __shared__ int s_bad;
  const int ncell = mb.nx * mb.ny;
  for (int k = threadIdx.x; k <= mb.T; k += blockDim.x) yg_soff[k] = a.soff[mb.toff + k];
  if (threadIdx.x == 0) s_bad = 0;
  __syncthreads();
  Ctx c;
  c.mb = &a.members[m];
  c.nx = mb.nx; c.ny = mb.ny; c.T = mb.T; c.ncell = ncell; c.vlen = mb.vlen;
  c.wrapx = a.wrapx; c.wrapy = a.wrapy;
  c.mask = a.mask + mb.coff;
  c.maxcnt = mb.maxcnt;
  c.fu = c.fu16 = 8ll * ncell;
  c.fd16 = 16ll * ncell;
  c.cb = 0;
  c.ce = ncell;
  c.grp = nullptr;
  c.pc = &c;
  c.lds = YG_DYN_LDS;
  {
    // line-to-wave mapping of the recurrences (exact or block ILU(0))
    const int g = threadIdx.x, lane = g & 63, w = g >> 6;
    const int wmax = (int)(blockDim.x >> 6);
    // a member's own schedule (autotuned) before the options' rule
    int nb = mb.nb != 0 ? mb.nb : a.opts.ilu_blocks;
    if (nb < 0) nb = MODE == YG_MODE_MIXED ? (mb.nx + 63) / 64 : 0;
    c.ncol = 1;
    c.mcol = mb.ny;
    c.rev = 0;
    if (nb == 0) {
      c.coupled = mb.nx > 64;
      c.nwl = (mb.nx + 63) / 64;
      c.lpw = 64;
      c.line = g < mb.nx ? g : -1;
    } else {
      nb = max(nb, (mb.nx + 63) / 64);
      nb = min(nb, min(wmax, mb.nx));
      const int lpw = (mb.nx + nb - 1) / nb;
      c.coupled = 0;
      c.lpw = lpw;
      c.nwl = (mb.nx + lpw - 1) / lpw;
      const int i = w * lpw + lane;
      c.line = (w < c.nwl && lane < lpw && i < mb.nx) ? i : -1;
      yg_tiles(mb.ncol != 0 ? mb.ncol : a.opts.ilu_cols, wmax / c.nwl, mb.ny, c);
      // reversed columns: FP32/FP64 factors, and the widest step of the
      // skew-1 wavefront (min(nx, ny) cells) fits the factorization's stage
      const bool f16 = MODE == YG_MODE_MIXED && a.opts.factor16 && a.F16;
      const int cap = factor_cap_of(YG_DYN_LDS, MODE == YG_MODE_FP64 ? 8 : 4);
      if (a.opts.ilu_rev && !f16 && min(mb.nx, mb.ny) <= cap) {
        c.rev = 1;
        c.T = mb.ny + mb.nx - 1;
        c.maxcnt = min(mb.nx, mb.ny);
        __syncthreads();   // the host's (skew 2) step table is replaced
        for (int k = threadIdx.x; k <= c.T; k += blockDim.x) yg_soff[k] = blk_soff1(k, mb.nx, mb.ny);
        __syncthreads();
      }
    }
  }
  __shared__ long long s_tm[8];
  if (threadIdx.x < 8) s_tm[threadIdx.x] = 0;
  c.tm = a.timers ? s_tm : nullptr;
  const long long t_kernel0 = wall_clock64();
  const size_t c2 = 2 * (size_t)mb.coff;
  const double* b64 = a.b + c2;
  double* x64 = a.x + c2;
  // b == 0 shortcut (solver_linear.F90:162-174)
  {
    double bm[1] = {0.0};
    bm[0] = count_nonzero(b64, 0, 2 * ncell);
    yg_block_sum<1>(bm, yg_red);
    if (bm[0] == 0.0) {
      for (int q = threadIdx.x; q < 2 * ncell; q += blockDim.x) x64[q] = 0.0;
      if (threadIdx.x == 0) { a.out_iters[m] = 0; a.out_status[m] = 0; a.out_resid[m] = 0.0; }
      return;
    }
  }
  YgKrylovOpts o = a.opts;
  if (MODE == YG_MODE_FP64) {
    const double* A = a.A64 + (size_t)YG_NA * mb.coff;
    double* F = a.F64 + (size_t)YG_NF * mb.coff;
    double* W = a.W64 + (size_t)YG_NW64 * mb.woff;
    if (o.initx_zeros) {
      for (int q = threadIdx.x; q < 2 * ncell; q += blockDim.x) x64[q] = 0.0;
      __syncthreads();
    }
    Precon<double> P;
    build_precon<double>(c, A, F, P, o.precon, &s_bad);
    SolveOut so = krylov<double>(c, A, P, b64, x64, W, o);
    if (s_bad && so.status == 0) so.status = 2;
    if (threadIdx.x == 0) { a.out_iters[m] = so.iters; a.out_status[m] = so.status; a.out_resid[m] = so.resid; }
    if (a.timers && threadIdx.x == 0) {
      s_tm[7] = wall_clock64() - t_kernel0;
      for (int k = 0; k < 8; ++k) a.timers[8 * m + k] += s_tm[k];
    }
  } else {
    // FP64 iterative refinement around an FP32 solve
    const double* A64 = a.A64 + (size_t)YG_NA * mb.coff;
    const float* A32 = a.A32 + (size_t)YG_NA * mb.coff;
    float* F = a.F32 + (size_t)YG_NF * mb.coff;
    float* W = a.W32 + (size_t)YG_NW32 * mb.woff;  // 9 inner Krylov vectors ...
    float* d = W + (size_t)9 * mb.vlen;             // ... the correction ...
    float* rin = W + (size_t)10 * mb.vlen;          // ... and the scaled residual
    double* r64 = a.W64 + (size_t)YG_NW64 * mb.woff; // FP64 residual ...
    double* y64 = r64 + mb.vlen;                      // ... and A x
    if (o.initx_zeros) {
      for (int q = threadIdx.x; q < 2 * ncell; q += blockDim.x) x64[q] = 0.0;
      __syncthreads();
    }
    Precon<float> P;
    const bool keepf = yg_keep_factors(a, m);
    if (keepf) { P.kind = 2; P.F = F; }   // the factors of the sequence's previous solve
    else build_precon<float>(c, A32, F, P, o.precon, &s_bad);
    if (o.factor16 && P.kind == 2 && !c.coupled && a.F16) {
      uint16_t* FB = a.F16 + (size_t)YG_NF16 * mb.coff;
      if (!keepf) factors_to_bf16(c, F, FB);
      P.FB = FB;
      P.fb_l = o.factor16 == 1;
    }
    // Refinement continues past the requested tolerance to ir_target (by
    // default a tenth of it): the FP32 inner solves then never leave the FP64
    // residual just under the tolerance, which would carry the
    // solver-tolerance error of a plain solve into the Picard iterate.  A
    // member far from Picard convergence (its last Picard residual above
    // ir_loose) solves to the tolerance itself, with looser inner solves: the
    // Picard iteration contracts that error away before it converges.  The
    // reported status refers to the requested tolerance.
    const bool loose = yg_loose(a, m);
    YgKrylovOpts oi = o;
    oi.tol = loose ? o.loose_inner : o.inner_tol;
    oi.conv_cond = 0;
    oi.initx_zeros = 1;
    double nrm0 = -1.0, res = 1.0;
    int iters = 0, status = 4;
    const double target = loose ? o.tol : o.ir_target > 0.0 ? o.ir_target : 1e-1 * o.tol;
    int nres = 0;   // FP64 residual passes
    const bool reliable = o.ir_delta > 0.0 && o.solver != 1;
    if (reliable) {
      YgKrylovOpts orl = o;
      orl.ir_target = target;
      const IrOut io = ir_bicgstab<false>(c, A64, A32, P, b64, x64, r64, y64, W, d, orl);
      iters = io.iters; status = io.status; res = io.res; nres = io.nres;
    }
    for (int ir = 0; !reliable && ir < o.ir_max; ++ir) {
      ++nres;
      double red[2];
      {
        const Sums2 z = spmv_dots<YG_SPMV_RES, double, double>(c, A64, x64, y64, b64, r64, nullptr);
        red[0] = z.s0;
        red[1] = z.s1;
      }
      const double rn = sqrt(red[0]);
      if (nrm0 < 0.0) nrm0 = o.conv_cond == 1 ? sqrt(red[1]) : rn;
      if (nrm0 == 0.0) { res = 0.0; status = 0; break; }
      res = rn / nrm0;
      status = res <= o.tol ? 0 : 4;
      if (res <= target) break;
      if (iters >= o.maxiter) break;
      const double sc = rn > 0.0 ? 1.0 / rn : 1.0;
      refine_scale(r64, rin, d, sc, 0, 2 * ncell);
      __syncthreads();
      oi.maxiter = o.maxiter - iters;
      SolveOut so = krylov<float>(c, A32, P, rin, d, W, oi);
      iters += so.iters;
      refine_update(x64, d, rn, 0, 2 * ncell);
      __syncthreads();
      if (so.status == 2 && so.iters <= 1) { status = 2; break; }
    }
    if (s_bad && status == 0) status = 2;
    if (threadIdx.x == 0) {
      a.out_iters[m] = iters; a.out_status[m] = status; a.out_resid[m] = res;
      if (a.out_sweeps) a.out_sweeps[m] = nres;
    }
    if (a.timers && threadIdx.x == 0) {
      s_tm[7] = wall_clock64() - t_kernel0;
      for (int k = 0; k < 8; ++k) a.timers[8 * m + k] += s_tm[k];
    }
  }
}

// Dynamic-LDS limits of both modes, once per process (before any launch, so
// that no attribute call happens while a stream is being captured).
template <int NT>
void yg_krylov_attrs_nt() {
  static bool done = false;
  if (done) return;
  (void)hipFuncSetAttribute((const void*)yg_krylov_kernel<YG_MODE_FP64, NT>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, YG_DYN_LDS);
  (void)hipFuncSetAttribute((const void*)yg_krylov_kernel<YG_MODE_MIXED, NT>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, YG_DYN_LDS);
  done = true;
}

// Launch of this translation unit's block size.
template <int NT>
void yg_launch_krylov_nt(const YgKrylovArgs& a, int mode, int nmembers, int block,
                         hipStream_t st) {
  yg_krylov_attrs_nt<NT>();
  if (mode == YG_MODE_FP64)
    hipLaunchKernelGGL((yg_krylov_kernel<YG_MODE_FP64, NT>), dim3(nmembers), dim3(block),
                       YG_DYN_LDS, st, a);
  else
    hipLaunchKernelGGL((yg_krylov_kernel<YG_MODE_MIXED, NT>), dim3(nmembers), dim3(block),
                       YG_DYN_LDS, st, a);
}

// ---------------------------------------------------------------------------
// The member-group kernel (mixed mode, block-Jacobi ILU(0), BiCGSTAB): G
// workgroups share each member.  Workgroup b is rank (b/8) % G of group
// (b/8/G)*8 + b%8, so the workgroups of a group run on one XCD (the groups
// of a launch spread evenly over the 8); group gi takes the members gi,
// gi + ngroups, ...  Rank r owns the line blocks [r nbr, (r+1) nbr) of its
// member (nbr = ceil(line blocks / G)): it factorizes them and runs their
// sweeps, and its SpMV rows and vector passes are the cells of their lines.
// Line blocks do not couple in block-Jacobi ILU(0), so the preconditioner
// needs no exchange; the SpMV reads the neighbouring ranks' edge lines after
// a group barrier.  The iteration is the one-block kernel's mixed mode
// (FP64 refinement around FP32 BiCGSTAB) with every sum a group sum.
// Dynamic LDS of the group kernel: 120 KB with 512 threads (one workgroup
// per CU); 56 KB with 256 threads, so that two workgroups (of different
// members) share a CU, each with at most 256 VGPRs per lane.
template <int NT> constexpr int yg_group_lds() { return NT == 512 ? YG_DYN_LDS : 56 * 1024; }

template <int NT>
__global__ __launch_bounds__(NT, 2) void yg_krylov_group_kernel(YgKrylovArgs a) {
  const int blk = blockIdx.x;
  int G, gi, rank, k0, k1, ks;   // this workgroup's group, its members k0, k0 + ks, ... < k1
  if (a.gtab) {
    // groups of different sizes: this workgroup's place from the table
    gi = a.gtab[3 * blk];
    if (gi < 0) return;   // no part in any group
    rank = a.gtab[3 * blk + 1];
    G = a.gtab[3 * blk + 2];
    k0 = a.gmoff[gi]; k1 = a.gmoff[gi + 1]; ks = 1;
  } else {
    G = a.group;
    const int slot = blk >> 3;
    gi = (slot / G) * 8 + (blk & 7);
    rank = slot % G;
    k0 = gi; k1 = a.nmembers; ks = gridDim.x / G;
  }
  unsigned char* gs = reinterpret_cast<unsigned char*>(a.gsync) + (size_t)gi * YG_GSYNC_BYTES;
  Grp grp;
  grp.cnt = reinterpret_cast<int*>(gs);
  grp.err = reinterpret_cast<int*>(gs) + 1;
  grp.part = reinterpret_cast<double*>(gs + 128);   // 2 x YG_GMAX x 4 doubles (ints 16..31: XCDs)
  grp.G = G;
  grp.rank = rank;
  __shared__ int s_bad;
  __shared__ long long s_tm[8];
  grp_start(grp, reinterpret_cast<int*>(gs) + 16);
  for (int k = k0; k < k1; k += ks) {
    const int m = a.gtab ? a.gmem[k] : k;
    if (a.skip && a.skip[m]) continue;
    const YgMember mb = a.members[m];
    const int nx = mb.nx, ny = mb.ny, ncell = nx * ny;
    // the member's line blocks (at most 64 lines each, at most one per wave
    // of the group) and this rank's lines [i0, i1)
    int nb = mb.nb != 0 ? mb.nb : a.opts.ilu_blocks;
    if (nb <= 0) nb = (nx + 63) / 64;
    nb = ((max(nb, 1) + G - 1) / G) * G;   // the same number of blocks on every rank
    nb = min(max(nb, (nx + 63) / 64), min(G * (NT / 64), nx));
    const int lpw = (nx + nb - 1) / nb;
    const int nwl = (nx + lpw - 1) / lpw;
    const int nbr = (nwl + G - 1) / G;
    const int i0 = min(nx, rank * nbr * lpw), i1 = min(nx, (rank + 1) * nbr * lpw);
    const int Lg = i1 - i0;
    // reversed columns (skew 1): FP32 factors, and the widest step (min(Lg,
    // ny) cells) fits the factorization's stage
    const int rev = a.opts.ilu_rev && !(a.opts.factor16 && a.F16) && Lg > 0 &&
                    min(Lg, ny) <= factor_cap_of(yg_group_lds<NT>(), 4);
    const int T = Lg > 0 ? ny + (rev ? 1 : 2) * (Lg - 1) : 0;
    __syncthreads();   // the previous member's step table is no longer read
    for (int k = threadIdx.x; k <= T; k += blockDim.x)
      yg_soff[k] = rev ? blk_soff1(k, Lg, ny) : blk_soff(k, Lg, ny);
    if (threadIdx.x == 0) { s_bad = 0; yg_gwait = 0; }
    if (threadIdx.x < 8) s_tm[threadIdx.x] = 0;
    __syncthreads();
    const long long t_kernel0 = wall_clock64();
    Ctx cm;   // the member: SpMV rows, vector passes and sums of this rank
    cm.mb = &a.members[m];
    cm.nx = nx; cm.ny = ny; cm.T = mb.T; cm.ncell = ncell; cm.vlen = mb.vlen;
    cm.wrapx = a.wrapx; cm.wrapy = a.wrapy;
    cm.mask = a.mask + mb.coff;
    cm.maxcnt = mb.maxcnt;
    cm.line = -1; cm.coupled = 0; cm.nwl = 0; cm.lpw = lpw;
    cm.ncol = 1; cm.mcol = ny;
    cm.rev = 0;
    cm.tm = a.timers ? s_tm : nullptr;
    cm.fu = cm.fu16 = 8ll * ncell;
    cm.fd16 = 16ll * ncell;
    cm.cb = i0 * ny; cm.ce = i1 * ny;
    cm.grp = &grp;
    cm.gv = grp;
    cm.lds = yg_group_lds<NT>();
    Ctx cr = cm;   // this rank's lines: factorization and sweeps
    cr.nx = Lg; cr.T = T; cr.ncell = Lg * ny;
    cr.mask = cm.mask + cm.cb;
    cr.maxcnt = rev ? min(Lg, ny) : min(Lg, (ny + 1) / 2);
    cr.rev = rev;
    cr.nwl = (Lg + lpw - 1) / lpw;
    {
      const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, i = w * lpw + lane;
      cr.line = (w < cr.nwl && lane < lpw && i < Lg) ? i : -1;
    }
    // the same tiles on every rank: as many as the fullest rank's waves allow
    yg_tiles(mb.ncol != 0 ? mb.ncol : a.opts.ilu_cols, (NT / 64) / nbr, ny, cr);
    cr.fu = 8ll * ncell + 4ll * cm.cb;     // the member's U part from this part's L part
    cr.fu16 = 8ll * ncell;
    cr.fd16 = 16ll * ncell - 4ll * cm.cb;
    cr.cb = 0; cr.ce = cr.ncell;
    cr.grp = nullptr;
    cr.pc = &cr;
    cm.pc = &cr;
    const size_t c2 = 2 * (size_t)mb.coff;
    const double* b64 = a.b + c2;
    double* x64 = a.x + c2;
    const int lo = 2 * cm.cb, hi = 2 * cm.ce;
    // b == 0 shortcut (solver_linear.F90:162-174)
    {
      double bm[1] = {0.0};
      bm[0] = count_nonzero(b64, lo, hi);
      csum<1>(cm, bm);
      if (bm[0] == 0.0) {
        for (int q = lo + threadIdx.x; q < hi; q += blockDim.x) x64[q] = 0.0;
        if (rank == 0 && threadIdx.x == 0) {
          a.out_iters[m] = 0; a.out_status[m] = 0; a.out_resid[m] = 0.0;
        }
        continue;
      }
    }
    const YgKrylovOpts o = a.opts;
    const double* A64 = a.A64 + (size_t)YG_NA * mb.coff;
    const float* A32 = a.A32 + (size_t)YG_NA * mb.coff;
    float* F = a.F32 + (size_t)YG_NF * mb.coff;
    float* W = a.W32 + (size_t)YG_NW32 * mb.woff;   // 9 inner Krylov vectors ...
    float* d = W + (size_t)9 * mb.vlen;              // ... the correction ...
    float* rin = W + (size_t)10 * mb.vlen;           // ... and the scaled residual
    double* r64 = a.W64 + (size_t)YG_NW64 * mb.woff; // FP64 residual ...
    double* y64 = r64 + mb.vlen;                      // ... and A x
    if (o.initx_zeros) {
      for (int q = lo + threadIdx.x; q < hi; q += blockDim.x) x64[q] = 0.0;
    }
    if (a.plain64) {
      // the plain FP64 path with Jacobi or no preconditioner: the
      // reference's Krylov solve from x0 (LIS semantics), over the group
      double* F64 = a.F64 + (size_t)YG_NF * mb.coff;
      Precon<double> P64;
      P64.kind = o.precon;
      P64.F = F64;
      if (o.precon == 1) {
        for (int q = cm.cb + threadIdx.x; q < cm.ce; q += blockDim.x) {
          F64[2 * q] = 1.0 / A64[(size_t)q * YG_NA + 0];
          F64[2 * q + 1] = 1.0 / A64[(size_t)q * YG_NA + 1];
        }
        __syncthreads();
      }
      const SolveOut so = o.solver == 1 ? cg_grp<double, double>(cm, A64, P64, b64, x64, r64, o, false)
                                        : bicgstab<double, double, true>(cm, A64, P64, b64, x64, r64, o, false);
      int st = so.status;
      if (yg_gfail) {
        st = 2;
        if (threadIdx.x == 0) __hip_atomic_store(a.group_fail, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      if (rank == 0 && threadIdx.x == 0) {
        a.out_iters[m] = so.iters; a.out_status[m] = st; a.out_resid[m] = so.resid;
        if (a.timers) {
          s_tm[5] = yg_gwait;
          s_tm[7] = wall_clock64() - t_kernel0;
          for (int k = 0; k < 8; ++k) a.timers[8 * m + k] += s_tm[k];
        }
      }
      continue;
    }
    Precon<float> P;
    if (o.precon == 2) {
      // a warm-started member's first Picard iteration keeps the factors of
      // its sequence's previous solve (same schedule, same cells)
      const bool keepf = yg_keep_factors(a, m);
      P.kind = 2;
      P.F = F + (size_t)8 * cm.cb;   // this part's L factors (rotated tile order)
      if (Lg > 0 && !keepf) ilu0_factor<float>(cr, A32 + (size_t)YG_NA * cm.cb, F + (size_t)8 * cm.cb, &s_bad);
      if (o.factor16 && a.F16) {
        uint16_t* FB = a.F16 + (size_t)YG_NF16 * mb.coff + (size_t)8 * cm.cb;
        if (Lg > 0 && !keepf) factors_to_bf16(cr, P.F, FB);
        P.FB = FB;
        P.fb_l = o.factor16 == 1;
      }
    } else {
      // Jacobi (the inverse diagonal of this rank's rows, by value in the
      // member's row order) or none
      P.kind = o.precon;
      P.F = F;
      if (o.precon == 1) {
        for (int q = cm.cb + threadIdx.x; q < cm.ce; q += blockDim.x) {
          F[2 * q] = (float)(1.0 / (double)A32[(size_t)q * YG_NA + 0]);
          F[2 * q + 1] = (float)(1.0 / (double)A32[(size_t)q * YG_NA + 1]);
        }
        __syncthreads();
      }
    }
    double bad[1] = {(double)s_bad};
    csum<1>(cm, bad);   // a zero pivot on any rank
    // a member far from Picard convergence solves to the tolerance itself
    const bool loose = yg_loose(a, m);
    YgKrylovOpts oi = o;
    oi.tol = loose ? o.loose_inner : o.inner_tol;
    oi.conv_cond = 0;
    oi.initx_zeros = 1;
    double nrm0 = -1.0, res = 1.0;
    int iters = 0, status = 4;
    const double target = loose ? o.tol : o.ir_target > 0.0 ? o.ir_target : 1e-1 * o.tol;
    int nres = 0;   // FP64 residual passes
    const bool reliable = o.ir_delta > 0.0 && o.solver != 1;
    if (reliable) {
      YgKrylovOpts orl = o;
      orl.ir_target = target;
      const IrOut io = ir_bicgstab<true>(cm, A64, A32, P, b64, x64, r64, y64, W, d, orl);
      iters = io.iters; status = io.status; res = io.res; nres = io.nres;
    }
    for (int ir = 0; !reliable && ir < o.ir_max; ++ir) {
      grp_barrier(grp);   // x's edge lines are current on every rank
      ++nres;
      double red[2];
      {
        const Sums2 z = spmv_dots<YG_SPMV_RES, double, double, true>(cm, A64, x64, y64, b64, r64, nullptr);
        red[0] = z.s0;
        red[1] = z.s1;
      }
      const double rn = sqrt(red[0]);
      if (nrm0 < 0.0) nrm0 = o.conv_cond == 1 ? sqrt(red[1]) : rn;
      if (nrm0 == 0.0) { res = 0.0; status = 0; break; }
      res = rn / nrm0;
      status = res <= o.tol ? 0 : 4;
      if (res <= target) break;
      if (iters >= o.maxiter) break;
      const double sc = rn > 0.0 ? 1.0 / rn : 1.0;
      refine_scale(r64, rin, d, sc, lo, hi);
      __syncthreads();
      oi.maxiter = o.maxiter - iters;
      SolveOut so = o.solver == 1 ? cg_grp<float, float>(cm, A32, P, rin, d, W, oi)
                                  : bicgstab<float, float, true>(cm, A32, P, rin, d, W, oi);
      iters += so.iters;
      refine_update(x64, d, rn, lo, hi);
      __syncthreads();
      if (so.status == 2 && so.iters <= 1) { status = 2; break; }
    }
    if (bad[0] != 0.0 && status == 0) status = 2;
    if (yg_gfail) {
      status = 2;
      if (threadIdx.x == 0) __hip_atomic_store(a.group_fail, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (rank == 0 && threadIdx.x == 0) {
      a.out_iters[m] = iters; a.out_status[m] = status; a.out_resid[m] = res;
      if (a.out_sweeps) a.out_sweeps[m] = nres;
      if (a.timers) {
        s_tm[5] = yg_gwait;
        s_tm[7] = wall_clock64() - t_kernel0;
        for (int k = 0; k < 8; ++k) a.timers[8 * m + k] += s_tm[k];
      }
    }
  }
}

namespace {
__global__ void k_clear_gsync(int* g, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) g[k] = 0;
}
}  // namespace

template <int NT>
void yg_krylov_group_attrs_nt() {
  static bool done = false;
  if (done) return;
  (void)hipFuncSetAttribute((const void*)yg_krylov_group_kernel<NT>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, yg_group_lds<NT>());
  done = true;
}

template <int NT>
void yg_launch_krylov_group_nt(const YgKrylovArgs& a, int ngroups, hipStream_t st) {
  yg_krylov_group_attrs_nt<NT>();
  const int ng = a.gtab ? a.ngrp : ngroups;
  const int nwords = ng * (YG_GSYNC_BYTES / 4);
  hipLaunchKernelGGL(k_clear_gsync, dim3((nwords + 255) / 256), dim3(256), 0, st, a.gsync, nwords);
  hipLaunchKernelGGL(yg_krylov_group_kernel<NT>, dim3(a.gtab ? a.nwg : ngroups * a.group), dim3(NT),
                     yg_group_lds<NT>(), st, a);
}
