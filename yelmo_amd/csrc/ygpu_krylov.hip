// ygpu_krylov.hip -- persistent per-system Krylov solve on the stencil matrix.
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
#include "ygpu_common.h"
#include "ygpu_krylov.h"

namespace {

// --------------------------------------------------------------------------
// per-block context
struct Ctx {
  const YgMember* mb;
  int nx, ny, T;
  int ncell;
  int wrapx, wrapy;
  const int* soff;            // LDS copy of the step offsets (T+1)
  const uint32_t* mask;       // per natural cell: presence bits of its two rows
  double* red;                // LDS reduction scratch
  int maxcnt;                 // most cells in one wavefront step
  long long* tm;              // phase timers (thread 0), may be null
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

// y = A x for the two rows of natural cell n = i*ny + j (the reference's
// unknown numbering: rows 2n, 2n+1).  Neighbours are at fixed offsets.
template <typename TA, typename TV>
YG_DEV void spmv_cell(const Ctx& c, const TA* __restrict__ A, const TV* __restrict__ x, int n,
                      double& yx, double& yy) {
  const int i = n / c.ny, j = n - i * c.ny;
  const uint32_t m = c.mask[n];
  double acc[2] = {0.0, 0.0};
#pragma unroll
  for (int comp = 0; comp < 2; ++comp) {
#pragma unroll
    for (int s = 0; s < YG_NSLOT; ++s) {
      if (!((m >> (16 * comp + s)) & 1u)) continue;
      const int8_t* sl = comp == 0 ? kSlotUx[s] : kSlotUy[s];
      int i2 = i + sl[0], j2 = j + sl[1];
      if (c.wrapx) i2 = (i2 + c.nx) % c.nx;
      if (c.wrapy) j2 = (j2 + c.ny) % c.ny;
      acc[comp] += (double)A[((size_t)s * c.ncell + n) * 2 + comp] *
                   (double)x[2 * (i2 * c.ny + j2) + sl[2]];
    }
  }
  yx = acc[0];
  yy = acc[1];
}

// Cell of a wavefront chunk handled by thread tid: the chunk's steps
// t0..t0+C-1 hold at most C*maxcnt cells; thread tid takes step offset
// tid / maxcnt and line offset tid % maxcnt.  Returns false if that slot is
// past the step's last cell.  Gives the skewed index q and natural index n.
YG_DEV bool chunk_cell(const Ctx& c, int t0, int t1, int tid, int& q, int& nat) {
  const int dt = tid / c.maxcnt, k = tid - dt * c.maxcnt;
  const int t = t0 + dt;
  if (t >= t1) return false;
  const int cnt = c.soff[t + 1] - c.soff[t];
  if (k >= cnt) return false;
  const int i = yg_glo(t, c.ny) + k, j = t - 2 * i;
  q = c.soff[t] + k;
  nat = i * c.ny + j;
  return true;
}

// ---- LDS staging of the wavefront sweeps ------------------------------------
// A wavefront sweep touches, at step t, the contiguous cells [soff[t],
// soff[t+1]) of the skewed order.  The steps are walked in chunks whose
// inputs fit one LDS buffer; while the lanes run the recurrence of one chunk
// out of LDS, every thread of the block already has the next chunk's global
// loads in flight (in registers), so the per-step cost is the recurrence and
// not a memory round trip.

#define YG_DYN_LDS (120 * 1024)

// Barrier of one wavefront step: waits for this wave's LDS traffic (the ring
// of boundary values, the staged outputs) and nothing else, so global loads
// and stores stay in flight across it.  The memory clobber keeps the
// compiler from moving LDS accesses across it.
YG_DEV void step_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// The stage buffers live in the kernel's dynamic LDS.  They are always
// addressed from this symbol (never through a stored pointer) so that the
// compiler emits LDS instructions for them.
extern __shared__ __align__(16) unsigned char yg_dyn_lds[];

template <typename T> YG_DEV T mask_to_stage(uint32_t m);
template <> YG_DEV float mask_to_stage<float>(uint32_t m) { return __uint_as_float(m); }
template <> YG_DEV double mask_to_stage<double>(uint32_t m) { return (double)m; }
template <typename T> YG_DEV uint32_t stage_to_mask(T v);
template <> YG_DEV uint32_t stage_to_mask<float>(float v) { return __float_as_uint(v); }
template <> YG_DEV uint32_t stage_to_mask<double>(double v) { return (uint32_t)v; }

// Neighbour lanes exchange values by DPP wave shifts (a VALU operation).
YG_DEV float dpp_shr1(float v) {   // lane i <- lane i-1, lane 0 <- 0
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xf, 0xf, false));
}
YG_DEV float dpp_shl1(float v) {   // lane i <- lane i+1, lane 63 <- 0
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xf, 0xf, false));
}
YG_DEV double dpp_shr1(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), 0x138, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x138, 0xf, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
YG_DEV double dpp_shl1(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), 0x130, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x130, 0xf, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// ---- ILU(0) factorization as a wavefront over grid lines -----------------
// lane i (one per line) handles cell (i, j = t - 2 i) at step t.  The U rows
// of the two most recent cells of each lane stay in registers; lane i reads
// lane i-1's through wave shuffles, and across waves through an LDS ring.
struct URow {   // U part of the two rows of one cell
  double x5, x3, x7, x1, x6, xd;  // ux row: slots 5,3,7,1,6, 1/diag
  double y6, y1, y3, yd;          // uy row: slots 6,1,3, 1/diag
};

YG_DEV void urow_zero(URow& u) {
  u.x5 = u.x3 = u.x7 = u.x1 = u.x6 = u.xd = 0.0;
  u.y6 = u.y1 = u.y3 = u.yd = 0.0;
}

YG_DEV double shup(double v) { return dpp_shr1(v); }

// TA is both the matrix and the factor type (FP64 path: double; the FP32
// inner solve of the mixed path: float)
template <typename TA>
YG_DEV void ilu0_factor(const Ctx& c, const TA* __restrict__ A, TA* __restrict__ F,
                        double (*ring)[3][10], int* bad) {
  PhaseTimer ptm(c, 0);
  const int g = threadIdx.x, lane = g & 63, w = g >> 6;
  const bool multi = c.nx > 64;
  const size_t nc = c.ncell;
  constexpr int NV = 19;  // 18 slot values + the presence mask
  const int capc = min((int)blockDim.x, (int)(YG_DYN_LDS / sizeof(TA)) / (2 * NV));
  const int C = max(1, capc / c.maxcnt);
  TA* S[2] = {reinterpret_cast<TA*>(yg_dyn_lds), reinterpret_cast<TA*>(yg_dyn_lds) + capc * NV};
  TA reg[NV];
  auto load = [&](int t0, int& cl) {
    const int t1 = min(c.T, t0 + C);
    cl = c.soff[t0];
    int q, nat;
    if (chunk_cell(c, t0, t1, g, q, nat)) {
#pragma unroll
      for (int v = 0; v < 18; ++v) reg[v] = A[((size_t)(v >> 1) * nc + nat) * 2 + (v & 1)];
      reg[18] = mask_to_stage<TA>(c.mask[nat]);
      return q - cl;
    }
    return -1;
  };
  auto stash = [&](TA* B, int l, int n) {
    if (l >= 0) {
#pragma unroll
      for (int v = 0; v < NV; ++v) B[v * n + l] = reg[v];
    }
  };
  URow o1, o2;  // own cells at steps t-1 and t-2
  urow_zero(o1);
  urow_zero(o2);
  const int nch = (c.T + C - 1) / C;
  {
    int cl;
    const int l = load(0, cl);
    stash(S[0], l, c.soff[min(c.T, C)] - cl);
  }
  __syncthreads();
  for (int ch = 0; ch < nch; ++ch) {
    const int t0 = ch * C, t1 = min(c.T, t0 + C);
    const TA* B = S[ch & 1];
    const int clo = c.soff[t0], n = c.soff[t1] - clo;
    for (int t = t0; t < t1; ++t) {
      // P = line i-1 cell j (its step t-2), Q = line i-1 cell j+1 (its step t-1)
      URow P, Q;
      P.x5 = shup(o2.x5); P.x3 = shup(o2.x3); P.x7 = shup(o2.x7); P.x1 = shup(o2.x1);
      P.x6 = shup(o2.x6); P.xd = shup(o2.xd); P.y6 = shup(o2.y6); P.y3 = shup(o2.y3);
      P.yd = shup(o2.yd); P.y1 = 0.0;
      Q.x7 = shup(o1.x7); Q.x1 = shup(o1.x1); Q.x6 = shup(o1.x6); Q.xd = shup(o1.xd);
      Q.x5 = Q.x3 = Q.y6 = Q.y1 = Q.y3 = Q.yd = 0.0;
      if (lane == 0) {
        if (w == 0 || !multi) {
          urow_zero(P);
          urow_zero(Q);
        } else {
          const double* rp = ring[w - 1][(t + 1) % 3];  // step t-2
          const double* rq = ring[w - 1][(t + 2) % 3];  // step t-1
          P.x5 = rp[0]; P.x3 = rp[1]; P.x7 = rp[2]; P.x1 = rp[3]; P.x6 = rp[4]; P.xd = rp[5];
          P.y6 = rp[6]; P.y1 = rp[7]; P.y3 = rp[8]; P.yd = rp[9];
          Q.x5 = rq[0]; Q.x3 = rq[1]; Q.x7 = rq[2]; Q.x1 = rq[3]; Q.x6 = rq[4]; Q.xd = rq[5];
          Q.y6 = rq[6]; Q.y1 = rq[7]; Q.y3 = rq[8]; Q.yd = rq[9];
        }
      }
      const int j = t - 2 * g;
      URow X;
      urow_zero(X);
      if (g < c.nx && j >= 0 && j < c.ny) {
        const int q = c.soff[t] + g - yg_glo(t, c.ny);
        const int l = q - clo;
        const uint32_t m = stage_to_mask<TA>(B[18 * n + l]);
        double a[9], b[9];
#pragma unroll
        for (int s = 0; s < 9; ++s) {
          a[s] = (double)B[(2 * s) * n + l];
          b[s] = (double)B[(2 * s + 1) * n + l];
        }
        const uint32_t mx = m & 0x1ffu, my = (m >> 16) & 0x1ffu;
#define PX(s) ((mx >> (s)) & 1u)
#define PY(s) ((my >> (s)) & 1u)
        const URow& O = o1;
        double lv;
        // ux row, lower entries in column order: ux(i-1,j), ux(i,j-1), uy(i,j-1)
        lv = a[2] * P.xd; a[2] = lv;
        if (PX(8)) a[8] -= lv * P.x7;
        if (PX(0)) a[0] -= lv * P.x1;
        if (PX(5)) a[5] -= lv * P.x6;
        lv = a[4] * O.xd; a[4] = lv;
        if (PX(8)) a[8] -= lv * O.x5;
        if (PX(0)) a[0] -= lv * O.x3;
        if (PX(7)) a[7] -= lv * O.x6;
        lv = a[8] * O.yd; a[8] = lv;
        if (PX(0)) a[0] -= lv * O.y6;
        if (PX(5)) a[5] -= lv * O.y1;
        if (PX(7)) a[7] -= lv * O.y3;
        X.xd = 1.0 / a[0];
        X.x5 = a[5]; X.x3 = a[3]; X.x7 = a[7]; X.x1 = a[1]; X.x6 = a[6];
        // uy row, lower entries: ux(i-1,j), uy(i-1,j), ux(i-1,j+1), uy(i,j-1), ux(i,j)
        lv = b[8] * P.xd; b[8] = lv;
        if (PY(4)) b[4] -= lv * P.x5;
        if (PY(7)) b[7] -= lv * P.x3;
        if (PY(2)) b[2] -= lv * P.x7;
        if (PY(5)) b[5] -= lv * P.x1;
        if (PY(0)) b[0] -= lv * P.x6;
        lv = b[4] * P.yd; b[4] = lv;
        if (PY(7)) b[7] -= lv * P.y6;
        if (PY(0)) b[0] -= lv * P.y3;
        lv = b[7] * Q.xd; b[7] = lv;
        if (PY(0)) b[0] -= lv * Q.x7;
        if (PY(6)) b[6] -= lv * Q.x1;
        if (PY(1)) b[1] -= lv * Q.x6;
        lv = b[2] * O.yd; b[2] = lv;
        if (PY(5)) b[5] -= lv * O.y6;
        if (PY(0)) b[0] -= lv * O.y1;
        lv = b[5] * X.xd; b[5] = lv;
        if (PY(0)) b[0] -= lv * X.x5;
        if (PY(6)) b[6] -= lv * X.x3;
        if (PY(3)) b[3] -= lv * X.x6;
        X.yd = 1.0 / b[0];
        X.y6 = b[6]; X.y1 = b[1]; X.y3 = b[3];
#undef PX
#undef PY
        if (!(a[0] != 0.0) || !(b[0] != 0.0)) *bad = 1;
        F[0 * nc + q] = (TA)a[2]; F[1 * nc + q] = (TA)a[4]; F[2 * nc + q] = (TA)a[8];
        F[3 * nc + q] = (TA)b[8]; F[4 * nc + q] = (TA)b[4]; F[5 * nc + q] = (TA)b[7];
        F[6 * nc + q] = (TA)b[2]; F[7 * nc + q] = (TA)b[5];
        F[8 * nc + q] = (TA)X.x5; F[9 * nc + q] = (TA)X.x3; F[10 * nc + q] = (TA)X.x7;
        F[11 * nc + q] = (TA)X.x1; F[12 * nc + q] = (TA)X.x6; F[13 * nc + q] = (TA)X.xd;
        F[14 * nc + q] = (TA)X.y6; F[15 * nc + q] = (TA)X.y1; F[16 * nc + q] = (TA)X.y3;
        F[17 * nc + q] = (TA)X.yd;
      }
      o2 = o1;
      o1 = X;
      if (multi) {
        if (lane == 63) {
          double* rr = ring[w][t % 3];
          rr[0] = X.x5; rr[1] = X.x3; rr[2] = X.x7; rr[3] = X.x1; rr[4] = X.x6; rr[5] = X.xd;
          rr[6] = X.y6; rr[7] = X.y1; rr[8] = X.y3; rr[9] = X.yd;
        }
        step_barrier();
      }
    }
    // the factorization runs once per linear solve: the next chunk is loaded
    // after the steps, which keeps its 19 values per cell out of the registers
    // that the recurrence needs
    if (ch + 1 < nch) {
      int cln = 0;
      const int lnext = load(t1, cln);
      stash(S[(ch + 1) & 1], lnext, c.soff[min(c.T, t1 + C)] - cln);
    }
    __syncthreads();
  }
}

// ---- triangular sweeps ---------------------------------------------------
// Forward: y = L^{-1} rhs; backward: z = U^{-1} y.  T is the arithmetic type
// of the sweep (the factor type).  The steps are walked in chunks of C
// steps.  Per chunk: every thread loads one cell's factor slots and right-
// hand side into registers for the next chunk while the lanes run the
// recurrence of the current chunk entirely in LDS (inputs staged, outputs
// collected); the chunk's outputs then leave in one coalesced store.  No
// global memory operation sits on the per-step critical path.
//

template <typename T, int NV>
YG_DEV int sweep_cap_cells() {
  const int c = (int)(YG_DYN_LDS / sizeof(T)) / (2 * NV + 2);
  return c < (int)blockDim.x ? c : (int)blockDim.x;
}

template <typename T>
YG_DEV void sweep_L(const Ctx& c, const T* __restrict__ F, const T* __restrict__ rhs,
                    T* __restrict__ y, double (*ring)[3][10]) {
  PhaseTimer ptm(c, 1);
  constexpr int NV = 10;  // L slots of both rows + the right-hand side
  const int g = threadIdx.x, lane = g & 63, w = g >> 6;
  const bool multi = c.nx > 64;
  const size_t nc = c.ncell;
  const int capc = sweep_cap_cells<T, NV>();
  const int C = max(1, capc / c.maxcnt);
  T* S[2] = {reinterpret_cast<T*>(yg_dyn_lds), reinterpret_cast<T*>(yg_dyn_lds) + capc * NV};
  T* O = reinterpret_cast<T*>(yg_dyn_lds) + 2 * capc * NV;
  T reg[NV];
  auto load = [&](int t0) {   // this thread's cell of the chunk at step t0 -> registers
    int q, nat;
    if (chunk_cell(c, t0, min(c.T, t0 + C), g, q, nat)) {
#pragma unroll
      for (int v = 0; v < 8; ++v) reg[v] = F[(size_t)v * nc + q];
      reg[8] = rhs[2 * nat];
      reg[9] = rhs[2 * nat + 1];
      return q - c.soff[t0];
    }
    return -1;
  };
  auto stash = [&](T* B, int l, int n) {
    if (l >= 0) {
#pragma unroll
      for (int v = 0; v < NV; ++v) B[v * n + l] = reg[v];
    }
  };
  T o1x = 0, o1y = 0, o2x = 0, o2y = 0;
  stash(S[0], load(0), c.soff[min(c.T, C)] - c.soff[0]);
  __syncthreads();
  const int nch = (c.T + C - 1) / C;
  for (int ch = 0; ch < nch; ++ch) {
    const int t0 = ch * C, t1 = min(c.T, t0 + C);
    const int lnext = ch + 1 < nch ? load(t1) : -1;
    const T* B = S[ch & 1];
    const int clo = c.soff[t0], n = c.soff[t1] - clo;
    for (int t = t0; t < t1; ++t) {
      T n1x = dpp_shr1(o1x), n2x = dpp_shr1(o2x), n2y = dpp_shr1(o2y);
      if (lane == 0 && w > 0 && multi) {
        n1x = (T)ring[w - 1][(t + 2) % 3][0];
        n2x = (T)ring[w - 1][(t + 1) % 3][0];
        n2y = (T)ring[w - 1][(t + 1) % 3][1];
      }
      const int j = t - 2 * g;
      T yx = 0, yy = 0;
      if (g < c.nx && j >= 0 && j < c.ny) {
        const int l = c.soff[t] + g - yg_glo(t, c.ny) - clo;
        yx = B[8 * n + l] - B[0 * n + l] * n2x - B[1 * n + l] * o1x - B[2 * n + l] * o1y;
        yy = B[9 * n + l] - B[3 * n + l] * n2x - B[4 * n + l] * n2y - B[5 * n + l] * n1x -
             B[6 * n + l] * o1y - B[7 * n + l] * yx;
        O[2 * l] = yx;
        O[2 * l + 1] = yy;
      }
      o2x = o1x; o2y = o1y; o1x = yx; o1y = yy;
      if (multi) {
        if (lane == 63) {
          ring[w][t % 3][0] = (double)yx;
          ring[w][t % 3][1] = (double)yy;
        }
        step_barrier();
      }
    }
    step_barrier();
    if (ch + 1 < nch) stash(S[(ch + 1) & 1], lnext, c.soff[min(c.T, t1 + C)] - c.soff[t1]);
    {
      int q, nat;
      if (chunk_cell(c, t0, t1, g, q, nat)) {
        y[2 * nat] = O[2 * (q - clo)];
        y[2 * nat + 1] = O[2 * (q - clo) + 1];
      }
    }
    step_barrier();
  }
  __syncthreads();
}

template <typename T>
YG_DEV void sweep_U(const Ctx& c, const T* __restrict__ F, const T* __restrict__ y,
                    T* __restrict__ z, double (*ring)[3][10]) {
  PhaseTimer ptm(c, 2);
  constexpr int NV = 12;  // U slots and 1/diag of both rows + y
  const int g = threadIdx.x, lane = g & 63, w = g >> 6;
  const int nw = (c.nx + 63) >> 6;
  const bool multi = c.nx > 64;
  const size_t nc = c.ncell;
  const int capc = sweep_cap_cells<T, NV>();
  const int C = max(1, capc / c.maxcnt);
  T* S[2] = {reinterpret_cast<T*>(yg_dyn_lds), reinterpret_cast<T*>(yg_dyn_lds) + capc * NV};
  T* O = reinterpret_cast<T*>(yg_dyn_lds) + 2 * capc * NV;
  T reg[NV];
  const int nch = (c.T + C - 1) / C;
  auto load = [&](int ch) {   // this thread's cell of chunk ch -> registers
    const int t0 = ch * C;
    int q, nat;
    if (chunk_cell(c, t0, min(c.T, t0 + C), g, q, nat)) {
#pragma unroll
      for (int v = 0; v < 10; ++v) reg[v] = F[(size_t)(8 + v) * nc + q];
      reg[10] = y[2 * nat];
      reg[11] = y[2 * nat + 1];
      return q - c.soff[t0];
    }
    return -1;
  };
  auto stash = [&](T* B, int l, int n) {
    if (l >= 0) {
#pragma unroll
      for (int v = 0; v < NV; ++v) B[v * n + l] = reg[v];
    }
  };
  auto ncells = [&](int ch) { return c.soff[min(c.T, ch * C + C)] - c.soff[ch * C]; };
  T o1x = 0, o1y = 0, o2x = 0, o2y = 0;  // own cells j+1 (step u-1), j+2 (u-2)
  stash(S[0], load(nch - 1), ncells(nch - 1));
  __syncthreads();
  for (int k = 0; k < nch; ++k) {
    const int ch = nch - 1 - k;
    const int t0 = ch * C, t1 = min(c.T, t0 + C);
    const int lnext = ch > 0 ? load(ch - 1) : -1;
    const T* B = S[k & 1];
    const int clo = c.soff[t0], n = c.soff[t1] - clo;
    for (int t = t1 - 1; t >= t0; --t) {
      const int u = c.T - 1 - t;
      // line i+1: cell j-1 (its step u-1) and cell j (its step u-2)
      T n1y = dpp_shl1(o1y), n2x = dpp_shl1(o2x), n2y = dpp_shl1(o2y);
      if (g >= c.nx - 1) {
        n1y = n2x = n2y = 0;
      } else if (lane == 63 && multi && w < nw - 1) {
        n1y = (T)ring[w + 1][(u + 2) % 3][1];
        n2x = (T)ring[w + 1][(u + 1) % 3][0];
        n2y = (T)ring[w + 1][(u + 1) % 3][1];
      }
      const int j = t - 2 * g;
      T zx = 0, zy = 0;
      if (g < c.nx && j >= 0 && j < c.ny) {
        const int l = c.soff[t] + g - yg_glo(t, c.ny) - clo;
        zy = (B[11 * n + l] - B[6 * n + l] * o1x - B[7 * n + l] * o1y - B[8 * n + l] * n2y) *
             B[9 * n + l];
        zx = (B[10 * n + l] - B[0 * n + l] * zy - B[1 * n + l] * o1x - B[2 * n + l] * n1y -
              B[3 * n + l] * n2x - B[4 * n + l] * n2y) * B[5 * n + l];
        O[2 * l] = zx;
        O[2 * l + 1] = zy;
      }
      o2x = o1x; o2y = o1y; o1x = zx; o1y = zy;
      if (multi) {
        if (lane == 0) {
          ring[w][u % 3][0] = (double)zx;
          ring[w][u % 3][1] = (double)zy;
        }
        step_barrier();
      }
    }
    step_barrier();
    if (ch > 0) stash(S[(k + 1) & 1], lnext, ncells(ch - 1));
    {
      int q, nat;
      if (chunk_cell(c, t0, t1, g, q, nat)) {
        z[2 * nat] = O[2 * (q - clo)];
        z[2 * nat + 1] = O[2 * (q - clo) + 1];
      }
    }
    step_barrier();
  }
  __syncthreads();
}

// ---- preconditioner --------------------------------------------------------
template <typename T>
struct Precon {
  int kind;        // 0 none, 1 jacobi, 2 ilu
  const T* F;      // ILU factors, or 1/diag pairs for jacobi
};

// out = M^-1 in; tmp is scratch of the same size
template <typename T>
YG_DEV void precon_apply(const Ctx& c, const Precon<T>& P, const T* in, T* tmp, T* out,
                         double (*ring)[3][10]) {
  if (P.kind == 2) {
    sweep_L<T>(c, P.F, in, tmp, ring);
    sweep_U<T>(c, P.F, tmp, out, ring);
  } else {
    for (int q = threadIdx.x; q < 2 * c.ncell; q += blockDim.x)
      out[q] = P.kind == 1 ? in[q] * P.F[q] : in[q];
    __syncthreads();
  }
}

// ---- BiCGSTAB / CG ------------------------------------------------------
struct SolveOut {
  int iters, status;
  double resid;
};

// y = A x for every cell, with per-thread partial sums of (w1 . y) and
// (w2 . y), (y . y) accumulated into acc (w1/w2 may be null)
template <typename TA, typename TV, typename F>
YG_DEV void spmv_all(const Ctx& c, const TA* __restrict__ A, const TV* __restrict__ x,
                     TV* __restrict__ y, F&& post) {
  PhaseTimer ptm(c, 3);
  const int B = blockDim.x;
  for (int q0 = threadIdx.x; q0 < c.ncell; q0 += 2 * B) {
    double ax0, ay0, ax1 = 0, ay1 = 0;
    const int q1 = q0 + B;
    spmv_cell(c, A, x, q0, ax0, ay0);
    if (q1 < c.ncell) spmv_cell(c, A, x, q1, ax1, ay1);
    y[2 * q0] = (TV)ax0;
    y[2 * q0 + 1] = (TV)ay0;
    post(q0);
    if (q1 < c.ncell) {
      y[2 * q1] = (TV)ax1;
      y[2 * q1 + 1] = (TV)ay1;
      post(q1);
    }
  }
}

template <typename TA, typename TV>
YG_DEV SolveOut bicgstab(const Ctx& c, const TA* A, const Precon<TV>& P, const TV* b, TV* x,
                         TV* W, const YgKrylovOpts& o, double (*ring)[3][10]) {
  const int n2 = 2 * c.ncell;
  TV* __restrict__ r = W;
  TV* __restrict__ rt = W + 1 * (size_t)n2;
  TV* __restrict__ p = W + 2 * (size_t)n2;
  TV* __restrict__ v = W + 3 * (size_t)n2;
  TV* __restrict__ ph = W + 4 * (size_t)n2;
  TV* __restrict__ s = W + 5 * (size_t)n2;
  TV* __restrict__ sh = W + 6 * (size_t)n2;
  TV* __restrict__ tt = W + 7 * (size_t)n2;
  TV* __restrict__ ytmp = W + 8 * (size_t)n2;
  SolveOut out{0, 4, 0.0};
  double red[3];
  // r = b - A x, rt = r
  red[0] = red[1] = red[2] = 0.0;
  spmv_all(c, A, x, tt, [&](int q) {
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const double rq = (double)b[2 * q + e] - (double)tt[2 * q + e];
      r[2 * q + e] = (TV)rq;
      rt[2 * q + e] = (TV)rq;
      red[0] += (double)r[2 * q + e] * (double)r[2 * q + e];
      red[1] += (double)b[2 * q + e] * (double)b[2 * q + e];
    }
  });
  yg_block_sum<3>(red, c.red);
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
      for (int q = threadIdx.x; q < n2; q += blockDim.x) p[q] = r[q];
    } else {
      const double beta = (rho / rho_old) * (alpha / omega);
      for (int q = threadIdx.x; q < n2; q += blockDim.x)
        p[q] = (TV)((double)r[q] + beta * ((double)p[q] - omega * (double)v[q]));
    }
    __syncthreads();
    precon_apply<TV>(c, P, p, ytmp, ph, ring);
    // v = A phat, (rt, v)
    red[0] = 0.0;
    spmv_all(c, A, ph, v, [&](int q) {
      red[0] += (double)rt[2 * q] * (double)v[2 * q] + (double)rt[2 * q + 1] * (double)v[2 * q + 1];
    });
    yg_block_sum<1>(*reinterpret_cast<double(*)[1]>(red), c.red);
    if (red[0] == 0.0) { out.status = 2; out.iters = it; return out; }
    alpha = rho / red[0];
    // s = r - alpha v and ||s||: converged on s ends the iteration early
    red[0] = 0.0;
    for (int q = threadIdx.x; q < n2; q += blockDim.x) {
      const TV sq = (TV)((double)r[q] - alpha * (double)v[q]);
      s[q] = sq;
      red[0] += (double)sq * (double)sq;
    }
    yg_block_sum<1>(*reinterpret_cast<double(*)[1]>(red), c.red);
    const double sn = sqrt(red[0]) * bnrm;
    if (sn <= o.tol) {
      for (int q = threadIdx.x; q < n2; q += blockDim.x) x[q] = (TV)((double)x[q] + alpha * (double)ph[q]);
      __syncthreads();
      out.resid = sn; out.iters = it; out.status = 0;
      return out;
    }
    precon_apply<TV>(c, P, s, ytmp, sh, ring);
    // t = A shat, (t,t), (t,s)
    red[0] = red[1] = 0.0;
    spmv_all(c, A, sh, tt, [&](int q) {
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        const double tq = (double)tt[2 * q + e];
        red[0] += tq * tq;
        red[1] += tq * (double)s[2 * q + e];
      }
    });
    yg_block_sum<2>(*reinterpret_cast<double(*)[2]>(red), c.red);
    if (red[0] == 0.0) { out.status = 2; out.iters = it; return out; }
    omega = red[1] / red[0];
    if (omega == 0.0) { out.status = 2; out.iters = it; return out; }
    // x += alpha phat + omega shat; r = s - omega t; (r,r), (rt,r)
    red[0] = red[1] = 0.0;
    for (int q = threadIdx.x; q < n2; q += blockDim.x) {
      x[q] = (TV)((double)x[q] + alpha * (double)ph[q] + omega * (double)sh[q]);
      const TV rq = (TV)((double)s[q] - omega * (double)tt[q]);
      r[q] = rq;
      red[0] += (double)rq * rq;
      red[1] += (double)rt[q] * rq;
    }
    yg_block_sum<2>(*reinterpret_cast<double(*)[2]>(red), c.red);
    out.resid = sqrt(red[0]) * bnrm;
    out.iters = it;
    if (out.resid <= o.tol) { out.status = 0; return out; }
    rho_old = rho;
    rho = red[1];
  }
  out.status = 4;
  return out;
}

template <typename TA, typename TV>
YG_DEV SolveOut cg(const Ctx& c, const TA* A, const Precon<TV>& P, const TV* b, TV* x, TV* W,
                   const YgKrylovOpts& o, double (*ring)[3][10]) {
  const int n2 = 2 * c.ncell;
  TV* r = W;
  TV* z = W + 1 * (size_t)n2;
  TV* p = W + 2 * (size_t)n2;
  TV* qv = W + 3 * (size_t)n2;
  TV* ytmp = W + 8 * (size_t)n2;
  SolveOut out{0, 4, 0.0};
  double red[2] = {0.0, 0.0};
  spmv_all(c, A, x, qv, [&](int q) {
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      r[2 * q + e] = (TV)((double)b[2 * q + e] - (double)qv[2 * q + e]);
      red[0] += (double)r[2 * q + e] * (double)r[2 * q + e];
      red[1] += (double)b[2 * q + e] * (double)b[2 * q + e];
    }
  });
  yg_block_sum<2>(red, c.red);
  const double nrm0 = o.conv_cond == 1 ? sqrt(red[1]) : sqrt(red[0]);
  if (nrm0 == 0.0) { out.status = 0; return out; }
  const double bnrm = 1.0 / nrm0;
  out.resid = sqrt(red[0]) * bnrm;
  if (out.resid <= o.tol) { out.status = 0; return out; }
  double rho_old = 0.0;
  for (int it = 1; it <= o.maxiter; ++it) {
    precon_apply<TV>(c, P, r, ytmp, z, ring);
    double rr[1] = {0.0};
    for (int q = threadIdx.x; q < n2; q += blockDim.x) rr[0] += (double)r[q] * (double)z[q];
    yg_block_sum<1>(rr, c.red);
    const double rho = rr[0];
    if (rho == 0.0) { out.status = 2; out.iters = it; return out; }
    const double beta = it == 1 ? 0.0 : rho / rho_old;
    for (int q = threadIdx.x; q < n2; q += blockDim.x)
      p[q] = it == 1 ? z[q] : (TV)((double)z[q] + beta * (double)p[q]);
    __syncthreads();
    double pq[1] = {0.0};
    spmv_all(c, A, p, qv, [&](int q) {
      pq[0] += (double)p[2 * q] * (double)qv[2 * q] + (double)p[2 * q + 1] * (double)qv[2 * q + 1];
    });
    yg_block_sum<1>(pq, c.red);
    if (pq[0] == 0.0) { out.status = 2; out.iters = it; return out; }
    const double alpha = rho / pq[0];
    double rn[1] = {0.0};
    for (int q = threadIdx.x; q < n2; q += blockDim.x) {
      x[q] = (TV)((double)x[q] + alpha * (double)p[q]);
      r[q] = (TV)((double)r[q] - alpha * (double)qv[q]);
      rn[0] += (double)r[q] * (double)r[q];
    }
    yg_block_sum<1>(rn, c.red);
    out.resid = sqrt(rn[0]) * bnrm;
    out.iters = it;
    if (out.resid <= o.tol) { out.status = 0; return out; }
    rho_old = rho;
  }
  out.status = 4;
  return out;
}

template <typename T>
YG_DEV void build_precon(const Ctx& c, const T* A, T* F, Precon<T>& P, int kind,
                         double (*ring)[3][10], int* bad) {
  P.kind = kind;
  P.F = F;
  if (kind == 2) {
    ilu0_factor<T>(c, A, F, ring, bad);
  } else if (kind == 1) {
    for (int q = threadIdx.x; q < c.ncell; q += blockDim.x) {
      F[2 * q] = (T)(1.0 / (double)A[((size_t)0 * c.ncell + q) * 2 + 0]);
      F[2 * q + 1] = (T)(1.0 / (double)A[((size_t)0 * c.ncell + q) * 2 + 1]);
    }
    __syncthreads();
  }
}

template <typename T>
YG_DEV SolveOut krylov(const Ctx& c, const T* A, const Precon<T>& P, const T* b, T* x, T* W,
                       const YgKrylovOpts& o, double (*ring)[3][10]) {
  return o.solver == 1 ? cg<T, T>(c, A, P, b, x, W, o, ring)
                       : bicgstab<T, T>(c, A, P, b, x, W, o, ring);
}

}  // namespace

// ---------------------------------------------------------------------------
// The kernel: one block per member.  Modes:
//   YG_MODE_FP64  : FP64 matrix, factors, vectors (the plain FP64 path)
//   YG_MODE_MIXED : FP64 residual b - A x and correction accumulation,
//                   FP32 matrix copy, factors and inner Krylov solve
template <int MODE>
__global__ __launch_bounds__(1024) void yg_krylov_kernel(YgKrylovArgs a) {
  const int m = blockIdx.x;
  if (a.skip && a.skip[m]) return;
  const YgMember mb = a.members[m];
  __shared__ int s_soff[YG_MAX_STEPS + 1];
  __shared__ double s_red[16 * 4];
  __shared__ double s_ring[16][3][10];
  __shared__ int s_bad;
  const int ncell = mb.nx * mb.ny;
  for (int k = threadIdx.x; k <= mb.T; k += blockDim.x) s_soff[k] = a.soff[mb.toff + k];
  if (threadIdx.x == 0) s_bad = 0;
  __syncthreads();
  Ctx c;
  c.mb = &a.members[m];
  c.nx = mb.nx; c.ny = mb.ny; c.T = mb.T; c.ncell = ncell;
  c.wrapx = a.wrapx; c.wrapy = a.wrapy;
  c.soff = s_soff;
  c.mask = a.mask + mb.coff;
  c.red = s_red;
  c.maxcnt = mb.maxcnt;
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
    for (int q = threadIdx.x; q < 2 * ncell; q += blockDim.x) bm[0] += b64[q] != 0.0 ? 1.0 : 0.0;
    yg_block_sum<1>(bm, s_red);
    if (bm[0] == 0.0) {
      for (int q = threadIdx.x; q < 2 * ncell; q += blockDim.x) x64[q] = 0.0;
      if (threadIdx.x == 0) { a.out_iters[m] = 0; a.out_status[m] = 0; a.out_resid[m] = 0.0; }
      return;
    }
  }
  YgKrylovOpts o = a.opts;
  if (MODE == YG_MODE_FP64) {
    const double* A = a.A64 + (size_t)18 * mb.coff;
    double* F = a.F64 + (size_t)YG_NF * mb.coff;
    double* W = a.W64 + (size_t)YG_NW64 * c2;
    if (o.initx_zeros) {
      for (int q = threadIdx.x; q < 2 * ncell; q += blockDim.x) x64[q] = 0.0;
      __syncthreads();
    }
    Precon<double> P;
    build_precon<double>(c, A, F, P, o.precon, s_ring, &s_bad);
    SolveOut so = krylov<double>(c, A, P, b64, x64, W, o, s_ring);
    if (s_bad && so.status == 0) so.status = 2;
    if (threadIdx.x == 0) { a.out_iters[m] = so.iters; a.out_status[m] = so.status; a.out_resid[m] = so.resid; }
    if (a.timers && threadIdx.x == 0) {
      s_tm[7] = wall_clock64() - t_kernel0;
      for (int k = 0; k < 8; ++k) a.timers[8 * m + k] += s_tm[k];
    }
  } else {
    // FP64 iterative refinement around an FP32 solve
    const double* A64 = a.A64 + (size_t)18 * mb.coff;
    const float* A32 = a.A32 + (size_t)18 * mb.coff;
    float* F = a.F32 + (size_t)YG_NF * mb.coff;
    float* W = a.W32 + (size_t)YG_NW32 * c2;    // 9 inner Krylov vectors ...
    float* d = W + (size_t)9 * 2 * ncell;       // ... the correction ...
    float* rin = W + (size_t)10 * 2 * ncell;    // ... and the scaled residual
    double* r64 = a.W64 + (size_t)YG_NW64 * c2; // FP64 residual
    if (o.initx_zeros) {
      for (int q = threadIdx.x; q < 2 * ncell; q += blockDim.x) x64[q] = 0.0;
      __syncthreads();
    }
    Precon<float> P;
    build_precon<float>(c, A32, F, P, o.precon, s_ring, &s_bad);
    YgKrylovOpts oi = o;
    oi.tol = o.inner_tol;
    oi.conv_cond = 0;
    oi.initx_zeros = 1;
    double nrm0 = -1.0, res = 1.0;
    int iters = 0, status = 4;
    for (int ir = 0; ir < o.ir_max; ++ir) {
      double red[2] = {0.0, 0.0};
      for (int q = threadIdx.x; q < ncell; q += blockDim.x) {
        double ax, ay;
        spmv_cell(c, A64, x64, q, ax, ay);
        const double rx = b64[2 * q] - ax, ry = b64[2 * q + 1] - ay;
        r64[2 * q] = rx; r64[2 * q + 1] = ry;
        red[0] += rx * rx + ry * ry;
        red[1] += b64[2 * q] * b64[2 * q] + b64[2 * q + 1] * b64[2 * q + 1];
      }
      yg_block_sum<2>(red, s_red);
      const double rn = sqrt(red[0]);
      if (nrm0 < 0.0) nrm0 = o.conv_cond == 1 ? sqrt(red[1]) : rn;
      if (nrm0 == 0.0) { res = 0.0; status = 0; break; }
      res = rn / nrm0;
      if (res <= o.tol) { status = 0; break; }
      if (iters >= o.maxiter) { status = 4; break; }
      const double sc = rn > 0.0 ? 1.0 / rn : 1.0;
      for (int q = threadIdx.x; q < 2 * ncell; q += blockDim.x) {
        rin[q] = (float)(r64[q] * sc);
        d[q] = 0.0f;
      }
      __syncthreads();
      oi.maxiter = o.maxiter - iters;
      SolveOut so = krylov<float>(c, A32, P, rin, d, W, oi, s_ring);
      iters += so.iters;
      for (int q = threadIdx.x; q < 2 * ncell; q += blockDim.x) x64[q] += rn * (double)d[q];
      __syncthreads();
      if (so.status == 2 && so.iters <= 1) { status = 2; break; }
    }
    if (s_bad && status == 0) status = 2;
    if (threadIdx.x == 0) { a.out_iters[m] = iters; a.out_status[m] = status; a.out_resid[m] = res; }
    if (a.timers && threadIdx.x == 0) {
      s_tm[7] = wall_clock64() - t_kernel0;
      for (int k = 0; k < 8; ++k) a.timers[8 * m + k] += s_tm[k];
    }
  }
}

template __global__ void yg_krylov_kernel<YG_MODE_FP64>(YgKrylovArgs);
template __global__ void yg_krylov_kernel<YG_MODE_MIXED>(YgKrylovArgs);

void yg_launch_krylov(const YgKrylovArgs& a, int mode, int nmembers, int block, hipStream_t st) {
  const size_t dyn = YG_DYN_LDS;
  static bool attr_set[2] = {false, false};
  if (mode == YG_MODE_FP64) {
    if (!attr_set[0]) {
      (void)hipFuncSetAttribute((const void*)yg_krylov_kernel<YG_MODE_FP64>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, YG_DYN_LDS);
      attr_set[0] = true;
    }
    hipLaunchKernelGGL(yg_krylov_kernel<YG_MODE_FP64>, dim3(nmembers), dim3(block), dyn, st, a);
  } else {
    if (!attr_set[1]) {
      (void)hipFuncSetAttribute((const void*)yg_krylov_kernel<YG_MODE_MIXED>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, YG_DYN_LDS);
      attr_set[1] = true;
    }
    hipLaunchKernelGGL(yg_krylov_kernel<YG_MODE_MIXED>, dim3(nmembers), dim3(block), dyn, st, a);
  }
}
