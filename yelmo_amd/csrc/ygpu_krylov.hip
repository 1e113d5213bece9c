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
  const uint32_t* mask;       // per cell presence bits
  const uint32_t* cellij;     // per cell (i << 16) | j
  double* red;                // LDS reduction scratch
};

template <typename TA, typename TV>
YG_DEV void spmv_cell(const Ctx& c, const TA* __restrict__ A, const TV* __restrict__ x, int q,
                      double& yx, double& yy) {
  const uint32_t ij = c.cellij[q];
  const int i = ij >> 16, j = ij & 0xffff;
  const uint32_t m = c.mask[q];
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
      const int q2 = yg_q(c.soff, i2, j2, c.ny);
      acc[comp] += (double)A[((size_t)s * c.ncell + q) * 2 + comp] * (double)x[2 * q2 + sl[2]];
    }
  }
  yx = acc[0];
  yy = acc[1];
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

YG_DEV double shup(double v) { return __shfl_up(v, 1, 64); }
YG_DEV double shdn(double v) { return __shfl_down(v, 1, 64); }

template <typename TA, typename TF>
YG_DEV void ilu0_factor(const Ctx& c, const TA* __restrict__ A, TF* __restrict__ F,
                        double (*ring)[3][10], int* bad) {
  const int g = threadIdx.x, lane = g & 63, w = g >> 6;
  const bool multi = c.nx > 64;
  const size_t nc = c.ncell;
  URow o1, o2;  // own cells at steps t-1 and t-2
  urow_zero(o1);
  urow_zero(o2);
  for (int t = 0; t < c.T; ++t) {
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
      const uint32_t m = c.mask[q];
      double a[9], b[9];
#pragma unroll
      for (int s = 0; s < 9; ++s) {
        a[s] = (double)A[((size_t)s * nc + q) * 2 + 0];
        b[s] = (double)A[((size_t)s * nc + q) * 2 + 1];
      }
      const uint32_t mx = m & 0x1ffu, my = (m >> 16) & 0x1ffu;
#define PX(s) ((mx >> (s)) & 1u)
#define PY(s) ((my >> (s)) & 1u)
      const URow& O = o1;
      double l;
      // ux row, lower entries in column order: ux(i-1,j), ux(i,j-1), uy(i,j-1)
      l = a[2] * P.xd; a[2] = l;
      if (PX(8)) a[8] -= l * P.x7;
      if (PX(0)) a[0] -= l * P.x1;
      if (PX(5)) a[5] -= l * P.x6;
      l = a[4] * O.xd; a[4] = l;
      if (PX(8)) a[8] -= l * O.x5;
      if (PX(0)) a[0] -= l * O.x3;
      if (PX(7)) a[7] -= l * O.x6;
      l = a[8] * O.yd; a[8] = l;
      if (PX(0)) a[0] -= l * O.y6;
      if (PX(5)) a[5] -= l * O.y1;
      if (PX(7)) a[7] -= l * O.y3;
      X.xd = 1.0 / a[0];
      X.x5 = a[5]; X.x3 = a[3]; X.x7 = a[7]; X.x1 = a[1]; X.x6 = a[6];
      // uy row, lower entries: ux(i-1,j), uy(i-1,j), ux(i-1,j+1), uy(i,j-1), ux(i,j)
      l = b[8] * P.xd; b[8] = l;
      if (PY(4)) b[4] -= l * P.x5;
      if (PY(7)) b[7] -= l * P.x3;
      if (PY(2)) b[2] -= l * P.x7;
      if (PY(5)) b[5] -= l * P.x1;
      if (PY(0)) b[0] -= l * P.x6;
      l = b[4] * P.yd; b[4] = l;
      if (PY(7)) b[7] -= l * P.y6;
      if (PY(0)) b[0] -= l * P.y3;
      l = b[7] * Q.xd; b[7] = l;
      if (PY(0)) b[0] -= l * Q.x7;
      if (PY(6)) b[6] -= l * Q.x1;
      if (PY(1)) b[1] -= l * Q.x6;
      l = b[2] * O.yd; b[2] = l;
      if (PY(5)) b[5] -= l * O.y6;
      if (PY(0)) b[0] -= l * O.y1;
      l = b[5] * X.xd; b[5] = l;
      if (PY(0)) b[0] -= l * X.x5;
      if (PY(6)) b[6] -= l * X.x3;
      if (PY(3)) b[3] -= l * X.x6;
      X.yd = 1.0 / b[0];
      X.y6 = b[6]; X.y1 = b[1]; X.y3 = b[3];
#undef PX
#undef PY
      if (!(a[0] != 0.0) || !(b[0] != 0.0)) *bad = 1;
      F[0 * nc + q] = (TF)a[2]; F[1 * nc + q] = (TF)a[4]; F[2 * nc + q] = (TF)a[8];
      F[3 * nc + q] = (TF)b[8]; F[4 * nc + q] = (TF)b[4]; F[5 * nc + q] = (TF)b[7];
      F[6 * nc + q] = (TF)b[2]; F[7 * nc + q] = (TF)b[5];
      F[8 * nc + q] = (TF)X.x5; F[9 * nc + q] = (TF)X.x3; F[10 * nc + q] = (TF)X.x7;
      F[11 * nc + q] = (TF)X.x1; F[12 * nc + q] = (TF)X.x6; F[13 * nc + q] = (TF)X.xd;
      F[14 * nc + q] = (TF)X.y6; F[15 * nc + q] = (TF)X.y1; F[16 * nc + q] = (TF)X.y3;
      F[17 * nc + q] = (TF)X.yd;
    }
    o2 = o1;
    o1 = X;
    if (multi) {
      if (lane == 63) {
        double* rr = ring[w][t % 3];
        rr[0] = X.x5; rr[1] = X.x3; rr[2] = X.x7; rr[3] = X.x1; rr[4] = X.x6; rr[5] = X.xd;
        rr[6] = X.y6; rr[7] = X.y1; rr[8] = X.y3; rr[9] = X.yd;
      }
      __syncthreads();
    }
  }
  __syncthreads();
}

// ---- triangular sweeps ---------------------------------------------------
// Forward: y = L^{-1} rhs, where rhs(q, bx, by) produces (and may store) the
// right-hand side of cell q.  Backward: z = U^{-1} y.  TV is the arithmetic
// type of the sweep.
template <typename TF, typename TV, typename RHS>
YG_DEV void sweep_L(const Ctx& c, const TF* __restrict__ F, RHS rhs, TV* __restrict__ y,
                    double (*ring)[3][10]) {
  const int g = threadIdx.x, lane = g & 63, w = g >> 6;
  const bool multi = c.nx > 64;
  const size_t nc = c.ncell;
  TV o1x = 0, o1y = 0, o2x = 0, o2y = 0;
  for (int t = 0; t < c.T; ++t) {
    TV n1x = __shfl_up(o1x, 1, 64), n2x = __shfl_up(o2x, 1, 64), n2y = __shfl_up(o2y, 1, 64);
    if (lane == 0) {
      if (w == 0 || !multi) {
        n1x = n2x = n2y = 0;
      } else {
        n1x = (TV)ring[w - 1][(t + 2) % 3][0];
        n2x = (TV)ring[w - 1][(t + 1) % 3][0];
        n2y = (TV)ring[w - 1][(t + 1) % 3][1];
      }
    }
    const int j = t - 2 * g;
    TV yx = 0, yy = 0;
    if (g < c.nx && j >= 0 && j < c.ny) {
      const int q = c.soff[t] + g - yg_glo(t, c.ny);
      TV bx, by;
      rhs(q, bx, by);
      yx = bx - (TV)F[0 * nc + q] * n2x - (TV)F[1 * nc + q] * o1x - (TV)F[2 * nc + q] * o1y;
      yy = by - (TV)F[3 * nc + q] * n2x - (TV)F[4 * nc + q] * n2y - (TV)F[5 * nc + q] * n1x -
           (TV)F[6 * nc + q] * o1y - (TV)F[7 * nc + q] * yx;
      y[2 * q] = yx;
      y[2 * q + 1] = yy;
    }
    o2x = o1x; o2y = o1y; o1x = yx; o1y = yy;
    if (multi) {
      if (lane == 63) {
        ring[w][t % 3][0] = (double)yx;
        ring[w][t % 3][1] = (double)yy;
      }
      __syncthreads();
    }
  }
  __syncthreads();
}

template <typename TF, typename TV>
YG_DEV void sweep_U(const Ctx& c, const TF* __restrict__ F, const TV* __restrict__ y,
                    TV* __restrict__ z, double (*ring)[3][10]) {
  const int g = threadIdx.x, lane = g & 63, w = g >> 6;
  const int nw = (c.nx + 63) >> 6;
  const bool multi = c.nx > 64;
  const size_t nc = c.ncell;
  TV o1x = 0, o1y = 0, o2x = 0, o2y = 0;  // own cells j+1 (step u-1), j+2 (u-2)
  for (int u = 0; u < c.T; ++u) {
    const int t = c.T - 1 - u;
    // line i+1: cell j-1 (its step u-1) and cell j (its step u-2)
    TV n1y = __shfl_down(o1y, 1, 64), n2x = __shfl_down(o2x, 1, 64), n2y = __shfl_down(o2y, 1, 64);
    if (lane == 63 || g == c.nx - 1) {
      if (g >= c.nx - 1 || w >= nw - 1 || !multi) {
        n1y = n2x = n2y = 0;
      } else {
        n1y = (TV)ring[w + 1][(u + 2) % 3][1];
        n2x = (TV)ring[w + 1][(u + 1) % 3][0];
        n2y = (TV)ring[w + 1][(u + 1) % 3][1];
      }
    }
    const int j = t - 2 * g;
    TV zx = 0, zy = 0;
    if (g < c.nx && j >= 0 && j < c.ny) {
      const int q = c.soff[t] + g - yg_glo(t, c.ny);
      zy = (y[2 * q + 1] - (TV)F[14 * nc + q] * o1x - (TV)F[15 * nc + q] * o1y -
            (TV)F[16 * nc + q] * n2y) * (TV)F[17 * nc + q];
      zx = (y[2 * q] - (TV)F[8 * nc + q] * zy - (TV)F[9 * nc + q] * o1x -
            (TV)F[10 * nc + q] * n1y - (TV)F[11 * nc + q] * n2x - (TV)F[12 * nc + q] * n2y) *
           (TV)F[13 * nc + q];
      z[2 * q] = zx;
      z[2 * q + 1] = zy;
    }
    o2x = o1x; o2y = o1y; o1x = zx; o1y = zy;
    if (multi) {
      if (lane == 0) {
        ring[w][u % 3][0] = (double)zx;
        ring[w][u % 3][1] = (double)zy;
      }
      __syncthreads();
    }
  }
  __syncthreads();
}

// ---- preconditioner dispatch --------------------------------------------
template <typename TF, typename TV>
struct Precon {
  int kind;                 // 0 none, 1 jacobi, 2 ilu
  const TF* F;              // ILU factors, or 1/diag pairs for jacobi
};

template <typename TF, typename TV, typename RHS>
YG_DEV void precon_apply(const Ctx& c, const Precon<TF, TV>& P, RHS rhs, TV* tmp, TV* out,
                         double (*ring)[3][10]) {
  if (P.kind == 2) {
    sweep_L<TF, TV>(c, P.F, rhs, tmp, ring);
    sweep_U<TF, TV>(c, P.F, tmp, out, ring);
  } else {
    for (int q = threadIdx.x; q < c.ncell; q += blockDim.x) {
      TV bx, by;
      rhs(q, bx, by);
      if (P.kind == 1) {
        bx *= (TV)P.F[2 * q];
        by *= (TV)P.F[2 * q + 1];
      }
      out[2 * q] = bx;
      out[2 * q + 1] = by;
    }
    __syncthreads();
  }
}

// ---- BiCGSTAB / CG ------------------------------------------------------
struct SolveOut {
  int iters, status;
  double resid;
};

template <typename TA, typename TF, typename TV>
YG_DEV SolveOut bicgstab(const Ctx& c, const TA* A, const Precon<TF, TV>& P, const TV* b, TV* x,
                         TV* W, const YgKrylovOpts& o, double (*ring)[3][10]) {
  const int n2 = 2 * c.ncell;
  TV* r = W;
  TV* rt = W + 1 * (size_t)n2;
  TV* p = W + 2 * (size_t)n2;
  TV* v = W + 3 * (size_t)n2;
  TV* ph = W + 4 * (size_t)n2;
  TV* s = W + 5 * (size_t)n2;
  TV* sh = W + 6 * (size_t)n2;
  TV* tt = W + 7 * (size_t)n2;
  TV* ytmp = W + 8 * (size_t)n2;
  SolveOut out{0, 4, 0.0};
  double red[3];
  // r = b - A x, rt = r
  red[0] = red[1] = red[2] = 0.0;
  for (int q = threadIdx.x; q < c.ncell; q += blockDim.x) {
    double ax, ay;
    spmv_cell(c, A, x, q, ax, ay);
    const double rx = (double)b[2 * q] - ax, ry = (double)b[2 * q + 1] - ay;
    r[2 * q] = (TV)rx; r[2 * q + 1] = (TV)ry;
    rt[2 * q] = (TV)rx; rt[2 * q + 1] = (TV)ry;
    red[0] += (double)r[2 * q] * r[2 * q] + (double)r[2 * q + 1] * r[2 * q + 1];
    red[1] += (double)b[2 * q] * b[2 * q] + (double)b[2 * q + 1] * b[2 * q + 1];
  }
  yg_block_sum<3>(red, c.red);
  const double nrm0 = o.conv_cond == 1 ? sqrt(red[1]) : sqrt(red[0]);
  if (nrm0 == 0.0) { out.status = 0; out.resid = 0.0; return out; }
  const double bnrm = 1.0 / nrm0;
  out.resid = sqrt(red[0]) * bnrm;
  if (out.resid <= o.tol) { out.status = 0; return out; }
  double rho = red[0], rho_old = 1.0, alpha = 1.0, omega = 1.0;  // (rt, r) = (r, r) at start
  for (int it = 1; it <= o.maxiter; ++it) {
    if (rho == 0.0) { out.status = 2; out.iters = it; return out; }
    const double beta = it == 1 ? 0.0 : (rho / rho_old) * (alpha / omega);
    const bool first = it == 1;
    // p = r + beta (p - omega v); phat = M^-1 p
    auto rhs_p = [&](int q, TV& bx, TV& by) {
      TV px, py;
      if (first) { px = r[2 * q]; py = r[2 * q + 1]; }
      else {
        px = (TV)((double)r[2 * q] + beta * ((double)p[2 * q] - omega * (double)v[2 * q]));
        py = (TV)((double)r[2 * q + 1] + beta * ((double)p[2 * q + 1] - omega * (double)v[2 * q + 1]));
      }
      p[2 * q] = px; p[2 * q + 1] = py;
      bx = px; by = py;
    };
    precon_apply<TF, TV>(c, P, rhs_p, ytmp, ph, ring);
    // v = A phat, (rt, v)
    red[0] = 0.0;
    for (int q = threadIdx.x; q < c.ncell; q += blockDim.x) {
      double ax, ay;
      spmv_cell(c, A, ph, q, ax, ay);
      v[2 * q] = (TV)ax; v[2 * q + 1] = (TV)ay;
      red[0] += (double)rt[2 * q] * (double)v[2 * q] + (double)rt[2 * q + 1] * (double)v[2 * q + 1];
    }
    yg_block_sum<1>(*reinterpret_cast<double(*)[1]>(red), c.red);
    if (red[0] == 0.0) { out.status = 2; out.iters = it; return out; }
    alpha = rho / red[0];
    // s = r - alpha v, ||s||; shat = M^-1 s
    double ss = 0.0;
    auto rhs_s = [&](int q, TV& bx, TV& by) {
      const TV sx = (TV)((double)r[2 * q] - alpha * (double)v[2 * q]);
      const TV sy = (TV)((double)r[2 * q + 1] - alpha * (double)v[2 * q + 1]);
      s[2 * q] = sx; s[2 * q + 1] = sy;
      ss += (double)sx * sx + (double)sy * sy;
      bx = sx; by = sy;
    };
    precon_apply<TF, TV>(c, P, rhs_s, ytmp, sh, ring);
    red[0] = ss;
    yg_block_sum<1>(*reinterpret_cast<double(*)[1]>(red), c.red);
    const double sn = sqrt(red[0]) * bnrm;
    if (sn <= o.tol) {
      for (int q = threadIdx.x; q < n2; q += blockDim.x) x[q] = (TV)((double)x[q] + alpha * (double)ph[q]);
      __syncthreads();
      out.resid = sn; out.iters = it; out.status = 0;
      return out;
    }
    // t = A shat, (t,t), (t,s)
    red[0] = red[1] = 0.0;
    for (int q = threadIdx.x; q < c.ncell; q += blockDim.x) {
      double ax, ay;
      spmv_cell(c, A, sh, q, ax, ay);
      tt[2 * q] = (TV)ax; tt[2 * q + 1] = (TV)ay;
      const double tx = (double)tt[2 * q], ty = (double)tt[2 * q + 1];
      red[0] += tx * tx + ty * ty;
      red[1] += tx * (double)s[2 * q] + ty * (double)s[2 * q + 1];
    }
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

template <typename TA, typename TF, typename TV>
YG_DEV SolveOut cg(const Ctx& c, const TA* A, const Precon<TF, TV>& P, const TV* b, TV* x,
                   TV* W, const YgKrylovOpts& o, double (*ring)[3][10]) {
  const int n2 = 2 * c.ncell;
  TV* r = W;
  TV* z = W + 1 * (size_t)n2;
  TV* p = W + 2 * (size_t)n2;
  TV* qv = W + 3 * (size_t)n2;
  TV* ytmp = W + 8 * (size_t)n2;
  SolveOut out{0, 4, 0.0};
  double red[2] = {0.0, 0.0};
  for (int q = threadIdx.x; q < c.ncell; q += blockDim.x) {
    double ax, ay;
    spmv_cell(c, A, x, q, ax, ay);
    r[2 * q] = (TV)((double)b[2 * q] - ax);
    r[2 * q + 1] = (TV)((double)b[2 * q + 1] - ay);
    red[0] += (double)r[2 * q] * r[2 * q] + (double)r[2 * q + 1] * r[2 * q + 1];
    red[1] += (double)b[2 * q] * b[2 * q] + (double)b[2 * q + 1] * b[2 * q + 1];
  }
  yg_block_sum<2>(red, c.red);
  const double nrm0 = o.conv_cond == 1 ? sqrt(red[1]) : sqrt(red[0]);
  if (nrm0 == 0.0) { out.status = 0; return out; }
  const double bnrm = 1.0 / nrm0;
  out.resid = sqrt(red[0]) * bnrm;
  if (out.resid <= o.tol) { out.status = 0; return out; }
  double rho_old = 0.0;
  for (int it = 1; it <= o.maxiter; ++it) {
    auto rhs_r = [&](int q, TV& bx, TV& by) { bx = r[2 * q]; by = r[2 * q + 1]; };
    precon_apply<TF, TV>(c, P, rhs_r, ytmp, z, ring);
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
    for (int q = threadIdx.x; q < c.ncell; q += blockDim.x) {
      double ax, ay;
      spmv_cell(c, A, p, q, ax, ay);
      qv[2 * q] = (TV)ax; qv[2 * q + 1] = (TV)ay;
      pq[0] += (double)p[2 * q] * ax + (double)p[2 * q + 1] * ay;
    }
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

template <typename TA, typename TF, typename TV>
YG_DEV void build_precon(const Ctx& c, const TA* A, TF* F, Precon<TF, TV>& P, int kind,
                         double (*ring)[3][10], int* bad) {
  P.kind = kind;
  P.F = F;
  if (kind == 2) {
    ilu0_factor<TA, TF>(c, A, F, ring, bad);
  } else if (kind == 1) {
    for (int q = threadIdx.x; q < c.ncell; q += blockDim.x) {
      F[2 * q] = (TF)(1.0 / (double)A[((size_t)0 * c.ncell + q) * 2 + 0]);
      F[2 * q + 1] = (TF)(1.0 / (double)A[((size_t)0 * c.ncell + q) * 2 + 1]);
    }
    __syncthreads();
  }
}

template <typename TA, typename TF, typename TV>
YG_DEV SolveOut krylov(const Ctx& c, const TA* A, const Precon<TF, TV>& P, const TV* b, TV* x,
                       TV* W, const YgKrylovOpts& o, double (*ring)[3][10]) {
  return o.solver == 1 ? cg<TA, TF, TV>(c, A, P, b, x, W, o, ring)
                       : bicgstab<TA, TF, TV>(c, A, P, b, x, W, o, ring);
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
  c.cellij = a.cellij + mb.coff;
  c.red = s_red;
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
    Precon<double, double> P;
    build_precon<double, double, double>(c, A, F, P, o.precon, s_ring, &s_bad);
    SolveOut so = krylov<double, double, double>(c, A, P, b64, x64, W, o, s_ring);
    if (s_bad && so.status == 0) so.status = 2;
    if (threadIdx.x == 0) { a.out_iters[m] = so.iters; a.out_status[m] = so.status; a.out_resid[m] = so.resid; }
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
    Precon<float, float> P;
    build_precon<float, float, float>(c, A32, F, P, o.precon, s_ring, &s_bad);
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
      SolveOut so = krylov<float, float, float>(c, A32, P, rin, d, W, oi, s_ring);
      iters += so.iters;
      for (int q = threadIdx.x; q < 2 * ncell; q += blockDim.x) x64[q] += rn * (double)d[q];
      __syncthreads();
      if (so.status == 2 && so.iters <= 1) { status = 2; break; }
    }
    if (s_bad && status == 0) status = 2;
    if (threadIdx.x == 0) { a.out_iters[m] = iters; a.out_status[m] = status; a.out_resid[m] = res; }
  }
}

template __global__ void yg_krylov_kernel<YG_MODE_FP64>(YgKrylovArgs);
template __global__ void yg_krylov_kernel<YG_MODE_MIXED>(YgKrylovArgs);

void yg_launch_krylov(const YgKrylovArgs& a, int mode, int nmembers, int block, hipStream_t st) {
  if (mode == YG_MODE_FP64)
    hipLaunchKernelGGL(yg_krylov_kernel<YG_MODE_FP64>, dim3(nmembers), dim3(block), 0, st, a);
  else
    hipLaunchKernelGGL(yg_krylov_kernel<YG_MODE_MIXED>, dim3(nmembers), dim3(block), 0, st, a);
}
