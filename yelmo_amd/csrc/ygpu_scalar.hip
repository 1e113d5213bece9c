// ygpu_scalar.hip -- the scalar five-point systems of ygpu_csr_solve: the
// implicit thickness advection of Yelmo (calc_advec2D "impl-lis",
// src/physics/solver_advection.f90:105-124, matrix from
// linear_solver_matrix_advection_csr_2D, :171-487), solved with LIS's
// BiCGSTAB and Jacobi (or no) preconditioner at the option string's
// tolerance (the reference passes "-i bicgstab -p jacobi -maxiter 1000
// -tol 1.0e-12 -initx_zeros false", :119).
//
// Rows are cells in linear_solver_init's order n = i ny + j (nvar = 1,
// solver_linear.F90:134-143).  Every row couples at most to (i, j-1),
// (i-1, j), (i, j), (i+1, j) and (i, j+1), with the neighbours wrapped
// around the domain edges as the assembly writes them (stored zeros stay
// zeros).  The matrix lives as five planar FP64 value arrays (one per
// stencil slot, coalesced), so the SpMV needs no index arrays.
//
// One system uses the whole device: every BiCGSTAB step is one grid-wide
// kernel (the step sequence of the oracle's restatement of LIS's BiCGSTAB),
// and every scalar the recurrences need is re-derived at the start of a
// kernel from the block partial sums of the kernels before it, summed in a
// fixed order, so every block takes the same decisions with the same bits.
// The host launches the steps of several iterations at a time and reads
// the termination flag once per group.
#include <chrono>

#include "ygpu_host.h"

namespace {

constexpr int kThreads = 256;
constexpr int kMaxBlocks = 1024;
constexpr int kIterPerLaunch = 8;   // BiCGSTAB iterations enqueued per flag read

// partial-sum slots
enum { P_RHO = 0, P_RTV, P_SS, P_TT, P_TS, P_RR, P_BB, NP };

struct ScState {
  double rho, rho_old, alpha, omega, inv_nrm0, resid;
  int it, status, done, iters;
};

struct ScArgs {
  int n, nx, ny, nb;
  const double* A;     // 5 planar slots of n values
  const double* dinv;  // Jacobi inverse diagonal (null: no preconditioner)
  const double* b;
  double *x, *r, *rt, *p, *v, *ph, *s, *sh, *t;
  double* part;        // NP x nb block partial sums
  ScState* st;
  double tol;
  int maxiter, conv_cond, initx_zeros;
};

__device__ __forceinline__ double block_sum(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  for (int k = 0; k < kThreads / 64; ++k) s += red[k];
  return s;
}

// sum of the nb partials of slot k (the same order in every block)
__device__ double part_sum(const ScArgs& a, int k, double* red) {
  double v = 0.0;
  for (int q = threadIdx.x; q < a.nb; q += kThreads) v += a.part[(size_t)k * a.nb + q];
  return block_sum(v, red);
}

__device__ __forceinline__ void put_part(const ScArgs& a, int k, double v) {
  if (threadIdx.x == 0) a.part[(size_t)k * a.nb + blockIdx.x] = v;
}

// y = A x at cell q (neighbours wrapped)
__device__ __forceinline__ double row_dot(const ScArgs& a, const double* __restrict__ x, int q) {
  const int n = a.n, ny = a.ny, nx = a.nx;
  const int i = q / ny, j = q - i * ny;
  const int jm = j > 0 ? q - 1 : q + ny - 1;
  const int im = i > 0 ? q - ny : q + (nx - 1) * ny;
  const int ip = i < nx - 1 ? q + ny : q - (nx - 1) * ny;
  const int jp = j < ny - 1 ? q + 1 : q - (ny - 1);
  const double* A = a.A;
  return A[q] * x[jm] + A[(size_t)n + q] * x[im] + A[(size_t)2 * n + q] * x[q] +
         A[(size_t)3 * n + q] * x[ip] + A[(size_t)4 * n + q] * x[jp];
}

__device__ __forceinline__ double precon(const ScArgs& a, double v, int q) {
  return a.dinv ? a.dinv[q] * v : v;
}

// r = b - A x0 (x0 = 0 with -initx_zeros true), rt = r; (r,r), (b,b)
__global__ void __launch_bounds__(kThreads) k_sc_init(ScArgs a) {
  __shared__ double red[kThreads / 64];
  double rr = 0.0, bb = 0.0;
  for (int q = blockIdx.x * kThreads + threadIdx.x; q < a.n; q += a.nb * kThreads) {
    const double ax = a.initx_zeros ? 0.0 : row_dot(a, a.x, q);
    const double rq = a.b[q] - ax;
    a.r[q] = rq;
    a.rt[q] = rq;
    rr += rq * rq;
    bb += a.b[q] * a.b[q];
  }
  rr = block_sum(rr, red);
  bb = block_sum(bb, red);
  put_part(a, P_RR, rr);
  put_part(a, P_RHO, rr);
  put_part(a, P_BB, bb);
}

// x = 0 for -initx_zeros true (after the residual, which read x as zero)
__global__ void __launch_bounds__(kThreads) k_sc_zero_x(ScArgs a) {
  for (int q = blockIdx.x * kThreads + threadIdx.x; q < a.n; q += a.nb * kThreads) a.x[q] = 0.0;
}

__global__ void __launch_bounds__(kThreads) k_sc_start(ScArgs a) {
  __shared__ double red[kThreads / 64];
  const double rr = part_sum(a, P_RR, red), bb = part_sum(a, P_BB, red);
  if (threadIdx.x) return;
  ScState& s = *a.st;
  const double nrm0 = a.conv_cond == 1 ? sqrt(bb) : sqrt(rr);
  s.rho_old = 1.0; s.alpha = 1.0; s.omega = 1.0; s.it = 1; s.iters = 0; s.status = 4; s.done = 0;
  if (nrm0 == 0.0) { s.inv_nrm0 = 0.0; s.resid = 0.0; s.status = 0; s.done = 1; return; }
  s.inv_nrm0 = 1.0 / nrm0;
  s.resid = sqrt(rr) * s.inv_nrm0;
  if (s.resid <= a.tol) { s.status = 0; s.done = 1; return; }
  if (a.maxiter < 1) { s.status = 4; s.done = 1; }
}

// p = r + beta (p - omega v) (p = r at the first iteration); phat = M^-1 p
__global__ void __launch_bounds__(kThreads) k_sc_p(ScArgs a) {
  __shared__ double red[kThreads / 64];
  if (a.st->done) return;
  const double rho = part_sum(a, P_RHO, red);
  const ScState s = *a.st;
  // the later steps of the iteration read rho here (a zero ends the solve)
  if (blockIdx.x == 0 && threadIdx.x == 0) a.st->rho = rho;
  if (rho == 0.0) return;
  const bool first = s.it == 1;
  const double beta = first ? 0.0 : (rho / s.rho_old) * (s.alpha / s.omega);
  for (int q = blockIdx.x * kThreads + threadIdx.x; q < a.n; q += a.nb * kThreads) {
    const double pq = first ? a.r[q] : a.r[q] + beta * (a.p[q] - s.omega * a.v[q]);
    a.p[q] = pq;
    a.ph[q] = precon(a, pq, q);
  }
}

// v = A phat; (rt, v)
__global__ void __launch_bounds__(kThreads) k_sc_v(ScArgs a) {
  __shared__ double red[kThreads / 64];
  if (a.st->done) return;
  if (a.st->rho == 0.0) return;
  double d = 0.0;
  for (int q = blockIdx.x * kThreads + threadIdx.x; q < a.n; q += a.nb * kThreads) {
    const double vq = row_dot(a, a.ph, q);
    a.v[q] = vq;
    d += a.rt[q] * vq;
  }
  put_part(a, P_RTV, block_sum(d, red));
}

// alpha = rho / (rt, v); s = r - alpha v; shat = M^-1 s; (s, s)
__global__ void __launch_bounds__(kThreads) k_sc_s(ScArgs a) {
  __shared__ double red[kThreads / 64];
  if (a.st->done) return;
  const double rho = a.st->rho;
  if (rho == 0.0) return;
  const double tmp = part_sum(a, P_RTV, red);
  if (tmp == 0.0) return;
  const double alpha = rho / tmp;
  if (blockIdx.x == 0 && threadIdx.x == 0) a.st->alpha = alpha;
  double ss = 0.0;
  for (int q = blockIdx.x * kThreads + threadIdx.x; q < a.n; q += a.nb * kThreads) {
    const double sq = a.r[q] - alpha * a.v[q];
    a.s[q] = sq;
    a.sh[q] = precon(a, sq, q);
    ss += sq * sq;
  }
  put_part(a, P_SS, block_sum(ss, red));
}

// converged on s: x += alpha phat; else t = A shat, (t, t), (t, s)
__global__ void __launch_bounds__(kThreads) k_sc_t(ScArgs a) {
  __shared__ double red[kThreads / 64];
  if (a.st->done) return;
  const double rho = a.st->rho;
  if (rho == 0.0) return;
  if (part_sum(a, P_RTV, red) == 0.0) return;
  const double alpha = a.st->alpha;
  const double sn = sqrt(part_sum(a, P_SS, red)) * a.st->inv_nrm0;
  if (sn <= a.tol) {
    for (int q = blockIdx.x * kThreads + threadIdx.x; q < a.n; q += a.nb * kThreads)
      a.x[q] += alpha * a.ph[q];
    return;
  }
  double tt = 0.0, ts = 0.0;
  for (int q = blockIdx.x * kThreads + threadIdx.x; q < a.n; q += a.nb * kThreads) {
    const double tq = row_dot(a, a.sh, q);
    a.t[q] = tq;
    tt += tq * tq;
    ts += tq * a.s[q];
  }
  tt = block_sum(tt, red);
  ts = block_sum(ts, red);
  put_part(a, P_TT, tt);
  put_part(a, P_TS, ts);
}

// omega = (t, s) / (t, t); x += alpha phat + omega shat; r = s - omega t;
// (r, r), (rt, r)
__global__ void __launch_bounds__(kThreads) k_sc_x(ScArgs a) {
  __shared__ double red[kThreads / 64];
  if (a.st->done) return;
  const double rho = a.st->rho;
  if (rho == 0.0) return;
  if (part_sum(a, P_RTV, red) == 0.0) return;
  if (sqrt(part_sum(a, P_SS, red)) * a.st->inv_nrm0 <= a.tol) return;
  const double tt = part_sum(a, P_TT, red);
  if (tt == 0.0) return;
  const double omega = part_sum(a, P_TS, red) / tt;
  if (omega == 0.0) return;
  const double alpha = a.st->alpha;
  if (blockIdx.x == 0 && threadIdx.x == 0) a.st->omega = omega;
  double rr = 0.0, rtr = 0.0;
  for (int q = blockIdx.x * kThreads + threadIdx.x; q < a.n; q += a.nb * kThreads) {
    a.x[q] += alpha * a.ph[q] + omega * a.sh[q];
    const double rq = a.s[q] - omega * a.t[q];
    a.r[q] = rq;
    rr += rq * rq;
    rtr += a.rt[q] * rq;
  }
  rr = block_sum(rr, red);
  rtr = block_sum(rtr, red);
  put_part(a, P_RR, rr);
  put_part(a, P_RHO, rtr);
}

// the iteration's outcome, in the order the recurrences test it
__global__ void __launch_bounds__(kThreads) k_sc_end(ScArgs a) {
  __shared__ double red[kThreads / 64];
  if (a.st->done) return;
  ScState& s = *a.st;
  const double rho = s.rho;
  const double tmp = rho == 0.0 ? 0.0 : part_sum(a, P_RTV, red);
  const double sn = (rho == 0.0 || tmp == 0.0) ? 0.0 : sqrt(part_sum(a, P_SS, red)) * s.inv_nrm0;
  const double tt = (rho == 0.0 || tmp == 0.0 || sn <= a.tol) ? 0.0 : part_sum(a, P_TT, red);
  const double ts = tt == 0.0 ? 0.0 : part_sum(a, P_TS, red);
  const double rr = part_sum(a, P_RR, red);
  if (threadIdx.x) return;
  const int it = s.it;
  if (rho == 0.0 || tmp == 0.0) { s.status = 2; s.iters = it; s.done = 1; return; }
  if (sn <= a.tol) { s.resid = sn; s.iters = it; s.status = 0; s.done = 1; return; }
  if (tt == 0.0 || ts / tt == 0.0) { s.status = 2; s.iters = it; s.done = 1; return; }
  s.resid = sqrt(rr) * s.inv_nrm0;
  s.iters = it;
  if (s.resid <= a.tol) { s.status = 0; s.done = 1; return; }
  s.rho_old = rho;
  s.it = it + 1;
  if (s.it > a.maxiter) { s.status = 4; s.done = 1; }
}

}  // namespace

namespace yg {

// CSR rows of at most five entries on the scalar five-point stencil (see
// the file header) -> planar records; ny from the most frequent column
// offset above 1 (the i +- 1 couplings) unless given.
static int scalar_records(ygpu_handle* h, int n, const int32_t* a_ptr1, const int32_t* a_idx1,
                          const double* a_val, int ny_hint, int& nx, int& ny,
                          std::vector<double>& A) {
  ny = ny_hint;
  if (ny <= 0) {
    std::map<int, int> hist;
    for (int r = 0; r < n; ++r)
      for (int k = a_ptr1[r] - 1; k < a_ptr1[r + 1] - 1; ++k) {
        const int d = std::abs((a_idx1[k] - 1) - r);
        if (d > 1) ++hist[d];
      }
    int best = 0;
    for (const auto& e : hist)
      if (e.second > best) { best = e.second; ny = e.first; }
  }
  if (ny < 3 || n % ny || n / ny < 3) {
    h->err = "cannot infer the grid of this five-point CSR pattern (give -ygpu_ny; nx, ny >= 3)";
    return YGPU_ERR_UNSUPPORTED;
  }
  nx = n / ny;
  A.assign((size_t)5 * n, 0.0);
  for (int r = 0; r < n; ++r) {
    const int i = r / ny, j = r % ny;
    for (int k = a_ptr1[r] - 1; k < a_ptr1[r + 1] - 1; ++k) {
      const int col = a_idx1[k] - 1;
      if (col < 0 || col >= n) { h->err = "column index out of range"; return YGPU_ERR_ARG; }
      int di = col / ny - i, dj = col % ny - j;
      if (di == nx - 1) di = -1; else if (di == -(nx - 1)) di = 1;
      if (dj == ny - 1) dj = -1; else if (dj == -(ny - 1)) dj = 1;
      int s = -1;
      if (di == 0 && dj == -1) s = 0;
      else if (di == -1 && dj == 0) s = 1;
      else if (di == 0 && dj == 0) s = 2;
      else if (di == 1 && dj == 0) s = 3;
      else if (di == 0 && dj == 1) s = 4;
      if (s < 0) { h->err = "entry outside the five-point stencil"; return YGPU_ERR_UNSUPPORTED; }
      A[(size_t)s * n + r] += a_val[k];
    }
  }
  return YGPU_OK;
}

int scalar_csr_solve(ygpu_handle* h, int32_t n, int32_t nnz, const int32_t* a_ptr1,
                     const int32_t* a_idx1, const double* a_val, const double* b, double* x_inout,
                     const Opts& o, int32_t* iters, int32_t* status, double* rel_resid,
                     double* seconds) {
  (void)nnz;
  if (o.k.solver != 4) {
    h->err = "five-point systems: BiCGSTAB only (-i bicgstab)";
    return YGPU_ERR_UNSUPPORTED;
  }
  if (o.k.precon == 2) {
    h->err = "five-point systems: Jacobi or no preconditioner (-p jacobi | none)";
    return YGPU_ERR_UNSUPPORTED;
  }
  int nx = 0, ny = 0;
  std::vector<double> A;
  int rc = scalar_records(h, n, a_ptr1, a_idx1, a_val, o.ny_hint, nx, ny, A);
  if (rc) return rc;
  std::vector<double> dinv;
  if (o.k.precon == 1) {
    dinv.resize(n);
    for (int q = 0; q < n; ++q) {
      const double d = A[(size_t)2 * n + q];
      if (d == 0.0) { h->err = "row without a diagonal entry"; return YGPU_ERR_UNSUPPORTED; }
      dinv[q] = 1.0 / d;
    }
  }
  if (hipSetDevice(h->device) != hipSuccess) { h->err = "hipSetDevice failed"; return YGPU_ERR_DEVICE; }
  Arena ar;
  ScArgs a{};
  a.n = n; a.nx = nx; a.ny = ny;
  a.nb = (int)std::min<long long>((n + kThreads - 1) / kThreads, kMaxBlocks);
  double *dA, *db, *dd = nullptr, *dpart, *vec;
  ScState* dst;
  YG_CHECK(h, ar.alloc(&dA, A.size()));
  YG_CHECK(h, ar.alloc(&db, n));
  YG_CHECK(h, ar.alloc(&vec, (size_t)10 * n));
  YG_CHECK(h, ar.alloc(&dpart, (size_t)NP * a.nb));
  YG_CHECK(h, ar.alloc(&dst, 1));
  if (!dinv.empty()) YG_CHECK(h, ar.alloc(&dd, n));
  hipStream_t st = h->stream;
  YG_CHECK(h, hipMemcpyAsync(dA, A.data(), sizeof(double) * A.size(), hipMemcpyHostToDevice, st));
  YG_CHECK(h, hipMemcpyAsync(db, b, sizeof(double) * n, hipMemcpyHostToDevice, st));
  YG_CHECK(h, hipMemcpyAsync(vec, x_inout, sizeof(double) * n, hipMemcpyHostToDevice, st));
  if (dd) YG_CHECK(h, hipMemcpyAsync(dd, dinv.data(), sizeof(double) * n, hipMemcpyHostToDevice, st));
  a.A = dA; a.dinv = dd; a.b = db;
  a.x = vec; a.r = vec + (size_t)n; a.rt = vec + (size_t)2 * n; a.p = vec + (size_t)3 * n;
  a.v = vec + (size_t)4 * n; a.ph = vec + (size_t)5 * n; a.s = vec + (size_t)6 * n;
  a.sh = vec + (size_t)7 * n; a.t = vec + (size_t)8 * n;
  a.part = dpart; a.st = dst;
  a.tol = o.k.tol; a.maxiter = o.k.maxiter; a.conv_cond = o.k.conv_cond;
  a.initx_zeros = o.k.initx_zeros;
  const dim3 grid(a.nb), blk(kThreads);
  hipEvent_t e0, e1;
  YG_CHECK(h, hipEventCreate(&e0));
  YG_CHECK(h, hipEventCreate(&e1));
  YG_CHECK(h, hipEventRecord(e0, st));
  hipLaunchKernelGGL(k_sc_init, grid, blk, 0, st, a);
  if (a.initx_zeros) hipLaunchKernelGGL(k_sc_zero_x, grid, blk, 0, st, a);
  hipLaunchKernelGGL(k_sc_start, dim3(1), blk, 0, st, a);
  YG_CHECK(h, hipGetLastError());
  ScState hs{};
  for (int done_iters = 0; done_iters <= o.k.maxiter; done_iters += kIterPerLaunch) {
    for (int k = 0; k < kIterPerLaunch; ++k) {
      hipLaunchKernelGGL(k_sc_p, grid, blk, 0, st, a);
      hipLaunchKernelGGL(k_sc_v, grid, blk, 0, st, a);
      hipLaunchKernelGGL(k_sc_s, grid, blk, 0, st, a);
      hipLaunchKernelGGL(k_sc_t, grid, blk, 0, st, a);
      hipLaunchKernelGGL(k_sc_x, grid, blk, 0, st, a);
      hipLaunchKernelGGL(k_sc_end, dim3(1), blk, 0, st, a);
    }
    YG_CHECK(h, hipGetLastError());
    YG_CHECK(h, hipMemcpyAsync(&hs, dst, sizeof(hs), hipMemcpyDeviceToHost, st));
    YG_CHECK(h, hipStreamSynchronize(st));
    if (hs.done) break;
  }
  YG_CHECK(h, hipEventRecord(e1, st));
  YG_CHECK(h, hipEventSynchronize(e1));
  float ms = 0.f;
  YG_CHECK(h, hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (seconds) *seconds = ms * 1e-3;
  YG_CHECK(h, hipMemcpy(x_inout, vec, sizeof(double) * n, hipMemcpyDeviceToHost));
  *iters = hs.iters; *status = hs.status; *rel_resid = hs.resid;
  return YGPU_OK;
}

}  // namespace yg
