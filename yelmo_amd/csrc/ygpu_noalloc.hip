// ygpu_noalloc.hip -- ygpu_csr_solve without device allocation.
//
// ygpu_csr_solve_noalloc solves the same CSR system as ygpu_csr_solve with
// the same kernels, but takes every device buffer it needs (the staged CSR
// arrays, b and x, the stencil records, the ILU(0) factors, the Krylov work
// vectors and the outputs) from one file-scope device array, `scratch`, of
// YGPU_NOALLOC_SCRATCH_BYTES.  It calls no hipMalloc/hipFree and no host heap
// allocation on its success path: the CSR -> record conversion runs on the
// device, the step table lives on the host stack, times come from the host
// clock.  Each device holds its own copy of the array, whose address
// yg_noalloc_init looks up once per device (ygpu_create calls it with the
// device current); host code uses only that address, never the array's name.
// Every call on a device shares the array: what must start at zero is
// cleared with fill kernels, and calls that overlap on one device race.
#include <atomic>
#include <chrono>
#include <mutex>

#include "ygpu_host.h"

namespace {

// 256-byte alignment of every region carved out of the array
constexpr size_t kAlign = 256;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[YGPU_NOALLOC_SCRATCH_BYTES];

std::mutex g_init_mu;
void* g_scratch[kMaxDevices] = {};
std::atomic<int> g_ready[kMaxDevices];

// ---- device side ----------------------------------------------------------
__global__ void k_na_zero(uint32_t* __restrict__ p, long long nwords) {
  for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < nwords;
       q += (long long)gridDim.x * blockDim.x)
    p[q] = 0u;
}

// largest column offset (col - row) of the matrix: the widest coupling of
// the C-grid stencil, uy(i+1,j) from ux(i,j), is 2 ny + 1 columns right
__global__ void k_na_maxd(const int32_t* __restrict__ a_ptr1, const int32_t* __restrict__ a_idx1,
                          int n, int* maxd) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  int m = 0;
  for (int k = a_ptr1[r] - 1; k < a_ptr1[r + 1] - 1; ++k) m = max(m, (a_idx1[k] - 1) - r);
  atomicMax(maxd, m);
}

// One thread per row: the row's entries go to its cell's record (value
// 2 s + comp of slot s), summed in CSR order as the host conversion of
// ygpu_csr_solve does; the row's presence bits are or'ed into the cell's
// mask.  err receives the first error code: 1 column out of range, 2 entry
// outside the stencil, 3 row without its diagonal (when one is needed).
__global__ void k_na_records(const int32_t* __restrict__ a_ptr1, const int32_t* __restrict__ a_idx1,
                             const double* __restrict__ a_val, int n, int ny, int need_diag,
                             double* __restrict__ A, uint32_t* __restrict__ mask, int* err) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int cell = r >> 1, comp = r & 1;
  const int i = cell / ny, j = cell - i * ny;
  uint32_t bits = 0u;
  for (int k = a_ptr1[r] - 1; k < a_ptr1[r + 1] - 1; ++k) {
    const int col = a_idx1[k] - 1;
    if (col < 0 || col >= n) { atomicCAS(err, 0, 1); return; }
    const int c2 = col >> 1, comp2 = col & 1;
    const int i2 = c2 / ny;
    const int di = i2 - i, dj = (c2 - i2 * ny) - j;
    int s = -1;
    for (int t = 0; t < YG_NSLOT; ++t) {
      const int8_t* e = comp ? kSlotUy[t] : kSlotUx[t];
      if (e[0] == di && e[1] == dj && e[2] == comp2) { s = t; break; }
    }
    if (s < 0) { atomicCAS(err, 0, 2); return; }
    A[(size_t)YG_NA * cell + 2 * s + comp] += a_val[k];
    bits |= 1u << (16 * comp + s);
  }
  atomicOr(&mask[cell], bits);
  if (need_diag && !(bits & (1u << (16 * comp)))) atomicCAS(err, 0, 3);
}

__global__ void k_na_to_float(const double* __restrict__ a, float* __restrict__ b, long long n) {
  for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < n;
       q += (long long)gridDim.x * blockDim.x)
    b[q] = (float)a[q];
}

// ---- host side ------------------------------------------------------------
// Regions of the array, carved in order, each aligned to kAlign.
struct Carve {
  unsigned char* base;
  size_t used = 0;
  template <typename T>
  T* take(size_t count) {
    used = (used + kAlign - 1) / kAlign * kAlign;
    T* p = reinterpret_cast<T*>(base + used);
    used += count * sizeof(T);
    return p;
  }
};

int zero(hipStream_t st, void* p, size_t bytes) {
  const long long words = (long long)((bytes + 3) / 4);
  const int blocks = (int)std::min<long long>((words + 255) / 256, 2048);
  hipLaunchKernelGGL(k_na_zero, dim3(std::max(blocks, 1)), dim3(256), 0, st, (uint32_t*)p, words);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace

// Looks up this device's copy of the scratch array (once per device, with
// the device current); 0 on success.
int yg_noalloc_init(int device) {
  if (device < 0 || device >= kMaxDevices) return -1;
  if (g_ready[device].load(std::memory_order_acquire)) return 0;
  std::lock_guard<std::mutex> lk(g_init_mu);
  if (g_ready[device].load(std::memory_order_relaxed)) return 0;
  void* ptr = nullptr;
  if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return -1;
  g_scratch[device] = ptr;
  g_ready[device].store(1, std::memory_order_release);
  return 0;
}

using namespace yg;

extern "C" int ygpu_csr_solve_noalloc(ygpu_handle* h, int32_t n, int32_t nnz,
                                      const int32_t* a_ptr1, const int32_t* a_idx1,
                                      const double* a_val, const double* b, double* x_inout,
                                      const char* opts, int32_t* iters, int32_t* status,
                                      double* rel_resid, double* seconds) {
  if (!h) return YGPU_ERR_ARG;
  h->err.clear();
  h->note.clear();
  if (n <= 0 || (n & 1) || !a_ptr1 || !a_idx1 || !a_val || !b || !x_inout || !iters || !status ||
      !rel_resid) {
    h->err = "invalid arguments (n must be even: two unknowns per cell)";
    return YGPU_ERR_ARG;
  }
  if (a_ptr1[0] != 1 || a_ptr1[n] - 1 != nnz || nnz < 0) {
    h->err = "a_ptr must be 1-based with a_ptr(n+1)-1 == nnz";
    return YGPU_ERR_ARG;
  }
  // the device kernels walk every row's entries: the row pointers must be
  // non-decreasing (with a_ptr(1) = 1 and a_ptr(n+1) = nnz + 1 above, every
  // entry index stays in [0, nnz))
  for (int r = 0; r < n; ++r)
    if (a_ptr1[r + 1] < a_ptr1[r]) {
      h->err = "a_ptr must be non-decreasing";
      return YGPU_ERR_ARG;
    }
  const int cells = n / 2;
  if (cells > YGPU_NOALLOC_MAX_CELLS || nnz > YG_NSLOT * n) {
    h->err = "ygpu_csr_solve_noalloc: system larger than the scratch holds (at most " +
             std::to_string(YGPU_NOALLOC_MAX_CELLS) + " cells, 9 entries per row)";
    return YGPU_ERR_ARG;
  }
  Opts o;
  std::string e;
  int rc = parse_opts(opts, o, e);
  if (rc) { h->err = e; return rc; }
  if (o.substituted) h->note = o.substituted;
  if (seconds) *seconds = 0.0;
  // b == 0 shortcut (solver_linear.F90:162-174)
  bool allzero = true;
  for (int q = 0; q < n; ++q) if (b[q] != 0.0) { allzero = false; break; }
  if (allzero) {
    for (int q = 0; q < n; ++q) x_inout[q] = 0.0;
    *iters = 0; *status = 0; *rel_resid = 0.0;
    return YGPU_OK;
  }
  if (hipSetDevice(h->device) != hipSuccess) { h->err = "hipSetDevice failed"; return YGPU_ERR_DEVICE; }
  if (yg_noalloc_init(h->device) != 0) {
    h->err = "hipGetSymbolAddress of the no-allocation scratch failed";
    return YGPU_ERR_DEVICE;
  }
  const int kmode = o.prec == 2 ? YG_MODE_MIXED : YG_MODE_FP64;
  const int vlen = (int)(((2LL * cells + 3) / 4) * 4);

  // ---- carve the array: CSR staging, b, x, records, factors, work vectors
  Carve cv{(unsigned char*)g_scratch[h->device]};
  int32_t* d_ptr = cv.take<int32_t>(n + 1);
  int32_t* d_idx = cv.take<int32_t>(std::max(nnz, 1));
  double* d_val = cv.take<double>(std::max(nnz, 1));
  double* d_b = cv.take<double>(n);
  double* d_x = cv.take<double>(n);
  double* d_A64 = cv.take<double>((size_t)YG_NA * cells);
  uint32_t* d_mask = cv.take<uint32_t>(cells);
  YgMember* d_mem = cv.take<YgMember>(1);
  int* d_soff = cv.take<int>(YG_MAX_STEPS + 1);
  int* d_flags = cv.take<int>(4);   // maxd, error code, iterations, status
  double* d_res = cv.take<double>(1);
  double* d_F64 = cv.take<double>((size_t)YG_NF * cells);
  double* d_W64 = cv.take<double>((size_t)YG_NW64 * vlen);
  float *d_A32 = nullptr, *d_F32 = nullptr, *d_W32 = nullptr;
  uint16_t* d_F16 = nullptr;
  if (kmode == YG_MODE_MIXED) {
    d_A32 = cv.take<float>((size_t)YG_NA * cells);
    d_F32 = cv.take<float>((size_t)YG_NF * cells);
    d_W32 = cv.take<float>((size_t)YG_NW32 * vlen);
    if (o.k.factor16) d_F16 = cv.take<uint16_t>((size_t)YG_NF16 * cells);
  }
  if (cv.used > (size_t)YGPU_NOALLOC_SCRATCH_BYTES) {
    h->err = "ygpu_csr_solve_noalloc: system larger than the scratch holds";
    return YGPU_ERR_ARG;
  }

  hipStream_t st = h->stream;
  YG_CHECK(h, hipMemcpyAsync(d_ptr, a_ptr1, sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice, st));
  YG_CHECK(h, hipMemcpyAsync(d_idx, a_idx1, sizeof(int32_t) * nnz, hipMemcpyHostToDevice, st));
  YG_CHECK(h, hipMemcpyAsync(d_val, a_val, sizeof(double) * nnz, hipMemcpyHostToDevice, st));
  YG_CHECK(h, hipMemcpyAsync(d_b, b, sizeof(double) * n, hipMemcpyHostToDevice, st));
  YG_CHECK(h, hipMemcpyAsync(d_x, x_inout, sizeof(double) * n, hipMemcpyHostToDevice, st));
  // the array keeps the last call's contents: clear the records, the mask
  // and the flags
  if (zero(st, d_A64, sizeof(double) * YG_NA * cells) || zero(st, d_mask, sizeof(uint32_t) * cells) ||
      zero(st, d_flags, sizeof(int) * 4)) {
    h->err = "fill kernel launch failed";
    return YGPU_ERR_DEVICE;
  }
  const int rb = (n + 255) / 256;
  int ny = o.ny_hint;
  if (ny <= 0) {
    hipLaunchKernelGGL(k_na_maxd, dim3(rb), dim3(256), 0, st, d_ptr, d_idx, n, d_flags);
    YG_CHECK(h, hipGetLastError());
    int maxd = 0;
    YG_CHECK(h, hipMemcpyAsync(&maxd, d_flags, sizeof(int), hipMemcpyDeviceToHost, st));
    YG_CHECK(h, hipStreamSynchronize(st));
    ny = maxd >= 2 ? maxd / 2 : cells;
  }
  if (ny <= 0 || cells % ny) { h->err = "cannot infer the grid of this CSR pattern (give -ygpu_ny)"; return YGPU_ERR_UNSUPPORTED; }
  const int nx = cells / ny;
  if (nx > YG_MAX_LINES || ny > 65535) { h->err = "grid too large"; return YGPU_ERR_UNSUPPORTED; }
  hipLaunchKernelGGL(k_na_records, dim3(rb), dim3(256), 0, st, d_ptr, d_idx, d_val, n, ny,
                     o.k.precon != 0 ? 1 : 0, d_A64, d_mask, d_flags + 1);
  YG_CHECK(h, hipGetLastError());
  int bad = 0;
  YG_CHECK(h, hipMemcpyAsync(&bad, d_flags + 1, sizeof(int), hipMemcpyDeviceToHost, st));
  YG_CHECK(h, hipStreamSynchronize(st));
  if (bad == 1) { h->err = "column index out of range"; return YGPU_ERR_ARG; }
  if (bad == 2) { h->err = "entry outside the C-grid SSA stencil"; return YGPU_ERR_UNSUPPORTED; }
  if (bad == 3) { h->err = "row without a diagonal entry"; return YGPU_ERR_UNSUPPORTED; }

  // skewed step table on the host stack (no heap)
  const int T = ny + 2 * (nx - 1);
  if (T > YG_MAX_STEPS) { h->err = "grid too large for the step table"; return YGPU_ERR_UNSUPPORTED; }
  int soff[YG_MAX_STEPS + 1];
  soff[0] = 0;
  int maxcnt = 0;
  for (int t = 0; t < T; ++t) {
    const int a = t - (ny - 1);
    const int glo = a <= 0 ? 0 : (a + 1) / 2;
    const int ghi = std::min(nx - 1, t / 2);
    const int cnt = std::max(0, ghi - glo + 1);
    soff[t + 1] = soff[t] + cnt;
    maxcnt = std::max(maxcnt, cnt);
  }
  if (maxcnt > YG_MAX_STEP_CELLS) {
    h->err = "a wavefront step holds more cells than the LDS stage";
    return YGPU_ERR_UNSUPPORTED;
  }
  YgMember M{};
  M.nx = nx; M.ny = ny; M.nz = 1; M.T = T; M.coff = 0; M.toff = 0; M.dx = M.dy = 1.0;
  M.woff = 0;
  M.vlen = vlen;
  M.maxcnt = maxcnt;
  YG_CHECK(h, hipMemcpyAsync(d_mem, &M, sizeof(M), hipMemcpyHostToDevice, st));
  YG_CHECK(h, hipMemcpyAsync(d_soff, soff, sizeof(int) * (T + 1), hipMemcpyHostToDevice, st));
  if (kmode == YG_MODE_MIXED) {
    const long long na = (long long)YG_NA * cells;
    hipLaunchKernelGGL(k_na_to_float, dim3((int)std::min<long long>((na + 255) / 256, 2048)),
                       dim3(256), 0, st, d_A64, d_A32, na);
    YG_CHECK(h, hipGetLastError());
  }
  YgKrylovArgs ka{};
  ka.members = d_mem; ka.soff = d_soff; ka.mask = d_mask;
  ka.A64 = d_A64; ka.A32 = d_A32; ka.F64 = d_F64; ka.F32 = d_F32; ka.F16 = d_F16;
  ka.b = d_b; ka.x = d_x;
  ka.W64 = d_W64; ka.W32 = d_W32; ka.total_cells = cells; ka.wrapx = 0; ka.wrapy = 0;
  ka.skip = nullptr; ka.out_iters = d_flags + 2; ka.out_status = d_flags + 3; ka.out_resid = d_res;
  ka.opts = o.k;
  const int block = nx <= 512 ? 512 : std::min(1024, ((nx + 63) / 64) * 64);
  YG_CHECK(h, hipStreamSynchronize(st));
  const auto t0 = std::chrono::steady_clock::now();
  yg_launch_krylov(ka, kmode, 1, block, st);
  YG_CHECK(h, hipGetLastError());
  YG_CHECK(h, hipStreamSynchronize(st));
  const auto t1 = std::chrono::steady_clock::now();
  if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
  int outs[2] = {0, 0};
  double res = 0.0;
  YG_CHECK(h, hipMemcpyAsync(x_inout, d_x, sizeof(double) * n, hipMemcpyDeviceToHost, st));
  YG_CHECK(h, hipMemcpyAsync(outs, d_flags + 2, sizeof(int) * 2, hipMemcpyDeviceToHost, st));
  YG_CHECK(h, hipMemcpyAsync(&res, d_res, sizeof(double), hipMemcpyDeviceToHost, st));
  YG_CHECK(h, hipStreamSynchronize(st));
  *iters = outs[0]; *status = outs[1]; *rel_resid = res;
  return YGPU_OK;
}
