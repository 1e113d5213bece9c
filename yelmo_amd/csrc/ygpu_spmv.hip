// ygpu_spmv.hip -- the stencil-record SpMV y = A x over every cell of a
// batch in one full-device launch (the matrix-vector product of the
// assembled systems, solver_ssa_ac.f90:73-848 pattern; LIS's lis_matvec at
// the CSR boundary).  One thread per cell computes both rows of its cell
// from the cell's 20-value record and the 7 neighbour velocity pairs.  It is
// HBM-bound: the record (80 B FP32 / 160 B FP64) is read once, through LDS,
// the neighbour pairs mostly come from cache, and y is written once.
#include "ygpu_spmv.h"

namespace {

template <typename T>
struct Pair { using type = float2; };
template <>
struct Pair<double> { using type = double2; };

constexpr int kSpmvBlock = 256;
constexpr int kSpmvSegs = 1024;   // member ranges searched in LDS (more: in global memory)

template <typename T>
__global__ __launch_bounds__(kSpmvBlock) void k_spmv_records(YgBatchGeom G, const T* __restrict__ A,
                                                             const T* __restrict__ x,
                                                             T* __restrict__ y, int wrapx,
                                                             int wrapy) {
  using P2 = typename Pair<T>::type;
  // The block's records are contiguous: stage them through LDS with 16-byte
  // loads of consecutive addresses across the lanes (a per-cell load would
  // put the lanes of one instruction 80 or 160 bytes apart)
  __shared__ uint4 rec[kSpmvBlock * YG_NA * sizeof(T) / 16];
  // the occupied-range list too, so that each thread's member search is a
  // search in LDS instead of a chain of dependent global loads
  __shared__ long long s_seg[kSpmvSegs];
  const long long g0 = (long long)blockIdx.x * kSpmvBlock;
  const long long nblk = min((long long)kSpmvBlock, G.total_cells - g0);
  const int nchunk = (int)(nblk * YG_NA * sizeof(T) / 16);
  const uint4* src = reinterpret_cast<const uint4*>(A + (size_t)YG_NA * g0);
  for (int k = threadIdx.x; k < nchunk; k += kSpmvBlock) rec[k] = src[k];
  const bool lds_seg = G.nseg <= kSpmvSegs && !G.segm;
  if (lds_seg)
    for (int k = threadIdx.x; k < G.nseg; k += kSpmvBlock) s_seg[k] = G.seg[k];
  __syncthreads();
  const long long gq = g0 + threadIdx.x;
  if (gq >= G.total_cells) return;
  int m;
  if (lds_seg) {
    int lo = 0, hi = G.nseg - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (s_seg[mid] <= gq) lo = mid; else hi = mid - 1;
    }
    m = lo;
    const YgMember& mm = G.members[m];
    if (gq < s_seg[lo] || gq - s_seg[lo] >= (long long)mm.nx * mm.ny) return;
  } else {
    m = yg_member_at(G, gq);
    if (m < 0) return;
  }
  const YgMember& mb = G.members[m];
  const long long c0 = G.coffs[m];
  const int nx = mb.nx, ny = mb.ny;
  const int q = (int)(gq - c0);
  const int i = q / ny, j = q - i * ny;
  // neighbour cells in the row order n = i*ny + j; an index outside a
  // non-periodic domain only meets a zero record value, so it is clamped
  const int ip = i + 1 < nx ? i + 1 : (wrapx ? 0 : i);
  const int im = i > 0 ? i - 1 : (wrapx ? nx - 1 : i);
  const int jp = j + 1 < ny ? j + 1 : (wrapy ? 0 : j);
  const int jm = j > 0 ? j - 1 : (wrapy ? ny - 1 : j);
  const P2* xp = reinterpret_cast<const P2*>(x + 2 * c0);
  const P2 xc = xp[q];
  const P2 xip = xp[ip * ny + j], xim = xp[im * ny + j];
  const P2 xjp = xp[i * ny + jp], xjm = xp[i * ny + jm];
  const P2 xipjm = xp[ip * ny + jm], ximjp = xp[im * ny + jp];
  const T* r = reinterpret_cast<const T*>(rec) + (size_t)YG_NA * threadIdx.x;
  T a[18];
#pragma unroll
  for (int e = 0; e < 18; ++e) a[e] = r[e];
  // slots (ygpu_common.h): ux row 0 ux(i,j) 1 ux(i+1,j) 2 ux(i-1,j) 3 ux(i,j+1)
  // 4 ux(i,j-1) 5 uy(i,j) 6 uy(i+1,j) 7 uy(i+1,j-1) 8 uy(i,j-1); uy row 0
  // uy(i,j) 1 uy(i,j+1) 2 uy(i,j-1) 3 uy(i+1,j) 4 uy(i-1,j) 5 ux(i,j)
  // 6 ux(i,j+1) 7 ux(i-1,j+1) 8 ux(i-1,j)
  const T yx = a[0] * xc.x + a[2] * xip.x + a[4] * xim.x + a[6] * xjp.x + a[8] * xjm.x +
               a[10] * xc.y + a[12] * xip.y + a[14] * xipjm.y + a[16] * xjm.y;
  const T yy = a[1] * xc.y + a[3] * xjp.y + a[5] * xjm.y + a[7] * xip.y + a[9] * xim.y +
               a[11] * xc.x + a[13] * xjp.x + a[15] * ximjp.x + a[17] * xim.x;
  P2 o;
  o.x = yx;
  o.y = yy;
  reinterpret_cast<P2*>(y + 2 * c0)[q] = o;
}

}  // namespace

template <typename T>
void yg_launch_spmv(const YgBatchGeom& G, const T* A, const T* x, T* y, int wrapx, int wrapy,
                    hipStream_t st) {
  const unsigned nb = (unsigned)((G.total_cells + kSpmvBlock - 1) / kSpmvBlock);
  hipLaunchKernelGGL(k_spmv_records<T>, dim3(nb), dim3(kSpmvBlock), 0, st, G, A, x, y, wrapx,
                     wrapy);
}

template void yg_launch_spmv<float>(const YgBatchGeom&, const float*, const float*, float*, int,
                                    int, hipStream_t);
template void yg_launch_spmv<double>(const YgBatchGeom&, const double*, const double*, double*,
                                     int, int, hipStream_t);

namespace {
__global__ void k_to_float(const double* __restrict__ src, float* __restrict__ dst, long long n) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) dst[k] = (float)src[k];
}
}  // namespace

void yg_launch_to_float(const double* src, float* dst, long long n, hipStream_t st) {
  hipLaunchKernelGGL(k_to_float, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, src, dst, n);
}
