// ygpu_spmv.hip -- the stencil-record SpMV y = A x over every cell of a
// batch in one full-device launch (the matrix-vector product of the
// assembled systems, solver_ssa_ac.f90:73-848 pattern; LIS's lis_matvec at
// the CSR boundary).  One thread per cell computes both rows of its cell
// from the cell's 20-value record and the 7 neighbour velocity pairs.  It is
// HBM-bound: the record (80 B FP32 / 160 B FP64) is read once, through LDS,
// the neighbour pairs mostly come from cache, and y is written once.  The
// members of a batch are contiguous; each block finds its members from the
// batch's chunk table (the member of every chunk's first cell).
#include "ygpu_spmv.h"

namespace {

template <typename T>
struct Pair { using type = float2; };
template <>
struct Pair<double> { using type = double2; };

constexpr int kSpmvBlock = YG_CHUNK;   // one block per chunk of the member table

template <typename T>
__global__ __launch_bounds__(kSpmvBlock) void k_spmv_records(YgBatchGeom G, const T* __restrict__ A,
                                                             const T* __restrict__ x,
                                                             T* __restrict__ y, int wrapx,
                                                             int wrapy) {
  using P2 = typename Pair<T>::type;
  // The block's records are contiguous: stage them through LDS with 16-byte
  // loads of consecutive addresses across the lanes (a per-cell load would
  // put the lanes of one instruction 80 or 160 bytes apart), non-temporal
  // (each record is read once; x's neighbour pairs are what the caches keep).
  __shared__ uint4 rec[kSpmvBlock * YG_NA * sizeof(T) / 16];
  const long long g0 = (long long)blockIdx.x * kSpmvBlock;
  // the members of the block's cells run from m0 (its first cell's) to at
  // most m1 (the next block's first cell's), from the batch's chunk table;
  // the first member's geometry is loaded before the records are waited for
  const int m0 = G.chunk_seg[blockIdx.x], m1 = G.chunk_seg[blockIdx.x + 1];
  const YgMember M0 = G.members[m0];
  const long long c00 = G.coffs[m0];
  const long long nblk = min((long long)kSpmvBlock, G.total_cells - g0);
  const int nchunk = (int)(nblk * YG_NA * sizeof(T) / 16);
  typedef unsigned V4u __attribute__((ext_vector_type(4)));
  const V4u* src = reinterpret_cast<const V4u*>(A + (size_t)YG_NA * g0);
  // every chunk of this thread (10 FP64, 5 FP32) is requested before the
  // first is waited for: one load in flight per thread would leave the CU's
  // memory pipeline mostly idle
  constexpr int kPer = YG_NA * (int)sizeof(T) / 16;
  V4u v[kPer];
  if (nblk == kSpmvBlock) {   // a whole block: no per-chunk bounds
#pragma unroll
    for (int u = 0; u < kPer; ++u) v[u] = __builtin_nontemporal_load(src + threadIdx.x + u * kSpmvBlock);
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int k = threadIdx.x + u * kSpmvBlock;
      rec[k] = make_uint4(v[u].x, v[u].y, v[u].z, v[u].w);
    }
  } else {
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int k = threadIdx.x + u * kSpmvBlock;
      if (k < nchunk) v[u] = __builtin_nontemporal_load(src + k);
    }
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int k = threadIdx.x + u * kSpmvBlock;
      if (k < nchunk) rec[k] = make_uint4(v[u].x, v[u].y, v[u].z, v[u].w);
    }
  }
  __syncthreads();
  const long long gq = g0 + threadIdx.x;
  if (gq >= G.total_cells) return;
  int m = m0, nx = M0.nx, ny = M0.ny;
  long long c0 = c00;
  if (m1 != m0) {   // the block spans several members
    while (m < m1 && gq >= G.coffs[m + 1]) ++m;
    const YgMember& mb = G.members[m];
    nx = mb.nx; ny = mb.ny; c0 = G.coffs[m];
  }
  const int q = (int)(gq - c0);
  if (q >= nx * ny) return;   // between members (none in a contiguous batch)
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

std::vector<int> yg_chunk_members(const long long* coffs, int nmem, long long total) {
  const long long nb = (total + YG_CHUNK - 1) / YG_CHUNK;
  std::vector<int> t(nb + 1);
  int m = 0;
  for (long long b = 0; b < nb; ++b) {
    const long long g = b * YG_CHUNK;
    while (m + 1 < nmem && g >= coffs[m + 1]) ++m;
    t[b] = m;
  }
  t[nb] = nmem - 1;
  return t;
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
