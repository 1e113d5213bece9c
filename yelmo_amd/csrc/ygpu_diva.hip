// ygpu_diva.hip -- the DIVA parts of the Picard iteration
// (calc_velocity_diva, velocity_diva.f90:59-429): vertical shear from the
// basal stress, and the 3D velocity after the loop.  The depth integral F2
// and beta_eff are formed inside the shared viscosity and staggering
// kernels, the basal stress and basal velocity inside the Picard epilogue.
#include "ygpu_diva.h"

namespace {

#define I2(i, j) ((i) + nx * (j))

YG_DEV void dv_nbr(int i, int j, int nx, int ny, int BC, int& ip1, int& jp1) {
  const bool wx = BC == YG_BC_PERIODIC || BC == YG_BC_PERIODIC_X;
  const bool wy = BC == YG_BC_MISMIP3D || BC == YG_BC_TROUGH || BC == YG_BC_PERIODIC ||
                  BC == YG_BC_PERIODIC_Y;
  ip1 = i < nx - 1 ? i + 1 : (wx ? 0 : nx - 1);
  jp1 = j < ny - 1 ? j + 1 : (wy ? 0 : ny - 1);
}

// calc_staggered_margin (velocity_diva.f90:1052-1082)
template <typename R>
YG_DEV R stag_margin(R v0, R v1, R f0, R f1) {
  if (f0 == R(1) && f1 < R(1)) return v0;
  if (f0 < R(1) && f1 == R(1)) return v1;
  return R(0.5) * (v0 + v1);
}

// calc_vertical_shear_3D (velocity_diva.f90:510-571)
template <typename R>
__global__ __launch_bounds__(256) void k_shear(YgBatchGeom G, YgFields<R> f, YgFieldParams p, const int* skip) {
  const long long gc = yg_xcd_cell();
  if (gc >= G.total_cells) return;
  const int m = yg_member_at(G, gc);
  if (m < 0 || (skip && skip[m])) return;
  const YgMember& mb = G.members[m];
  const int nx = mb.nx, ny = mb.ny, nz = mb.nz;
  const long long c0 = G.coffs[m], z0 = G.zoffs[m], nc = (long long)nx * ny;
  const int loc = (int)(gc - c0);
  const int i = loc % nx, j = loc / nx;
  int ip1, jp1;
  dv_nbr(i, j, nx, ny, p.BC, ip1, jp1);
  const R* fi = f.f_ice + c0;
  const R* visc = f.visc + z0;
  const R* zeta = f.zeta + (long long)m * f.zstride;
  const R tbx = f.taub_x[c0 + loc], tby = f.taub_y[c0 + loc];
  for (int k = 0; k < nz; ++k) {
    R va = stag_margin<R>(visc[I2(i, j) + k * nc], visc[I2(ip1, j) + k * nc], fi[I2(i, j)], fi[I2(ip1, j)]);
    f.duxdz[z0 + loc + k * nc] = va != R(0) ? (tbx / va) * (R(1) - zeta[k]) : R(0);
    va = stag_margin<R>(visc[I2(i, j) + k * nc], visc[I2(i, jp1) + k * nc], fi[I2(i, j)], fi[I2(i, jp1)]);
    f.duydz[z0 + loc + k * nc] = va != R(0) ? (tby / va) * (R(1) - zeta[k]) : R(0);
  }
}

// calc_vel_horizontal_3D (velocity_diva.f90:431-508)
template <typename R>
__global__ __launch_bounds__(256) void k_vel3d(YgBatchGeom G, YgFields<R> f, YgDivaFields<R> d, YgFieldParams p,
                        const int* skip) {
  const long long gc = yg_xcd_cell();
  if (gc >= G.total_cells) return;
  const int m = yg_member_at(G, gc);
  if (m < 0 || (skip && skip[m])) return;
  const YgMember& mb = G.members[m];
  const int nx = mb.nx, ny = mb.ny, nz = mb.nz;
  const long long c0 = G.coffs[m], z0 = G.zoffs[m], nc = (long long)nx * ny;
  const int loc = (int)(gc - c0);
  const int i = loc % nx, j = loc / nx;
  int ip1, jp1;
  dv_nbr(i, j, nx, ny, p.BC, ip1, jp1);
  const R* fi = f.f_ice + c0;
  const R* H = f.H_ice + c0;
  const R* visc = f.visc + z0;
  const R* zeta = f.zeta + (long long)m * f.zstride;
  // F1(k) = integral_0^zeta_k (H/visc)(1 - zeta) dzeta at the three cells
  const int cs[3] = {I2(i, j), I2(ip1, j), I2(i, jp1)};
  R F1[3] = {R(0), R(0), R(0)};
  R prev[3];
  for (int s = 0; s < 3; ++s)
    prev[s] = fi[cs[s]] == R(1) ? (H[cs[s]] / visc[cs[s]]) * (R(1) - zeta[0]) : R(0);
  for (int k = 0; k < nz; ++k) {
    if (k > 0) {
      for (int s = 0; s < 3; ++s) {
        if (fi[cs[s]] != R(1)) continue;
        const R cur = (H[cs[s]] / visc[cs[s] + k * nc]) * (R(1) - zeta[k]);
        const R mid = R(0.5) * (cur + prev[s]);
        F1[s] = F1[s] + mid * (zeta[k] - zeta[k - 1]);
        prev[s] = cur;
      }
    }
    const R fx = stag_margin<R>(F1[0], F1[1], fi[cs[0]], fi[cs[1]]);
    const R fy = stag_margin<R>(F1[0], F1[2], fi[cs[0]], fi[cs[2]]);
    d.ux3[z0 + loc + k * nc] = f.ux[gc] + f.taub_x[gc] * fx;
    d.uy3[z0 + loc + k * nc] = f.uy[gc] + f.taub_y[gc] * fy;
  }
}

}  // namespace

static inline unsigned dnblk(long long n, int b) { return yg_xcd_grid(n, b); }

template <typename R>
void yg_launch_diva_pre(const YgBatchGeom& G, const YgFields<R>& f, const YgDivaFields<R>&,
                        const YgFieldParams& p, const int* skip, hipStream_t st) {
  hipLaunchKernelGGL(k_shear<R>, dim3(dnblk(G.total_cells, 256)), dim3(256), 0, st, G, f, p, skip);
}

template <typename R>
void yg_launch_diva_finish(const YgBatchGeom& G, const YgFields<R>& f, const YgDivaFields<R>& d,
                           const YgFieldParams& p, const int* skip, hipStream_t st) {
  hipLaunchKernelGGL(k_vel3d<R>, dim3(dnblk(G.total_cells, 256)), dim3(256), 0, st, G, f, d, p,
                     skip);
}

#define YG_DINST(R)                                                                              \
  template void yg_launch_diva_pre<R>(const YgBatchGeom&, const YgFields<R>&,                    \
                                      const YgDivaFields<R>&, const YgFieldParams&, const int*,   \
                                      hipStream_t);                                              \
  template void yg_launch_diva_finish<R>(const YgBatchGeom&, const YgFields<R>&,                 \
                                         const YgDivaFields<R>&, const YgFieldParams&,             \
                                         const int*, hipStream_t);
YG_DINST(float)
YG_DINST(double)
