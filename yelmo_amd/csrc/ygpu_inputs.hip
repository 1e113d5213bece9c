// ygpu_inputs.hip -- the per-solve inputs of the velocity solve on the
// device: what Yelmo's calc_ydyn (src/yelmo_dynamics.f90:105-122, 464) and
// the topography step before it (src/yelmo_topography.f90:1373-1374, 1472)
// compute from the geometry before calling calc_velocity_ssa/_diva:
//
//   calc_ice_front        (src/physics/topography.f90:724-818)
//   calc_gradient_acx/acy (src/yelmo_tools.f90:341-478), of z_srf
//   calc_driving_stress   (src/physics/velocity_general.f90:1118-1205)
//   calc_lateral_bc_stress_2D / calc_lateral_bc_stress (:1453-1609)
//   set_ssa_masks + front_face_mask (src/physics/solver_ssa_ac.f90:851-996)
//
// Two per-cell kernels, one thread per cell, fields in the reference's
// Fortran order.  The reference writes the ice-free codes of calc_ice_front
// from each front cell into its open neighbours; here every cell gathers
// its own code: a full cell with an open neighbour is a front, and an open
// cell is beside a front exactly when one of its neighbours is full (that
// neighbour then has this open cell as a neighbour, the neighbour relation
// being symmetric under the clamped and the wrapped borders alike).  The
// gradients' border copies read the inner face's difference directly.
// Arithmetic follows the reference's kinds (wp, with the double literals of
// calc_lateral_bc_stress) without contracted multiply-adds, so the results
// are the reference's operations in the reference's order.
#include <algorithm>

#include "ygpu_host.h"

using yg::Arena;
using yg::bc_code;
using yg::YgDynArgs;

namespace {

#define YGI_FRNT_ICE_FREE (-1)
#define YGI_FRNT_ICE_FREE_LAND (-2)
#define YGI_FRNT_NONE 0
#define YGI_FRNT_FLOAT 1
#define YGI_FRNT_MARINE 2
#define YGI_FRNT_GRND 3


__device__ __forceinline__ void inbr(int i, int j, int nx, int ny, int BC, int& im1, int& ip1,
                                     int& jm1, int& jp1) {
  const bool wx = BC == YG_BC_PERIODIC || BC == YG_BC_PERIODIC_X;
  const bool wy = BC == YG_BC_MISMIP3D || BC == YG_BC_TROUGH || BC == YG_BC_PERIODIC ||
                  BC == YG_BC_PERIODIC_Y;
  if (wx) { im1 = i == 0 ? nx - 1 : i - 1; ip1 = i == nx - 1 ? 0 : i + 1; }
  else { im1 = i > 0 ? i - 1 : 0; ip1 = i < nx - 1 ? i + 1 : nx - 1; }
  if (wy) { jm1 = j == 0 ? ny - 1 : j - 1; jp1 = j == ny - 1 ? 0 : j + 1; }
  else { jm1 = j > 0 ? j - 1 : 0; jp1 = j < ny - 1 ? j + 1 : ny - 1; }
}

// the (unclipped) difference quotient of z_srf on the face after cell (i, j)
// in direction dir, before the border copy
template <typename R>
__device__ __forceinline__ R face_diff(const yg::YgDynArgs<R>& a, int i, int j, int dir) {
#pragma clang fp contract(off)
  int im1, ip1, jm1, jp1;
  inbr(i, j, a.nx, a.ny, a.BC, im1, ip1, jm1, jp1);
  const int c = i + a.nx * j;
  const int q = dir == 0 ? ip1 + a.nx * j : i + a.nx * jp1;
  return (a.zs[q] - a.zs[c]) / (dir == 0 ? a.dx : a.dy);
}

// calc_gradient_acx/acy with zero_outside = .false. (the z_srf slopes of
// yelmo_topography.f90:1373-1374): border copy, background slope, clip
template <typename R>
__device__ __forceinline__ R slope(const yg::YgDynArgs<R>& a, int i, int j, int dir) {
#pragma clang fp contract(off)
  const int b = a.bnd;
  if (dir == 0) {
    const bool copy = b == YGPU_BND_INFINITE || b == YGPU_BND_MISMIP3D || b == YGPU_BND_TROUGH ||
                      b == YGPU_BND_MASK || b == YGPU_BND_PERIODIC_Y;
    if (copy && i == 0) i = 1;
    else if (copy && i == a.nx - 1) i = a.nx - 2;
  } else {
    const bool copy = b == YGPU_BND_INFINITE || b == YGPU_BND_MASK || b == YGPU_BND_PERIODIC_X;
    if (copy && j == 0) j = 1;
    else if (copy && j == a.ny - 1) j = a.ny - 2;
  }
  R v = face_diff<R>(a, i, j, dir) + (dir == 0 ? a.slope_x : a.slope_y);
  if (v > a.grad_lim) v = a.grad_lim;
  if (v < -a.grad_lim) v = -a.grad_lim;
  return v;
}

template <typename R>
__device__ __forceinline__ R hmid(R Ha, R Hb, R fa, R fb) {
#pragma clang fp contract(off)
  if (fa == R(1) && fb < R(1)) return R(0.5) * (Ha + R(0));
  if (fa < R(1) && fb == R(1)) return R(0.5) * (R(0) + Hb);
  return R(0.5) * (Ha + Hb);
}

// ice front codes, slopes and driving stress
template <typename R>
__global__ void k_front_slope_taud(yg::YgDynArgs<R> a) {
#pragma clang fp contract(off)
  const long long n = (long long)a.nx * a.ny;
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < n;
       c += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(c % a.nx), j = (int)(c / a.nx);
    int im1, ip1, jm1, jp1;
    inbr(i, j, a.nx, a.ny, a.BC, im1, ip1, jm1, jp1);
    const long long nb[4] = {im1 + (long long)a.nx * j, ip1 + (long long)a.nx * j,
                             i + (long long)a.nx * jm1, i + (long long)a.nx * jp1};
    int code = YGI_FRNT_NONE;
    if (a.fi[c] == R(1)) {
      int nopen = 0;
      for (int k = 0; k < 4; ++k) nopen += a.fi[nb[k]] < R(1);
      if (nopen > 0) {
        if (a.fg[c] > R(0) && a.zsl[c] <= a.zb[c]) code = YGI_FRNT_GRND;
        else if (a.fg[c] > R(0)) code = YGI_FRNT_MARINE;
        else code = YGI_FRNT_FLOAT;
      }
    } else if (a.fi[c] < R(1)) {
      bool beside = false;
      for (int k = 0; k < 4; ++k) beside |= nb[k] != c && a.fi[nb[k]] == R(1);
      if (beside) code = a.zb[c] < a.zsl[c] ? YGI_FRNT_ICE_FREE : YGI_FRNT_ICE_FREE_LAND;
    }
    a.mf[c] = code;
    const R gx = slope<R>(a, i, j, 0), gy = slope<R>(a, i, j, 1);
    if (a.sx) a.sx[c] = gx;
    if (a.sy) a.sy[c] = gy;
    const R rhog = a.rho_ice * a.g;
    R t = rhog * hmid<R>(a.H[c], a.H[nb[1]], a.fi[c], a.fi[nb[1]]) * gx;
    if (t > a.taud_lim) t = a.taud_lim;
    if (t < -a.taud_lim) t = -a.taud_lim;
    a.tdx[c] = t;
    t = rhog * hmid<R>(a.H[c], a.H[nb[3]], a.fi[c], a.fi[nb[3]]) * gy;
    if (t > a.taud_lim) t = a.taud_lim;
    if (t < -a.taud_lim) t = -a.taud_lim;
    a.tdy[c] = t;
  }
}

// calc_lateral_bc_stress with the reference's kinds (wp = sp values, the
// double literals 1.d0 and 0.5d0)
template <typename R>
__device__ __forceinline__ R lateral_stress(R H, R zs, R zsl, R rho_ice, R rho_sw, R g) {
#pragma clang fp contract(off)
  const double q = (double)((zs - zsl) / H);
  const R f_sub = (R)(1.0 - (q < 1.0 ? q : 1.0));
  const R H_ocn = H * f_sub;
  const R H2 = H * H, O2 = H_ocn * H_ocn;
  return (R)(0.5 * (double)rho_ice * (double)g * (double)H2 -
             0.5 * (double)rho_sw * (double)g * (double)O2);
}

// front_face_mask (solver_ssa_ac.f90:950-996); lateral_bc 0 none,
// 1 floating, 2 marine, 3 all
__device__ __forceinline__ int face_code(int a, int b, int lateral_bc) {
  int ice, fr;
  if (a > 0 && b < 0) { ice = a; fr = b; }
  else if (a < 0 && b > 0) { ice = b; fr = a; }
  else return -1;
  if (fr == YGI_FRNT_ICE_FREE_LAND) {
    if (ice == YGI_FRNT_FLOAT) return 0;
    ice = YGI_FRNT_GRND;
  }
  switch (lateral_bc) {
    case 0: return 4;
    case 1: return ice == YGI_FRNT_FLOAT ? 3 : 4;
    case 2: return (ice == YGI_FRNT_FLOAT || ice == YGI_FRNT_MARINE) ? 3 : 4;
    case 3: return 3;
  }
  return -1;
}

// lateral front stress and SSA face masks of the faces after each cell
template <typename R>
__global__ void k_taul_masks(yg::YgDynArgs<R> a) {
#pragma clang fp contract(off)
  const long long n = (long long)a.nx * a.ny;
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < n;
       c += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(c % a.nx), j = (int)(c / a.nx);
    int im1, ip1, jm1, jp1;
    inbr(i, j, a.nx, a.ny, a.BC, im1, ip1, jm1, jp1);
    for (int dir = 0; dir < 2; ++dir) {
      const long long q = dir == 0 ? ip1 + (long long)a.nx * j : i + (long long)a.nx * jp1;
      const int ma = a.mf[c], mb = a.mf[q];
      R t = R(0);
      if ((ma > 0 && mb < 0) || (ma < 0 && mb > 0)) {
        const long long k = ma < 0 ? q : c;   // the ice-covered cell
        const R H = a.H[k], zs = a.zs[k];
        R zsl = a.zsl[k];
        if (ma == YGI_FRNT_ICE_FREE_LAND || mb == YGI_FRNT_ICE_FREE_LAND) zsl = zs - H;
        t = lateral_stress<R>(H, zs, zsl, a.rho_ice, a.rho_sw, a.g);
      }
      // set_ssa_masks: active faces by ice cover and grounding, then the
      // front faces' codes
      int m = 0;
      if (a.fi[c] == R(1) || a.fi[q] == R(1)) m = (a.fg[c] > R(0) || a.fg[q] > R(0)) ? 1 : 2;
      const int ml = face_code(ma, mb, a.lateral_bc);
      if (ml >= 0) m = ml;
      if (dir == 0) { a.tlx[c] = t; a.mx[c] = m; }
      else { a.tly[c] = t; a.my[c] = m; }
    }
  }
}

}  // namespace

namespace yg {

template <typename R>
int launch_dyn_inputs(const YgDynArgs<R>& a, hipStream_t st) {
  const long long n = (long long)a.nx * a.ny;
  const int threads = 256;
  const int blocks = (int)std::min<long long>((n + threads - 1) / threads, 65535);
  hipLaunchKernelGGL(k_front_slope_taud<R>, dim3(blocks), dim3(threads), 0, st, a);
  if (hipGetLastError() != hipSuccess) return YGPU_ERR_DEVICE;
  hipLaunchKernelGGL(k_taul_masks<R>, dim3(blocks), dim3(threads), 0, st, a);
  if (hipGetLastError() != hipSuccess) return YGPU_ERR_DEVICE;
  return YGPU_OK;
}
template int launch_dyn_inputs<float>(const YgDynArgs<float>&, hipStream_t);
template int launch_dyn_inputs<double>(const YgDynArgs<double>&, hipStream_t);

}  // namespace yg

extern "C" {

int ygpu_dyn_inputs(ygpu_handle* h, const ygpu_ssa_params* par, int32_t nx, int32_t ny,
                    float dx, float dy, float taud_lim, float grad_lim, float slope_bg_x,
                    float slope_bg_y, int32_t lateral_bc, const float* H_ice, const float* f_ice,
                    const float* f_grnd, const float* z_bed, const float* z_sl,
                    const float* z_srf, int32_t* mask_frnt, float* dzsdx, float* dzsdy,
                    float* taud_acx, float* taud_acy, float* taul_int_acx, float* taul_int_acy,
                    int32_t* ssa_mask_acx, int32_t* ssa_mask_acy) {
  if (!h) return YGPU_ERR_ARG;
  h->err.clear(); h->note.clear();
  if (!par || nx < 2 || ny < 2 || !H_ice || !f_ice || !f_grnd || !z_bed || !z_sl || !z_srf ||
      lateral_bc < 0 || lateral_bc > 3 || !(dx > 0.0f) || !(dy > 0.0f) ||
      par->boundaries < YGPU_BND_ZEROS || par->boundaries > YGPU_BND_PERIODIC_Y) {
    h->err = "ygpu_dyn_inputs: invalid arguments";
    return YGPU_ERR_ARG;
  }
  if (hipSetDevice(h->device) != hipSuccess) { h->err = "hipSetDevice failed"; return YGPU_ERR_DEVICE; }
  const size_t n = (size_t)nx * ny;
  Arena ar;
  float* d_in;                 // H, f_ice, f_grnd, z_bed, z_sl, z_srf
  float* d_out;                // dzsdx, dzsdy, taud x/y, taul x/y
  int* d_int;                  // mask_frnt, ssa masks x/y
  YG_CHECK(h, ar.alloc(&d_in, 6 * n));
  YG_CHECK(h, ar.alloc(&d_out, 6 * n));
  YG_CHECK(h, ar.alloc(&d_int, 3 * n));
  hipStream_t st = h->stream;
  const float* src[6] = {H_ice, f_ice, f_grnd, z_bed, z_sl, z_srf};
  for (int k = 0; k < 6; ++k)
    YG_CHECK(h, hipMemcpyAsync(d_in + k * n, src[k], n * sizeof(float), hipMemcpyHostToDevice, st));
  yg::YgDynArgs<float> a{};
  a.nx = nx; a.ny = ny; a.bnd = par->boundaries; a.BC = bc_code(par->boundaries);
  a.lateral_bc = lateral_bc;
  a.taud_lim = taud_lim; a.grad_lim = grad_lim; a.slope_x = slope_bg_x; a.slope_y = slope_bg_y;
  a.dx = dx; a.dy = dy;
  a.rho_ice = (float)par->rho_ice; a.rho_sw = (float)par->rho_sw; a.g = (float)par->g;
  a.H = d_in; a.fi = d_in + n; a.fg = d_in + 2 * n; a.zb = d_in + 3 * n; a.zsl = d_in + 4 * n;
  a.zs = d_in + 5 * n;
  a.sx = d_out; a.sy = d_out + n; a.tdx = d_out + 2 * n; a.tdy = d_out + 3 * n;
  a.tlx = d_out + 4 * n; a.tly = d_out + 5 * n;
  a.mf = d_int; a.mx = d_int + n; a.my = d_int + 2 * n;
  if (yg::launch_dyn_inputs<float>(a, st)) {
    h->err = "ygpu_dyn_inputs: kernel launch failed";
    return YGPU_ERR_DEVICE;
  }
  float* fo[6] = {dzsdx, dzsdy, taud_acx, taud_acy, taul_int_acx, taul_int_acy};
  for (int k = 0; k < 6; ++k)
    if (fo[k]) YG_CHECK(h, hipMemcpyAsync(fo[k], d_out + k * n, n * sizeof(float), hipMemcpyDeviceToHost, st));
  int32_t* io[3] = {mask_frnt, ssa_mask_acx, ssa_mask_acy};
  for (int k = 0; k < 3; ++k)
    if (io[k]) YG_CHECK(h, hipMemcpyAsync(io[k], d_int + k * n, n * sizeof(int32_t), hipMemcpyDeviceToHost, st));
  YG_CHECK(h, hipStreamSynchronize(st));
  return YGPU_OK;
}

}  // extern "C"
