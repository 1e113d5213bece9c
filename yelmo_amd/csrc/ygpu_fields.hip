// ygpu_fields.hip -- grid-wide field kernels of one Picard iteration.
//
// Every kernel runs over all cells of all members of a batch (thread per
// cell, member found from the packed cell offsets), so the whole GPU works on
// these phases even when the Krylov solve owns one CU per member.  Arrays are
// in the reference's Fortran order (i fastest) inside each member.  The
// arithmetic follows the reference routines cited at each kernel; the order
// of operations is kept so FP64 results agree with the FP64 oracle to
// rounding.
#include "ygpu_fields.h"
#include "ygpu_krylov.h"

namespace {

#define I2(i, j) ((i) + nx * (j))

template <typename R> YG_DEV R ypow(R a, R b);
template <> YG_DEV float ypow<float>(float a, float b) { return powf(a, b); }
template <> YG_DEV double ypow<double>(double a, double b) { return pow(a, b); }
template <typename R> YG_DEV R ysqrt(R a);
template <> YG_DEV float ysqrt<float>(float a) { return sqrtf(a); }
template <> YG_DEV double ysqrt<double>(double a) { return sqrt(a); }
template <typename R> YG_DEV R yexp(R a);
template <> YG_DEV float yexp<float>(float a) { return expf(a); }
template <> YG_DEV double yexp<double>(double a) { return exp(a); }
template <typename R> YG_DEV R ylog(R a);
template <> YG_DEV float ylog<float>(float a) { return logf(a); }
template <> YG_DEV double ylog<double>(double a) { return log(a); }
template <typename R> YG_DEV R yabs(R a) { return a < R(0) ? -a : a; }
// a^b as exp(b log a) for a > 0: a few units in the last place of
// |b log a| against pow's correctly rounded result, at a fraction of its cost
// (the viscosity and the friction law take it at every quadrature node)
template <typename R> YG_DEV R ypow_el(R a, R b) { return yexp<R>(b * ylog<R>(a)); }

// default-real literal of the reference (single precision), converted to R
#define RS(x) ((R)(float)(x))

// get_neighbor_indices_bc_codes (yelmo_tools.f90:118-201), 0-based
YG_DEV void nbr(int i, int j, int nx, int ny, int BC, int& im1, int& ip1, int& jm1, int& jp1) {
  const bool wx = BC == YG_BC_PERIODIC || BC == YG_BC_PERIODIC_X;
  const bool wy = BC == YG_BC_MISMIP3D || BC == YG_BC_TROUGH || BC == YG_BC_PERIODIC ||
                  BC == YG_BC_PERIODIC_Y;
  if (wx) { im1 = i == 0 ? nx - 1 : i - 1; ip1 = i == nx - 1 ? 0 : i + 1; }
  else { im1 = i > 0 ? i - 1 : 0; ip1 = i < nx - 1 ? i + 1 : nx - 1; }
  if (wy) { jm1 = j == 0 ? ny - 1 : j - 1; jp1 = j == ny - 1 ? 0 : j + 1; }
  else { jm1 = j > 0 ? j - 1 : 0; jp1 = j < ny - 1 ? j + 1 : ny - 1; }
}

struct CellCtx {
  int m, nx, ny, nz, i, j;
  long long c0;   // member cell offset
  long long z0;   // member 3D offset
  double dx, dy;
};

YG_DEV bool cell_of(const YgBatchGeom& G, long long gc, const int* skip, CellCtx& cc) {
  if (gc >= G.total_cells) return false;
  const int m = yg_member_at(G, gc);
  if (m < 0 || (skip && skip[m])) return false;
  const YgMember& mb = G.members[m];
  cc.m = m; cc.nx = mb.nx; cc.ny = mb.ny; cc.nz = mb.nz;
  cc.c0 = G.coffs[m]; cc.z0 = G.zoffs[m];
  const int loc = (int)(gc - cc.c0);
  cc.i = loc % mb.nx; cc.j = loc / mb.nx;
  cc.dx = mb.dx; cc.dy = mb.dy;
  return true;
}

// ---- strain rates (deformation.f90:1316-1507) ---------------------------
template <typename R>
YG_DEV R deriv_active(R vm, R v0, R vp, bool am, bool ap, R h) {
  if (am && ap) return (vp - vm) / (R(2) * h);
  if (ap) return (vp - v0) / h;
  if (am) return (v0 - vm) / h;
  return R(0);
}

template <typename R>
YG_DEV bool act_x(const R* fi, int nx, int ny, int BC, int i, int j) {
  int im1, ip1, jm1, jp1;
  nbr(i, j, nx, ny, BC, im1, ip1, jm1, jp1);
  return fi[I2(i, j)] == R(1) || fi[I2(ip1, j)] == R(1);
}
template <typename R>
YG_DEV bool act_y(const R* fi, int nx, int ny, int BC, int i, int j) {
  int im1, ip1, jm1, jp1;
  nbr(i, j, nx, ny, BC, im1, ip1, jm1, jp1);
  return fi[I2(i, j)] == R(1) || fi[I2(i, jp1)] == R(1);
}

template <typename R>
__global__ __launch_bounds__(256) void k_strain(YgBatchGeom G, YgFields<R> f, YgFieldParams p, const int* skip) {
  const long long gc = yg_xcd_cell();
  CellCtx cc;
  if (!cell_of(G, gc, skip, cc)) return;
  const int nx = cc.nx, ny = cc.ny, i = cc.i, j = cc.j, BC = p.BC;
  const R* ux = (p.diva ? f.ubar_x : f.ux) + cc.c0;
  const R* uy = (p.diva ? f.ubar_y : f.uy) + cc.c0;
  const R* fi = f.f_ice + cc.c0;
  const R dx = (R)cc.dx, dy = (R)cc.dy;
  int im1, ip1, jm1, jp1, im2, ip2, jm2, jp2, d1, d2, d3;
  nbr(i, j, nx, ny, BC, im1, ip1, jm1, jp1);
  const int c = I2(i, j);
  R dudx = (ux[I2(ip1, j)] - ux[I2(im1, j)]) / (R(2) * dx);
  const R dudy = deriv_active<R>(ux[I2(i, jm1)], ux[c], ux[I2(i, jp1)], act_x(fi, nx, ny, BC, i, jm1),
                                 act_x(fi, nx, ny, BC, i, jp1), dy);
  const R dvdx = deriv_active<R>(uy[I2(im1, j)], uy[c], uy[I2(ip1, j)], act_y(fi, nx, ny, BC, im1, j),
                                 act_y(fi, nx, ny, BC, ip1, j), dx);
  R dvdy = (uy[I2(i, jp1)] - uy[I2(i, jm1)]) / (R(2) * dy);
  if (fi[c] == R(1) && fi[I2(ip1, j)] < R(1)) {
    nbr(im1, j, nx, ny, BC, im2, d1, d2, d3);
    if (fi[I2(im1, j)] == R(1) && im2 != im1)
      dudx = (R(1) * ux[I2(im2, j)] - R(4) * ux[I2(im1, j)] + R(3) * ux[c]) / (R(2) * dx);
    else
      dudx = (ux[c] - ux[I2(im1, j)]) / dx;
  } else if (fi[c] < R(1) && fi[I2(ip1, j)] == R(1)) {
    nbr(ip1, j, nx, ny, BC, d1, ip2, d2, d3);
    if (ip2 != ip1 && fi[I2(ip2, j)] == R(1))
      dudx = -(R(1) * ux[I2(ip2, j)] - R(4) * ux[I2(ip1, j)] + R(3) * ux[c]) / (R(2) * dx);
    else
      dudx = (ux[I2(ip1, j)] - ux[c]) / dx;
  }
  if (fi[c] == R(1) && fi[I2(i, jp1)] < R(1)) {
    nbr(i, jm1, nx, ny, BC, d1, d2, jm2, d3);
    if (fi[I2(i, jm1)] == R(1) && jm2 != jm1)
      dvdy = (R(1) * uy[I2(i, jm2)] - R(4) * uy[I2(i, jm1)] + R(3) * uy[c]) / (R(2) * dy);
    else
      dvdy = (uy[c] - uy[I2(i, jm1)]) / dy;
  } else if (fi[c] < R(1) && fi[I2(i, jp1)] == R(1)) {
    nbr(i, jp1, nx, ny, BC, d1, d2, d3, jp2);
    if (jp2 != jp1 && fi[I2(i, jp2)] == R(1))
      dvdy = -(R(1) * uy[I2(i, jp2)] - R(4) * uy[I2(i, jp1)] + R(3) * uy[c]) / (R(2) * dy);
    else
      dvdy = (uy[I2(i, jp1)] - uy[c]) / dy;
  }
  const long long g = cc.c0 + c;
  f.dudx[g] = dudx; f.dudy[g] = dudy; f.dvdx[g] = dvdx; f.dvdy[g] = dvdy;
  // state of the previous Picard iteration (velocity_ssa.f90:175-193)
  f.ux_nm1[g] = ux[c];
  f.uy_nm1[g] = uy[c];
}

// ---- 2x2 Gauss quadrature in a cell ----------------------------------------
// Node values are bilinear interpolations of the cell-corner values at
// xi, eta = +-1/sqrt(3); corner values are means over the active points of
// the field's own grid (fesm-utils gaussian_quadrature, used at
// velocity_ssa.f90:454-471).  Corners and nodes ordered SW, SE, NE, NW.
template <typename R>
YG_DEV void corners_to_nodes(const R c[4], R n[4]) {
  const double A = 0.57735026918962576451;
  const double xs[4] = {-1, 1, 1, -1}, ys[4] = {-1, -1, 1, 1};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const double xi = xs[q] * A, eta = ys[q] * A;
    const double w0 = 0.25 * (1 - xi) * (1 - eta), w1 = 0.25 * (1 + xi) * (1 - eta);
    const double w2 = 0.25 * (1 + xi) * (1 + eta), w3 = 0.25 * (1 - xi) * (1 + eta);
    n[q] = (R)(w0 * (double)c[0] + w1 * (double)c[1] + w2 * (double)c[2] + w3 * (double)c[3]);
  }
}

template <typename R, int K>
YG_DEV R mean_act(const R (&v)[K], const bool (&a)[K]) {
  int cnt = 0;
  R s = R(0);
#pragma unroll
  for (int q = 0; q < K; ++q)
    if (a[q]) { s += v[q]; cnt++; }
  return cnt > 0 ? s / (R)cnt : R(0);
}

// The two points of each corner (SW, SE, NE, NW) that the corner value of an
// acx (kind 1) or acy (kind 2) field averages
YG_DEV void gq_pts(int kind, int (&cs)[4][2], int nx, int i, int j, int im1, int ip1, int jm1,
                   int jp1) {
  if (kind == 1) {
    cs[0][0] = I2(im1, jm1); cs[0][1] = I2(im1, j);
    cs[1][0] = I2(i, jm1);   cs[1][1] = I2(i, j);
    cs[2][0] = I2(i, j);     cs[2][1] = I2(i, jp1);
    cs[3][0] = I2(im1, j);   cs[3][1] = I2(im1, jp1);
  } else {
    cs[0][0] = I2(im1, jm1); cs[0][1] = I2(i, jm1);
    cs[1][0] = I2(i, jm1);   cs[1][1] = I2(ip1, jm1);
    cs[2][0] = I2(i, j);     cs[2][1] = I2(ip1, j);
    cs[3][0] = I2(im1, j);   cs[3][1] = I2(i, j);
  }
}

// The activity flags of those 8 points (bit 2 corner + point) from a test on
// their grid coordinates, for fields that share them (several fields, or one
// field on many levels of a cell)
template <typename ACT>
YG_DEV uint32_t gq_mask_ij(int kind, ACT act, int i, int j, int im1, int ip1, int jm1, int jp1) {
  int ci[4][2], cj[4][2];
  if (kind == 1) {
    ci[0][0] = im1; cj[0][0] = jm1; ci[0][1] = im1; cj[0][1] = j;
    ci[1][0] = i;   cj[1][0] = jm1; ci[1][1] = i;   cj[1][1] = j;
    ci[2][0] = i;   cj[2][0] = j;   ci[2][1] = i;   cj[2][1] = jp1;
    ci[3][0] = im1; cj[3][0] = j;   ci[3][1] = im1; cj[3][1] = jp1;
  } else {
    ci[0][0] = im1; cj[0][0] = jm1; ci[0][1] = i;   cj[0][1] = jm1;
    ci[1][0] = i;   cj[1][0] = jm1; ci[1][1] = ip1; cj[1][1] = jm1;
    ci[2][0] = i;   cj[2][0] = j;   ci[2][1] = ip1; cj[2][1] = j;
    ci[3][0] = im1; cj[3][0] = j;   ci[3][1] = i;   cj[3][1] = j;
  }
  uint32_t m = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int k = 0; k < 2; ++k)
      if (act(ci[q][k], cj[q][k])) m |= 1u << (2 * q + k);
  return m;
}

// gq_nodes of an acx or acy field with the activity flags of gq_mask_ij
template <typename R>
YG_DEV void gq_nodes_m(int kind, R n[4], const R* v, uint32_t m, int nx, int i, int j, int im1,
                       int ip1, int jm1, int jp1) {
  int cs[4][2];
  gq_pts(kind, cs, nx, i, j, im1, ip1, jm1, jp1);
  R c[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    R vv[2];
    bool aa[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) { vv[k] = v[cs[q][k]]; aa[k] = (m >> (2 * q + k)) & 1u; }
    c[q] = mean_act<R, 2>(vv, aa);
  }
  corners_to_nodes<R>(c, n);
}

// kind 0 aa (act: flags or all), 1 acx, 2 acy
template <typename R, typename ACT>
YG_DEV void gq_nodes(int kind, R n[4], const R* v, ACT act, int nx, int i, int j, int im1,
                     int ip1, int jm1, int jp1) {
  R c[4];
  if (kind == 0) {
    const int cs[4][4] = {{I2(im1, jm1), I2(i, jm1), I2(im1, j), I2(i, j)},
                          {I2(i, jm1), I2(ip1, jm1), I2(i, j), I2(ip1, j)},
                          {I2(i, j), I2(ip1, j), I2(i, jp1), I2(ip1, jp1)},
                          {I2(im1, j), I2(i, j), I2(im1, jp1), I2(i, jp1)}};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      R vv[4];
      bool aa[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) { vv[k] = v[cs[q][k]]; aa[k] = act(cs[q][k]); }
      c[q] = mean_act<R, 4>(vv, aa);
    }
  } else {
    int cs[4][2];
    gq_pts(kind, cs, nx, i, j, im1, ip1, jm1, jp1);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      R vv[2];
      bool aa[2];
#pragma unroll
      for (int k = 0; k < 2; ++k) { vv[k] = v[cs[q][k]]; aa[k] = act(cs[q][k]); }
      c[q] = mean_act<R, 2>(vv, aa);
    }
  }
  corners_to_nodes<R>(c, n);
}

// ---- effective viscosity, relaxation and depth integral -----------------
// calc_visc_eff_3D_nodes / _aa (velocity_ssa.f90:359-586; DIVA forms
// velocity_diva.f90:573-822 add 1/4 (du/dz)^2 + 1/4 (dv/dz)^2),
// picard_relax_visc (velocity_general.f90:2188-2213), calc_visc_eff_int
// (velocity_general.f90:2215-2261).
// FORM 1: the SSA form of visc_method 1, FORM 2: the DIVA form of visc_method
// 1, both with ATT^(-1/n) at the nodes filled once per solve (the common
// cases), each compiled without the other forms' code so that it needs few
// registers (the DIVA form without spills needs the 256-thread bound it is
// launched with); FORM 0: every form.
template <typename R, int FORM>
__global__ __launch_bounds__(FORM == 2 ? 256 : 1024) void k_visc(YgBatchGeom G, YgFields<R> f, YgFieldParams p, const int* skip) {
  const long long gc = yg_xcd_cell();
  CellCtx cc;
  if (!cell_of(G, gc, skip, cc)) return;
  const int nx = cc.nx, ny = cc.ny, nz = cc.nz, i = cc.i, j = cc.j, BC = p.BC;
  const long long nc = (long long)nx * ny;
  const int c = I2(i, j);
  const R* fi = f.f_ice + cc.c0;
  const R visc_min = R(1e5);
  R* visc = f.visc + cc.z0;
  const R* ATT = f.ATT + cc.z0;
  const R* zeta = f.zeta + (long long)cc.m * f.zstride;
  const R rel = (R)p.iter_rel;
  const R n_glen = (R)p.n_glen;
  const bool ice = fi[c] == R(1);
  const int vm = FORM ? 1 : p.visc_method;
  const bool diva = FORM ? FORM == 2 : (bool)p.diva;
  const R eps_0 = (R)p.eps_0, eps_0_sq = eps_0 * eps_0;
  int im1, ip1, jm1, jp1;
  nbr(i, j, nx, ny, BC, im1, ip1, jm1, jp1);
  const R dx = (R)cc.dx, dy = (R)cc.dy;
  const R p1 = (RS(1.0) - n_glen) / (RS(2.0) * n_glen);
  const R p2 = -RS(1.0) / n_glen;
  // per-cell terms of the level values
  R eps_sq_n[4], epow[4];        // visc_method 1: strain rates at the quadrature nodes
  R eps_sq_h = R(0), eps_sq_aa = R(0);   // visc_method 2: the cell-centre strain rate
  const R* attp = nullptr;
  uint32_t mdzx = 0, mdzy = 0;   // DIVA, visc_method 1: activity of the shear points
  if (vm == 1 && ice) {
    const R* dudx = f.dudx + cc.c0;
    const R* dudy = f.dudy + cc.c0;
    const R* dvdx = f.dvdx + cc.c0;
    const R* dvdy = f.dvdy + cc.c0;
    // activity of the acx and acy points, shared by the strain-rate fields
    // and (DIVA) the shear fields of every level
    const uint32_t max_ = gq_mask_ij(1, [&](int qi, int qj) { return act_x<R>(fi, nx, ny, BC, qi, qj); },
                                     i, j, im1, ip1, jm1, jp1);
    const uint32_t may_ = gq_mask_ij(2, [&](int qi, int qj) { return act_y<R>(fi, nx, ny, BC, qi, qj); },
                                     i, j, im1, ip1, jm1, jp1);
    R dudxn[4], dudyn[4], dvdxn[4], dvdyn[4];
    gq_nodes_m<R>(1, dudxn, dudx, max_, nx, i, j, im1, ip1, jm1, jp1);
    gq_nodes_m<R>(1, dudyn, dudy, max_, nx, i, j, im1, ip1, jm1, jp1);
    gq_nodes_m<R>(2, dvdxn, dvdx, may_, nx, i, j, im1, ip1, jm1, jp1);
    gq_nodes_m<R>(2, dvdyn, dvdy, may_, nx, i, j, im1, ip1, jm1, jp1);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const R s = dudyn[q] + dvdxn[q];
      eps_sq_n[q] = dudxn[q] * dudxn[q] + dvdyn[q] * dvdyn[q] + dudxn[q] * dvdyn[q] +
                    RS(0.25) * (s * s);
      if (!diva) eps_sq_n[q] = eps_sq_n[q] + eps_0_sq;
    }
    // SSA: the strain-rate term of a node is the same on every level, so
    // its power is taken once per node (the same value the level loop
    // would compute nz times)
    if (!diva) {
#pragma unroll
      for (int q = 0; q < 4; ++q) epow[q] = ypow_el<R>(eps_sq_n[q], p1);
    } else {
      // the shear fields' activity flags are the same on every level
      mdzx = max_;
      mdzy = may_;
    }
    attp = (FORM || f.attp) ? f.attp + 4 * cc.z0 + c : nullptr;
  } else if (vm == 2 && ice) {
    const R* ux = (diva ? f.ubar_x : f.ux) + cc.c0;
    const R* uy = (diva ? f.ubar_y : f.uy) + cc.c0;
    const R dudx_aa = (ux[c] - ux[I2(im1, j)]) / dx;
    const R dvdy_aa = (uy[c] - uy[I2(i, jm1)]) / dy;
    const R dudy_aa_1 = (ux[I2(i, jp1)] - ux[I2(i, jm1)]) / (R(2) * dy);
    const R dudy_aa_2 = (ux[I2(im1, jp1)] - ux[I2(im1, jm1)]) / (R(2) * dy);
    const R dudy_aa = R(0.5) * (dudy_aa_1 + dudy_aa_2);
    const R dvdx_aa_1 = (uy[I2(ip1, j)] - uy[I2(im1, j)]) / (R(2) * dx);
    const R dvdx_aa_2 = (uy[I2(ip1, jm1)] - uy[I2(im1, jm1)]) / (R(2) * dx);
    const R dvdx_aa = R(0.5) * (dvdx_aa_1 + dvdx_aa_2);
    const R s = dudy_aa + dvdx_aa;
    eps_sq_h = dudx_aa * dudx_aa + dvdy_aa * dvdy_aa + dudx_aa * dvdy_aa + R(0.25) * (s * s);
    eps_sq_aa = eps_sq_h + eps_0_sq;
  }
  // The viscosity of level k before relaxation
  auto level_value = [&](int k) -> R {
    if (vm == 0) return (R)p.visc_const;
    if (!ice) return visc_min;
    if (vm == 1) {
      auto all = [&](int) { return true; };
      // ATT^p2 at the nodes: filled once per solve, or computed here
      R ATTp[4];
      if (attp) {
#pragma unroll
        for (int q = 0; q < 4; ++q) ATTp[q] = attp[(q * nz + k) * nc];
      } else {
        R ATTn[4];
        gq_nodes<R>(0, ATTn, ATT + k * nc, all, nx, i, j, im1, ip1, jm1, jp1);
#pragma unroll
        for (int q = 0; q < 4; ++q) ATTp[q] = ypow<R>(ATTn[q], p2);
      }
      R ep[4];
      if (diva) {
        R dzx[4], dzy[4];
        // DIVA: vertical shear at the nodes from the acx/acy shear fields
        gq_nodes_m<R>(1, dzx, f.duxdz + cc.z0 + k * nc, mdzx, nx, i, j, im1, ip1, jm1, jp1);
        gq_nodes_m<R>(2, dzy, f.duydz + cc.z0 + k * nc, mdzy, nx, i, j, im1, ip1, jm1, jp1);
#pragma unroll
        for (int q = 0; q < 4; ++q)
          ep[q] = ypow_el<R>(eps_sq_n[q] + R(0.25) * (dzx[q] * dzx[q]) +
                                 R(0.25) * (dzy[q] * dzy[q]) + eps_0_sq, p1);
      } else {
#pragma unroll
        for (int q = 0; q < 4; ++q) ep[q] = epow[q];
      }
      R sum = R(0);
#pragma unroll
      for (int q = 0; q < 4; ++q) sum += (RS(0.5) * ep[q] * ATTp[q]) * R(1);
      return sum / R(4);
    }
    R e = eps_sq_aa;
    if (diva) {
      // DIVA aa form: shear of the cell's own faces, averaged to the centre
      const R* dzx = f.duxdz + cc.z0 + k * nc;
      const R* dzy = f.duydz + cc.z0 + k * nc;
      const R ddx = R(0.5) * (dzx[c] + dzx[I2(im1, j)]);
      const R ddy = R(0.5) * (dzy[c] + dzy[I2(i, jm1)]);
      e = eps_sq_h + R(0.25) * (ddx * ddx) + R(0.25) * (ddy * ddy) + eps_0_sq;
    }
    return R(0.5) * ypow<R>(e, p1) * ypow<R>(ATT[c + k * nc], p2);
  };
  // One pass over the levels: the previous iterate, the new value, its
  // relaxation against the previous iterate, the depth integral (trapezoid
  // from the base, yelmo_tools.f90:1092-1122) and, for DIVA, calc_F_integral
  // with n = 2 (velocity_diva.f90:824-869), each level's relaxed value held in
  // registers for the next level's trapezoid.
  const bool relax = vm != 0 || diva;
  const R H = f.H_ice[cc.c0 + c];
  R s = R(0), F2 = R(0), vprev = R(0), fprev = R(0);
  for (int k = 0; k < nz; ++k) {
    const long long q = c + k * nc;
    const R old = visc[q];
    R v = level_value(k);
    if (relax) v = yexp<R>((RS(1.0) - rel) * ylog<R>(old) + rel * ylog<R>(v));
    visc[q] = v;
    if (ice) {
      const R fk = diva ? (H / v) * ((R(1) - zeta[k]) * (R(1) - zeta[k])) : R(0);
      if (k > 0) {
        const R mid = R(0.5) * (v + vprev);
        s = s + mid * (zeta[k] - zeta[k - 1]);
        if (diva) F2 = F2 + R(0.5) * (fk + fprev) * (zeta[k] - zeta[k - 1]);
      }
      vprev = v;
      fprev = fk;
    }
  }
  R vi = ice ? s * H : visc_min;
  if (vi < visc_min) vi = visc_min;
  f.visc_int[cc.c0 + c] = vi;
  if (diva) f.F2[cc.c0 + c] = F2;
}

// ---- basal friction on aa-nodes (basal_dragging.f90:384-579, 949-1292) ---
template <typename R>
YG_DEV R beta_point(const YgFields<R>& f, const YgFieldParams& p, const CellCtx& cc, int i, int j) {
  const int nx = cc.nx, ny = cc.ny, BC = p.BC;
  const int c = I2(i, j);
  const R* fi = f.f_ice + cc.c0;
  const R* fg = f.f_grnd + cc.c0;
  const R* cb = f.c_bed + cc.c0;
  const R* ux = (p.diva ? f.ux : f.ux) + cc.c0;   // basal velocity (ux_b)
  const R* uy = (p.diva ? f.uy : f.uy) + cc.c0;
  const R q = (R)p.beta_q, u0 = (R)p.beta_u0;
  const int bm = p.beta_method;
  R beta;
  if (bm == 0) {
    beta = (R)p.beta_const;
  } else {
    const bool coulomb = bm == 3 || bm == 5;
    const bool simple = bm == 4 || bm == 5;
    const R qq = bm == 1 ? R(1) : q;
    const R ub_min = R(1e-3), ub_sq_min = ub_min * ub_min;
    int im1, ip1, jm1, jp1;
    nbr(i, j, nx, ny, BC, im1, ip1, jm1, jp1);
    if (fi[c] == R(1)) {
      R cbn[4], uxn[4], uyn[4];
      if (simple) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          cbn[k] = cb[c];
          uxn[k] = RS(0.5) * (ux[c] + ux[I2(im1, j)]);
          uyn[k] = RS(0.5) * (uy[c] + uy[I2(i, jm1)]);
        }
      } else {
        auto aa = [&](int qq_) { return fi[qq_] == R(1) && fg[qq_] > R(0); };
        auto ax = [&](int qi, int qj) { int a1, a2, a3, a4; nbr(qi, qj, nx, ny, BC, a1, a2, a3, a4); return fi[I2(qi, qj)] == R(1) || fi[I2(a2, qj)] == R(1); };
        auto ay = [&](int qi, int qj) { int a1, a2, a3, a4; nbr(qi, qj, nx, ny, BC, a1, a2, a3, a4); return fi[I2(qi, qj)] == R(1) || fi[I2(qi, a4)] == R(1); };
        gq_nodes<R>(0, cbn, cb, aa, nx, i, j, im1, ip1, jm1, jp1);
        gq_nodes_m<R>(1, uxn, ux, gq_mask_ij(1, ax, i, j, im1, ip1, jm1, jp1), nx, i, j, im1, ip1, jm1, jp1);
        gq_nodes_m<R>(2, uyn, uy, gq_mask_ij(2, ay, i, j, im1, ip1, jm1, jp1), nx, i, j, im1, ip1, jm1, jp1);
      }
      R sum = R(0);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const R uxyn = ysqrt<R>(uxn[k] * uxn[k] + uyn[k] * uyn[k] + ub_sq_min);
        const R ratio = coulomb ? uxyn / (uxyn + u0) : uxyn / u0;
        sum += (cbn[k] * ypow_el<R>(ratio, qq) * (R(1) / uxyn)) * R(1);
      }
      beta = sum / R(4);
    } else {
      const R uxy_b = ub_min;
      const R ratio = coulomb ? uxy_b / (uxy_b + u0) : uxy_b / u0;
      beta = cb[c] * ypow_el<R>(ratio, qq) * (R(1) / uxy_b);
    }
  }
  const R q_frz = (bm == 0 || bm == 1) ? R(1) : q;
  beta = beta * ypow<R>(f.f_slide[cc.c0 + c], -q_frz);
  if (p.beta_gl_scale == 0) {
    int im1, ip1, jm1, jp1;
    nbr(i, j, nx, ny, BC, im1, ip1, jm1, jp1);
    if (fg[c] > R(0) && (fg[I2(im1, j)] == R(0) || fg[I2(ip1, j)] == R(0) ||
                         fg[I2(i, jm1)] == R(0) || fg[I2(i, jp1)] == R(0)))
      beta = beta * (R)p.beta_gl_f;
  } else if (p.beta_gl_scale == 1) {
    const R lim = (R)p.H_grnd_lim;
    const R hg = f.H_grnd[cc.c0 + c];
    R fs = (hg < lim ? hg : lim) / lim;
    if (fs < R(0)) fs = R(0);
    beta = beta * fs;
  } else if (p.beta_gl_scale == 2) {
    // scale_beta_gl_zstar with norm (basal_dragging.f90:1296-1352): the
    // thickness above flotation over the effective thickness (calc_H_eff
    // with set_frac_zero, topography.f90:851-874)
    const R fic = fi[c];
    if (fic > R(0)) {
      const R H = f.H_ice[cc.c0 + c];
      R H_eff = H / fic;
      if (fic < R(1)) H_eff = R(0);
      const R zb = f.z_bed[cc.c0 + c], zs = f.z_sl[cc.c0 + c];
      const R rho_sw_ice = (R)p.rho_sw / (R)p.rho_ice;
      R fs;
      if (zb > zs) fs = H_eff;
      else {
        fs = H_eff - (zs - zb) * rho_sw_ice;
        if (fs < R(0)) fs = R(0);
      }
      if (H_eff > R(0)) fs = fs / H_eff;
      beta = beta * fs;
    }
  } else if (p.beta_gl_scale == 3) {
    beta = fg[c] * beta;
  }
  if (fg[c] == R(0)) beta = R(0);
  return beta;
}

template <typename R>
__global__ __launch_bounds__(256) void k_beta(YgBatchGeom G, YgFields<R> f, YgFieldParams p, const int* skip) {
  const long long gc = yg_xcd_cell();
  CellCtx cc;
  if (!cell_of(G, gc, skip, cc)) return;
  if (p.beta_method == -1) return;
  const int nx = cc.nx, ny = cc.ny;
  int i = cc.i, j = cc.j;
  // border copies of calc_beta (basal_dragging.f90:544-563)
  if (p.bnd == YG_B_MISMIP3D) {
    if (i == 0) i = 1;
  } else if (p.bnd == YG_B_INFINITE || p.bnd == YG_B_MASK) {
    if (i == 0) i = 1;
    if (i == nx - 1) i = nx - 2;
    if (j == 0) j = 1;
    if (j == ny - 1) j = ny - 2;
  }
  R beta = beta_point<R>(f, p, cc, i, j);
  const R bmin = (R)p.beta_min;
  if (beta > R(0) && beta < bmin) beta = bmin;
  f.beta[cc.c0 + I2(cc.i, cc.j)] = beta;
}

// ---- staggering, beta_min, matrix friction and limit drag -----------------
// calc_beta_gl_flux_weight (basal_dragging.f90:1760-1845): node a grounded,
// node b floating; the grounded share of the summed segment speeds of 100
// segments, thickness and flux linear between the nodes
template <typename R>
YG_DEV R beta_gl_flux_weight(R beta_a, R u_a, R u_b, R H_a, R H_b, R f_grnd_ac) {
  const R q_a = u_a * H_a, q_b = u_b * H_b;
  R uu_grnd = R(0), uu_tot = R(0);
  for (int k = 1; k <= 100; ++k) {
    const R lambda = ((R)k - R(0.5)) / (R)100;
    const R H_now = H_a * lambda + H_b * (R(1) - lambda);
    const R q_now = q_a * lambda + q_b * (R(1) - lambda);
    const R u_now = H_now > R(0) ? q_now / H_now : R(0);
    uu_tot = uu_tot + yabs<R>(u_now);
    if (lambda > R(1) - f_grnd_ac) uu_grnd = uu_grnd + yabs<R>(u_now);
  }
  const R weight = uu_tot > R(0) ? uu_grnd / uu_tot : R(0);
  return beta_a * weight;
}

// stagger_beta (basal_dragging.f90:581-704), set_beta_min_grounded (:706-780),
// calc_vel_lim_drag (solver_ssa_ac.f90:1095-1199), DIVA calc_beta_eff
// (velocity_diva.f90:912-958).
template <typename R>
YG_DEV void stag_point(const YgFields<R>& f, const YgFieldParams& p, const CellCtx& cc, int i,
                       int j, R& bx, R& by) {
  const int nx = cc.nx, ny = cc.ny;
  const R* fi = f.f_ice + cc.c0;
  const R* fg = f.f_grnd + cc.c0;
  const R* be = f.beta + cc.c0;
  int im1, ip1, jm1, jp1;
  nbr(i, j, nx, ny, p.BC, im1, ip1, jm1, jp1);
  const int c = I2(i, j), cx = I2(ip1, j), cy = I2(i, jp1);
  if (fg[c] == R(0) && fg[cx] == R(0)) bx = R(0);
  else if (fi[c] == R(1) && fi[cx] < R(1)) bx = be[c];
  else if (fi[c] < R(1) && fi[cx] == R(1)) bx = be[cx];
  else bx = R(0.5) * (be[c] + be[cx]);
  if (fg[c] == R(0) && fg[cy] == R(0)) by = R(0);
  else if (fi[c] == R(1) && fi[cy] < R(1)) by = be[c];
  else if (fi[c] < R(1) && fi[cy] == R(1)) by = be[cy];
  else by = R(0.5) * (be[c] + be[cy]);
  const int st = p.beta_gl_stag;
  if (st == 1 || st == 2) {
    if (fi[c] == R(1) && fi[cx] == R(1)) {
      if (fg[c] > R(0) && fg[cx] == R(0)) bx = st == 1 ? be[c] : be[cx];
      else if (fg[c] == R(0) && fg[cx] > R(0)) bx = st == 1 ? be[cx] : be[c];
    }
    if (fi[c] == R(1) && fi[cy] == R(1)) {
      if (fg[c] > R(0) && fg[cy] == R(0)) by = st == 1 ? be[c] : be[cy];
      else if (fg[c] == R(0) && fg[cy] > R(0)) by = st == 1 ? be[cy] : be[c];
    }
  } else if (st == 3) {
    // stagger_beta_aa_gl_subgrid (basal_dragging.f90:1578-1658): weight
    // f_grnd_ac^2 on the grounded side
    if (fi[c] == R(1) && fi[cx] == R(1)) {
      const R g = f.f_grnd_acx[cc.c0 + c], wt = g * g;
      if (fg[c] > R(0) && fg[cx] == R(0)) bx = wt * be[c] + (R(1) - wt) * be[cx];
      else if (fg[c] == R(0) && fg[cx] > R(0)) bx = (R(1) - wt) * be[c] + wt * be[cx];
    }
    if (fi[c] == R(1) && fi[cy] == R(1)) {
      const R g = f.f_grnd_acy[cc.c0 + c], wt = g * g;
      if (fg[c] > R(0) && fg[cy] == R(0)) by = wt * be[c] + (R(1) - wt) * be[cy];
      else if (fg[c] == R(0) && fg[cy] > R(0)) by = (R(1) - wt) * be[c] + wt * be[cy];
    }
  } else if (st == 4) {
    // stagger_beta_aa_gl_subgrid_flux (basal_dragging.f90:1660-1758) with the
    // iterate's velocity (the basal one for SSA, the depth average for DIVA)
    const R* u = (p.diva ? f.ubar_x : f.ux) + cc.c0;
    const R* v = (p.diva ? f.ubar_y : f.uy) + cc.c0;
    const R* H = f.H_ice + cc.c0;
    if (fi[c] == R(1) && fi[cx] == R(1)) {
      const R g = f.f_grnd_acx[cc.c0 + c];
      if (fg[c] > R(0) && fg[cx] == R(0))
        bx = beta_gl_flux_weight<R>(be[c], R(0.5) * (u[I2(im1, j)] + u[c]),
                                    R(0.5) * (u[I2(ip1, j)] + u[c]), H[c], H[cx], g);
      else if (fg[c] == R(0) && fg[cx] > R(0))
        bx = beta_gl_flux_weight<R>(be[cx], R(0.5) * (u[I2(ip1, j)] + u[c]),
                                    R(0.5) * (u[I2(im1, j)] + u[c]), H[cx], H[c], g);
    }
    if (fi[c] == R(1) && fi[cy] == R(1)) {
      const R g = f.f_grnd_acy[cc.c0 + c];
      if (fg[c] > R(0) && fg[cy] == R(0))
        by = beta_gl_flux_weight<R>(be[c], R(0.5) * (v[I2(i, jm1)] + v[c]),
                                    R(0.5) * (v[I2(i, jp1)] + v[c]), H[c], H[cy], g);
      else if (fg[c] == R(0) && fg[cy] > R(0))
        by = beta_gl_flux_weight<R>(be[cy], R(0.5) * (v[I2(i, jp1)] + v[c]),
                                    R(0.5) * (v[I2(i, jm1)] + v[c]), H[cy], H[c], g);
    }
  }
}

template <typename R>
YG_DEV void lim_drag_face(R& k, R& r, R u, R u_cross, int mask, R s0, R inv_du, R tau_c) {
  k = R(0);
  r = R(0);
  if (mask < 1) return;
  const R s = ysqrt<R>(u * u + u_cross * u_cross);
  if (s <= s0) return;
  const R x = (s - s0) * inv_du;
  const R tau = tau_c * x * x;
  const R dtauds = R(2) * tau_c * x * inv_du;
  const R b = tau / s;
  const R w = (u / s) * (u / s);
  k = b * (R(1) - w) + dtauds * w;
  r = (k - b) * u;
}

template <typename R>
__global__ __launch_bounds__(256) void k_stagger(YgBatchGeom G, YgFields<R> f, YgFieldParams p, const int* skip) {
  const long long gc = yg_xcd_cell();
  CellCtx cc;
  if (!cell_of(G, gc, skip, cc)) return;
  const int nx = cc.nx, ny = cc.ny, i = cc.i, j = cc.j;
  const long long g = cc.c0 + I2(i, j);
  R bx, by;
  if (p.beta_gl_stag == -1) {
    bx = f.beta_acx[g];
    by = f.beta_acy[g];
  } else {
    // border copies (basal_dragging.f90:652-683): final(i,j) = staggered(map(i), map(j))
    int ix = i, jx = j, iy = i, jy = j;
    const int b = p.bnd;
    if (b == YG_B_INFINITE || b == YG_B_MISMIP3D || b == YG_B_MASK) {
      if (ix == 0) ix = 1; else if (ix == nx - 2 || ix == nx - 1) ix = nx - 3;
      if (iy == 0) iy = 1; else if (iy == nx - 1) iy = nx - 2;
    }
    if (b == YG_B_INFINITE || b == YG_B_MASK) {
      if (jx == 0) jx = 1; else if (jx == ny - 1) jx = ny - 2;
      if (jy == 0) jy = 1; else if (jy == ny - 2 || jy == ny - 1) jy = ny - 3;
    }
    R t1, t2;
    stag_point<R>(f, p, cc, ix, jx, bx, t1);
    stag_point<R>(f, p, cc, iy, jy, t2, by);
    if (p.beta_method != -1) {
      const R bmin = (R)p.beta_min;
      if (bx > R(0) && bx < bmin) bx = bmin;
      if (by > R(0) && by < bmin) by = bmin;
    }
  }
  f.beta_acx[g] = bx;
  f.beta_acy[g] = by;
  R mbx = bx, mby = by;
  if (p.diva) {
    // stagger_F_integral (velocity_diva.f90:871-910) and calc_beta_eff (:912-958)
    int im1, ip1, jm1, jp1;
    nbr(i, j, nx, ny, p.BC, im1, ip1, jm1, jp1);
    const R* F2 = f.F2 + cc.c0;
    const R* fi = f.f_ice + cc.c0;
    auto smg = [](R v0, R v1, R f0, R f1) {
      if (f0 == R(1) && f1 < R(1)) return v0;
      if (f0 < R(1) && f1 == R(1)) return v1;
      return R(0.5) * (v0 + v1);
    };
    const int c = I2(i, j);
    const R F2x = smg(F2[c], F2[I2(ip1, j)], fi[c], fi[I2(ip1, j)]);
    const R F2y = smg(F2[c], F2[I2(i, jp1)], fi[c], fi[I2(i, jp1)]);
    f.F2_acx[g] = F2x;
    f.F2_acy[g] = F2y;
    auto beff = [&](R b, R F) {
      if (p.no_slip) return F > R(0) ? R(1) / F : R(0);
      return b / (R(1) + b * F);
    };
    mbx = beff(bx, F2x);
    mby = beff(by, F2y);
    f.beff_acx[g] = mbx;
    f.beff_acy[g] = mby;
    f.beta_eff[g] = beff(f.beta[g], F2[c]);
  }
  // set_beta_min_grounded on the copy used by the matrix
  if (f.mx[g] == 1 && mbx == R(0)) mbx = (R)p.beta_min;
  if (f.my[g] == 1 && mby == R(0)) mby = (R)p.beta_min;
  f.bss_x[g] = mbx;
  f.bss_y[g] = mby;
  // limit drag
  R kx = R(0), ky = R(0), rx = R(0), ry = R(0);
  if (p.vel_lim_method == 1) {
    const R* ux = (p.diva ? f.ubar_x : f.ux) + cc.c0;
    const R* uy = (p.diva ? f.ubar_y : f.uy) + cc.c0;
    int im1, ip1, jm1, jp1;
    nbr(i, j, nx, ny, p.BC, im1, ip1, jm1, jp1);
    const R u_max = (R)p.vel_max, tau_c = (R)p.vel_lim_tau;
    const R s0 = R(0.8) * u_max;
    const R inv_du = R(1) / (u_max - s0);
    const int c = I2(i, j);
    R uc = R(0.25) * (uy[c] + uy[I2(ip1, j)] + uy[I2(i, jm1)] + uy[I2(ip1, jm1)]);
    lim_drag_face<R>(kx, rx, ux[c], uc, f.mx[g], s0, inv_du, tau_c);
    uc = R(0.25) * (ux[c] + ux[I2(im1, j)] + ux[I2(i, jp1)] + ux[I2(im1, jp1)]);
    lim_drag_face<R>(ky, ry, uy[c], uc, f.my[g], s0, inv_du, tau_c);
  }
  f.lkx[g] = kx; f.lky[g] = ky; f.lrx[g] = rx; f.lry[g] = ry;
}

// ---- fused residual-form assembly (solver_ssa_ac.f90:73-848, 1000-1067) --
template <typename R>
YG_DEV R visc_ab(const R* N, const R* fi, const R* zb, const R* zs, int nx, int ny, int BC, int i,
                 int j) {
  int im1, ip1, jm1, jp1;
  nbr(i, j, nx, ny, BC, im1, ip1, jm1, jp1);
  const int cs[4] = {I2(i, j), I2(ip1, j), I2(i, jp1), I2(ip1, jp1)};
  int k = 0;
  bool any_ocn = false;
  R v[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const bool ice = fi[cs[q]] == R(1);
    any_ocn |= (!ice) && (zb[cs[q]] < zs[cs[q]]);
    k += ice ? 1 : 0;
    v[q] = ice ? N[cs[q]] : R(0);
  }
  if (any_ocn || k == 0) return R(0);
  return ((v[0] + v[1]) + (v[2] + v[3])) / (R)k;
}

// A cell's record (value 2s + comp is slot s of row comp, then 2 zero pad
// values) as 16-byte stores: the records are 16-byte aligned, and one
// 16-byte store per two or four values instead of one store per value cuts
// the kernel's store instructions by more than half.
YG_DEV void store_record(double* A64, float* A32, size_t base, const double (&a)[2][9]) {
  if (A64) {
    double2* d = reinterpret_cast<double2*>(A64 + base);
#pragma unroll
    for (int s = 0; s < 9; ++s) d[s] = make_double2(a[0][s], a[1][s]);
    d[9] = make_double2(0.0, 0.0);
  }
  if (A32) {
    float4* d = reinterpret_cast<float4*>(A32 + base);
#pragma unroll
    for (int t = 0; t < 4; ++t)
      d[t] = make_float4((float)a[0][2 * t], (float)a[1][2 * t], (float)a[0][2 * t + 1],
                         (float)a[1][2 * t + 1]);
    d[4] = make_float4((float)a[0][8], (float)a[1][8], 0.0f, 0.0f);
  }
}

template <typename R>
__global__ __launch_bounds__(256) void k_assemble(YgBatchGeom G, YgFields<R> f, YgFieldParams p, const int* skip,
                           double* A64, float* A32, double* bvec, double* x0, uint32_t* maskv) {
  // one thread per cell, in the fields' order (i fastest, so the field reads
  // of a wave are coalesced); the cell's unknown pair, record, b and x0 sit
  // at its index in the reference's row order n = i*ny + j
  const long long gt = yg_xcd_cell();
  if (gt >= G.total_cells) return;
  const int m = yg_member_at(G, gt);
  if (m < 0 || (skip && skip[m])) return;
  const YgMember& mb = G.members[m];
  const long long c0 = G.coffs[m];
  const int q = (int)(gt - c0);
  const int nx = mb.nx, ny = mb.ny;
  const int i = q % nx, j = q / nx;
  const long long gq = c0 + (long long)i * ny + j;
  const int ncell = nx * ny;
  // border types (solver_ssa_ac.f90:148-196): 0 no-slip, 1 free-slip, 2 periodic
  int bc[4];  // 0 right, 1 upper, 2 left, 3 lower
  switch (p.bnd) {
    case YG_B_MISMIP3D:
    case YG_B_TROUGH: bc[0] = 1; bc[1] = 2; bc[2] = 0; bc[3] = 2; break;
    case YG_B_PERIODIC: bc[0] = bc[1] = bc[2] = bc[3] = 2; break;
    case YG_B_PERIODIC_X: bc[0] = 2; bc[1] = 1; bc[2] = 2; bc[3] = 1; break;
    case YG_B_PERIODIC_Y: bc[0] = 1; bc[1] = 2; bc[2] = 1; bc[3] = 2; break;
    case YG_B_INFINITE:
    case YG_B_MASK: bc[0] = bc[1] = bc[2] = bc[3] = 1; break;
    default: bc[0] = bc[1] = bc[2] = bc[3] = 0; break;
  }
  const R dx = (R)mb.dx, dy = (R)mb.dy;
  const R inv_dx = R(1) / dx, inv_dxdx = R(1) / (dx * dx);
  const R inv_dy = R(1) / dy, inv_dydy = R(1) / (dy * dy);
  const R inv_dxdy = R(1) / (dx * dy);
  const R* N = f.visc_int + c0;
  const R* fi = f.f_ice + c0;
  const R* zb = f.z_bed + c0;
  const R* zs = f.z_sl + c0;
  const R* ux = (p.diva ? f.ubar_x : f.ux) + c0;
  const R* uy = (p.diva ? f.ubar_y : f.uy) + c0;
  const int c = I2(i, j);
  const int im1 = i == 0 ? nx - 1 : i - 1, ip1 = i == nx - 1 ? 0 : i + 1;
  const int jm1 = j == 0 ? ny - 1 : j - 1, jp1 = j == ny - 1 ? 0 : j + 1;
  const int BC = p.BC;
  double a[2][9];
  uint32_t pm[2] = {0u, 0u};
#pragma unroll
  for (int s = 0; s < 9; ++s) { a[0][s] = 0.0; a[1][s] = 0.0; }
  double bb[2], xx[2];
  const int32_t mxc = f.mx[c0 + c], myc = f.my[c0 + c];
#define PUT(row, s, val) do { a[row][s] = (double)(val); pm[row] |= 1u << (s); } while (0)
  // ---- ux row ----
  if (mxc == 0) {
    PUT(0, 0, 1.0); bb[0] = 0.0; xx[0] = 0.0;
  } else if (mxc == -1) {
    PUT(0, 0, 1.0); bb[0] = ux[c]; xx[0] = ux[c];
  } else if (i == 0 && bc[2] != 2) {
    if (bc[2] == 1) { PUT(0, 0, 1.0); PUT(0, 1, -1.0); bb[0] = 0.0; xx[0] = ux[c]; }
    else { PUT(0, 0, 1.0); bb[0] = 0.0; xx[0] = 0.0; }
  } else if (i == nx - 1 && bc[0] != 2) {
    if (bc[0] == 1) { PUT(0, 0, 1.0); PUT(0, 2, -1.0); bb[0] = 0.0; xx[0] = ux[c]; }
    else { PUT(0, 0, 1.0); bb[0] = 0.0; xx[0] = 0.0; }
  } else if (j == 0 && bc[3] != 2) {
    if (bc[3] == 1) { PUT(0, 0, 1.0); PUT(0, 3, -1.0); bb[0] = 0.0; xx[0] = ux[c]; }
    else { PUT(0, 0, 1.0); bb[0] = 0.0; xx[0] = 0.0; }
  } else if (j == ny - 1 && bc[1] != 2) {
    if (bc[1] == 1) { PUT(0, 0, 1.0); PUT(0, 4, -1.0); bb[0] = 0.0; xx[0] = ux[c]; }
    else { PUT(0, 0, 1.0); bb[0] = 0.0; xx[0] = 0.0; }
  } else if (mxc == 3) {
    if (fi[c] == R(1) && fi[I2(ip1, j)] < R(1)) {
      const R Nn = N[c];
      PUT(0, 2, -R(4) * inv_dx * Nn); PUT(0, 8, -R(2) * inv_dy * Nn);
      PUT(0, 0, R(4) * inv_dx * Nn);  PUT(0, 5, R(2) * inv_dy * Nn);
    } else {
      const R Nn = N[I2(ip1, j)];
      PUT(0, 0, -R(4) * inv_dx * Nn); PUT(0, 7, -R(2) * inv_dy * Nn);
      PUT(0, 1, R(4) * inv_dx * Nn);  PUT(0, 6, R(2) * inv_dy * Nn);
    }
    bb[0] = f.taul_acx[c0 + c]; xx[0] = ux[c];
  } else {
    const R Nab_c = visc_ab<R>(N, fi, zb, zs, nx, ny, BC, i, j);
    const R Nab_jm = visc_ab<R>(N, fi, zb, zs, nx, ny, BC, i, jm1);
    const R w_face = mxc == 4 ? R(0.5) : R(1);
    R beta_now = w_face * f.bss_x[c0 + c];
    R lim_r = R(0);
    beta_now = beta_now + w_face * f.lkx[c0 + c];
    lim_r = w_face * f.lrx[c0 + c];
    PUT(0, 0, -R(4) * inv_dxdx * (N[I2(ip1, j)] + N[c]) - R(1) * inv_dydy * (Nab_c + Nab_jm) - beta_now);
    PUT(0, 1, R(4) * inv_dxdx * N[I2(ip1, j)]);
    PUT(0, 2, R(4) * inv_dxdx * N[c]);
    PUT(0, 3, R(1) * inv_dydy * Nab_c);
    PUT(0, 4, R(1) * inv_dydy * Nab_jm);
    PUT(0, 5, -R(2) * inv_dxdy * N[c] - R(1) * inv_dxdy * Nab_c);
    PUT(0, 6, R(2) * inv_dxdy * N[I2(ip1, j)] + R(1) * inv_dxdy * Nab_c);
    PUT(0, 7, -R(2) * inv_dxdy * N[I2(ip1, j)] - R(1) * inv_dxdy * Nab_jm);
    PUT(0, 8, R(2) * inv_dxdy * N[c] + R(1) * inv_dxdy * Nab_jm);
    bb[0] = f.taud_acx[c0 + c] - lim_r; xx[0] = ux[c];
  }
  // ---- uy row ----
  if (myc == 0) {
    PUT(1, 0, 1.0); bb[1] = 0.0; xx[1] = 0.0;
  } else if (myc == -1) {
    PUT(1, 0, 1.0); bb[1] = uy[c]; xx[1] = uy[c];
  } else if (j == 0 && bc[3] != 2) {
    if (bc[3] == 1) { PUT(1, 0, 1.0); PUT(1, 1, -1.0); bb[1] = 0.0; xx[1] = uy[c]; }
    else { PUT(1, 0, 1.0); bb[1] = 0.0; xx[1] = 0.0; }
  } else if (j == ny - 1 && bc[1] != 2) {
    if (bc[1] == 1) { PUT(1, 0, 1.0); PUT(1, 2, -1.0); bb[1] = 0.0; xx[1] = uy[c]; }
    else { PUT(1, 0, 1.0); bb[1] = 0.0; xx[1] = 0.0; }
  } else if (i == 0 && bc[2] != 2) {
    if (bc[2] == 1) { PUT(1, 0, 1.0); PUT(1, 3, -1.0); bb[1] = 0.0; xx[1] = uy[c]; }
    else { PUT(1, 0, 1.0); bb[1] = 0.0; xx[1] = 0.0; }
  } else if (i == nx - 1 && bc[0] != 2) {
    if (bc[0] == 1) { PUT(1, 0, 1.0); PUT(1, 4, -1.0); bb[1] = 0.0; xx[1] = uy[c]; }
    else { PUT(1, 0, 1.0); bb[1] = 0.0; xx[1] = 0.0; }
  } else if (myc == 3) {
    if (fi[c] == R(1) && fi[I2(i, jp1)] < R(1)) {
      const R Nn = N[c];
      PUT(1, 8, -R(2) * inv_dx * Nn); PUT(1, 2, -R(4) * inv_dy * Nn);
      PUT(1, 5, R(2) * inv_dx * Nn);  PUT(1, 0, R(4) * inv_dy * Nn);
    } else {
      const R Nn = N[I2(i, jp1)];
      PUT(1, 7, -R(2) * inv_dx * Nn); PUT(1, 0, -R(4) * inv_dy * Nn);
      PUT(1, 6, R(2) * inv_dx * Nn);  PUT(1, 1, R(4) * inv_dy * Nn);
    }
    bb[1] = f.taul_acy[c0 + c]; xx[1] = uy[c];
  } else {
    const R Nab_c = visc_ab<R>(N, fi, zb, zs, nx, ny, BC, i, j);
    const R Nab_im = visc_ab<R>(N, fi, zb, zs, nx, ny, BC, im1, j);
    const R w_face = myc == 4 ? R(0.5) : R(1);
    R beta_now = w_face * f.bss_y[c0 + c];
    beta_now = beta_now + w_face * f.lky[c0 + c];
    const R lim_r = w_face * f.lry[c0 + c];
    PUT(1, 0, -R(4) * inv_dydy * (N[I2(i, jp1)] + N[c]) - R(1) * inv_dxdx * (Nab_c + Nab_im) - beta_now);
    PUT(1, 1, R(4) * inv_dydy * N[I2(i, jp1)]);
    PUT(1, 2, R(4) * inv_dydy * N[c]);
    PUT(1, 3, R(1) * inv_dxdx * Nab_c);
    PUT(1, 4, R(1) * inv_dxdx * Nab_im);
    PUT(1, 5, -R(2) * inv_dxdy * N[c] - R(1) * inv_dxdy * Nab_c);
    PUT(1, 6, R(2) * inv_dxdy * N[I2(i, jp1)] + R(1) * inv_dxdy * Nab_c);
    PUT(1, 7, -R(2) * inv_dxdy * N[I2(i, jp1)] - R(1) * inv_dxdy * Nab_im);
    PUT(1, 8, R(2) * inv_dxdy * N[c] + R(1) * inv_dxdy * Nab_im);
    bb[1] = f.taud_acy[c0 + c] - lim_r; xx[1] = uy[c];
  }
#undef PUT
  const size_t base = (size_t)YG_NA * gq;
  store_record(A64, A32, base, a);
  bvec[2 * gq] = bb[0]; bvec[2 * gq + 1] = bb[1];
  x0[2 * gq] = xx[0]; x0[2 * gq + 1] = xx[1];
  maskv[gq] = pm[0] | (pm[1] << 16);
}

// ---- energy-formulation assembly (solver_ssa_ac_energy.f90:70-708) --------
// The matrix is the Hessian of the discrete SSA energy (cell membrane,
// corner shear and face drag terms, each a small local Hessian) with
// Dirichlet unknowns lifted to the right-hand side and free-slip edge
// unknowns folded into their inner root.  Every free row has the 9-slot
// pattern of the residual assembly (inner rows equal -A_residual dx dy), so
// it goes into the same stencil records.  One thread per cell gathers both
// rows from the elements around the unknown and around the edge unknowns
// tied to it, in the reference's order of terms.
namespace en {

enum { FREE = 0, DIR = 1, TIED = 2 };

struct Geo {
  int nx, ny;
  bool per[4], fre[4];  // sides 0 right, 1 upper, 2 left, 3 lower
  bool per_x, per_y;
};

template <typename R>
struct In {
  const int32_t *mx, *my;
  const R *ux, *uy;
};

// classify_ux / classify_uy / edge (:318-386)
template <typename R>
YG_DEV int classify(const Geo& g, const In<R>& in, int i, int j, int comp, int& pi, int& pj,
                    double& val) {
  const int nx = g.nx, ny = g.ny;
  const int c = I2(i, j);
  const int mask = comp ? in.my[c] : in.mx[c];
  val = 0.0;
  if (mask == 0) return DIR;
  if (mask == -1) { val = (double)(comp ? in.uy[c] : in.ux[c]); return DIR; }
  int side = -1;
  const bool xl = i == 0 && !g.per[2], xr = i == nx - 1 && !g.per[0];
  const bool yl = j == 0 && !g.per[3], yu = j == ny - 1 && !g.per[1];
  if (comp == 0) {
    if (xl) { side = 2; pi = 1; pj = j; }
    else if (xr) { side = 0; pi = nx - 2; pj = j; }
    else if (yl) { side = 3; pi = i; pj = 1; }
    else if (yu) { side = 1; pi = i; pj = ny - 2; }
  } else {
    if (yl) { side = 3; pi = i; pj = 1; }
    else if (yu) { side = 1; pi = i; pj = ny - 2; }
    else if (xl) { side = 2; pi = 1; pj = j; }
    else if (xr) { side = 0; pi = nx - 2; pj = j; }
  }
  if (side < 0) return FREE;
  return g.fre[side] ? TIED : DIR;
}

// Kind after following a tied chain (:212-235): FREE (root = itself), DIR
// (value of the chain's end), TIED (root ri, rj), -1 for a chain that does
// not end within 8 steps.
template <typename R>
YG_DEV int resolve(const Geo& g, const In<R>& in, int i, int j, int comp, int& ri, int& rj,
                   double& val) {
  int pi = i, pj = j;
  const int k = classify<R>(g, in, i, j, comp, pi, pj, val);
  ri = i; rj = j;
  if (k != TIED) return k;
  for (int depth = 0; depth < 8; ++depth) {
    int qi = pi, qj = pj;
    double v;
    const int k2 = classify<R>(g, in, pi, pj, comp, qi, qj, v);
    if (k2 == FREE) { ri = pi; rj = pj; return TIED; }
    if (k2 == DIR) { val = v; return DIR; }
    pi = qi; pj = qj;
  }
  return -1;
}

// wrap_index (:420-443)
YG_DEV bool wrap(int& k, int nk, bool periodic, bool free_lo, bool free_hi) {
  if (k < 0) {
    if (periodic) k += nk;
    else if (free_lo) k = 0;
    else return false;
  } else if (k >= nk) {
    if (periodic) k -= nk;
    else if (free_hi) k = nk - 1;
    else return false;
  }
  return true;
}

// dof_ux / dof_uy (:395-418): false for a zero ghost
YG_DEV bool dof(const Geo& g, int& i, int& j) {
  return wrap(i, g.nx, g.per_x, g.fre[2], g.fre[0]) && wrap(j, g.ny, g.per_y, g.fre[3], g.fre[1]);
}

// element_index (:445-457)
YG_DEV bool elem(int& k, int nk, bool periodic) {
  if (k < 0 || k >= nk) {
    if (!periodic) return false;
    k = ((k % nk) + nk) % nk;
  }
  return true;
}

// stencil slot of unknown (ci, cj, ccomp) in the row of (i, j, comp)
YG_DEV int slot(const Geo& g, int comp, int i, int j, int ci, int cj, int ccomp) {
  int di = ci - i, dj = cj - j;
  if (g.per_x) { if (di > 1) di -= g.nx; else if (di < -1) di += g.nx; }
  if (g.per_y) { if (dj > 1) dj -= g.ny; else if (dj < -1) dj += g.ny; }
#pragma unroll
  for (int s = 0; s < YG_NSLOT; ++s) {
    const int8_t* t = comp ? kSlotUy[s] : kSlotUx[s];
    if (t[0] == di && t[1] == dj && t[2] == ccomp) return s;
  }
  return -1;
}

struct Row {
  double a[YG_NSLOT];
  double b;
  bool bad;
};

YG_DEV void add_slot(Row& r, int s, double v) {
#pragma unroll
  for (int t = 0; t < YG_NSLOT; ++t)
    if (t == s) r.a[t] += v;
}

// add_element's inner loop for one column (:626-651)
template <typename R>
YG_DEV void add_col(Row& r, const Geo& g, const In<R>& in, int comp, int i, int j, int ci, int cj,
                    int ccomp, double h) {
  int ri, rj;
  double v;
  const int k = resolve<R>(g, in, ci, cj, ccomp, ri, rj, v);
  if (k == DIR) { r.b -= h * v; return; }
  const int s = k < 0 ? -1 : slot(g, comp, i, j, ri, rj, ccomp);
  if (s < 0) { r.bad = true; return; }
  add_slot(r, s, h);
}

// add_element (:626-651) for an element with dofs (dc, di, dj, valid)
template <typename R>
YG_DEV void add_element(Row& r, const Geo& g, const In<R>& in, int comp, int i, int j, int mc,
                        int mi, int mj, const int dc[4], const int di[4], const int dj[4],
                        const bool dv[4], const double H[4][4]) {
#pragma unroll
  for (int l = 0; l < 4; ++l) {
    if (!dv[l] || dc[l] != mc || di[l] != mi || dj[l] != mj) continue;
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (dv[k]) add_col<R>(r, g, in, comp, i, j, di[k], dj[k], dc[k], H[l][k]);
  }
}

}  // namespace en

// add_member (:488-564): every term of unknown (mc, mi, mj), the row's root
// or an unknown tied to it, into the row of (comp, i, j)
template <typename R>
YG_DEV void en_member(en::Row& r, const en::Geo& g, const en::In<R>& in, const YgFields<R>& f,
                      long long c0, const R* N, const R* fi, const R* zb, const R* zs, int BC,
                      double dx, double dy, int comp, int i, int j, int mi, int mj) {
  const int nx = g.nx, ny = g.ny;
  const int c = I2(mi, mj);
  const int ip1 = mi + 1 == nx ? 0 : mi + 1, jp1 = mj + 1 == ny ? 0 : mj + 1;
  const double dxdy = dx * dy, dyodx = dy / dx, dxody = dx / dy;
  const int mask = comp ? in.my[c] : in.mx[c];
  const double w = (mask == 3 || mask == 4) ? 0.5 : 1.0;
  R beta_now = comp ? f.bss_y[c0 + c] : f.bss_x[c0 + c];
  beta_now = beta_now + (comp ? f.lky[c0 + c] : f.lkx[c0 + c]);
  r.b += w * (double)(comp ? f.lry[c0 + c] : f.lrx[c0 + c]) * dxdy;
  r.a[0] += w * (double)beta_now * dxdy;
  if (mask == 3) {
    const bool out = comp ? (fi[c] == R(1) && fi[I2(mi, jp1)] < R(1))
                          : (fi[c] == R(1) && fi[I2(ip1, mj)] < R(1));
    const double tl = (double)(comp ? f.taul_acy[c0 + c] : f.taul_acx[c0 + c]) * (comp ? dx : dy);
    r.b += out ? tl : -tl;
  } else {
    r.b -= (double)(comp ? f.taud_acy[c0 + c] : f.taud_acx[c0 + c]) * dxdy;
  }
  // elements holding the member: two cells, two corners
  int ce[4][2];
  if (comp == 0) {
    ce[0][0] = mi; ce[0][1] = mj; ce[1][0] = mi + 1; ce[1][1] = mj;
    ce[2][0] = mi; ce[2][1] = mj; ce[3][0] = mi; ce[3][1] = mj - 1;
  } else {
    ce[0][0] = mi; ce[0][1] = mj; ce[1][0] = mi; ce[1][1] = mj + 1;
    ce[2][0] = mi; ce[2][1] = mj; ce[3][0] = mi - 1; ce[3][1] = mj;
  }
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    int ci = ce[e][0], cj = ce[e][1];
    if (!en::elem(ci, nx, g.per_x) || !en::elem(cj, ny, g.per_y)) continue;
    int dc[4], di[4], dj[4];
    bool dv[4];
    double H[4][4];
    if (e < 2) {
      // cell_dofs (:566-596): membrane term of cell (ci, cj)
      const int oi[4] = {ci - 1, ci, ci, ci}, oj[4] = {cj, cj, cj - 1, cj};
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        dc[k] = k < 2 ? 0 : 1; di[k] = oi[k]; dj[k] = oj[k];
        dv[k] = en::dof(g, di[k], dj[k]);
      }
      const double Nc = (double)N[I2(ci, cj)];
      const double hx = 4.0 * Nc * dyodx, hy = 4.0 * Nc * dxody, hc = 2.0 * Nc;
      const double h[4][4] = {{hx, -hx, hc, -hc}, {-hx, hx, -hc, hc},
                              {hc, -hc, hy, -hy}, {-hc, hc, -hy, hy}};
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) H[a][b] = h[a][b];
    } else {
      // corner_dofs (:598-624): shear term of corner (ci, cj)
      const int oi[4] = {ci, ci, ci, ci + 1}, oj[4] = {cj, cj + 1, cj, cj};
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        dc[k] = k < 2 ? 0 : 1; di[k] = oi[k]; dj[k] = oj[k];
        dv[k] = en::dof(g, di[k], dj[k]);
      }
      const double Nc = (double)visc_ab<R>(N, fi, zb, zs, nx, ny, BC, ci, cj);
      const double hx = Nc * dxody, hy = Nc * dyodx;
      const double h[4][4] = {{hx, -hx, Nc, -Nc}, {-hx, hx, -Nc, Nc},
                              {Nc, -Nc, hy, -hy}, {-Nc, Nc, -hy, hy}};
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) H[a][b] = h[a][b];
    }
    en::add_element<R>(r, g, in, comp, i, j, comp, mi, mj, dc, di, dj, dv, H);
  }
}

template <typename R>
__global__ __launch_bounds__(256) void k_assemble_energy(YgBatchGeom G, YgFields<R> f, YgFieldParams p, const int* skip,
                                  double* A64, float* A32, double* bvec, double* x0,
                                  uint32_t* maskv) {
  // one thread per cell in the fields' order; outputs at the cell's index
  // in the reference's row order n = i*ny + j
  const long long gt = yg_xcd_cell();
  if (gt >= G.total_cells) return;
  const int m = yg_member_at(G, gt);
  if (m < 0 || (skip && skip[m])) return;
  const YgMember& mb = G.members[m];
  const long long c0 = G.coffs[m];
  const int q = (int)(gt - c0);
  const int nx = mb.nx, ny = mb.ny;
  const int i = q % nx, j = q / nx;
  const long long gq = c0 + (long long)i * ny + j;
  en::Geo g;
  g.nx = nx; g.ny = ny;
  // border types (:144-167)
  int b[4];  // 0 no-slip, 1 free-slip, 2 periodic
  switch (p.bnd) {
    case YG_B_MISMIP3D:
    case YG_B_TROUGH: b[0] = 1; b[1] = 2; b[2] = 0; b[3] = 2; break;
    case YG_B_PERIODIC: b[0] = b[1] = b[2] = b[3] = 2; break;
    case YG_B_PERIODIC_X: b[0] = 2; b[1] = 1; b[2] = 2; b[3] = 1; break;
    case YG_B_PERIODIC_Y: b[0] = 1; b[1] = 2; b[2] = 1; b[3] = 2; break;
    case YG_B_INFINITE:
    case YG_B_MASK: b[0] = b[1] = b[2] = b[3] = 1; break;
    default: b[0] = b[1] = b[2] = b[3] = 0; break;
  }
#pragma unroll
  for (int s = 0; s < 4; ++s) { g.per[s] = b[s] == 2; g.fre[s] = b[s] == 1; }
  g.per_x = g.per[0] && g.per[2];
  g.per_y = g.per[1] && g.per[3];
  en::In<R> in;
  in.mx = f.mx + c0; in.my = f.my + c0;
  in.ux = (p.diva ? f.ubar_x : f.ux) + c0;
  in.uy = (p.diva ? f.ubar_y : f.uy) + c0;
  const R* N = f.visc_int + c0;
  const R* fi = f.f_ice + c0;
  const R* zb = f.z_bed + c0;
  const R* zs = f.z_sl + c0;
  const double dx = (double)(R)mb.dx, dy = (double)(R)mb.dy;
  double a[2][YG_NSLOT], bb[2], xx[2];
  uint32_t pm[2] = {0u, 0u};
  bool bad = false;
#pragma unroll
  for (int comp = 0; comp < 2; ++comp) {
    int ri, rj;
    double v;
    const int k = en::resolve<R>(g, in, i, j, comp, ri, rj, v);
    en::Row r;
#pragma unroll
    for (int s = 0; s < YG_NSLOT; ++s) r.a[s] = 0.0;
    r.b = 0.0;
    r.bad = false;
    int cf = -1;
    if (k == en::FREE) {
      en_member<R>(r, g, in, f, c0, N, fi, zb, zs, p.BC, dx, dy, comp, i, j, i, j);
      // unknowns of this component tied to (i, j): edge neighbours, visited
      // in increasing unknown order
      for (int ii = i - 1; ii <= i + 1; ++ii)
        for (int jj = j - 1; jj <= j + 1; ++jj) {
          if ((ii == i && jj == j) || ii < 0 || ii >= nx || jj < 0 || jj >= ny) continue;
          int ti, tj;
          double tv;
          if (en::resolve<R>(g, in, ii, jj, comp, ti, tj, tv) == en::TIED && ti == i && tj == j)
            en_member<R>(r, g, in, f, c0, N, fi, zb, zs, p.BC, dx, dy, comp, i, j, ii, jj);
        }
      xx[comp] = (double)(comp ? in.uy[I2(i, j)] : in.ux[I2(i, j)]);
      bb[comp] = r.b;
    } else {
      r.a[0] = 1.0;
      if (k == en::DIR) { bb[comp] = v; xx[comp] = v; }
      else if (k == en::TIED) { bb[comp] = 0.0; xx[comp] = 0.0; cf = 2 * (ri * ny + rj) + comp; }
      else { bb[comp] = 0.0; xx[comp] = 0.0; r.bad = true; }
    }
    bad |= r.bad;
#pragma unroll
    for (int s = 0; s < YG_NSLOT; ++s) {
      a[comp][s] = r.a[s];
      if (r.a[s] != 0.0) pm[comp] |= 1u << s;
    }
    f.cfrom[2 * gq + comp] = cf;
  }
  if (bad) f.asm_err[m] = 1;
  const size_t base = (size_t)YG_NA * gq;
  store_record(A64, A32, base, a);
  bvec[2 * gq] = bb[0]; bvec[2 * gq + 1] = bb[1];
  x0[2 * gq] = xx[0]; x0[2 * gq + 1] = xx[1];
  maskv[gq] = pm[0] | (pm[1] << 16);
}

// ---- Picard epilogue (ps.eparts blocks per member) -------------------------
// linear_solver_save_velocity (solver_ssa_ac.f90:33-71), ssa_vel_clip
// (:1069-1093), picard_relax_vel, picard_calc_convergence_l2 and _l1rel
// (velocity_general.f90:1907-2186).
// Each member counts its own Picard iterations (iter_now), so members that
// entered a batch at different times (served requests) advance together.
// For DIVA the basal stress and basal velocity of the member follow in the
// same pass (calc_basal_stress, velocity_diva.f90:1011-1050; calc_vel_basal,
// :960-1009).
// A member's cells are split over ps.eparts blocks.  Each block publishes its
// block-reduced sums; the block that arrives last (a per-member counter, reset
// to zero by that block) adds the parts in part order, so the result does not
// depend on which block finishes when, and takes the convergence decision.
template <typename R>
__global__ __launch_bounds__(1024) void k_epilogue(YgBatchGeom G, YgFields<R> f, YgFieldParams p,
                                                   const double* x, const int* lin_iters,
                                                   const int* lin_status, int iter_max,
                                                   YgPicardState ps) {
  const int EP = ps.eparts > 1 && ps.epart && ps.eticket ? ps.eparts : 1;
  const int m = blockIdx.x / EP, part = blockIdx.x - m * EP;
  if (ps.done[m]) return;
  const int iter = ps.iter_now[m] + 1;
  const YgMember mb = G.members[m];
  const int nx = mb.nx, ny = mb.ny, ncell = nx * ny;
  const long long c0 = G.coffs[m];
  R* ux = (p.diva ? f.ubar_x : f.ux) + c0;
  R* uy = (p.diva ? f.ubar_y : f.uy) + c0;
  const R* uxp = f.ux_nm1 + c0;
  const R* uyp = f.uy_nm1 + c0;
  const int32_t* mx = f.mx + c0;
  const int32_t* my = f.my + c0;
  const R rel = (R)p.iter_rel, ulim = (R)p.vel_max;
  const R vel_tol = RS(1e-5), tol_uf = R(1e-15);
  const R etol = RS(1e-2), etol2 = RS(1e-5);
  __shared__ double red[16 * 4];
  double acc[4] = {0.0, 0.0, 0.0, 0.0};  // res1, res2, nchk, nonfinite
  const int chunk = (ncell + EP - 1) / EP, c_lo = part * chunk, c_hi = min(ncell, c_lo + chunk);
  for (int c = c_lo + threadIdx.x; c < c_hi; c += blockDim.x) {
    const int i = c % nx, j = c / nx;
    const long long n = c0 + (long long)i * ny + j;
    long long sx = 2 * n, sy = 2 * n + 1;
    if (f.cfrom) {
      // rows tied to another row take its solution (solver_ssa_ac.f90:51-59)
      const int qx = f.cfrom[sx], qy = f.cfrom[sy];
      if (qx >= 0) sx = 2 * c0 + qx;
      if (qy >= 0) sy = 2 * c0 + qy;
    }
    R vx = (R)x[sx], vy = (R)x[sy];
    if (yabs<R>(vx) < tol_uf) vx = R(0);
    if (yabs<R>(vy) < tol_uf) vy = R(0);
    if (vx > ulim) vx = ulim;
    if (vx < -ulim) vx = -ulim;
    if (vy > ulim) vy = ulim;
    if (vy < -ulim) vy = -ulim;
    vx = uxp[c] + rel * (vx - uxp[c]);
    vy = uyp[c] + rel * (vy - uyp[c]);
    ux[c] = vx;
    uy[c] = vy;
    if (yabs<R>(vx) > vel_tol && mx[c] > 0) {
      double tt = (double)(vx - uxp[c]);
      if (fabs(tt) < (double)tol_uf) tt = 0.0;
      acc[0] += tt * tt;
      tt = (double)uxp[c];
      if (fabs(tt) < (double)tol_uf) tt = 0.0;
      acc[1] += tt * tt;
      acc[2] += 1.0;
    }
    if (yabs<R>(vy) > vel_tol && my[c] > 0) {
      double tt = (double)(vy - uyp[c]);
      if (fabs(tt) < (double)tol_uf) tt = 0.0;
      acc[0] += tt * tt;
      tt = (double)uyp[c];
      if (fabs(tt) < (double)tol_uf) tt = 0.0;
      acc[1] += tt * tt;
      acc[2] += 1.0;
    }
    if ((mx[c] > 0 && !isfinite((double)vx)) || (my[c] > 0 && !isfinite((double)vy))) acc[3] += 1.0;
    f.err_x[c0 + c] = yabs<R>(vx) > etol ? R(2) * yabs<R>(vx - uxp[c]) / yabs<R>(vx + uxp[c] + etol2) : R(0);
    f.err_y[c0 + c] = yabs<R>(vy) > etol ? R(2) * yabs<R>(vy - uyp[c]) / yabs<R>(vy + uyp[c] + etol2) : R(0);
    if (p.diva) {
      const R tol = (R)(float)1e-15;   // TOL_UNDERFLOW (yelmo_defs.f90:64)
      R tx = f.beff_acx[c0 + c] * vx;
      if (yabs<R>(tx) < tol) tx = R(0);
      R ty = f.beff_acy[c0 + c] * vy;
      if (yabs<R>(ty) < tol) ty = R(0);
      f.taub_x[c0 + c] = tx;
      f.taub_y[c0 + c] = ty;
      // the basal velocity is the DIVA state's ux/uy (ux_b, uy_b)
      f.ux[c0 + c] = p.no_slip ? R(0) : vx - tx * f.F2_acx[c0 + c];
      f.uy[c0 + c] = p.no_slip ? R(0) : vy - ty * f.F2_acy[c0 + c];
    }
  }
  yg_block_sum<4>(acc, red);
  if (threadIdx.x == 0 && EP > 1) {
    double* mine = ps.epart + ((size_t)m * EP + part) * 4;
    for (int k = 0; k < 4; ++k) mine[k] = acc[k];
    __threadfence();   // the part is visible before the arrival is counted
    if (atomicAdd(&ps.eticket[m], 1) != EP - 1) return;
    __threadfence();   // every other part is visible to the last block
    const double* parts = ps.epart + (size_t)m * EP * 4;
    for (int k = 0; k < 4; ++k) acc[k] = 0.0;
    for (int q = 0; q < EP; ++q)
      for (int k = 0; k < 4; ++k) acc[k] += parts[q * 4 + k];
    ps.eticket[m] = 0;
  }
  if (threadIdx.x == 0) {
    const R resid = acc[2] > 0.0 ? (R)(sqrt(acc[0]) / (sqrt(acc[1]) + 1e-10)) : R(0);
    int d = 0;
    if (resid < (R)p.iter_conv) d = 1;
    else if (iter == iter_max) d = 2;
    if (acc[3] > 0.0 || !isfinite((double)resid)) d = -1;
    ps.resid[m] = (double)resid;
    ps.iter_now[m] = iter;
    ps.lin_iter[m] += lin_iters[m];
    if (lin_status[m] != 0) ps.lin_fail[m] += 1;
    ps.done[m] = d;
  }
}

// ---- after the loop: count_vel_lim_faces (solver_ssa_ac.f90:1201-1262) and
// calc_basal_stress (velocity_ssa.f90:588-617) ------------------------------
template <typename R>
__global__ __launch_bounds__(1024) void k_finish(YgBatchGeom G, YgFields<R> f, YgFieldParams p,
                                                 YgPicardState ps, const int* skip) {
  const int m = blockIdx.x;
  if (skip && skip[m]) return;
  const YgMember mb = G.members[m];
  const int nx = mb.nx, ny = mb.ny, ncell = nx * ny;
  const long long c0 = G.coffs[m];
  const R* ux = (p.diva ? f.ubar_x : f.ux) + c0;
  const R* uy = (p.diva ? f.ubar_y : f.uy) + c0;
  const int32_t* mx = f.mx + c0;
  const int32_t* my = f.my + c0;
  const R u_max = (R)p.vel_max, s0 = R(0.8) * u_max;
  const bool drag = p.vel_lim_method == 1;
  __shared__ double red[16];
  double cnt[1] = {0.0};
  for (int c = threadIdx.x; c < ncell; c += blockDim.x) {
    const int i = c % nx, j = c / nx;
    int im1, ip1, jm1, jp1;
    nbr(i, j, nx, ny, p.BC, im1, ip1, jm1, jp1);
    if (mx[c] >= 1) {
      if (drag) {
        const R uc = R(0.25) * (uy[c] + uy[I2(ip1, j)] + uy[I2(i, jm1)] + uy[I2(ip1, jm1)]);
        if (ux[c] * ux[c] + uc * uc > s0 * s0) cnt[0] += 1.0;
      } else if (yabs<R>(ux[c]) >= u_max) cnt[0] += 1.0;
    }
    if (my[c] >= 1) {
      if (drag) {
        const R uc = R(0.25) * (ux[c] + ux[I2(im1, j)] + ux[I2(i, jp1)] + ux[I2(im1, jp1)]);
        if (uy[c] * uy[c] + uc * uc > s0 * s0) cnt[0] += 1.0;
      } else if (yabs<R>(uy[c]) >= u_max) cnt[0] += 1.0;
    }
    if (!p.diva) {
      R tx = f.bss_x[c0 + c] * ux[c], ty = f.bss_y[c0 + c] * uy[c];
      if (yabs<R>(tx) < R(1e-5)) tx = R(0);
      if (yabs<R>(ty) < R(1e-5)) ty = R(0);
      f.taub_x[c0 + c] = tx;
      f.taub_y[c0 + c] = ty;
    }
  }
  yg_block_sum<1>(cnt, red);
  if (threadIdx.x == 0) ps.lim_n[m] = (int)cnt[0];
}

// ---- before the loop: set_inactive_margins (velocity_general.f90:1687-1750)
template <typename R>
__global__ __launch_bounds__(256) void k_init(YgBatchGeom G, YgFields<R> f, YgFieldParams p, const int* skip) {
  const long long gc = yg_xcd_cell();
  CellCtx cc;
  if (!cell_of(G, gc, skip, cc)) return;
  const int nx = cc.nx, ny = cc.ny, i = cc.i, j = cc.j;
  int im1, ip1, jm1, jp1;
  nbr(i, j, nx, ny, p.BC, im1, ip1, jm1, jp1);
  const R* fi = f.f_ice + cc.c0;
  const long long g = cc.c0 + I2(i, j);
  const R fa = fi[I2(i, j)];
  R fb = fi[I2(ip1, j)];
  R* ux = (p.diva ? f.ubar_x : f.ux);
  R* uy = (p.diva ? f.ubar_y : f.uy);
  if (!p.diva) {
    if ((fb == R(0) && fa < R(1)) || (fa == R(0) && fb < R(1))) ux[g] = R(0);
    fb = fi[I2(i, jp1)];
    if ((fb == R(0) && fa < R(1)) || (fa == R(0) && fb < R(1))) uy[g] = R(0);
  }
  f.err_x[g] = R(1);
  f.err_y[g] = R(1);
  // ATT^(-1/n) at the quadrature nodes for visc_method 1 (the same values
  // k_visc computes from ATT; ATT does not change during a solve)
  if (f.attp && p.visc_method == 1) {
    const int nz = cc.nz;
    const long long nc = (long long)nx * ny;
    const int c = I2(i, j);
    const R p2 = -RS(1.0) / (R)p.n_glen;
    const R* ATT = f.ATT + cc.z0;
    R* attp = f.attp + 4 * cc.z0 + c;
    auto all = [&](int) { return true; };
    for (int k = 0; k < nz; ++k) {
      R ATTn[4];
      gq_nodes<R>(0, ATTn, ATT + k * nc, all, nx, i, j, im1, ip1, jm1, jp1);
#pragma unroll
      for (int q = 0; q < 4; ++q) attp[(q * nz + k) * nc] = ypow<R>(ATTn[q], p2);
    }
  }
}

}  // namespace

// grids of the per-cell kernels: a multiple of 8 blocks (yg_xcd_cell)
static inline unsigned nblk(long long n, int b) { return yg_xcd_grid(n, b); }

template <typename R>
void yg_launch_init(const YgBatchGeom& G, const YgFields<R>& f, const YgFieldParams& p,
                    const int* skip, hipStream_t st) {
  hipLaunchKernelGGL(k_init<R>, dim3(nblk(G.total_cells, 256)), dim3(256), 0, st, G, f, p, skip);
}

template <typename R>
void yg_launch_fields_pre(const YgBatchGeom& G, const YgFields<R>& f, const YgFieldParams& p,
                          const int* skip, hipStream_t st) {
  const dim3 grid(nblk(G.total_cells, 256)), blk(256);
  hipLaunchKernelGGL(k_strain<R>, grid, blk, 0, st, G, f, p, skip);
  if (p.visc_method == 1 && f.attp && !p.diva)
    hipLaunchKernelGGL((k_visc<R, 1>), grid, blk, 0, st, G, f, p, skip);
  else if (p.visc_method == 1 && f.attp)
    hipLaunchKernelGGL((k_visc<R, 2>), grid, blk, 0, st, G, f, p, skip);
  else
    hipLaunchKernelGGL((k_visc<R, 0>), grid, blk, 0, st, G, f, p, skip);
  hipLaunchKernelGGL(k_beta<R>, grid, blk, 0, st, G, f, p, skip);
  hipLaunchKernelGGL(k_stagger<R>, grid, blk, 0, st, G, f, p, skip);
}

template <typename R>
void yg_launch_assemble(const YgBatchGeom& G, const YgFields<R>& f, const YgFieldParams& p,
                        const int* skip, double* A64, float* A32, double* b, double* x0,
                        uint32_t* mask, hipStream_t st) {
  if (p.energy)
    hipLaunchKernelGGL(k_assemble_energy<R>, dim3(nblk(G.total_cells, 256)), dim3(256), 0, st, G,
                       f, p, skip, A64, A32, b, x0, mask);
  else
    hipLaunchKernelGGL(k_assemble<R>, dim3(nblk(G.total_cells, 256)), dim3(256), 0, st, G, f, p,
                       skip, A64, A32, b, x0, mask);
}

template <typename R>
void yg_launch_epilogue(const YgBatchGeom& G, const YgFields<R>& f, const YgFieldParams& p,
                        const double* x, const int* lin_iters, const int* lin_status,
                        int iter_max, YgPicardState ps, hipStream_t st) {
  const int ep = ps.eparts > 1 && ps.epart && ps.eticket ? ps.eparts : 1;
  hipLaunchKernelGGL(k_epilogue<R>, dim3(G.nmem * ep), dim3(1024), 0, st, G, f, p, x, lin_iters,
                     lin_status, iter_max, ps);
}

template <typename R>
void yg_launch_finish(const YgBatchGeom& G, const YgFields<R>& f, const YgFieldParams& p,
                      YgPicardState ps, const int* skip, hipStream_t st) {
  hipLaunchKernelGGL(k_finish<R>, dim3(G.nmem), dim3(1024), 0, st, G, f, p, ps, skip);
}


#define YG_INST(R)                                                                              \
  template void yg_launch_init<R>(const YgBatchGeom&, const YgFields<R>&, const YgFieldParams&, \
                                  const int*, hipStream_t);                                     \
  template void yg_launch_fields_pre<R>(const YgBatchGeom&, const YgFields<R>&,                 \
                                        const YgFieldParams&, const int*, hipStream_t);     \
  template void yg_launch_assemble<R>(const YgBatchGeom&, const YgFields<R>&,                   \
                                      const YgFieldParams&, const int*, double*, float*,         \
                                      double*, double*, uint32_t*, hipStream_t);                 \
  template void yg_launch_epilogue<R>(const YgBatchGeom&, const YgFields<R>&,                   \
                                      const YgFieldParams&, const double*, const int*,           \
                                      const int*, int, YgPicardState, hipStream_t);              \
  template void yg_launch_finish<R>(const YgBatchGeom&, const YgFields<R>&,                     \
                                    const YgFieldParams&, YgPicardState, const int*,            \
                                    hipStream_t);
YG_INST(float)
YG_INST(double)
