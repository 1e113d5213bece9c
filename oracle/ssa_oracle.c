/*
 * ssa_oracle.c -- CPU restatement of Yelmo's SSA velocity solve.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle that the HIP library is
 * checked against.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it; the product library never links or calls it.
 *
 * Built twice (oracle/Makefile):
 *   liboracle_sp.so  wp = float   -- the reference as shipped (yelmo_defs.f90:42
 *                                   wp = sp): FP32 fields, FP64 CSR + solve
 *   liboracle_dp.so  wp = double  -- the all-FP64 build (wp = dp)
 *
 * Every function cites the reference routine it restates (paths relative to
 * the reference tree).  Arrays are Fortran column-major (nx,ny[,nz]) with i
 * fastest: a(i,j,k) -> a[i + nx*(j + ny*k)], indices 0-based here.
 *
 * Parity status (the pinning checks are the tests in tests/test_oracle_pins.py):
 *   - assembly, masks, strain rates, aa-viscosity, power/Coulomb beta at the
 *     cell centre (beta_method 4/5), staggering, limit drag, Picard norms:
 *     pinned by closed forms / invariants / manufactured solutions in tests/.
 *   - gq2D quadrature (visc_method 1, beta_method 1-3): fesm-utils
 *     gaussian_quadrature is not vendored; the node interpolation below is a
 *     restatement from the reference's own description -- parity unpinned.
 *   - LIS Krylov semantics (BiCGSTAB/CG, ILU(0)/Jacobi, conv_cond default,
 *     early exit on s): LIS is not vendored -- parity unpinned beyond the
 *     residual criterion and the dense-LU solution of the same system.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#ifdef ORACLE_DP
typedef double wp;
#define WPOW pow
#define WSQRT sqrt
#define WEXP exp
#define WLOG log
#define WABS fabs
#else
typedef float wp;
#define WPOW powf
#define WSQRT sqrtf
#define WEXP expf
#define WLOG logf
#define WABS fabsf
#endif

#define R(x) ((wp)(x))            /* x_wp literal                            */
#define RS(x) ((wp)(float)(x))    /* default-real (single) literal, to wp    */

/* Planted mistakes, for the pinning tests only: oracle_plant = k switches on
 * mistake k (a plausible slip in the restatement), so a test can show that
 * its check fails on it.  0 (the default) is the restatement itself.       */
int oracle_plant = 0;
#define PLANT(k) (oracle_plant == (k))
enum {
    PL_GQ_NODE_POS = 1,      /* Gauss nodes at +-1/2 instead of +-1/sqrt(3)        */
    PL_GQ_ACX_CORNER = 2,    /* acx corner NE takes face (i, j-1) for (i, j+1)     */
    PL_GQ_ACT_IGNORED = 3,   /* corner means ignore the active-face flags          */
    PL_GLEN_SHEAR = 4,       /* 1/4 (du/dy + dv/dx)^2 without the 1/4              */
    PL_LIM_W = 5,            /* w = u/s instead of (u/s)^2                        */
    PL_LIM_R = 6,            /* r = k u instead of (k - b) u                      */
    PL_LIM_CROSS = 7,        /* cross velocity at the acx face from uy(i, j+1)    */
    PL_CONV_VEL_TOL = 8,     /* faces with |u| <= vel_tol counted in the L2 norm  */
    PL_CONV_MASK = 9,        /* faces with mask 0 counted in the L2 norm          */
    PL_CONV_DU_REG = 10,     /* no du_reg in the denominator                      */
    PL_CONV_NORM2 = 11,      /* norm_method 2 (2 |du| / |u + u_prev|)             */
    PL_L1REL_TOL = 12,       /* l1rel without the + tol in the denominator        */
    PL_RELAX_VEL = 13,       /* u_prev + (1 - rel)(u - u_prev)                    */
    PL_RELAX_VISC = 14,      /* linear instead of log-space viscosity relaxation  */
    PL_CLIP = 15,            /* clip only positive velocities                     */
    PL_GL_UPSTREAM = 16,     /* grounding-line face takes the floating side       */
    PL_BETA_MIN_MASK = 17,   /* beta_min on every free face (mask >= 1)           */
    PL_DIVA_SHEAR = 18,      /* DIVA viscosity without the 1/4 (dv/dz)^2 term     */
    PL_BETA_Q = 19,          /* friction law with |u|^q instead of |u|^(q-1)      */
    PL_GQ_ATT_ACT = 20,      /* rate factor corners over grounded cells only      */
    PL_ADV_INFLOW = 21,      /* advection: the left face takes the right velocity */
    PL_ADV_DIAG = 22,        /* advection: outflow through the top face dropped    */
    PL_ZSTAR_NORM = 23,      /* zstar scaling not normalised by H_eff              */
    PL_ZSTAR_RHO = 24,       /* zstar flotation depth with rho_ice/rho_sw          */
    PL_SUBGRID_WT = 25,      /* subgrid staggering weight f_grnd_ac, not squared   */
    PL_FLUX_SIDE = 26,       /* flux weight counts segments with lambda < f_grnd   */
    PL_FRONT_FREE_BED = 27,  /* ice-free front code from the front cell's bed      */
    PL_GRAD_BORDER = 28,     /* no border copy of the gradient                     */
    PL_TAUD_MARGIN = 29,     /* margin faces take the mean thickness               */
    PL_TAUL_LAND = 30,       /* sea-level back-pressure also against land          */
    PL_TAUL_RHO = 31         /* rho_sw in the ice term of the front stress         */
};

/* ------------------------------------------------------------------ */
/* boundaries strings (yelmo_tools.f90:248-271, solver_ssa_ac.f90:148-196) */
enum {
    OB_ZEROS = 0, OB_INFINITE = 1, OB_MASK = 2, OB_MISMIP3D = 3,
    OB_TROUGH = 4, OB_PERIODIC = 5, OB_PERIODIC_X = 6, OB_PERIODIC_Y = 7
};
/* BND_* codes (yelmo_tools.f90:14-20) */
enum { BND_ZEROS = 0, BND_INFINITE = 1, BND_MISMIP3D = 2, BND_TROUGH = 3,
       BND_PERIODIC = 4, BND_PERIODIC_X = 5, BND_PERIODIC_Y = 6 };

/* boundary_code (yelmo_tools.f90:248-271) */
static int boundary_code(int bnd)
{
    switch (bnd) {
    case OB_ZEROS: return BND_ZEROS;
    case OB_INFINITE: return BND_INFINITE;
    case OB_MASK: return BND_INFINITE;
    case OB_MISMIP3D: return BND_MISMIP3D;
    case OB_TROUGH: return BND_TROUGH;
    case OB_PERIODIC: return BND_PERIODIC;
    case OB_PERIODIC_X: return BND_PERIODIC_X;
    case OB_PERIODIC_Y: return BND_PERIODIC_Y;
    }
    return BND_ZEROS;
}

/* get_neighbor_indices_bc_codes (yelmo_tools.f90:118-201), 0-based */
static void nbr(int i, int j, int nx, int ny, int BC,
                int *im1, int *ip1, int *jm1, int *jp1)
{
    int wrapx = (BC == BND_PERIODIC || BC == BND_PERIODIC_X);
    int wrapy = (BC == BND_MISMIP3D || BC == BND_TROUGH || BC == BND_PERIODIC ||
                 BC == BND_PERIODIC_Y);
    if (wrapx) {
        *im1 = (i == 0) ? nx - 1 : i - 1;
        *ip1 = (i == nx - 1) ? 0 : i + 1;
    } else {
        *im1 = i > 0 ? i - 1 : 0;
        *ip1 = i < nx - 1 ? i + 1 : nx - 1;
    }
    if (wrapy) {
        *jm1 = (j == 0) ? ny - 1 : j - 1;
        *jp1 = (j == ny - 1) ? 0 : j + 1;
    } else {
        *jm1 = j > 0 ? j - 1 : 0;
        *jp1 = j < ny - 1 ? j + 1 : ny - 1;
    }
}

#define IX(i, j) ((i) + nx * (j))

/* ------------------------------------------------------------------ */
/* Parameters: mirrors ssa_param_class (velocity_ssa.f90:22-52).  Layout
 * is shared with the Python test driver (oracle/oracle.py).               */
typedef struct {
    int32_t ssa_solver;      /* 0 residual, 1 energy                       */
    int32_t boundaries;      /* OB_*                                       */
    int32_t visc_method;     /* 0, 1, 2                                    */
    int32_t beta_method;     /* -1 .. 5                                    */
    int32_t beta_gl_scale;   /* 0 .. 3                                     */
    int32_t beta_gl_stag;    /* -1 .. 4                                    */
    int32_t vel_lim_method;  /* 0 clip, 1 drag                             */
    int32_t ssa_iter_max;
    int32_t ssa_write_log;
    int32_t pad0;
    double visc_const, beta_const, beta_q, beta_u0, beta_gl_f, H_grnd_lim,
        beta_min, eps_0, ssa_vel_lim_tau, ssa_vel_max, ssa_iter_rel,
        ssa_iter_conv, rho_ice, rho_sw, g;
} oracle_ssa_params;

/* Krylov settings parsed from the LIS option string
 * (solver_linear.F90:350; grammar in input/yelmo_defaults.nml:160-161)    */
typedef struct {
    int solver;      /* 1 cg, 4 bicgstab                                   */
    int precon;      /* 0 none, 1 jacobi, 2 ilu                            */
    int maxiter;
    double tol;
    int initx_zeros;
    int conv_cond;   /* 0: ||r||/||r0||, 1: ||r||/||b||                     */
    int ilu_blocks;  /* ILU(0) over this many contiguous row blocks         */
    int ilu_nowrap_ny; /* > 0: ILU(0) without the couplings that wrap around a
                          periodic domain (C-grid rows, ny cells per line)     */
} lis_opts;

static int parse_lis_opts(const char *s, lis_opts *o)
{
    o->solver = 4; o->precon = 0; o->maxiter = 1000; o->tol = 1e-12;
    o->initx_zeros = 1; o->conv_cond = 0; o->ilu_blocks = 1; o->ilu_nowrap_ny = 0;
    if (!s) return 0;
    char buf[512];
    strncpy(buf, s, sizeof(buf) - 1);
    buf[sizeof(buf) - 1] = 0;
    char *save = NULL;
    char *tok = strtok_r(buf, " \t", &save);
    while (tok) {
        char *val = strtok_r(NULL, " \t", &save);
        if (!val) return -1;
        if (!strcmp(tok, "-i")) {
            if (!strcmp(val, "cg") || !strcmp(val, "1")) o->solver = 1;
            else if (!strcmp(val, "bicgstab") || !strcmp(val, "4")) o->solver = 4;
            /* LIS's MINRES and BiCGSafe (named by the reference's namelists,
             * input/yelmo_defaults.nml:160) solve the same system here with
             * BiCGSTAB, as the library does */
            else if (!strcmp(val, "minres") || !strcmp(val, "bicgsafe")) o->solver = 4;
            else return -2;
        } else if (!strcmp(tok, "-p")) {
            if (!strcmp(val, "none") || !strcmp(val, "0")) o->precon = 0;
            else if (!strcmp(val, "jacobi") || !strcmp(val, "1")) o->precon = 1;
            else if (!strcmp(val, "ilu") || !strcmp(val, "2")) o->precon = 2;
            else return -3;
        } else if (!strcmp(tok, "-maxiter")) {
            o->maxiter = atoi(val);
        } else if (!strcmp(tok, "-tol")) {
            o->tol = atof(val);
        } else if (!strcmp(tok, "-initx_zeros")) {
            o->initx_zeros = (!strcmp(val, "true") || !strcmp(val, "1"));
        } else if (!strcmp(tok, "-conv_cond")) {
            o->conv_cond = atoi(val);
        } else if (!strcmp(tok, "-ilu_nowrap")) {
            o->ilu_nowrap_ny = atoi(val);
        } else if (!strcmp(tok, "-ilu_blocks")) {
            o->ilu_blocks = atoi(val);
            if (o->ilu_blocks < 1) o->ilu_blocks = 1;
        } else if (!strcmp(tok, "-omp_num_threads") || !strcmp(tok, "-print") ||
                   !strcmp(tok, "-scale") || !strcmp(tok, "-f")) {
            /* accepted, no effect on the arithmetic restated here */
        } else {
            return -4;
        }
        tok = strtok_r(NULL, " \t", &save);
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* calc_active_faces (deformation.f90:1316-1344)                        */
static void calc_active_faces(uint8_t *act_acx, uint8_t *act_acy, const wp *f_ice,
                              int nx, int ny, int BC)
{
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
            int im1, ip1, jm1, jp1;
            nbr(i, j, nx, ny, BC, &im1, &ip1, &jm1, &jp1);
            act_acx[IX(i, j)] = f_ice[IX(i, j)] == R(1) || f_ice[IX(ip1, j)] == R(1);
            act_acy[IX(i, j)] = f_ice[IX(i, j)] == R(1) || f_ice[IX(i, jp1)] == R(1);
        }
}

/* deriv_active (deformation.f90:1346-1368) */
static wp deriv_active(wp v_m, wp v_0, wp v_p, int act_m, int act_p, wp h)
{
    if (act_m && act_p) return (v_p - v_m) / (R(2) * h);
    if (act_p) return (v_p - v_0) / h;
    if (act_m) return (v_0 - v_m) / h;
    return R(0);
}

/* calc_strain_rate_horizontal_2D (deformation.f90:1393-1507) */
void oracle_strain_rate_2D(wp *dudx, wp *dudy, wp *dvdx, wp *dvdy, const wp *ux,
                           const wp *uy, const wp *f_ice, wp dx, wp dy, int nx, int ny,
                           int bnd)
{
    int BC = boundary_code(bnd);
    uint8_t *act_acx = malloc(nx * ny), *act_acy = malloc(nx * ny);
    calc_active_faces(act_acx, act_acy, f_ice, nx, ny, BC);
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
            int im1, ip1, jm1, jp1, im2, ip2, jm2, jp2, d1, d2, d3;
            nbr(i, j, nx, ny, BC, &im1, &ip1, &jm1, &jp1);
            int c = IX(i, j);
            dudx[c] = (ux[IX(ip1, j)] - ux[IX(im1, j)]) / (R(2) * dx);
            dudy[c] = deriv_active(ux[IX(i, jm1)], ux[c], ux[IX(i, jp1)],
                                   act_acx[IX(i, jm1)], act_acx[IX(i, jp1)], dy);
            dvdx[c] = deriv_active(uy[IX(im1, j)], uy[c], uy[IX(ip1, j)],
                                   act_acy[IX(im1, j)], act_acy[IX(ip1, j)], dx);
            dvdy[c] = (uy[IX(i, jp1)] - uy[IX(i, jm1)]) / (R(2) * dy);

            if (f_ice[c] == R(1) && f_ice[IX(ip1, j)] < R(1)) {
                nbr(im1, j, nx, ny, BC, &im2, &d1, &d2, &d3);
                if (f_ice[IX(im1, j)] == R(1) && im2 != im1)
                    dudx[c] = (R(1) * ux[IX(im2, j)] - R(4) * ux[IX(im1, j)] + R(3) * ux[c]) /
                              (R(2) * dx);
                else
                    dudx[c] = (ux[c] - ux[IX(im1, j)]) / dx;
            } else if (f_ice[c] < R(1) && f_ice[IX(ip1, j)] == R(1)) {
                nbr(ip1, j, nx, ny, BC, &d1, &ip2, &d2, &d3);
                if (ip2 != ip1 && f_ice[IX(ip2, j)] == R(1))
                    dudx[c] = -(R(1) * ux[IX(ip2, j)] - R(4) * ux[IX(ip1, j)] + R(3) * ux[c]) /
                              (R(2) * dx);
                else
                    dudx[c] = (ux[IX(ip1, j)] - ux[c]) / dx;
            }

            if (f_ice[c] == R(1) && f_ice[IX(i, jp1)] < R(1)) {
                nbr(i, jm1, nx, ny, BC, &d1, &d2, &jm2, &d3);
                if (f_ice[IX(i, jm1)] == R(1) && jm2 != jm1)
                    dvdy[c] = (R(1) * uy[IX(i, jm2)] - R(4) * uy[IX(i, jm1)] + R(3) * uy[c]) /
                              (R(2) * dy);
                else
                    dvdy[c] = (uy[c] - uy[IX(i, jm1)]) / dy;
            } else if (f_ice[c] < R(1) && f_ice[IX(i, jp1)] == R(1)) {
                nbr(i, jp1, nx, ny, BC, &d1, &d2, &d3, &jp2);
                if (jp2 != jp1 && f_ice[IX(i, jp2)] == R(1))
                    dvdy[c] = -(R(1) * uy[IX(i, jp2)] - R(4) * uy[IX(i, jp1)] + R(3) * uy[c]) /
                              (R(2) * dy);
                else
                    dvdy[c] = (uy[IX(i, jp1)] - uy[c]) / dy;
            }
        }
    free(act_acx);
    free(act_acy);
}

/* ------------------------------------------------------------------ */
/* 2x2 Gauss quadrature inside a cell (fesm-utils gaussian_quadrature,
 * not vendored; used at velocity_ssa.f90:454-471, basal_dragging.f90:1039-1052).
 * Restated from the reference's description: node values are bilinear
 * interpolations of the four cell-corner values at xi,eta = +-1/sqrt(3);
 * corner values are means of the adjacent points of the field's grid over
 * the points flagged active (act), zero if none; all node weights equal, so
 * the weighted node mean equals the mean of the corner values
 * (thermodynamics.f90:859-861).  PARITY UNPINNED.                          */
static const double GQ_A = 0.57735026918962576451; /* 1/sqrt(3) */

static wp mean_act(const wp *v, const uint8_t *a, int n)
{
    int cnt = 0;
    wp s = R(0);
    for (int q = 0; q < n; q++)
        if (!a || a[q] || PLANT(PL_GQ_ACT_IGNORED)) { s += v[q]; cnt++; }
    return cnt > 0 ? s / (wp)cnt : R(0);
}

/* corners: 0 SW, 1 SE, 2 NE, 3 NW; nodes in the same order */
static void gq_corners_to_nodes(const wp c[4], wp n[4])
{
    static const double xs[4] = {-1, 1, 1, -1}, ys[4] = {-1, -1, 1, 1};
    const double a = PLANT(PL_GQ_NODE_POS) ? 0.5 : GQ_A;
    for (int q = 0; q < 4; q++) {
        double xi = xs[q] * a, eta = ys[q] * a;
        double w0 = 0.25 * (1 - xi) * (1 - eta), w1 = 0.25 * (1 + xi) * (1 - eta);
        double w2 = 0.25 * (1 + xi) * (1 + eta), w3 = 0.25 * (1 - xi) * (1 + eta);
        n[q] = (wp)(w0 * c[0] + w1 * c[1] + w2 * c[2] + w3 * c[3]);
    }
}

static void gq_nodes_aa(wp n[4], const wp *v, const uint8_t *act, int nx, int i, int j,
                        int im1, int ip1, int jm1, int jp1)
{
    int sw[4] = {IX(im1, jm1), IX(i, jm1), IX(im1, j), IX(i, j)};
    int se[4] = {IX(i, jm1), IX(ip1, jm1), IX(i, j), IX(ip1, j)};
    int ne[4] = {IX(i, j), IX(ip1, j), IX(i, jp1), IX(ip1, jp1)};
    int nw[4] = {IX(im1, j), IX(i, j), IX(im1, jp1), IX(i, jp1)};
    int *cs[4] = {sw, se, ne, nw};
    wp c[4];
    for (int q = 0; q < 4; q++) {
        wp vv[4];
        uint8_t aa[4];
        for (int p = 0; p < 4; p++) { vv[p] = v[cs[q][p]]; aa[p] = act ? act[cs[q][p]] : 1; }
        c[q] = mean_act(vv, aa, 4);
    }
    gq_corners_to_nodes(c, n);
}

static void gq_nodes_acx(wp n[4], const wp *v, const uint8_t *act, int nx, int i, int j,
                         int im1, int ip1, int jm1, int jp1)
{
    (void)ip1;
    int sw[2] = {IX(im1, jm1), IX(im1, j)};
    int se[2] = {IX(i, jm1), IX(i, j)};
    int ne[2] = {IX(i, j), PLANT(PL_GQ_ACX_CORNER) ? IX(i, jm1) : IX(i, jp1)};
    int nw[2] = {IX(im1, j), IX(im1, jp1)};
    int *cs[4] = {sw, se, ne, nw};
    wp c[4];
    for (int q = 0; q < 4; q++) {
        wp vv[2];
        uint8_t aa[2];
        for (int p = 0; p < 2; p++) { vv[p] = v[cs[q][p]]; aa[p] = act ? act[cs[q][p]] : 1; }
        c[q] = mean_act(vv, aa, 2);
    }
    gq_corners_to_nodes(c, n);
}

static void gq_nodes_acy(wp n[4], const wp *v, const uint8_t *act, int nx, int i, int j,
                         int im1, int ip1, int jm1, int jp1)
{
    (void)jp1;
    int sw[2] = {IX(im1, jm1), IX(i, jm1)};
    int se[2] = {IX(i, jm1), IX(ip1, jm1)};
    int ne[2] = {IX(i, j), IX(ip1, j)};
    int nw[2] = {IX(im1, j), IX(i, j)};
    int *cs[4] = {sw, se, ne, nw};
    wp c[4];
    for (int q = 0; q < 4; q++) {
        wp vv[2];
        uint8_t aa[2];
        for (int p = 0; p < 2; p++) { vv[p] = v[cs[q][p]]; aa[p] = act ? act[cs[q][p]] : 1; }
        c[q] = mean_act(vv, aa, 2);
    }
    gq_corners_to_nodes(c, n);
}

/* The four node values of cell (i, j) (0-based) of field v on the aa
 * (kind 0), acx (1) or acy (2) grid, with the active flags act (may be
 * NULL: all active), as gq_nodes_* give them to the kernels above.       */
void oracle_gq2D_nodes(wp n[4], const wp *v, const uint8_t *act, int kind, int nx, int ny,
                       int i, int j, int bnd)
{
    int im1, ip1, jm1, jp1;
    nbr(i, j, nx, ny, boundary_code(bnd), &im1, &ip1, &jm1, &jp1);
    if (kind == 0) gq_nodes_aa(n, v, act, nx, i, j, im1, ip1, jm1, jp1);
    else if (kind == 1) gq_nodes_acx(n, v, act, nx, i, j, im1, ip1, jm1, jp1);
    else gq_nodes_acy(n, v, act, nx, i, j, im1, ip1, jm1, jp1);
}

/* ------------------------------------------------------------------ */
/* calc_visc_eff_3D_nodes (velocity_ssa.f90:359-488)                    */
void oracle_visc_eff_3D_nodes(wp *visc, const wp *ux, const wp *uy, const wp *ATT,
                              const wp *H_ice, const wp *f_ice, const wp *zeta, wp dx,
                              wp dy, wp n_glen, wp eps_0, int nx, int ny, int nz, int bnd)
{
    (void)H_ice; (void)zeta;
    const wp visc_min = R(1e5);
    int BC = boundary_code(bnd);
    int nc = nx * ny;
    uint8_t *act_acx = malloc(nc), *act_acy = malloc(nc);
    calc_active_faces(act_acx, act_acy, f_ice, nx, ny, BC);
    wp *dudx = malloc(sizeof(wp) * nc), *dudy = malloc(sizeof(wp) * nc);
    wp *dvdx = malloc(sizeof(wp) * nc), *dvdy = malloc(sizeof(wp) * nc);
    wp p1 = (RS(1.0) - n_glen) / (RS(2.0) * n_glen);
    wp p2 = -RS(1.0) / n_glen;
    wp eps_0_sq = eps_0 * eps_0;
    oracle_strain_rate_2D(dudx, dudy, dvdx, dvdy, ux, uy, f_ice, dx, dy, nx, ny, bnd);
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) {
            int c = IX(i, j);
            if (f_ice[c] == R(1)) {
                int im1, ip1, jm1, jp1;
                nbr(i, j, nx, ny, BC, &im1, &ip1, &jm1, &jp1);
                wp dudxn[4], dudyn[4], dvdxn[4], dvdyn[4], eps_sq_n[4];
                gq_nodes_acx(dudxn, dudx, act_acx, nx, i, j, im1, ip1, jm1, jp1);
                gq_nodes_acx(dudyn, dudy, act_acx, nx, i, j, im1, ip1, jm1, jp1);
                gq_nodes_acy(dvdxn, dvdx, act_acy, nx, i, j, im1, ip1, jm1, jp1);
                gq_nodes_acy(dvdyn, dvdy, act_acy, nx, i, j, im1, ip1, jm1, jp1);
                for (int q = 0; q < 4; q++) {
                    wp s = dudyn[q] + dvdxn[q];
                    eps_sq_n[q] = dudxn[q] * dudxn[q] + dvdyn[q] * dvdyn[q] +
                                  dudxn[q] * dvdyn[q] +
                                  (PLANT(PL_GLEN_SHEAR) ? R(1) : RS(0.25)) * (s * s) + eps_0_sq;
                }
                for (int k = 0; k < nz; k++) {
                    wp ATTn[4];
                    gq_nodes_aa(ATTn, ATT + (size_t)k * nc, PLANT(PL_GQ_ATT_ACT) ? act_acx : NULL,
                                nx, i, j, im1, ip1, jm1, jp1);
                    wp sum = R(0);
                    for (int q = 0; q < 4; q++)
                        sum += (RS(0.5) * WPOW(eps_sq_n[q], p1) * WPOW(ATTn[q], p2)) * R(1);
                    visc[c + (size_t)k * nc] = sum / R(4);
                }
            } else {
                for (int k = 0; k < nz; k++) visc[c + (size_t)k * nc] = visc_min;
            }
        }
    free(act_acx); free(act_acy); free(dudx); free(dudy); free(dvdx); free(dvdy);
}

/* calc_visc_eff_3D_aa (velocity_ssa.f90:490-586) */
void oracle_visc_eff_3D_aa(wp *visc, const wp *ux, const wp *uy, const wp *ATT,
                           const wp *f_ice, wp dx, wp dy, wp n_glen, wp eps_0, int nx,
                           int ny, int nz, int bnd)
{
    const wp visc_min = R(1e5);
    int BC = boundary_code(bnd);
    int nc = nx * ny;
    wp p1 = (R(1) - n_glen) / (R(2) * n_glen);
    wp p2 = -R(1) / n_glen;
    wp eps_0_sq = eps_0 * eps_0;
    for (size_t q = 0; q < (size_t)nc * nz; q++) visc[q] = visc_min;
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
            int c = IX(i, j);
            if (f_ice[c] != R(1)) continue;
            int im1, ip1, jm1, jp1;
            nbr(i, j, nx, ny, BC, &im1, &ip1, &jm1, &jp1);
            wp dudx_aa = (ux[c] - ux[IX(im1, j)]) / dx;
            wp dvdy_aa = (uy[c] - uy[IX(i, jm1)]) / dy;
            wp dudy_aa_1 = (ux[IX(i, jp1)] - ux[IX(i, jm1)]) / (R(2) * dy);
            wp dudy_aa_2 = (ux[IX(im1, jp1)] - ux[IX(im1, jm1)]) / (R(2) * dy);
            wp dudy_aa = R(0.5) * (dudy_aa_1 + dudy_aa_2);
            wp dvdx_aa_1 = (uy[IX(ip1, j)] - uy[IX(im1, j)]) / (R(2) * dx);
            wp dvdx_aa_2 = (uy[IX(ip1, jm1)] - uy[IX(im1, jm1)]) / (R(2) * dx);
            wp dvdx_aa = R(0.5) * (dvdx_aa_1 + dvdx_aa_2);
            wp s = dudy_aa + dvdx_aa;
            wp eps_sq_aa = dudx_aa * dudx_aa + dvdy_aa * dvdy_aa + dudx_aa * dvdy_aa +
                           R(0.25) * (s * s) + eps_0_sq;
            for (int k = 0; k < nz; k++)
                visc[c + (size_t)k * nc] =
                    R(0.5) * WPOW(eps_sq_aa, p1) * WPOW(ATT[c + (size_t)k * nc], p2);
        }
}

/* picard_relax_visc (velocity_general.f90:2188-2213) */
static void picard_relax_visc(wp *visc, const wp *visc_prev, wp rel, size_t n)
{
    if (PLANT(PL_RELAX_VISC)) {
        for (size_t q = 0; q < n; q++) visc[q] = visc_prev[q] + rel * (visc[q] - visc_prev[q]);
        return;
    }
    for (size_t q = 0; q < n; q++)
        visc[q] = WEXP((RS(1.0) - rel) * WLOG(visc_prev[q]) + rel * WLOG(visc[q]));
}

void oracle_picard_relax_visc(wp *visc, const wp *visc_prev, wp rel, int n)
{
    picard_relax_visc(visc, visc_prev, rel, (size_t)n);
}

/* picard_relax_vel (velocity_general.f90:2160-2186) */
void oracle_picard_relax_vel(wp *ux, wp *uy, const wp *ux_prev, const wp *uy_prev, wp rel, int n)
{
    const wp r = PLANT(PL_RELAX_VEL) ? R(1) - rel : rel;
    for (int c = 0; c < n; c++) {
        ux[c] = ux_prev[c] + r * (ux[c] - ux_prev[c]);
        uy[c] = uy_prev[c] + r * (uy[c] - uy_prev[c]);
    }
}

/* ssa_vel_clip / limit_vel (solver_ssa_ac.f90:1069-1093,1264-1280): each
 * component clipped to [-u_max, u_max]; NaN compares false both ways and
 * passes through                                                           */
void oracle_ssa_vel_clip(wp *ux, wp *uy, wp u_max, int n)
{
    for (int c = 0; c < n; c++) {
        if (ux[c] > u_max) ux[c] = u_max;
        if (ux[c] < -u_max && !PLANT(PL_CLIP)) ux[c] = -u_max;
        if (uy[c] > u_max) uy[c] = u_max;
        if (uy[c] < -u_max && !PLANT(PL_CLIP)) uy[c] = -u_max;
    }
}

/* integrate_trapezoid1D_pt (yelmo_tools.f90:1092-1122) */
static wp trapz_pt(const wp *v, size_t stride, const wp *zeta, int nz)
{
    wp s = R(0);
    for (int k = 1; k < nz; k++) {
        wp mid = R(0.5) * (v[k * stride] + v[(k - 1) * stride]);
        s = s + mid * (zeta[k] - zeta[k - 1]);
    }
    return s;
}

/* calc_visc_eff_int (velocity_general.f90:2215-2261) */
void oracle_visc_eff_int(wp *visc_int, const wp *visc, const wp *H_ice, const wp *f_ice,
                         const wp *zeta, int nx, int ny, int nz)
{
    const wp visc_min = R(1e5);
    int nc = nx * ny;
    for (int c = 0; c < nc; c++) {
        if (f_ice[c] == R(1))
            visc_int[c] = trapz_pt(visc + c, nc, zeta, nz) * H_ice[c];
        else
            visc_int[c] = visc_min;
        if (visc_int[c] < visc_min) visc_int[c] = visc_min;
    }
}

/* ------------------------------------------------------------------ */
/* calc_beta_aa_power_plastic / _reg_coulomb (basal_dragging.f90:949-1196);
 * coulomb = 0 for the power law, 1 for the regularized Coulomb law.        */
static void beta_aa_law(wp *beta, const wp *ux_b, const wp *uy_b, const wp *c_bed,
                        const wp *f_ice, const wp *f_grnd, wp q, wp u_0, int nx, int ny,
                        int BC, int simple_stagger, int coulomb)
{
    const wp ub_min = R(1e-3), ub_sq_min = ub_min * ub_min;
    int nc = nx * ny;
    uint8_t *act_acx = malloc(nc), *act_acy = malloc(nc), *act_aa = malloc(nc);
    calc_active_faces(act_acx, act_acy, f_ice, nx, ny, BC);
    for (int c = 0; c < nc; c++) act_aa[c] = (f_ice[c] == R(1) && f_grnd[c] > R(0));
    for (int c = 0; c < nc; c++) beta[c] = R(0);
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
            int c = IX(i, j), im1, ip1, jm1, jp1;
            nbr(i, j, nx, ny, BC, &im1, &ip1, &jm1, &jp1);
            if (f_ice[c] == R(1)) {
                wp cbn[4], uxn[4], uyn[4];
                if (simple_stagger) {
                    for (int p = 0; p < 4; p++) {
                        cbn[p] = c_bed[c];
                        uxn[p] = RS(0.5) * (ux_b[c] + ux_b[IX(im1, j)]);
                        uyn[p] = RS(0.5) * (uy_b[c] + uy_b[IX(i, jm1)]);
                    }
                } else {
                    gq_nodes_aa(cbn, c_bed, act_aa, nx, i, j, im1, ip1, jm1, jp1);
                    gq_nodes_acx(uxn, ux_b, act_acx, nx, i, j, im1, ip1, jm1, jp1);
                    gq_nodes_acy(uyn, uy_b, act_acy, nx, i, j, im1, ip1, jm1, jp1);
                }
                wp sum = R(0);
                for (int p = 0; p < 4; p++) {
                    wp uxyn = WSQRT(uxn[p] * uxn[p] + uyn[p] * uyn[p] + ub_sq_min);
                    wp ratio = coulomb ? uxyn / (uxyn + u_0) : uxyn / u_0;
                    wp betan = cbn[p] * WPOW(ratio, q) * (PLANT(PL_BETA_Q) ? R(1) : R(1) / uxyn);
                    sum += betan * R(1);
                }
                beta[c] = sum / R(4);
            } else {
                wp uxy_b = ub_min;
                wp ratio = coulomb ? uxy_b / (uxy_b + u_0) : uxy_b / u_0;
                beta[c] = c_bed[c] * WPOW(ratio, q) * (R(1) / uxy_b);
            }
        }
    free(act_acx); free(act_acy); free(act_aa);
}

/* calc_beta (basal_dragging.f90:384-579) */
int oracle_calc_beta(wp *beta, const wp *c_bed, const wp *f_slide, const wp *ux_b,
                     const wp *uy_b, const wp *H_ice, const wp *f_ice, const wp *H_grnd,
                     const wp *f_grnd, const wp *z_bed, const wp *z_sl,
                     const oracle_ssa_params *par, int nx, int ny)
{
    int BC = boundary_code(par->boundaries);
    int nc = nx * ny;
    wp q = (wp)par->beta_q, u0 = (wp)par->beta_u0;
    if (par->beta_method == -1) return 0;
    switch (par->beta_method) {
    case 0: for (int c = 0; c < nc; c++) beta[c] = (wp)par->beta_const; break;
    case 1: beta_aa_law(beta, ux_b, uy_b, c_bed, f_ice, f_grnd, R(1), u0, nx, ny, BC, 0, 0); break;
    case 2: beta_aa_law(beta, ux_b, uy_b, c_bed, f_ice, f_grnd, q, u0, nx, ny, BC, 0, 0); break;
    case 3: beta_aa_law(beta, ux_b, uy_b, c_bed, f_ice, f_grnd, q, u0, nx, ny, BC, 0, 1); break;
    case 4: beta_aa_law(beta, ux_b, uy_b, c_bed, f_ice, f_grnd, q, u0, nx, ny, BC, 1, 0); break;
    case 5: beta_aa_law(beta, ux_b, uy_b, c_bed, f_ice, f_grnd, q, u0, nx, ny, BC, 1, 1); break;
    default: return -1;
    }
    wp q_frz = (par->beta_method == 0 || par->beta_method == 1) ? R(1) : q;
    for (int c = 0; c < nc; c++) beta[c] = beta[c] * WPOW(f_slide[c], -q_frz);
    switch (par->beta_gl_scale) {
    case 0: { /* scale_beta_gl_fraction (basal_dragging.f90:1200-1245) */
        wp f_gl = (wp)par->beta_gl_f;
        wp *tmp = malloc(sizeof(wp) * nc);
        memcpy(tmp, beta, sizeof(wp) * nc);
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                int im1, ip1, jm1, jp1;
                nbr(i, j, nx, ny, BC, &im1, &ip1, &jm1, &jp1);
                if (f_grnd[IX(i, j)] > R(0) &&
                    (f_grnd[IX(im1, j)] == R(0) || f_grnd[IX(ip1, j)] == R(0) ||
                     f_grnd[IX(i, jm1)] == R(0) || f_grnd[IX(i, jp1)] == R(0)))
                    tmp[IX(i, j)] = beta[IX(i, j)] * f_gl;
            }
        memcpy(beta, tmp, sizeof(wp) * nc);
        free(tmp);
        break;
    }
    case 1: { /* scale_beta_gl_Hgrnd (basal_dragging.f90:1247-1292) */
        wp lim = (wp)par->H_grnd_lim;
        for (int c = 0; c < nc; c++) {
            wp hg = H_grnd[c] < lim ? H_grnd[c] : lim;
            wp f = hg / lim;
            if (f < R(0)) f = R(0);
            beta[c] = beta[c] * f;
        }
        break;
    }
    case 2: { /* scale_beta_gl_zstar with norm (basal_dragging.f90:1296-1352);
               * calc_H_eff with set_frac_zero (topography.f90:851-874)     */
        const wp rho_sw_ice = PLANT(PL_ZSTAR_RHO) ? (wp)par->rho_ice / (wp)par->rho_sw
                                                  : (wp)par->rho_sw / (wp)par->rho_ice;
        for (int c = 0; c < nc; c++) {
            if (!(f_ice[c] > R(0))) continue;
            wp H_eff = H_ice[c] / f_ice[c];
            if (f_ice[c] < R(1)) H_eff = R(0);
            wp f_scale;
            if (z_bed[c] > z_sl[c]) f_scale = H_eff;
            else {
                f_scale = H_eff - (z_sl[c] - z_bed[c]) * rho_sw_ice;
                if (f_scale < R(0)) f_scale = R(0);
            }
            if (H_eff > R(0) && !PLANT(PL_ZSTAR_NORM)) f_scale = f_scale / H_eff;
            beta[c] = beta[c] * f_scale;
        }
        break;
    }
    case 3:
        for (int c = 0; c < nc; c++) beta[c] = f_grnd[c] * beta[c];
        break;
    default: return -2;
    }
    for (int c = 0; c < nc; c++)
        if (f_grnd[c] == R(0)) beta[c] = R(0);
    if (par->boundaries == OB_MISMIP3D) {
        for (int j = 0; j < ny; j++) beta[IX(0, j)] = beta[IX(1, j)];
    } else if (par->boundaries == OB_INFINITE || par->boundaries == OB_MASK) {
        for (int j = 0; j < ny; j++) { beta[IX(0, j)] = beta[IX(1, j)]; beta[IX(nx - 1, j)] = beta[IX(nx - 2, j)]; }
        for (int i = 0; i < nx; i++) { beta[IX(i, 0)] = beta[IX(i, 1)]; beta[IX(i, ny - 1)] = beta[IX(i, ny - 2)]; }
    }
    wp bmin = (wp)par->beta_min;
    for (int c = 0; c < nc; c++)
        if (beta[c] > R(0) && beta[c] < bmin) beta[c] = bmin;
    return 0;
}

/* calc_beta_gl_flux_weight (basal_dragging.f90:1760-1845): node a grounded,
 * node b floating, 100 segments with thickness and flux linear in between */
wp oracle_beta_gl_flux_weight(wp beta_a, wp u_a, wp u_b, wp H_a, wp H_b, wp f_grnd_ac)
{
    const wp q_a = u_a * H_a, q_b = u_b * H_b;
    wp uu_grnd = R(0), uu_tot = R(0);
    for (int k = 1; k <= 100; k++) {
        const wp lambda = ((wp)k - R(0.5)) / (wp)100;
        const wp H_now = H_a * lambda + H_b * (R(1) - lambda);
        const wp q_now = q_a * lambda + q_b * (R(1) - lambda);
        const wp u_now = H_now > R(0) ? q_now / H_now : R(0);
        const wp au = u_now < R(0) ? -u_now : u_now;
        uu_tot += au;
        const int grounded = PLANT(PL_FLUX_SIDE) ? lambda < f_grnd_ac : lambda > R(1) - f_grnd_ac;
        if (grounded) uu_grnd += au;
    }
    const wp weight = uu_tot > R(0) ? uu_grnd / uu_tot : R(0);
    return beta_a * weight;
}

/* stagger_beta (basal_dragging.f90:581-704) with stagger_beta_aa_mean
 * (:1360-1437), stagger_beta_aa_gl_upstream / _downstream (:1439-1576),
 * stagger_beta_aa_gl_subgrid (:1578-1658) and _subgrid_flux (:1660-1758).
 * ux/uy: the velocity the caller staggers with (ux_b for SSA, ux_bar for
 * DIVA; velocity_ssa.f90:247, velocity_diva.f90:271).                      */
int oracle_stagger_beta(wp *beta_acx, wp *beta_acy, const wp *beta, const wp *H_ice,
                        const wp *f_ice, const wp *ux, const wp *uy, const wp *f_grnd,
                        const wp *f_grnd_acx, const wp *f_grnd_acy,
                        const oracle_ssa_params *par, int nx, int ny)
{
    int BC = boundary_code(par->boundaries);
    int stag = par->beta_gl_stag;
    if (stag != -1) {
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                int im1, ip1, jm1, jp1, c = IX(i, j);
                nbr(i, j, nx, ny, BC, &im1, &ip1, &jm1, &jp1);
                if (f_grnd[c] == R(0) && f_grnd[IX(ip1, j)] == R(0)) beta_acx[c] = R(0);
                else if (f_ice[c] == R(1) && f_ice[IX(ip1, j)] < R(1)) beta_acx[c] = beta[c];
                else if (f_ice[c] < R(1) && f_ice[IX(ip1, j)] == R(1)) beta_acx[c] = beta[IX(ip1, j)];
                else beta_acx[c] = R(0.5) * (beta[c] + beta[IX(ip1, j)]);
                if (f_grnd[c] == R(0) && f_grnd[IX(i, jp1)] == R(0)) beta_acy[c] = R(0);
                else if (f_ice[c] == R(1) && f_ice[IX(i, jp1)] < R(1)) beta_acy[c] = beta[c];
                else if (f_ice[c] < R(1) && f_ice[IX(i, jp1)] == R(1)) beta_acy[c] = beta[IX(i, jp1)];
                else beta_acy[c] = R(0.5) * (beta[c] + beta[IX(i, jp1)]);
            }
        if (stag == 1 || stag == 2) {
            for (int j = 0; j < ny; j++)
                for (int i = 0; i < nx; i++) {
                    int im1, ip1, jm1, jp1, c = IX(i, j);
                    nbr(i, j, nx, ny, BC, &im1, &ip1, &jm1, &jp1);
                    const int up = (stag == 1) != PLANT(PL_GL_UPSTREAM);
                    if (f_ice[c] == R(1) && f_ice[IX(ip1, j)] == R(1)) {
                        if (f_grnd[c] > R(0) && f_grnd[IX(ip1, j)] == R(0))
                            beta_acx[c] = up ? beta[c] : beta[IX(ip1, j)];
                        else if (f_grnd[c] == R(0) && f_grnd[IX(ip1, j)] > R(0))
                            beta_acx[c] = up ? beta[IX(ip1, j)] : beta[c];
                    }
                    if (f_ice[c] == R(1) && f_ice[IX(i, jp1)] == R(1)) {
                        if (f_grnd[c] > R(0) && f_grnd[IX(i, jp1)] == R(0))
                            beta_acy[c] = up ? beta[c] : beta[IX(i, jp1)];
                        else if (f_grnd[c] == R(0) && f_grnd[IX(i, jp1)] > R(0))
                            beta_acy[c] = up ? beta[IX(i, jp1)] : beta[c];
                    }
                }
        } else if (stag == 3) {
            for (int j = 0; j < ny; j++)
                for (int i = 0; i < nx; i++) {
                    int im1, ip1, jm1, jp1, c = IX(i, j);
                    nbr(i, j, nx, ny, BC, &im1, &ip1, &jm1, &jp1);
                    const int cx = IX(ip1, j), cy = IX(i, jp1);
                    if (f_ice[c] == R(1) && f_ice[cx] == R(1)) {
                        const wp wt = PLANT(PL_SUBGRID_WT) ? f_grnd_acx[c] : f_grnd_acx[c] * f_grnd_acx[c];
                        if (f_grnd[c] > R(0) && f_grnd[cx] == R(0))
                            beta_acx[c] = wt * beta[c] + (R(1) - wt) * beta[cx];
                        else if (f_grnd[c] == R(0) && f_grnd[cx] > R(0))
                            beta_acx[c] = (R(1) - wt) * beta[c] + wt * beta[cx];
                    }
                    if (f_ice[c] == R(1) && f_ice[cy] == R(1)) {
                        const wp wt = PLANT(PL_SUBGRID_WT) ? f_grnd_acy[c] : f_grnd_acy[c] * f_grnd_acy[c];
                        if (f_grnd[c] > R(0) && f_grnd[cy] == R(0))
                            beta_acy[c] = wt * beta[c] + (R(1) - wt) * beta[cy];
                        else if (f_grnd[c] == R(0) && f_grnd[cy] > R(0))
                            beta_acy[c] = (R(1) - wt) * beta[c] + wt * beta[cy];
                    }
                }
        } else if (stag == 4) {
            for (int j = 0; j < ny; j++)
                for (int i = 0; i < nx; i++) {
                    int im1, ip1, jm1, jp1, c = IX(i, j);
                    nbr(i, j, nx, ny, BC, &im1, &ip1, &jm1, &jp1);
                    const int cx = IX(ip1, j), cy = IX(i, jp1);
                    if (f_ice[c] == R(1) && f_ice[cx] == R(1)) {
                        if (f_grnd[c] > R(0) && f_grnd[cx] == R(0))
                            beta_acx[c] = oracle_beta_gl_flux_weight(
                                beta[c], R(0.5) * (ux[IX(im1, j)] + ux[c]), R(0.5) * (ux[cx] + ux[c]),
                                H_ice[c], H_ice[cx], f_grnd_acx[c]);
                        else if (f_grnd[c] == R(0) && f_grnd[cx] > R(0))
                            beta_acx[c] = oracle_beta_gl_flux_weight(
                                beta[cx], R(0.5) * (ux[cx] + ux[c]), R(0.5) * (ux[IX(im1, j)] + ux[c]),
                                H_ice[cx], H_ice[c], f_grnd_acx[c]);
                    }
                    if (f_ice[c] == R(1) && f_ice[cy] == R(1)) {
                        if (f_grnd[c] > R(0) && f_grnd[cy] == R(0))
                            beta_acy[c] = oracle_beta_gl_flux_weight(
                                beta[c], R(0.5) * (uy[IX(i, jm1)] + uy[c]), R(0.5) * (uy[cy] + uy[c]),
                                H_ice[c], H_ice[cy], f_grnd_acy[c]);
                        else if (f_grnd[c] == R(0) && f_grnd[cy] > R(0))
                            beta_acy[c] = oracle_beta_gl_flux_weight(
                                beta[cy], R(0.5) * (uy[cy] + uy[c]), R(0.5) * (uy[IX(i, jm1)] + uy[c]),
                                H_ice[cy], H_ice[c], f_grnd_acy[c]);
                    }
                }
        } else if (stag != 0) {
            return -1;
        }
        int b = par->boundaries;
        if (b == OB_INFINITE || b == OB_MISMIP3D || b == OB_MASK) {
            for (int j = 0; j < ny; j++) {
                beta_acx[IX(0, j)] = beta_acx[IX(1, j)];
                beta_acx[IX(nx - 2, j)] = beta_acx[IX(nx - 3, j)];
                beta_acx[IX(nx - 1, j)] = beta_acx[IX(nx - 3, j)];
                beta_acy[IX(0, j)] = beta_acy[IX(1, j)];
                beta_acy[IX(nx - 1, j)] = beta_acy[IX(nx - 2, j)];
            }
        }
        if (b == OB_INFINITE || b == OB_MASK) {
            for (int i = 0; i < nx; i++) {
                beta_acx[IX(i, 0)] = beta_acx[IX(i, 1)];
                beta_acx[IX(i, ny - 1)] = beta_acx[IX(i, ny - 2)];
                beta_acy[IX(i, 0)] = beta_acy[IX(i, 1)];
                beta_acy[IX(i, ny - 2)] = beta_acy[IX(i, ny - 3)];
                beta_acy[IX(i, ny - 1)] = beta_acy[IX(i, ny - 3)];
            }
        }
    }
    int limit_beta = par->beta_method != -1 && par->beta_gl_stag != -1;
    if (limit_beta) {
        wp bmin = (wp)par->beta_min;
        for (int c = 0; c < nx * ny; c++) {
            if (beta_acx[c] > R(0) && beta_acx[c] < bmin) beta_acx[c] = bmin;
            if (beta_acy[c] > R(0) && beta_acy[c] < bmin) beta_acy[c] = bmin;
        }
    }
    return 0;
}

/* set_beta_min_grounded (basal_dragging.f90:706-780): the warning branch
 * only prints; the arithmetic is the two where-statements.                */
static void set_beta_min_grounded(wp *bx, wp *by, const int32_t *mx, const int32_t *my,
                                  wp beta_min, int nc)
{
    const int any = PLANT(PL_BETA_MIN_MASK);
    for (int c = 0; c < nc; c++) {
        if ((any ? mx[c] >= 1 : mx[c] == 1) && bx[c] == R(0)) bx[c] = beta_min;
        if ((any ? my[c] >= 1 : my[c] == 1) && by[c] == R(0)) by[c] = beta_min;
    }
}

void oracle_set_beta_min_grounded(wp *bx, wp *by, const int32_t *mx, const int32_t *my,
                                  wp beta_min, int nc)
{
    set_beta_min_grounded(bx, by, mx, my, beta_min, nc);
}

/* calc_vel_lim_drag (solver_ssa_ac.f90:1095-1199) */
static void lim_drag_face(wp *k, wp *r, wp u, wp u_cross, int mask, wp s0, wp inv_du, wp tau_c)
{
    *k = R(0);
    *r = R(0);
    if (mask < 1) return;
    wp s = WSQRT(u * u + u_cross * u_cross);
    if (s <= s0) return;
    wp x = (s - s0) * inv_du;
    wp tau = tau_c * x * x;
    wp dtauds = R(2) * tau_c * x * inv_du;
    wp b = tau / s;
    wp w = PLANT(PL_LIM_W) ? u / s : (u / s) * (u / s);
    *k = b * (R(1) - w) + dtauds * w;
    *r = PLANT(PL_LIM_R) ? *k * u : (*k - b) * u;
}

void oracle_vel_lim_drag(wp *kx, wp *ky, wp *rx, wp *ry, const wp *ux, const wp *uy,
                         const int32_t *mx, const int32_t *my, wp u_max, wp tau_c, int nx,
                         int ny, int bnd)
{
    int BC = boundary_code(bnd);
    wp s0 = R(0.8) * u_max;
    wp inv_du = R(1) / (u_max - s0);
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
            int im1, ip1, jm1, jp1, c = IX(i, j);
            nbr(i, j, nx, ny, BC, &im1, &ip1, &jm1, &jp1);
            const int jx = PLANT(PL_LIM_CROSS) ? jp1 : jm1;
            wp u_cross = R(0.25) * (uy[c] + uy[IX(ip1, j)] + uy[IX(i, jx)] + uy[IX(ip1, jx)]);
            lim_drag_face(&kx[c], &rx[c], ux[c], u_cross, mx[c], s0, inv_du, tau_c);
            u_cross = R(0.25) * (ux[c] + ux[IX(im1, j)] + ux[IX(i, jp1)] + ux[IX(im1, jp1)]);
            lim_drag_face(&ky[c], &ry[c], uy[c], u_cross, my[c], s0, inv_du, tau_c);
        }
}

/* count_vel_lim_faces (solver_ssa_ac.f90:1201-1262) */
int oracle_count_vel_lim_faces(const wp *ux, const wp *uy, const int32_t *mx,
                               const int32_t *my, wp u_max, int drag, int nx, int ny, int bnd)
{
    int BC = boundary_code(bnd), n = 0;
    wp s0 = R(0.8) * u_max;
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
            int im1, ip1, jm1, jp1, c = IX(i, j);
            nbr(i, j, nx, ny, BC, &im1, &ip1, &jm1, &jp1);
            if (mx[c] >= 1) {
                if (drag) {
                    wp uc = R(0.25) * (uy[c] + uy[IX(ip1, j)] + uy[IX(i, jm1)] + uy[IX(ip1, jm1)]);
                    if (ux[c] * ux[c] + uc * uc > s0 * s0) n++;
                } else if (WABS(ux[c]) >= u_max) n++;
            }
            if (my[c] >= 1) {
                if (drag) {
                    wp uc = R(0.25) * (ux[c] + ux[IX(im1, j)] + ux[IX(i, jp1)] + ux[IX(im1, jp1)]);
                    if (uy[c] * uy[c] + uc * uc > s0 * s0) n++;
                } else if (WABS(uy[c]) >= u_max) n++;
            }
        }
    return n;
}

/* stagger_visc_aa_ab (solver_ssa_ac.f90:1000-1067) */
void oracle_stagger_visc_aa_ab(wp *visc_ab, const wp *visc, const wp *f_ice, const wp *z_bed,
                               const wp *z_sl, int nx, int ny, int bnd)
{
    int BC = boundary_code(bnd);
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
            int im1, ip1, jm1, jp1;
            nbr(i, j, nx, ny, BC, &im1, &ip1, &jm1, &jp1);
            int cs[4] = {IX(i, j), IX(ip1, j), IX(i, jp1), IX(ip1, jp1)};
            int k = 0, any_ocn = 0;
            wp v[4];
            for (int p = 0; p < 4; p++) {
                int ice = f_ice[cs[p]] == R(1);
                int ocn = !ice && (z_bed[cs[p]] < z_sl[cs[p]]);
                k += ice;
                any_ocn |= ocn;
                v[p] = ice ? visc[cs[p]] : R(0);
            }
            if (any_ocn || k == 0) visc_ab[IX(i, j)] = R(0);
            else visc_ab[IX(i, j)] = ((v[0] + v[1]) + (v[2] + v[3])) / (wp)k;
        }
}

/* ------------------------------------------------------------------ */
/* Residual-form CSR assembly: linear_solver_matrix_ssa_ac_csr_2D
 * (solver_ssa_ac.f90:73-848).  Unknown ordering from linear_solver_init
 * (solver_linear.F90:134-143): cell n = i*ny + j (i outer, j inner,
 * 0-based), ux -> row 2n, uy -> row 2n+1.  CSR arrays are written 1-based
 * exactly as the reference hands them to linear_solver_matrix_solve.       */
typedef struct {
    int nmax;
    int32_t *a_ptr;   /* nmax+1, 1-based */
    int32_t *a_index; /* 18*cells capacity, 1-based */
    double *a_value;
    double *b_value;
    double *x_value;
} oracle_lgs;

static void bcs_of(int bnd, int per[4], int fre[4])
{
    /* sides: 0 right (x), 1 upper (y), 2 left (x), 3 lower (y) */
    const int NS = 0, FS = 1, PER = 2;
    int b[4];
    switch (bnd) {
    case OB_MISMIP3D:
    case OB_TROUGH: b[0] = FS; b[1] = PER; b[2] = NS; b[3] = PER; break;
    case OB_PERIODIC: b[0] = b[1] = b[2] = b[3] = PER; break;
    case OB_PERIODIC_X: b[0] = PER; b[1] = FS; b[2] = PER; b[3] = FS; break;
    case OB_PERIODIC_Y: b[0] = FS; b[1] = PER; b[2] = FS; b[3] = PER; break;
    case OB_INFINITE:
    case OB_MASK: b[0] = b[1] = b[2] = b[3] = FS; break;
    default: b[0] = b[1] = b[2] = b[3] = NS; break;
    }
    for (int s = 0; s < 4; s++) { per[s] = b[s] == PER; fre[s] = b[s] == FS; }
}

int oracle_assemble_ssa(oracle_lgs *lgs, const wp *ux, const wp *uy, const wp *beta_acx,
                        const wp *beta_acy, const wp *N_aa, const int32_t *mx,
                        const int32_t *my, const wp *H_ice, const wp *f_ice, const wp *z_bed,
                        const wp *z_sl, const wp *taud_acx, const wp *taud_acy,
                        const wp *taul_acx, const wp *taul_acy, wp dx, wp dy, int bnd,
                        const wp *lim_k_acx, const wp *lim_k_acy, const wp *lim_r_acx,
                        const wp *lim_r_acy, wp *N_ab_out, int nx, int ny)
{
    (void)H_ice;
    int per[4], fre[4];
    bcs_of(bnd, per, fre);
    int use_lim = lim_k_acx && lim_k_acy && lim_r_acx && lim_r_acy;
    wp inv_dx = R(1) / dx, inv_dxdx = R(1) / (dx * dx);
    wp inv_dy = R(1) / dy, inv_dydy = R(1) / (dy * dy);
    wp inv_dxdy = R(1) / (dx * dy);
    wp *N_ab = N_ab_out ? N_ab_out : malloc(sizeof(wp) * nx * ny);
    oracle_stagger_visc_aa_ab(N_ab, N_aa, f_ice, z_bed, z_sl, nx, ny, bnd);

    int32_t *ap = lgs->a_ptr, *ai = lgs->a_index;
    double *av = lgs->a_value, *bv = lgs->b_value, *xv = lgs->x_value;
    int k = 0; /* number of entries written so far */
    ap[0] = 1;
#define COLX(ii, jj) (2 * ((ii) * ny + (jj)) + 1) /* 1-based column of ux(ii,jj) */
#define COLY(ii, jj) (2 * ((ii) * ny + (jj)) + 2) /* 1-based column of uy(ii,jj) */
#define PUT(col, val) do { ai[k] = (col); av[k] = (double)(val); k++; } while (0)
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) {
            int n = i * ny + j;
            int c = IX(i, j);
            int im1 = i == 0 ? nx - 1 : i - 1, ip1 = i == nx - 1 ? 0 : i + 1;
            int jm1 = j == 0 ? ny - 1 : j - 1, jp1 = j == ny - 1 ? 0 : j + 1;
            /* ------ ux row ------ */
            int nr = 2 * n; /* 0-based row */
            if (mx[c] == 0) {
                PUT(nr + 1, 1.0); bv[nr] = 0.0; xv[nr] = 0.0;
            } else if (mx[c] == -1) {
                PUT(nr + 1, 1.0); bv[nr] = ux[c]; xv[nr] = ux[c];
            } else if (i == 0 && !per[2]) {
                if (fre[2]) { PUT(COLX(i, j), 1.0); PUT(COLX(ip1, j), -1.0); bv[nr] = 0.0; xv[nr] = ux[c]; }
                else { PUT(nr + 1, 1.0); bv[nr] = 0.0; xv[nr] = 0.0; }
            } else if (i == nx - 1 && !per[0]) {
                if (fre[0]) { PUT(COLX(i, j), 1.0); PUT(COLX(nx - 2, j), -1.0); bv[nr] = 0.0; xv[nr] = ux[c]; }
                else { PUT(nr + 1, 1.0); bv[nr] = 0.0; xv[nr] = 0.0; }
            } else if (j == 0 && !per[3]) {
                if (fre[3]) { PUT(COLX(i, j), 1.0); PUT(COLX(i, jp1), -1.0); bv[nr] = 0.0; xv[nr] = ux[c]; }
                else { PUT(nr + 1, 1.0); bv[nr] = 0.0; xv[nr] = 0.0; }
            } else if (j == ny - 1 && !per[1]) {
                if (fre[1]) { PUT(COLX(i, j), 1.0); PUT(COLX(i, ny - 2), -1.0); bv[nr] = 0.0; xv[nr] = ux[c]; }
                else { PUT(nr + 1, 1.0); bv[nr] = 0.0; xv[nr] = 0.0; }
            } else if (mx[c] == 3) {
                if (f_ice[c] == R(1) && f_ice[IX(ip1, j)] < R(1)) {
                    wp Nn = N_aa[c];
                    PUT(COLX(im1, j), -R(4) * inv_dx * Nn);
                    PUT(COLY(i, jm1), -R(2) * inv_dy * Nn);
                    PUT(COLX(i, j), R(4) * inv_dx * Nn);
                    PUT(COLY(i, j), R(2) * inv_dy * Nn);
                } else {
                    wp Nn = N_aa[IX(ip1, j)];
                    PUT(COLX(i, j), -R(4) * inv_dx * Nn);
                    PUT(COLY(ip1, jm1), -R(2) * inv_dy * Nn);
                    PUT(COLX(ip1, j), R(4) * inv_dx * Nn);
                    PUT(COLY(ip1, j), R(2) * inv_dy * Nn);
                }
                bv[nr] = taul_acx[c]; xv[nr] = ux[c];
            } else {
                wp w_face = mx[c] == 4 ? R(0.5) : R(1);
                wp beta_now = w_face * beta_acx[c], lim_r_now = R(0);
                if (use_lim) { beta_now = beta_now + w_face * lim_k_acx[c]; lim_r_now = w_face * lim_r_acx[c]; }
                PUT(COLX(i, j), -R(4) * inv_dxdx * (N_aa[IX(ip1, j)] + N_aa[c]) -
                                    R(1) * inv_dydy * (N_ab[c] + N_ab[IX(i, jm1)]) - beta_now);
                PUT(COLX(ip1, j), R(4) * inv_dxdx * N_aa[IX(ip1, j)]);
                PUT(COLX(im1, j), R(4) * inv_dxdx * N_aa[c]);
                PUT(COLX(i, jp1), R(1) * inv_dydy * N_ab[c]);
                PUT(COLX(i, jm1), R(1) * inv_dydy * N_ab[IX(i, jm1)]);
                PUT(COLY(i, j), -R(2) * inv_dxdy * N_aa[c] - R(1) * inv_dxdy * N_ab[c]);
                PUT(COLY(ip1, j), R(2) * inv_dxdy * N_aa[IX(ip1, j)] + R(1) * inv_dxdy * N_ab[c]);
                PUT(COLY(ip1, jm1), -R(2) * inv_dxdy * N_aa[IX(ip1, j)] - R(1) * inv_dxdy * N_ab[IX(i, jm1)]);
                PUT(COLY(i, jm1), R(2) * inv_dxdy * N_aa[c] + R(1) * inv_dxdy * N_ab[IX(i, jm1)]);
                bv[nr] = taud_acx[c] - lim_r_now; xv[nr] = ux[c];
            }
            ap[nr + 1] = k + 1;
            /* ------ uy row ------ */
            nr = 2 * n + 1;
            if (my[c] == 0) {
                PUT(nr + 1, 1.0); bv[nr] = 0.0; xv[nr] = 0.0;
            } else if (my[c] == -1) {
                PUT(nr + 1, 1.0); bv[nr] = uy[c]; xv[nr] = uy[c];
            } else if (j == 0 && !per[3]) {
                if (fre[3]) { PUT(COLY(i, j), 1.0); PUT(COLY(i, jp1), -1.0); bv[nr] = 0.0; xv[nr] = uy[c]; }
                else { PUT(nr + 1, 1.0); bv[nr] = 0.0; xv[nr] = 0.0; }
            } else if (j == ny - 1 && !per[1]) {
                if (fre[1]) { PUT(COLY(i, j), 1.0); PUT(COLY(i, ny - 2), -1.0); bv[nr] = 0.0; xv[nr] = uy[c]; }
                else { PUT(nr + 1, 1.0); bv[nr] = 0.0; xv[nr] = 0.0; }
            } else if (i == 0 && !per[2]) {
                if (fre[2]) { PUT(COLY(i, j), 1.0); PUT(COLY(ip1, j), -1.0); bv[nr] = 0.0; xv[nr] = uy[c]; }
                else { PUT(nr + 1, 1.0); bv[nr] = 0.0; xv[nr] = 0.0; }
            } else if (i == nx - 1 && !per[0]) {
                if (fre[0]) { PUT(COLY(i, j), 1.0); PUT(COLY(nx - 2, j), -1.0); bv[nr] = 0.0; xv[nr] = uy[c]; }
                else { PUT(nr + 1, 1.0); bv[nr] = 0.0; xv[nr] = 0.0; }
            } else if (my[c] == 3) {
                if (f_ice[c] == R(1) && f_ice[IX(i, jp1)] < R(1)) {
                    wp Nn = N_aa[c];
                    PUT(COLX(im1, j), -R(2) * inv_dx * Nn);
                    PUT(COLY(i, jm1), -R(4) * inv_dy * Nn);
                    PUT(COLX(i, j), R(2) * inv_dx * Nn);
                    PUT(COLY(i, j), R(4) * inv_dy * Nn);
                } else {
                    wp Nn = N_aa[IX(i, jp1)];
                    PUT(COLX(im1, jp1), -R(2) * inv_dx * Nn);
                    PUT(COLY(i, j), -R(4) * inv_dy * Nn);
                    PUT(COLX(i, jp1), R(2) * inv_dx * Nn);
                    PUT(COLY(i, jp1), R(4) * inv_dy * Nn);
                }
                bv[nr] = taul_acy[c]; xv[nr] = uy[c];
            } else {
                wp w_face = my[c] == 4 ? R(0.5) : R(1);
                wp beta_now = w_face * beta_acy[c], lim_r_now = R(0);
                if (use_lim) { beta_now = beta_now + w_face * lim_k_acy[c]; lim_r_now = w_face * lim_r_acy[c]; }
                PUT(COLY(i, j), -R(4) * inv_dydy * (N_aa[IX(i, jp1)] + N_aa[c]) -
                                    R(1) * inv_dxdx * (N_ab[c] + N_ab[IX(im1, j)]) - beta_now);
                PUT(COLY(i, jp1), R(4) * inv_dydy * N_aa[IX(i, jp1)]);
                PUT(COLY(i, jm1), R(4) * inv_dydy * N_aa[c]);
                PUT(COLY(ip1, j), R(1) * inv_dxdx * N_ab[c]);
                PUT(COLY(im1, j), R(1) * inv_dxdx * N_ab[IX(im1, j)]);
                PUT(COLX(i, j), -R(2) * inv_dxdy * N_aa[c] - R(1) * inv_dxdy * N_ab[c]);
                PUT(COLX(i, jp1), R(2) * inv_dxdy * N_aa[IX(i, jp1)] + R(1) * inv_dxdy * N_ab[c]);
                PUT(COLX(im1, jp1), -R(2) * inv_dxdy * N_aa[IX(i, jp1)] - R(1) * inv_dxdy * N_ab[IX(im1, j)]);
                PUT(COLX(im1, j), R(2) * inv_dxdy * N_aa[c] + R(1) * inv_dxdy * N_ab[IX(im1, j)]);
                bv[nr] = taud_acy[c] - lim_r_now; xv[nr] = uy[c];
            }
            ap[nr + 1] = k + 1;
        }
#undef PUT
#undef COLX
#undef COLY
    if (!N_ab_out) free(N_ab);
    return k;
}

/* the energy formulation (solver_ssa_ac_energy.f90) */
#include "energy_oracle.inc"

/* The assembler par->ssa_solver selects (velocity_ssa.f90:270-283,
 * velocity_diva.f90:323-336); copy_from is -1 everywhere for the residual
 * form.  Returns the entry count or a negative error.                      */
static int assemble_selected(const oracle_ssa_params *par, oracle_lgs *L, int cap,
                             int32_t *copy_from, const wp *ux, const wp *uy, const wp *bx,
                             const wp *by, const wp *N_aa, const int32_t *mx, const int32_t *my,
                             const wp *H_ice, const wp *f_ice, const wp *z_bed, const wp *z_sl,
                             const wp *taud_acx, const wp *taud_acy, const wp *taul_acx,
                             const wp *taul_acy, wp dx, wp dy, const wp *lkx, const wp *lky,
                             const wp *lrx, const wp *lry, int nx, int ny)
{
    if (par->ssa_solver == 1)
        return oracle_assemble_ssa_energy(L, cap, copy_from, ux, uy, bx, by, N_aa, mx, my, H_ice,
                                          f_ice, z_bed, z_sl, taud_acx, taud_acy, taul_acx,
                                          taul_acy, dx, dy, par->boundaries, lkx, lky, lrx, lry,
                                          nx, ny);
    for (int q = 0; q < 2 * nx * ny; q++) copy_from[q] = -1;
    return oracle_assemble_ssa(L, ux, uy, bx, by, N_aa, mx, my, H_ice, f_ice, z_bed, z_sl,
                               taud_acx, taud_acy, taul_acx, taul_acy, dx, dy, par->boundaries,
                               lkx, lky, lrx, lry, NULL, nx, ny);
}

/* linear_solver_save_velocity (solver_ssa_ac.f90:33-71) without the
 * underflow clip: rows tied to another row take that row's solution        */
static void save_velocity(wp *ux, wp *uy, const oracle_lgs *L, const int32_t *copy_from, int nx,
                          int ny)
{
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) {
            const int n = i * ny + j;
            int r = 2 * n;
            if (copy_from[r] >= 0) r = copy_from[r];
            ux[IX(i, j)] = (wp)L->x_value[r];
            r = 2 * n + 1;
            if (copy_from[r] >= 0) r = copy_from[r];
            uy[IX(i, j)] = (wp)L->x_value[r];
        }
}

/* CSR capacity of an assembly: 18 per cell (solver_linear.F90:85), or the
 * energy form's row limit                                                   */
static int csr_capacity(const oracle_ssa_params *par, int nc)
{
    return par->ssa_solver == 1 ? EN_NNZ_ROW_MAX * 2 * nc : 18 * nc;
}

/* ------------------------------------------------------------------ */
/* FP64 Krylov solve behind linear_solver_matrix_solve
 * (solver_linear.F90:151-400).  LIS is not vendored: BiCGSTAB / CG with
 * ILU(0) or Jacobi are restated from their published algorithms, with
 * LIS's defaults (conv_cond 0: ||r_k||/||r_0||; BiCGSTAB early exit on s;
 * right preconditioning).  PARITY UNPINNED against LIS itself.            */
typedef struct {
    int n;
    const int32_t *ptr; /* 0-based */
    const int32_t *col;
    const double *val;
} csr0;

static void spmv(const csr0 *A, const double *x, double *y)
{
    for (int r = 0; r < A->n; r++) {
        double s = 0.0;
        for (int k = A->ptr[r]; k < A->ptr[r + 1]; k++) s += A->val[k] * x[A->col[k]];
        y[r] = s;
    }
}

static double dot(const double *a, const double *b, int n)
{
    double s = 0.0;
    for (int q = 0; q < n; q++) s += a[q] * b[q];
    return s;
}

/* ILU(0) on the stored pattern, rows split into nb contiguous blocks with
 * couplings between blocks dropped (nb = 1: the sequential factorization).
 * Entries of each row are processed in ascending column order (IKJ).       */
typedef struct {
    int n;
    int32_t *ptr, *col; /* sorted by column within each row */
    double *val;        /* L (unit diag, strictly lower) and U (with diag)   */
    int32_t *diag;      /* position of the diagonal in each row              */
    double *dinv;
} ilu0;

static int cmp_int2(const void *a, const void *b)
{
    const int32_t *x = a, *y = b;
    return x[0] < y[0] ? -1 : x[0] > y[0];
}

static int ilu0_build(ilu0 *F, const csr0 *A, int nb, int nowrap_ny)
{
    int n = A->n;
    F->n = n;
    F->ptr = malloc(sizeof(int32_t) * (n + 1));
    int nnz = A->ptr[n];
    F->col = malloc(sizeof(int32_t) * nnz);
    F->val = malloc(sizeof(double) * nnz);
    F->diag = malloc(sizeof(int32_t) * n);
    F->dinv = malloc(sizeof(double) * n);
    int32_t *tmp = malloc(sizeof(int32_t) * 2 * 64);
    int *pos = malloc(sizeof(int) * n);
    for (int q = 0; q < n; q++) pos[q] = -1;
    int m = 0;
    F->ptr[0] = 0;
    for (int r = 0; r < n; r++) {
        int lo = A->ptr[r], hi = A->ptr[r + 1];
        int b0 = (int)((long long)n * (long long)((long long)r * nb / n) / nb);
        /* block bounds for row r */
        int blk = (int)((long long)r * nb / n);
        int bs = (int)((long long)n * blk / nb), be = (int)((long long)n * (blk + 1) / nb);
        (void)b0;
        int cnt = 0;
        if (hi - lo > 64) { free(tmp); free(pos); return -1; }
        for (int k = lo; k < hi; k++) {
            int c = A->col[k];
            if (c < bs || c >= be) continue;
            if (nowrap_ny > 0) {
                /* a coupling more than one line or column away wraps */
                const int cr = r / 2, cc = c / 2;
                const int di = cc / nowrap_ny - cr / nowrap_ny, dj = cc % nowrap_ny - cr % nowrap_ny;
                if (di > 1 || di < -1 || dj > 1 || dj < -1) continue;
            }
            tmp[2 * cnt] = c;
            tmp[2 * cnt + 1] = k;
            cnt++;
        }
        qsort(tmp, cnt, sizeof(int32_t) * 2, cmp_int2);
        F->diag[r] = -1;
        for (int q = 0; q < cnt; q++) {
            if (q > 0 && tmp[2 * q] == tmp[2 * q - 2]) {
                /* duplicate column: summed into the stored entry */
                F->val[m - 1] += A->val[tmp[2 * q + 1]];
                continue;
            }
            F->col[m] = tmp[2 * q];
            F->val[m] = A->val[tmp[2 * q + 1]];
            if (tmp[2 * q] == r) F->diag[r] = m;
            m++;
        }
        F->ptr[r + 1] = m;
        if (F->diag[r] < 0) { free(tmp); free(pos); return -2; }
    }
    /* numeric factorization, IKJ */
    for (int r = 0; r < n; r++) {
        for (int k = F->ptr[r]; k < F->ptr[r + 1]; k++) pos[F->col[k]] = k;
        for (int k = F->ptr[r]; k < F->diag[r]; k++) {
            int kr = F->col[k];
            double lik = F->val[k] * F->dinv[kr];
            F->val[k] = lik;
            for (int q = F->diag[kr] + 1; q < F->ptr[kr + 1]; q++) {
                int p = pos[F->col[q]];
                if (p >= 0 && p > k) F->val[p] -= lik * F->val[q];
            }
        }
        double d = F->val[F->diag[r]];
        F->dinv[r] = 1.0 / d;
        for (int k = F->ptr[r]; k < F->ptr[r + 1]; k++) pos[F->col[k]] = -1;
    }
    free(tmp);
    free(pos);
    return 0;
}

static void ilu0_free(ilu0 *F)
{
    free(F->ptr); free(F->col); free(F->val); free(F->diag); free(F->dinv);
}

/* z = (LU)^{-1} v */
static void ilu0_apply(const ilu0 *F, const double *v, double *z)
{
    int n = F->n;
    for (int r = 0; r < n; r++) {
        double s = v[r];
        for (int k = F->ptr[r]; k < F->diag[r]; k++) s -= F->val[k] * z[F->col[k]];
        z[r] = s;
    }
    for (int r = n - 1; r >= 0; r--) {
        double s = z[r];
        for (int k = F->diag[r] + 1; k < F->ptr[r + 1]; k++) s -= F->val[k] * z[F->col[k]];
        z[r] = s * F->dinv[r];
    }
}

typedef struct {
    int kind; /* 0 none, 1 jacobi, 2 ilu */
    double *dinv;
    ilu0 F;
} precon;

static void precon_apply(const precon *P, const double *v, double *z, int n)
{
    if (P->kind == 2) ilu0_apply(&P->F, v, z);
    else if (P->kind == 1) for (int q = 0; q < n; q++) z[q] = v[q] * P->dinv[q];
    else memcpy(z, v, sizeof(double) * n);
}

/* returns LIS status: 0 converged, 2 breakdown, 4 maxiter */
static int bicgstab(const csr0 *A, const double *b, double *x, const precon *P,
                    const lis_opts *o, int *iters, double *resid)
{
    int n = A->n;
    double *r = calloc(n, 8), *rt = calloc(n, 8), *p = calloc(n, 8), *v = calloc(n, 8);
    double *ph = calloc(n, 8), *s = calloc(n, 8), *sh = calloc(n, 8), *t = calloc(n, 8);
    int status = 4;
    *iters = 0;
    if (o->initx_zeros) memset(x, 0, sizeof(double) * n);
    spmv(A, x, t);
    for (int q = 0; q < n; q++) r[q] = b[q] - t[q];
    double nrm0 = o->conv_cond == 1 ? sqrt(dot(b, b, n)) : sqrt(dot(r, r, n));
    double rn = sqrt(dot(r, r, n));
    if (nrm0 == 0.0) { *resid = 0.0; status = 0; goto done; }
    double bnrm = 1.0 / nrm0;
    *resid = rn * bnrm;
    if (*resid <= o->tol) { status = 0; goto done; }
    memcpy(rt, r, sizeof(double) * n);
    double rho_old = 1.0, alpha = 1.0, omega = 1.0;
    for (int it = 1; it <= o->maxiter; it++) {
        double rho = dot(rt, r, n);
        if (rho == 0.0) { status = 2; *iters = it; goto done; }
        if (it == 1) memcpy(p, r, sizeof(double) * n);
        else {
            double beta = (rho / rho_old) * (alpha / omega);
            for (int q = 0; q < n; q++) p[q] = r[q] + beta * (p[q] - omega * v[q]);
        }
        precon_apply(P, p, ph, n);
        spmv(A, ph, v);
        double tmp = dot(rt, v, n);
        if (tmp == 0.0) { status = 2; *iters = it; goto done; }
        alpha = rho / tmp;
        for (int q = 0; q < n; q++) s[q] = r[q] - alpha * v[q];
        double sn = sqrt(dot(s, s, n)) * bnrm;
        if (sn <= o->tol) {
            for (int q = 0; q < n; q++) x[q] += alpha * ph[q];
            *resid = sn; *iters = it; status = 0; goto done;
        }
        precon_apply(P, s, sh, n);
        spmv(A, sh, t);
        double tt = dot(t, t, n);
        if (tt == 0.0) { status = 2; *iters = it; goto done; }
        omega = dot(t, s, n) / tt;
        if (omega == 0.0) { status = 2; *iters = it; goto done; }
        for (int q = 0; q < n; q++) {
            x[q] += alpha * ph[q] + omega * sh[q];
            r[q] = s[q] - omega * t[q];
        }
        *resid = sqrt(dot(r, r, n)) * bnrm;
        *iters = it;
        if (*resid <= o->tol) { status = 0; goto done; }
        rho_old = rho;
    }
    status = 4;
done:
    free(r); free(rt); free(p); free(v); free(ph); free(s); free(sh); free(t);
    return status;
}

static int cg(const csr0 *A, const double *b, double *x, const precon *P, const lis_opts *o,
              int *iters, double *resid)
{
    int n = A->n;
    double *r = calloc(n, 8), *z = calloc(n, 8), *p = calloc(n, 8), *q = calloc(n, 8);
    int status = 4;
    *iters = 0;
    if (o->initx_zeros) memset(x, 0, sizeof(double) * n);
    spmv(A, x, q);
    for (int k = 0; k < n; k++) r[k] = b[k] - q[k];
    double nrm0 = o->conv_cond == 1 ? sqrt(dot(b, b, n)) : sqrt(dot(r, r, n));
    if (nrm0 == 0.0) { *resid = 0.0; status = 0; goto done; }
    double bnrm = 1.0 / nrm0;
    *resid = sqrt(dot(r, r, n)) * bnrm;
    if (*resid <= o->tol) { status = 0; goto done; }
    double rho_old = 0.0;
    for (int it = 1; it <= o->maxiter; it++) {
        precon_apply(P, r, z, n);
        double rho = dot(r, z, n);
        if (rho == 0.0) { status = 2; *iters = it; goto done; }
        if (it == 1) memcpy(p, z, sizeof(double) * n);
        else {
            double beta = rho / rho_old;
            for (int k = 0; k < n; k++) p[k] = z[k] + beta * p[k];
        }
        spmv(A, p, q);
        double pq = dot(p, q, n);
        if (pq == 0.0) { status = 2; *iters = it; goto done; }
        double alpha = rho / pq;
        for (int k = 0; k < n; k++) { x[k] += alpha * p[k]; r[k] -= alpha * q[k]; }
        *resid = sqrt(dot(r, r, n)) * bnrm;
        *iters = it;
        if (*resid <= o->tol) { status = 0; goto done; }
        rho_old = rho;
    }
    status = 4;
done:
    free(r); free(z); free(p); free(q);
    return status;
}

/* linear_solver_matrix_solve (solver_linear.F90:151-190) + the LIS call
 * sequence (:194-400).  CSR arrays are 1-based as in lgs.                 */
int oracle_csr_solve(int n, const int32_t *a_ptr1, const int32_t *a_idx1, const double *a_val,
                     const double *b, double *x, const char *opts, int32_t *iters,
                     int32_t *status, double *rel_resid, double *seconds)
{
    lis_opts o;
    if (parse_lis_opts(opts, &o) != 0) return -10;
    int allzero = 1;
    for (int q = 0; q < n; q++) if (b[q] != 0.0) { allzero = 0; break; }
    if (allzero) {
        memset(x, 0, sizeof(double) * n);
        *iters = 0; *status = 0; *rel_resid = 0.0; *seconds = 0.0;
        return 0;
    }
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    int nnz = a_ptr1[n] - 1;
    int32_t *ptr = malloc(sizeof(int32_t) * (n + 1));
    int32_t *col = malloc(sizeof(int32_t) * (nnz > 0 ? nnz : 1));
    for (int q = 0; q <= n; q++) ptr[q] = a_ptr1[q] - 1;
    for (int q = 0; q < nnz; q++) col[q] = a_idx1[q] - 1;
    csr0 A = {n, ptr, col, a_val};
    precon P;
    memset(&P, 0, sizeof(P));
    P.kind = o.precon;
    int rc = 0;
    if (P.kind == 1) {
        P.dinv = malloc(sizeof(double) * n);
        for (int r = 0; r < n; r++) {
            double d = 0.0;
            for (int k = ptr[r]; k < ptr[r + 1]; k++) if (col[k] == r) d += a_val[k];
            P.dinv[r] = 1.0 / d;
        }
    } else if (P.kind == 2) {
        if (ilu0_build(&P.F, &A, o.ilu_blocks, o.ilu_nowrap_ny) != 0) { rc = -11; goto out; }
    }
    int it = 0;
    double res = 0.0;
    int st = (o.solver == 1) ? cg(&A, b, x, &P, &o, &it, &res) : bicgstab(&A, b, x, &P, &o, &it, &res);
    *iters = it; *status = st; *rel_resid = res;
out:
    if (P.kind == 1) free(P.dinv);
    if (P.kind == 2 && rc == 0) ilu0_free(&P.F);
    free(ptr); free(col);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    *seconds = (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec);
    return rc;
}

/* Exposed for the sweep-level test of the GPU triangular solves: applies
 * (LU)^{-1} of the ILU(0) factorization of a 1-based CSR matrix.           */
int oracle_ilu0_apply(int n, const int32_t *a_ptr1, const int32_t *a_idx1, const double *a_val,
                      int nb, const double *v, double *z)
{
    int nnz = a_ptr1[n] - 1;
    int32_t *ptr = malloc(sizeof(int32_t) * (n + 1));
    int32_t *col = malloc(sizeof(int32_t) * nnz);
    for (int q = 0; q <= n; q++) ptr[q] = a_ptr1[q] - 1;
    for (int q = 0; q < nnz; q++) col[q] = a_idx1[q] - 1;
    csr0 A = {n, ptr, col, a_val};
    ilu0 F;
    int rc = ilu0_build(&F, &A, nb, 0);
    if (rc == 0) { ilu0_apply(&F, v, z); ilu0_free(&F); }
    free(ptr); free(col);
    return rc;
}

/* The ILU(0) factors of a 1-based CSR matrix as a sorted 0-based CSR with
 * unit-lower L and U (with diagonal) in one array; for the pattern-identity
 * test (L U)_ij = A_ij on pattern(A).                                      */
int oracle_ilu0_factors(int n, const int32_t *a_ptr1, const int32_t *a_idx1, const double *a_val,
                        int nb, int32_t *f_ptr, int32_t *f_col, double *f_val)
{
    int nnz = a_ptr1[n] - 1;
    int32_t *ptr = malloc(sizeof(int32_t) * (n + 1));
    int32_t *col = malloc(sizeof(int32_t) * nnz);
    for (int q = 0; q <= n; q++) ptr[q] = a_ptr1[q] - 1;
    for (int q = 0; q < nnz; q++) col[q] = a_idx1[q] - 1;
    csr0 A = {n, ptr, col, a_val};
    ilu0 F;
    int rc = ilu0_build(&F, &A, nb, 0);
    if (rc == 0) {
        memcpy(f_ptr, F.ptr, sizeof(int32_t) * (n + 1));
        memcpy(f_col, F.col, sizeof(int32_t) * F.ptr[n]);
        memcpy(f_val, F.val, sizeof(double) * F.ptr[n]);
        ilu0_free(&F);
    }
    free(ptr); free(col);
    return rc;
}

/* ------------------------------------------------------------------ */
/* set_ssa_masks + front_face_mask (solver_ssa_ac.f90:851-996)          */
enum { FRNT_ICE_FREE = -1, FRNT_ICE_FREE_LAND = -2, FRNT_NONE = 0, FRNT_FLOAT = 1,
       FRNT_MARINE = 2, FRNT_GRND = 3 };
/* lateral_bc: 0 none, 1 floating, 2 marine, 3 all */
static int front_face_mask(int code_a, int code_b, int lateral_bc)
{
    int code_ice, code_free;
    if (code_a > 0 && code_b < 0) { code_ice = code_a; code_free = code_b; }
    else if (code_a < 0 && code_b > 0) { code_ice = code_b; code_free = code_a; }
    else return -1;
    if (code_free == FRNT_ICE_FREE_LAND) {
        if (code_ice == FRNT_FLOAT) return 0;
        code_ice = FRNT_GRND;
    }
    switch (lateral_bc) {
    case 0: return 4;
    case 1: return code_ice == FRNT_FLOAT ? 3 : 4;
    case 2: return (code_ice == FRNT_FLOAT || code_ice == FRNT_MARINE) ? 3 : 4;
    case 3: return 3;
    }
    return -1;
}

void oracle_set_ssa_masks(int32_t *mx, int32_t *my, const int32_t *mask_frnt, const wp *f_ice,
                          const wp *f_grnd, int lateral_bc, int nx, int ny, int bnd)
{
    int BC = boundary_code(bnd);
    for (int c = 0; c < nx * ny; c++) { mx[c] = 0; my[c] = 0; }
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
            int im1, ip1, jm1, jp1, c = IX(i, j);
            nbr(i, j, nx, ny, BC, &im1, &ip1, &jm1, &jp1);
            if (f_ice[c] == R(1) || f_ice[IX(ip1, j)] == R(1))
                mx[c] = (f_grnd[c] > R(0) || f_grnd[IX(ip1, j)] > R(0)) ? 1 : 2;
            int ml = front_face_mask(mask_frnt[c], mask_frnt[IX(ip1, j)], lateral_bc);
            if (ml >= 0) mx[c] = ml;
            if (f_ice[c] == R(1) || f_ice[IX(i, jp1)] == R(1))
                my[c] = (f_grnd[c] > RS(0.0) || f_grnd[IX(i, jp1)] > RS(0.0)) ? 1 : 2;
            ml = front_face_mask(mask_frnt[c], mask_frnt[IX(i, jp1)], lateral_bc);
            if (ml >= 0) my[c] = ml;
        }
}

/* set_inactive_margins without a prescribed front
 * (velocity_general.f90:1687-1750)                                        */
static void set_inactive_margins(wp *ux, wp *uy, const wp *f_ice, int nx, int ny, int BC)
{
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
            int im1, ip1, jm1, jp1, c = IX(i, j);
            nbr(i, j, nx, ny, BC, &im1, &ip1, &jm1, &jp1);
            wp fa = f_ice[c], fb = f_ice[IX(ip1, j)];
            if ((fb == R(0) && fa < R(1)) || (fa == R(0) && fb < R(1))) ux[c] = R(0);
            fb = f_ice[IX(i, jp1)];
            if ((fb == R(0) && fa < R(1)) || (fa == R(0) && fb < R(1))) uy[c] = R(0);
        }
}

/* picard_calc_convergence_l2 (velocity_general.f90:1950-2158); returns the
 * residual, sets *conv; -1 in *conv for a non-finite solution.            */
static wp picard_conv_l2(int *conv, const wp *ux, const wp *uy, const wp *uxp, const wp *uyp,
                         const int32_t *mx, const int32_t *my, wp tol, int iter, int iter_max,
                         int nx, int ny, int *ncheck)
{
    const double du_reg = 1e-10;
    const wp vel_tol = RS(1e-5);
    const wp tol_uf = R(1e-15);
    double res1 = 0, res2 = 0, res3 = 0;
    int nchk = 0;
    double *r1 = calloc(ny, 8), *r2 = calloc(ny, 8), *r3 = calloc(ny, 8);
    for (int j = 0; j < ny; j++) {
        double a = 0, b = 0, e = 0;
        for (int i = 0; i < nx; i++) {
            int c = IX(i, j);
            const wp vt = PLANT(PL_CONV_VEL_TOL) ? R(-1) : vel_tol;
            const int m0 = PLANT(PL_CONV_MASK) ? -1 : 0;
            if (WABS(ux[c]) > vt && mx[c] > m0) {
                nchk++;
                double t = (double)(ux[c] - uxp[c]);
                if (fabs(t) < (double)tol_uf) t = 0.0;
                a += t * t;
                t = (double)uxp[c];
                if (fabs(t) < (double)tol_uf) t = 0.0;
                b += t * t;
                t = (double)(ux[c] + uxp[c]);
                e += t * t;
            }
            if (WABS(uy[c]) > vt && my[c] > m0) {
                nchk++;
                double t = (double)(uy[c] - uyp[c]);
                if (fabs(t) < (double)tol_uf) t = 0.0;
                a += t * t;
                t = (double)uyp[c];
                if (fabs(t) < (double)tol_uf) t = 0.0;
                b += t * t;
                t = (double)(uy[c] + uyp[c]);
                e += t * t;
            }
        }
        r1[j] = a;
        r2[j] = b;
        r3[j] = e;
    }
    for (int j = 0; j < ny; j++) { res1 += r1[j]; res2 += r2[j]; res3 += r3[j]; }
    free(r1); free(r2); free(r3);
    /* norm_method = 1 (:1976): ||du|| / (||u_prev|| + du_reg) */
    const double reg = PLANT(PL_CONV_DU_REG) ? 0.0 : du_reg;
    wp resid = nchk > 0 ? (wp)(PLANT(PL_CONV_NORM2) ? 2.0 * sqrt(res1) / (sqrt(res3) + reg)
                                                    : sqrt(res1) / (sqrt(res2) + reg))
                        : R(0);
    *ncheck = nchk;
    if (resid < tol) *conv = 1;
    else if (iter == iter_max) *conv = 2;
    else *conv = 0;
    int bad = !isfinite((double)resid);
    for (int c = 0; c < nx * ny && !bad; c++)
        if ((mx[c] > 0 && !isfinite((double)ux[c])) || (my[c] > 0 && !isfinite((double)uy[c]))) bad = 1;
    if (bad) *conv = -1;
    return resid;
}

int oracle_picard_conv_l2(int *conv, wp *resid, const wp *ux, const wp *uy, const wp *uxp,
                          const wp *uyp, const int32_t *mx, const int32_t *my, wp tol, int iter,
                          int iter_max, int nx, int ny)
{
    int nchk = 0;
    *resid = picard_conv_l2(conv, ux, uy, uxp, uyp, mx, my, tol, iter, iter_max, nx, ny, &nchk);
    return nchk;
}

/* picard_calc_convergence_l1rel_matrix (velocity_general.f90:1907-1948) */
static void picard_l1rel(wp *ex, wp *ey, const wp *ux, const wp *uy, const wp *uxp,
                         const wp *uyp, int nc)
{
    const wp vtol = RS(1e-2), tol = PLANT(PL_L1REL_TOL) ? R(0) : RS(1e-5);
    for (int c = 0; c < nc; c++) {
        ex[c] = WABS(ux[c]) > vtol ? R(2) * WABS(ux[c] - uxp[c]) / WABS(ux[c] + uxp[c] + tol) : R(0);
        ey[c] = WABS(uy[c]) > vtol ? R(2) * WABS(uy[c] - uyp[c]) / WABS(uy[c] + uyp[c] + tol) : R(0);
    }
}

void oracle_picard_l1rel(wp *ex, wp *ey, const wp *ux, const wp *uy, const wp *uxp,
                         const wp *uyp, int nc)
{
    picard_l1rel(ex, ey, ux, uy, uxp, uyp, nc);
}

/* ------------------------------------------------------------------ */
/* calc_velocity_ssa (velocity_ssa.f90:60-357).  Argument order follows the
 * reference; scalar outputs go to out_i[4] = {iter_now, lin_iter, lin_fail,
 * lim_n}; resid_hist (length ssa_iter_max) gets the Picard residual of
 * each iteration and lin_hist the Krylov count of each.  Returns 0, or <0
 * for a fatal condition (non-finite solution: the reference's error stop). */
int oracle_calc_velocity_ssa(
    wp *ux_b, wp *uy_b, wp *taub_acx, wp *taub_acy, wp *visc_eff, wp *visc_eff_int,
    int32_t *ssa_mask_acx, int32_t *ssa_mask_acy, wp *ssa_err_acx, wp *ssa_err_acy,
    int32_t *out_i, wp *beta, wp *beta_acx, wp *beta_acy, const wp *c_bed, const wp *f_slide,
    const wp *taud_acx, const wp *taud_acy, const wp *taul_int_acx, const wp *taul_int_acy,
    const wp *H_ice, const wp *f_ice, const wp *H_grnd, const wp *f_grnd, const wp *f_grnd_acx,
    const wp *f_grnd_acy, const wp *ATT, const wp *zeta_aa, const wp *z_sl, const wp *z_bed,
    const wp *z_srf, wp dx, wp dy, wp n_glen, const oracle_ssa_params *par, const char *lis_opt,
    int nx, int ny, int nz, double *resid_hist, int32_t *lin_hist, double *solve_seconds)
{
    (void)z_srf;
    int nc = nx * ny;
    int BC = boundary_code(par->boundaries);
    size_t n3 = (size_t)nc * nz;
    wp *visc_nm1 = malloc(sizeof(wp) * n3);
    wp *ux_nm1 = malloc(sizeof(wp) * nc), *uy_nm1 = malloc(sizeof(wp) * nc);
    wp *bsx = malloc(sizeof(wp) * nc), *bsy = malloc(sizeof(wp) * nc);
    wp *lkx = calloc(nc, sizeof(wp)), *lky = calloc(nc, sizeof(wp));
    wp *lrx = calloc(nc, sizeof(wp)), *lry = calloc(nc, sizeof(wp));
    oracle_lgs L;
    L.nmax = 2 * nc;
    L.a_ptr = calloc(L.nmax + 1, sizeof(int32_t));
    const int cap = csr_capacity(par, nc);
    L.a_index = calloc((size_t)cap, sizeof(int32_t));
    L.a_value = calloc((size_t)cap, sizeof(double));
    L.b_value = calloc(L.nmax, sizeof(double));
    L.x_value = calloc(L.nmax, sizeof(double));
    int32_t *cfrom = malloc(sizeof(int32_t) * L.nmax);
    int rc = 0;
    memcpy(bsx, beta_acx, sizeof(wp) * nc);
    memcpy(bsy, beta_acy, sizeof(wp) * nc);
    for (int c = 0; c < nc; c++) { ssa_err_acx[c] = R(1); ssa_err_acy[c] = R(1); }
    set_inactive_margins(ux_b, uy_b, f_ice, nx, ny, BC);
    int lin_iter = 0, lin_fail = 0, iter_now = 0;
    wp rel = (wp)par->ssa_iter_rel;
    double tsolve = 0.0;
    for (int iter = 1; iter <= par->ssa_iter_max; iter++) {
        memcpy(visc_nm1, visc_eff, sizeof(wp) * n3);
        memcpy(ux_nm1, ux_b, sizeof(wp) * nc);
        memcpy(uy_nm1, uy_b, sizeof(wp) * nc);
        switch (par->visc_method) {
        case 0: for (size_t q = 0; q < n3; q++) visc_eff[q] = (wp)par->visc_const; break;
        case 1:
            oracle_visc_eff_3D_nodes(visc_eff, ux_b, uy_b, ATT, H_ice, f_ice, zeta_aa, dx, dy,
                                     n_glen, (wp)par->eps_0, nx, ny, nz, par->boundaries);
            picard_relax_visc(visc_eff, visc_nm1, rel, n3);
            break;
        case 2:
            oracle_visc_eff_3D_aa(visc_eff, ux_b, uy_b, ATT, f_ice, dx, dy, n_glen,
                                  (wp)par->eps_0, nx, ny, nz, par->boundaries);
            picard_relax_visc(visc_eff, visc_nm1, rel, n3);
            break;
        default: rc = -2; goto out;
        }
        oracle_visc_eff_int(visc_eff_int, visc_eff, H_ice, f_ice, zeta_aa, nx, ny, nz);
        if (oracle_calc_beta(beta, c_bed, f_slide, ux_b, uy_b, H_ice, f_ice, H_grnd, f_grnd, z_bed,
                             z_sl, par, nx, ny) != 0) { rc = -3; goto out; }
        if (oracle_stagger_beta(beta_acx, beta_acy, beta, H_ice, f_ice, ux_b, uy_b, f_grnd,
                                f_grnd_acx, f_grnd_acy, par, nx, ny) != 0) {
            rc = -4; goto out;
        }
        memcpy(bsx, beta_acx, sizeof(wp) * nc);
        memcpy(bsy, beta_acy, sizeof(wp) * nc);
        set_beta_min_grounded(bsx, bsy, ssa_mask_acx, ssa_mask_acy, (wp)par->beta_min, nc);
        if (par->vel_lim_method == 1)
            oracle_vel_lim_drag(lkx, lky, lrx, lry, ux_b, uy_b, ssa_mask_acx, ssa_mask_acy,
                                (wp)par->ssa_vel_max, (wp)par->ssa_vel_lim_tau, nx, ny,
                                par->boundaries);
        if (assemble_selected(par, &L, cap, cfrom, ux_b, uy_b, bsx, bsy, visc_eff_int,
                              ssa_mask_acx, ssa_mask_acy, H_ice, f_ice, z_bed, z_sl, taud_acx,
                              taud_acy, taul_int_acx, taul_int_acy, dx, dy, lkx, lky, lrx, lry,
                              nx, ny) < 0) { rc = -7; goto out; }
        int32_t it = 0, st = 0;
        double res = 0, secs = 0;
#ifdef ORACLE_DEBUG_DUMP
        {
            FILE *fd = fopen("/tmp/oracle_csr.bin", "wb");
            int nn = L.nmax, nnz = L.a_ptr[L.nmax] - 1;
            fwrite(&nn, 4, 1, fd); fwrite(&nnz, 4, 1, fd);
            fwrite(L.a_ptr, 4, nn + 1, fd); fwrite(L.a_index, 4, nnz, fd);
            fwrite(L.a_value, 8, nnz, fd); fwrite(L.b_value, 8, nn, fd); fwrite(L.x_value, 8, nn, fd);
            fclose(fd);
        }
#endif
        int src = oracle_csr_solve(L.nmax, L.a_ptr, L.a_index, L.a_value, L.b_value, L.x_value,
                                   lis_opt, &it, &st, &res, &secs);
        if (src != 0) { rc = -5; goto out; }
        tsolve += secs;
        lin_iter += it;
        if (st != 0) lin_fail++;
        if (lin_hist) lin_hist[iter - 1] = it;
        /* linear_solver_save_velocity (solver_ssa_ac.f90:33-71) */
        save_velocity(ux_b, uy_b, &L, cfrom, nx, ny);
        for (int c = 0; c < nc; c++) {
            if (WABS(ux_b[c]) < R(1e-15)) ux_b[c] = R(0);
            if (WABS(uy_b[c]) < R(1e-15)) uy_b[c] = R(0);
        }
        oracle_ssa_vel_clip(ux_b, uy_b, (wp)par->ssa_vel_max, nc);
        oracle_picard_relax_vel(ux_b, uy_b, ux_nm1, uy_nm1, rel, nc);
        int conv = 0, nchk = 0;
        wp resid = picard_conv_l2(&conv, ux_b, uy_b, ux_nm1, uy_nm1, ssa_mask_acx, ssa_mask_acy,
                                  (wp)par->ssa_iter_conv, iter, par->ssa_iter_max, nx, ny, &nchk);
        if (resid_hist) resid_hist[iter - 1] = (double)resid;
        if (par->ssa_write_log)
            printf("ssa: %s %4d %12.4g | %8d | lin %6d %d\n", conv == 1 ? "C" : conv == 2 ? "X" : " ",
                   iter, (double)resid, nchk, it, st);
        if (conv < 0) { rc = -6; iter_now = iter; goto out; }
        picard_l1rel(ssa_err_acx, ssa_err_acy, ux_b, uy_b, ux_nm1, uy_nm1, nc);
        iter_now = iter;
        if (conv) break;
    }
    out_i[3] = oracle_count_vel_lim_faces(ux_b, uy_b, ssa_mask_acx, ssa_mask_acy,
                                          (wp)par->ssa_vel_max, par->vel_lim_method == 1, nx, ny,
                                          par->boundaries);
    /* calc_basal_stress (velocity_ssa.f90:588-617) */
    for (int c = 0; c < nc; c++) {
        taub_acx[c] = bsx[c] * ux_b[c];
        taub_acy[c] = bsy[c] * uy_b[c];
        if (WABS(taub_acx[c]) < R(1e-5)) taub_acx[c] = R(0);
        if (WABS(taub_acy[c]) < R(1e-5)) taub_acy[c] = R(0);
    }
    if (par->visc_method == 0)
        oracle_visc_eff_3D_nodes(visc_eff, ux_b, uy_b, ATT, H_ice, f_ice, zeta_aa, dx, dy, n_glen,
                                 (wp)par->eps_0, nx, ny, nz, par->boundaries);
out:
    out_i[0] = iter_now;
    out_i[1] = lin_iter;
    out_i[2] = lin_fail;
    if (solve_seconds) *solve_seconds = tsolve;
    free(visc_nm1); free(ux_nm1); free(uy_nm1); free(bsx); free(bsy);
    free(lkx); free(lky); free(lrx); free(lry); free(cfrom);
    free(L.a_ptr); free(L.a_index); free(L.a_value); free(L.b_value); free(L.x_value);
    return rc;
}

/* the DIVA solve (calc_velocity_diva) and its helpers */
#include "diva_oracle.inc"

/* the implicit thickness advection system (calc_advec2D, "impl-lis") */
#include "advection_oracle.inc"
#include "inputs_oracle.inc"

/* size of wp, so the Python driver can check which build it loaded */
int oracle_wp_bytes(void) { return (int)sizeof(wp); }
