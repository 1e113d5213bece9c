// ygpu_host.h -- host-side state shared by the C-ABI translation units:
// the handle, the LIS option grammar, and the device residency of a set of
// domains (Batch) that the batch, session and server entry points build on.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <tuple>
#include <string>
#include <utility>
#include <vector>

#include "../../include/ygpu.h"
#include "ygpu_diva.h"
#include "ygpu_fields.h"
#include "ygpu_gtable.h"
#include "ygpu_krylov.h"
#include "ygpu_spmv.h"

struct ygpu_handle {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  std::string note;   // what the last call did differently from what it was asked
  bool stats_on = false;
  ygpu_stats stats{};
  // Krylov schedules of the mixed path per grid and workgroups per member
  // (nx, ny, G): ILU(0) line blocks and threads per member block, from
  // ygpu_autotune or ygpu_set_schedule
  struct Sched {
    int nb, block, ncol;   // line blocks, threads per block, column blocks (0: options)
  };
  std::map<std::tuple<int, int, int>, Sched> sched;
};

#define YG_CHECK(h, call)                                                          \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess) {                                                        \
      (h)->err = std::string(#call) + ": " + hipGetErrorString(e_);                \
      return YGPU_ERR_DEVICE;                                                      \
    }                                                                              \
  } while (0)

namespace yg {

// ---- LIS option string (solver_linear.F90:350) ----------------------------
struct Opts {
  YgKrylovOpts k;
  int ny_hint = 0;
  int prec = -1;  // -1 unset, 0 fp64, 2 mixed
  int group = 0;  // workgroups per member: 0 automatic, 1 one (no groups), 2, 4 or 8
  int group_threads = 0;   // threads per group workgroup: 512 (one per CU), 256 (two), 0 automatic
  const char* substituted = nullptr;   // solver named in -i that another one stands in for
  int fields32 = 1;   // mixed mode: field kernels in FP32 (-ygpu_fields fp32, the default)
  int pattern = 0;    // CSR pattern: 0 automatic, 1 scalar five-point, 2 C-grid SSA stencil
  // -ygpu_inputs device: the driving stress, lateral front stress and SSA
  // masks of each solve derived on the device from its geometry (z_srf
  // required; the caller's arrays for them may be null and receive the
  // derived values when given), with calc_ydyn's parameters
  int dev_inputs = 0;
  double taud_lim = 2e5, grad_lim = 0.5, slope_x = 0.0, slope_y = 0.0;
  int lat_bc = 3;
};

int parse_opts(const char* s, Opts& o, std::string& err);
// Algorithmic HBM bytes of one member's Krylov launch (defined with parse_opts)
// pF: bytes per factor value (BF16 line-block factors: 40 bytes per cell / 18)
double krylov_bytes(double C, int iters, int pA, double pF, int pv, bool ilu);
// Algorithmic HBM bytes of the mixed mode's FP64 refinement for one member
// of C cells that made nres residual passes (and nres - 1 corrections)
double refinement_bytes(double C, int nres);
// skewed step table of an nx x ny member: soff[t] = cells before step t
void step_table(int nx, int ny, std::vector<int>& soff, int& T);
// boundaries enum of the C-ABI -> neighbour-index code of the kernels
int bc_code(int bnd);
// ygpu_csr_solve on a scalar five-point system (ygpu_scalar.hip)
int scalar_csr_solve(ygpu_handle* h, int32_t n, int32_t nnz, const int32_t* a_ptr1,
                     const int32_t* a_idx1, const double* a_val, const double* b, double* x_inout,
                     const Opts& o, int32_t* iters, int32_t* status, double* rel_resid,
                     double* seconds);
// Field kernels in FP32: the reference mode (the reference's wp = sp) and
// the mixed mode unless -ygpu_fields fp64; FP64 in the FP64 mode.
bool fields32(const ygpu_ssa_params* par);
// The per-solve inputs of one domain on the device (ygpu_inputs.hip): ice
// front codes, surface slopes, driving stress, lateral front stress and SSA
// face masks from the geometry.  Outputs sx, sy may be null.
template <typename R>
struct YgDynArgs {
  int nx, ny, bnd, BC, lateral_bc;
  R taud_lim, grad_lim, slope_x, slope_y, dx, dy, rho_ice, rho_sw, g;
  const R *H, *fi, *fg, *zb, *zsl, *zs;
  int* mf;
  R *sx, *sy, *tdx, *tdy, *tlx, *tly;
  int *mx, *my;
};
template <typename R>
int launch_dyn_inputs(const YgDynArgs<R>& a, hipStream_t st);
}  // namespace yg
// looks up the device's no-allocation scratch array (ygpu_noalloc.hip); 0 ok
int yg_noalloc_init(int device);
namespace yg {

// ---- device allocations of one batch --------------------------------------
struct Arena {
  std::vector<void*> ptrs;
  size_t bytes = 0;
  ~Arena() { for (void* p : ptrs) (void)hipFree(p); }
  template <typename T>
  hipError_t alloc(T** p, size_t n) {
    *p = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void**)p, n * sizeof(T));
    if (e == hipSuccess) { ptrs.push_back(*p); bytes += n * sizeof(T); }
    return e;
  }
};

template <typename R>
struct Batch {
  ygpu_handle* h = nullptr;
  Arena ar;
  int nmem = 0, diva = 0, prec = 0, kmode = 0, block = 1024;
  // member groups: workgroups per member (1: the one-block kernel) and groups
  int group = 1, ngroups = 0, gthreads = 512;
  int *d_gsync = nullptr, *d_gfail = nullptr;
  // member groups of different sizes (automatic, members of uneven sizes):
  // every member's group size and the launch table (ygpu_krylov.h); group
  // reports the largest size and ngroups the groups
  bool vargroup = false;
  std::vector<int> mgroup, h_gtab, h_gmoff, h_gmem;
  int *d_gtab = nullptr, *d_gmoff = nullptr, *d_gmem = nullptr;
  int gnwg = 0;
  long long total_cells = 0, total_z = 0, total_v = 0;
  int zstride = 0;
  std::vector<YgMember> mem;
  std::vector<long long> coffs, zoffs;
  std::vector<int> soff;
  YgFieldParams fp{};
  Opts opts;
  int iter_max = 1;
  // device
  YgMember* d_mem = nullptr;
  long long *d_coffs = nullptr, *d_zoffs = nullptr;
  int* d_chunk = nullptr;   // member of every YG_CHUNK-cell chunk's first cell
  int* d_soff = nullptr;
  uint32_t* d_mask = nullptr;
  double *d_A64 = nullptr, *d_b = nullptr, *d_x = nullptr, *d_F64 = nullptr, *d_W64 = nullptr;
  float *d_A32 = nullptr, *d_F32 = nullptr, *d_W32 = nullptr;
  uint16_t* d_F16 = nullptr;   // BF16 line-block factors (-ygpu_factor bf16)
  int *d_done = nullptr, *d_iter_now = nullptr, *d_lin_iter = nullptr, *d_lin_fail = nullptr,
      *d_lim_n = nullptr, *d_kit = nullptr, *d_kst = nullptr, *d_ksw = nullptr;
  int* d_warm = nullptr;   // per member: factors of the previous solve resident (keep_factors)
  double *d_resid = nullptr, *d_kres = nullptr;
  double* d_epart = nullptr;   // epilogue partial sums, YG_EPI_BLOCKS x 4 per member
  int* d_eticket = nullptr;    // epilogue arrival counters
  YgFields<R> f{};
  YgDivaFields<R> dv{};
  // -ygpu_inputs device: the surface elevation and the front codes
  R* d_zsrf = nullptr;
  int* d_mf = nullptr;
  double grav = 9.81;

  YgBatchGeom geom() const {
    YgBatchGeom G;
    G.members = d_mem; G.coffs = d_coffs; G.zoffs = d_zoffs; G.soff = d_soff;
    G.nmem = nmem; G.total_cells = total_cells;
    G.seg = d_coffs; G.segm = nullptr; G.nseg = nmem;
    G.chunk_seg = d_chunk;
    return G;
  }

  // parameters shared by every member (ssa_param_class / diva_param_class)
  int setup_params(ygpu_handle* hh, const ygpu_ssa_params* par, int is_diva) {
    h = hh;
    diva = is_diva;
    std::string e;
    int rc = parse_opts(par->ssa_lis_opt, opts, e);
    if (rc) { h->err = e; return rc; }
    if (opts.substituted) h->note = opts.substituted;
    // mixed mode: linear solves of a member whose Picard residual is above
    // 10 ssa_iter_conv stop at the tolerance itself (-ygpu_ir_loose).  The
    // Picard iteration contracts (the residual falls by about 0.82 per
    // iteration on the ensemble), so such a member runs another dozen
    // iterations on tight solves, over which the looser solve's error
    // shrinks tenfold.
    if (opts.k.ir_loose < 0.0) opts.k.ir_loose = 10.0 * par->ssa_iter_conv;
    prec = par->precision;
    kmode = prec == YGPU_PREC_MIXED ? YG_MODE_MIXED : YG_MODE_FP64;
    if (par->ssa_solver != 0 && par->ssa_solver != 1) { h->err = "unknown ssa_solver"; return YGPU_ERR_ARG; }
    const int BC = bc_code(par->boundaries);
    if (BC < 0) { h->err = "unknown boundaries code"; return YGPU_ERR_ARG; }
    // ILU(0) on a periodic domain: the wavefront factorization and sweeps
    // see no cell beyond the domain edge, so the couplings that wrap around
    // it stay out of the factors (the SpMV keeps them: the solution meets
    // the same tolerance; iteration counts differ from an ILU(0) with them)
    if (opts.k.precon == 2 && (BC == YG_BC_PERIODIC || BC == YG_BC_PERIODIC_X ||
                               BC == YG_BC_PERIODIC_Y || BC == YG_BC_MISMIP3D || BC == YG_BC_TROUGH) &&
        h->note.empty())
      h->note = "ILU(0) without the couplings that wrap around the periodic domain";
    fp.bnd = par->boundaries; fp.BC = BC;
    fp.visc_method = par->visc_method; fp.beta_method = par->beta_method;
    fp.beta_gl_scale = par->beta_gl_scale; fp.beta_gl_stag = par->beta_gl_stag;
    fp.vel_lim_method = par->ssa_vel_lim_method; fp.diva = is_diva; fp.no_slip = par->no_slip;
    fp.energy = par->ssa_solver == 1;
    fp.visc_const = par->visc_const; fp.beta_const = par->beta_const; fp.beta_q = par->beta_q;
    fp.beta_u0 = par->beta_u0; fp.beta_gl_f = par->beta_gl_f; fp.H_grnd_lim = par->H_grnd_lim;
    fp.beta_min = par->beta_min; fp.eps_0 = par->eps_0; fp.vel_lim_tau = par->ssa_vel_lim_tau;
    fp.vel_max = par->ssa_vel_max; fp.iter_rel = par->ssa_iter_rel; fp.iter_conv = par->ssa_iter_conv;
    fp.rho_ice = par->rho_ice; fp.rho_sw = par->rho_sw;
    grav = par->g;
    if (par->visc_method < 0 || par->visc_method > 2 || par->beta_method < -1 || par->beta_method > 5 ||
        par->beta_gl_scale < 0 || par->beta_gl_scale > 3 ||
        par->beta_gl_stag < -1 || par->beta_gl_stag > 4) {
      h->err = "parameter value not supported";
      return YGPU_ERR_UNSUPPORTED;
    }
    iter_max = par->ssa_iter_max;
    if (iter_max < 1) { h->err = "ssa_iter_max < 1"; return YGPU_ERR_ARG; }
    return YGPU_OK;
  }

  // Checks one domain against the kernels' limits and fills its member
  // record except the offsets (coff, toff, woff); st receives its step table.
  int describe(const ygpu_ssa_problem& P, int k, YgMember& M, std::vector<int>& st) {
    if (P.nx < 3 || P.ny < 3 || P.nz < 2 || P.nx > YG_MAX_LINES || P.ny > 65535 ||
        (P.real_bytes != 4 && P.real_bytes != 8)) {
      h->err = "domain " + std::to_string(k) + ": unsupported size or real_bytes";
      return YGPU_ERR_ARG;
    }
    M.nx = P.nx; M.ny = P.ny; M.nz = P.nz;
    int T;
    step_table(P.nx, P.ny, st, T);
    if (T > YG_MAX_STEPS) { h->err = "domain too large for the step table"; return YGPU_ERR_ARG; }
    M.T = T;
    M.maxcnt = 0;
    for (int t = 0; t < T; ++t) M.maxcnt = std::max(M.maxcnt, st[t + 1] - st[t]);
    if (M.maxcnt > YG_MAX_STEP_CELLS) {
      h->err = "domain " + std::to_string(k) + ": a wavefront step holds more cells than the LDS stage";
      return YGPU_ERR_ARG;
    }
    M.dx = P.dx; M.dy = P.dy;
    M.vlen = (int)(((2LL * P.nx * P.ny + 3) / 4) * 4);
    M.coff = M.toff = M.woff = 0;
    M.nb = 0;
    M.ncol = 0;
    tuned_block = 0;
    if (kmode == YG_MODE_MIXED && opts.k.precon == 2 && opts.k.ilu_blocks < 0) {
      auto it = h->sched.find({P.nx, P.ny, vargroup ? mgroup[k] : group});
      if (it != h->sched.end()) {
        M.nb = it->second.nb;
        M.ncol = it->second.ncol;
        tuned_block = it->second.block;
      }
    }
    return YGPU_OK;
  }

  int tuned_block = 0;   // threads the last described member's schedule asks for

  // threads per member block: 512 up to 512 grid lines, else 1024
  static int pick_block(const ygpu_ssa_params* par, int maxnx) {
    int b = par->block > 0 ? par->block : (maxnx <= 512 ? 512 : 1024);
    b = std::max(b, ((maxnx + 63) / 64) * 64);
    return std::min(b, 1024);
  }

  int setup(ygpu_handle* hh, const ygpu_ssa_params* par, int np, const ygpu_ssa_problem* probs,
            int is_diva) {
    int rc = setup_params(hh, par, is_diva);
    if (rc) return rc;
    nmem = np;
    mem.resize(np);
    coffs.assign(np + 1, 0);
    zoffs.assign(np + 1, 0);
    soff.clear();
    int maxnx = 0, want_block = 0;
    long long maxcells = 0;
    for (int k = 0; k < np; ++k) {
      maxnx = std::max(maxnx, probs[k].nx);
      maxcells = std::max(maxcells, (long long)probs[k].nx * probs[k].ny);
    }
    std::vector<std::pair<int, int>> sizes(np);
    for (int k = 0; k < np; ++k) sizes[k] = {probs[k].nx, probs[k].ny};
    if ((rc = pick_group(maxnx, maxcells, sizes))) return rc;   // before the schedules are looked up
    for (int k = 0; k < np; ++k) {
      const ygpu_ssa_problem& P = probs[k];
      YgMember& M = mem[k];
      std::vector<int> st;
      if ((rc = describe(P, k, M, st))) return rc;
      want_block = std::max(want_block, tuned_block);
      if (k > 0 && P.n_glen != probs[0].n_glen) {
        h->err = "domains of one call must share n_glen";
        return YGPU_ERR_ARG;
      }
      M.coff = coffs[k];
      M.toff = (long long)soff.size();
      soff.insert(soff.end(), st.begin(), st.end());
      M.woff = total_v;
      total_v += M.vlen;
      coffs[k + 1] = coffs[k] + (long long)P.nx * P.ny;
      zoffs[k + 1] = zoffs[k] + (long long)P.nx * P.ny * P.nz;
      zstride = std::max(zstride, P.nz);
      fp.n_glen = P.n_glen;
    }
    total_cells = coffs[np];
    total_z = zoffs[np];
    block = pick_block(par, maxnx);
    if (par->block <= 0 && want_block > block) block = std::min(want_block, 1024);
    return allocate();
  }

  // Workgroups per member: the member-group kernel covers the mixed mode
  // with BiCGSTAB and block-Jacobi ILU(0).  Automatic: a member of nx x ny
  // cells gets the largest G of 16, 8, 4, 2 that leaves every workgroup at
  // least 4096 of its cells, with 256-thread workgroups (up to two per CU).
  // When every member gets the same G, the groups (a multiple of 8) are as
  // many as the CUs hold at once, at most one per member, and take the
  // members in turn.  With one workgroup per CU either way (64 x 161^2,
  // G = 4), four waves run the SpMV 13 % faster than eight (301 against
  // 292 solves/s): fewer records in flight per CU evict fewer of the
  // neighbour velocity pairs that the stencil re-reads from the L1.
  // Members of uneven sizes get groups of their own sizes (a member's
  // group size depends on its grid alone), placed by a table (vargroup).
  // The solves the member-group kernel runs: mixed precision, BiCGSTAB with
  // line-block ILU(0), Jacobi or no preconditioner, or CG with Jacobi or
  // none; plain FP64, BiCGSTAB or CG with Jacobi or none (the FP64 path's
  // ILU(0) is the exact one, a wavefront over all lines of a member).
  static bool group_solver_fits(int kmode, const YgKrylovOpts& k) {
    if (k.solver != 1 && k.solver != 4) return false;
    if (kmode == YG_MODE_FP64) return k.precon != 2;
    if (k.solver == 4) return k.precon != 2 || k.ilu_blocks != 0;
    return k.precon != 2;
  }

  static int auto_group(int nx, int ny, int thr) {
    const long long cells = (long long)nx * ny;
    const int waves = thr / 64;
    int g = 1;
    for (int c : {16, 8, 4, 2})
      if (cells / c >= 4096 && c <= nx) { g = c; break; }
    while (g < 16 && nx > 64 * g * waves) g *= 2;   // the group's waves sweep at most 64 lines each
    return nx > 64 * g * waves ? 0 : g;            // 0: no group size covers the lines
  }

  int pick_group(int maxnx, long long maxcells, const std::vector<std::pair<int, int>>& sizes) {
    group = 1;
    ngroups = 0;
    vargroup = false;
    const int want = opts.group;
    if (want == 1) return YGPU_OK;
    const bool fits = group_solver_fits(kmode, opts.k);
    if (want != 0 && want != 2 && want != 4 && want != 8 && want != 16) {
      h->err = "-ygpu_group takes 0, 1, 2, 4, 8 or 16";
      return YGPU_ERR_ARG;
    }
    if (!fits) {
      if (want == 0) return YGPU_OK;
      h->err = "-ygpu_group: member groups run BiCGSTAB or CG with Jacobi or no preconditioner, "
               "and in mixed precision BiCGSTAB with line-block ILU(0)";
      return YGPU_ERR_UNSUPPORTED;
    }
    int ncu0 = 0;
    YG_CHECK(h, hipDeviceGetAttribute(&ncu0, hipDeviceAttributeMultiprocessorCount, h->device));
    const int groups8 = ((nmem + 7) / 8) * 8;
    // do all groups of G workgroups of `thr` threads fit the CUs at once
    auto all_fit = [&](int g, int thr) { return groups8 * g <= ncu0 * (512 / thr); };
    int G = want, thr = opts.group_threads;
    if (G == 0) {
      if (thr == 0) thr = YG_GROUP_THREADS;
      mgroup.assign(nmem, 1);
      bool same = true, covered = true;
      for (int k = 0; k < nmem; ++k) {
        mgroup[k] = auto_group(sizes[k].first, sizes[k].second, thr);
        same &= mgroup[k] == mgroup[0];
        covered &= mgroup[k] > 0;
      }
      if (!same && covered) return build_group_table(ncu0, thr);
      G = covered ? mgroup[0] : 1;
      if (G == 1) {
        // the size rule of one group size for the largest member (members
        // too narrow for their own groups)
        for (int g : {8, 4, 2}) {
          if (maxcells / g < 4096) continue;
          int t = thr;
          if (maxnx > 64 * g * (t / 64)) {
            if (opts.group_threads == 0 && maxnx <= 64 * g * 8) t = 512;
            else continue;
          }
          G = g;
          thr = t;
          break;
        }
      }
      if (G == 1) return YGPU_OK;
    } else if (thr == 0) {
      thr = all_fit(G, 256) ? 256 : 512;
    }
    gthreads = thr;
    const int ncu = ncu0 * (512 / gthreads);   // workgroups the CUs hold at once
    if (8 * G > ncu) { h->err = "-ygpu_group: more workgroups than CUs"; return YGPU_ERR_ARG; }
    if (maxnx > 64 * G * (gthreads / 64)) {
      h->err = "-ygpu_group: a member has more lines than the group's waves can sweep";
      return YGPU_ERR_ARG;
    }
    group = G;
    ngroups = std::min(groups8, (ncu / G) / 8 * 8);
    return YGPU_OK;
  }

  // The launch table of groups of different sizes (mgroup): group_table
  // (ygpu_gtable.h) with the XCDs' slots at `thr` threads per workgroup.
  int build_group_table(int ncu0, int thr) {
    const int per_xcd = (ncu0 / 8) * (512 / thr);
    std::string e;
    if (!group_table(mgroup, per_xcd, h_gtab, h_gmoff, h_gmem, gnwg, e)) {
      h->err = e;
      return YGPU_ERR_ARG;
    }
    vargroup = true;
    group = *std::max_element(mgroup.begin(), mgroup.end());
    ngroups = (int)h_gmoff.size() - 1;
    gthreads = thr;
    return YGPU_OK;
  }

  int allocate() {
    const long long C = total_cells, Z = total_z;
    YG_CHECK(h, ar.alloc(&d_mem, nmem));
    YG_CHECK(h, ar.alloc(&d_coffs, nmem + 1));
    YG_CHECK(h, ar.alloc(&d_zoffs, nmem + 1));
    YG_CHECK(h, ar.alloc(&d_soff, soff.size()));
    YG_CHECK(h, ar.alloc(&d_mask, C));
    YG_CHECK(h, ar.alloc(&d_A64, (size_t)YG_NA * C));
    YG_CHECK(h, ar.alloc(&d_b, 2 * C));
    YG_CHECK(h, ar.alloc(&d_x, 2 * C));
    YG_CHECK(h, ar.alloc(&d_W64, (size_t)YG_NW64 * total_v));
    if (kmode == YG_MODE_MIXED) {
      YG_CHECK(h, ar.alloc(&d_A32, (size_t)YG_NA * C));
      YG_CHECK(h, ar.alloc(&d_F32, (size_t)YG_NF * C));
      YG_CHECK(h, ar.alloc(&d_W32, (size_t)YG_NW32 * total_v));
      if (opts.k.factor16) YG_CHECK(h, ar.alloc(&d_F16, (size_t)YG_NF16 * C));
    } else {
      YG_CHECK(h, ar.alloc(&d_F64, (size_t)YG_NF * C));
    }
    int* ints[9];
    for (int q = 0; q < 9; ++q) YG_CHECK(h, ar.alloc(&ints[q], nmem));
    d_done = ints[0]; d_iter_now = ints[1]; d_lin_iter = ints[2]; d_lin_fail = ints[3];
    d_lim_n = ints[4]; d_kit = ints[5]; d_kst = ints[6]; d_ksw = ints[7];
    d_warm = ints[8];
    YG_CHECK(h, hipMemsetAsync(d_warm, 0, sizeof(int) * nmem, h->stream));
    YG_CHECK(h, ar.alloc(&d_resid, nmem));
    YG_CHECK(h, ar.alloc(&d_kres, nmem));
    YG_CHECK(h, ar.alloc(&d_epart, (size_t)nmem * YG_EPI_BLOCKS * 4));
    YG_CHECK(h, ar.alloc(&d_eticket, nmem));
    YG_CHECK(h, hipMemsetAsync(d_eticket, 0, sizeof(int) * nmem, h->stream));
    if (group > 1) {
      YG_CHECK(h, ar.alloc(&d_gsync, (size_t)ngroups * (YG_GSYNC_BYTES / 4)));
      YG_CHECK(h, ar.alloc(&d_gfail, 1));
    }
    if (vargroup) {
      YG_CHECK(h, ar.alloc(&d_gtab, h_gtab.size()));
      YG_CHECK(h, ar.alloc(&d_gmoff, h_gmoff.size()));
      YG_CHECK(h, ar.alloc(&d_gmem, h_gmem.size()));
      YG_CHECK(h, hipMemcpy(d_gtab, h_gtab.data(), sizeof(int) * h_gtab.size(), hipMemcpyHostToDevice));
      YG_CHECK(h, hipMemcpy(d_gmoff, h_gmoff.data(), sizeof(int) * h_gmoff.size(), hipMemcpyHostToDevice));
      YG_CHECK(h, hipMemcpy(d_gmem, h_gmem.data(), sizeof(int) * h_gmem.size(), hipMemcpyHostToDevice));
    }
    // fields
    R** r2d[] = {(R**)&f.c_bed, (R**)&f.f_slide, (R**)&f.taud_acx, (R**)&f.taud_acy,
                 (R**)&f.taul_acx, (R**)&f.taul_acy, (R**)&f.H_ice, (R**)&f.f_ice,
                 (R**)&f.H_grnd, (R**)&f.f_grnd, (R**)&f.z_bed, (R**)&f.z_sl,
                 (R**)&f.f_grnd_acx, (R**)&f.f_grnd_acy, &f.ux, &f.uy,
                 &f.ux_nm1, &f.uy_nm1, &f.visc_int, &f.beta, &f.beta_acx, &f.beta_acy, &f.bss_x,
                 &f.bss_y, &f.lkx, &f.lky, &f.lrx, &f.lry, &f.dudx, &f.dudy, &f.dvdx, &f.dvdy,
                 &f.err_x, &f.err_y, &f.taub_x, &f.taub_y};
    for (R** p : r2d) YG_CHECK(h, ar.alloc(p, C));
    YG_CHECK(h, ar.alloc((R**)&f.ATT, Z));
    YG_CHECK(h, ar.alloc(&f.visc, Z));
    if (fp.visc_method == 1) YG_CHECK(h, ar.alloc(&f.attp, 4 * Z));
    YG_CHECK(h, ar.alloc((R**)&f.zeta, (size_t)zstride * nmem));
    f.zstride = zstride;
    YG_CHECK(h, ar.alloc((int32_t**)&f.mx, C));
    YG_CHECK(h, ar.alloc((int32_t**)&f.my, C));
    if (opts.dev_inputs) {
      YG_CHECK(h, ar.alloc(&d_zsrf, C));
      YG_CHECK(h, ar.alloc(&d_mf, C));
    }
    if (fp.energy) {
      YG_CHECK(h, ar.alloc(&f.cfrom, 2 * C));
      YG_CHECK(h, ar.alloc(&f.asm_err, nmem));
      YG_CHECK(h, hipMemsetAsync(f.asm_err, 0, sizeof(int) * nmem, h->stream));
    }
    if (diva) {
      R** d2[] = {&f.F2, &f.F2_acx, &f.F2_acy, &f.beff_acx, &f.beff_acy, &f.beta_eff, &f.ubar_x,
                  &f.ubar_y};
      for (R** p : d2) YG_CHECK(h, ar.alloc(p, C));
      YG_CHECK(h, ar.alloc(&f.duxdz, Z));
      YG_CHECK(h, ar.alloc(&f.duydz, Z));
      YG_CHECK(h, ar.alloc(&dv.ux3, Z));
      YG_CHECK(h, ar.alloc(&dv.uy3, Z));
    }
    // geometry tables
    std::vector<long long> mc(coffs.begin(), coffs.end()), mz(zoffs.begin(), zoffs.end());
    YG_CHECK(h, hipMemcpyAsync(d_mem, mem.data(), sizeof(YgMember) * nmem, hipMemcpyHostToDevice, h->stream));
    YG_CHECK(h, hipMemcpyAsync(d_coffs, mc.data(), sizeof(long long) * (nmem + 1), hipMemcpyHostToDevice, h->stream));
    YG_CHECK(h, hipMemcpyAsync(d_zoffs, mz.data(), sizeof(long long) * (nmem + 1), hipMemcpyHostToDevice, h->stream));
    YG_CHECK(h, hipMemcpyAsync(d_soff, soff.data(), sizeof(int) * soff.size(), hipMemcpyHostToDevice, h->stream));
    const std::vector<int> ch = yg_chunk_members(coffs.data(), nmem, total_cells);
    YG_CHECK(h, ar.alloc(&d_chunk, ch.size()));
    YG_CHECK(h, hipMemcpyAsync(d_chunk, ch.data(), sizeof(int) * ch.size(), hipMemcpyHostToDevice, h->stream));
    YG_CHECK(h, hipStreamSynchronize(h->stream));
    return YGPU_OK;
  }

  // pack one field of every domain (host float or double) into R and upload
  int up2(R* dst, const void* const* src, int nz_mult, bool may_null, R fill = R(0)) {
    const long long n = nz_mult ? total_z : total_cells;
    std::vector<R> st(n);
    for (int k = 0; k < nmem; ++k) {
      const long long off = nz_mult ? zoffs[k] : coffs[k];
      const long long cnt = (nz_mult ? zoffs[k + 1] : coffs[k + 1]) - off;
      const void* s = src[k];
      if (!s) {
        if (!may_null) { h->err = "missing input array"; return YGPU_ERR_ARG; }
        for (long long q = 0; q < cnt; ++q) st[off + q] = fill;
      } else if (rb[k] == 4) {
        const float* p = (const float*)s;
        for (long long q = 0; q < cnt; ++q) st[off + q] = (R)p[q];
      } else {
        const double* p = (const double*)s;
        for (long long q = 0; q < cnt; ++q) st[off + q] = (R)p[q];
      }
    }
    YG_CHECK(h, hipMemcpyAsync(dst, st.data(), sizeof(R) * n, hipMemcpyHostToDevice, h->stream));
    YG_CHECK(h, hipStreamSynchronize(h->stream));
    return YGPU_OK;
  }

  int down2(const R* src, void* const* dst, int nz_mult) {
    const long long n = nz_mult ? total_z : total_cells;
    std::vector<R> st(n);
    YG_CHECK(h, hipMemcpyAsync(st.data(), src, sizeof(R) * n, hipMemcpyDeviceToHost, h->stream));
    YG_CHECK(h, hipStreamSynchronize(h->stream));
    for (int k = 0; k < nmem; ++k) {
      const long long off = nz_mult ? zoffs[k] : coffs[k];
      const long long cnt = (nz_mult ? zoffs[k + 1] : coffs[k + 1]) - off;
      void* d = dst[k];
      if (!d) continue;
      if (rb[k] == 4) {
        float* p = (float*)d;
        for (long long q = 0; q < cnt; ++q) p[q] = (float)st[off + q];
      } else {
        double* p = (double*)d;
        for (long long q = 0; q < cnt; ++q) p[q] = (double)st[off + q];
      }
    }
    return YGPU_OK;
  }

  std::vector<int> rb;

#define YG_FIELD_OF(arr, member) [&] { std::vector<const void*> v(nmem); for (int k = 0; k < nmem; ++k) v[k] = (arr)[k].member; return v; }()
#define YG_FIELDW_OF(arr, member) [&] { std::vector<void*> v(nmem); for (int k = 0; k < nmem; ++k) v[k] = (arr)[k].member; return v; }()
#define YG_FIELD(member) [&] { std::vector<const void*> v(nmem); for (int k = 0; k < nmem; ++k) v[k] = probs[k].member; return v; }()
#define YG_FIELDW(member) [&] { std::vector<void*> v(nmem); for (int k = 0; k < nmem; ++k) v[k] = probs[k].member; return v; }()

  int upload_inputs(const ygpu_ssa_problem* probs) {
    rb.resize(nmem);
    for (int k = 0; k < nmem; ++k) rb[k] = probs[k].real_bytes;
    int rc;
#define UP(dst, member) if ((rc = up2((R*)(dst), YG_FIELD(member).data(), 0, false))) return rc;
    UP(f.c_bed, c_bed) UP(f.f_slide, f_slide) UP(f.H_ice, H_ice) UP(f.f_ice, f_ice)
    UP(f.H_grnd, H_grnd) UP(f.f_grnd, f_grnd) UP(f.z_bed, z_bed) UP(f.z_sl, z_sl)
    if (opts.dev_inputs) {
      UP(d_zsrf, z_srf)
    } else {
      UP(f.taud_acx, taud_acx) UP(f.taud_acy, taud_acy)
      UP(f.taul_acx, taul_int_acx) UP(f.taul_acy, taul_int_acy)
    }
#undef UP
    // the faces' grounded fraction (read by the subgrid staggering only)
    if ((rc = up2((R*)f.f_grnd_acx, YG_FIELD(f_grnd_acx).data(), 0, true))) return rc;
    if ((rc = up2((R*)f.f_grnd_acy, YG_FIELD(f_grnd_acy).data(), 0, true))) return rc;
    if ((rc = up2((R*)f.ATT, YG_FIELD(ATT).data(), 1, false))) return rc;
    // zeta: nz per member, padded to zstride
    std::vector<R> zs((size_t)zstride * nmem, R(0));
    for (int k = 0; k < nmem; ++k) {
      if (!probs[k].zeta_aa) { h->err = "missing zeta_aa"; return YGPU_ERR_ARG; }
      for (int q = 0; q < probs[k].nz; ++q)
        zs[(size_t)k * zstride + q] = rb[k] == 4 ? (R)((const float*)probs[k].zeta_aa)[q]
                                                 : (R)((const double*)probs[k].zeta_aa)[q];
    }
    YG_CHECK(h, hipMemcpyAsync((R*)f.zeta, zs.data(), sizeof(R) * zs.size(), hipMemcpyHostToDevice, h->stream));
    if (opts.dev_inputs) return derive_inputs(probs);
    // masks
    std::vector<int32_t> mxs(total_cells), mys(total_cells);
    for (int k = 0; k < nmem; ++k) {
      if (!probs[k].ssa_mask_acx || !probs[k].ssa_mask_acy) { h->err = "missing ssa masks"; return YGPU_ERR_ARG; }
      memcpy(&mxs[coffs[k]], probs[k].ssa_mask_acx, sizeof(int32_t) * (coffs[k + 1] - coffs[k]));
      memcpy(&mys[coffs[k]], probs[k].ssa_mask_acy, sizeof(int32_t) * (coffs[k + 1] - coffs[k]));
    }
    YG_CHECK(h, hipMemcpyAsync((int32_t*)f.mx, mxs.data(), sizeof(int32_t) * total_cells, hipMemcpyHostToDevice, h->stream));
    YG_CHECK(h, hipMemcpyAsync((int32_t*)f.my, mys.data(), sizeof(int32_t) * total_cells, hipMemcpyHostToDevice, h->stream));
    YG_CHECK(h, hipStreamSynchronize(h->stream));
    return YGPU_OK;
  }

  // -ygpu_inputs device: calc_ydyn's inputs of every member (driving stress,
  // lateral front stress, SSA face masks; yelmo_dynamics.f90:105-122,464)
  // from the resident geometry, in place of uploaded ones
  int derive_inputs(const ygpu_ssa_problem* probs) {
    for (int k = 0; k < nmem; ++k) {
      const long long c0 = coffs[k];
      YgDynArgs<R> a{};
      a.nx = probs[k].nx; a.ny = probs[k].ny; a.bnd = fp.bnd; a.BC = fp.BC;
      a.lateral_bc = opts.lat_bc;
      a.taud_lim = (R)opts.taud_lim; a.grad_lim = (R)opts.grad_lim;
      a.slope_x = (R)opts.slope_x; a.slope_y = (R)opts.slope_y;
      a.dx = (R)probs[k].dx; a.dy = (R)probs[k].dy;
      a.rho_ice = (R)fp.rho_ice; a.rho_sw = (R)fp.rho_sw; a.g = (R)grav;
      a.H = f.H_ice + c0; a.fi = f.f_ice + c0; a.fg = f.f_grnd + c0; a.zb = f.z_bed + c0;
      a.zsl = f.z_sl + c0; a.zs = d_zsrf + c0;
      a.mf = d_mf + c0; a.sx = nullptr; a.sy = nullptr;
      a.tdx = (R*)f.taud_acx + c0; a.tdy = (R*)f.taud_acy + c0;
      a.tlx = (R*)f.taul_acx + c0; a.tly = (R*)f.taul_acy + c0;
      a.mx = (int*)f.mx + c0; a.my = (int*)f.my + c0;
      if (launch_dyn_inputs<R>(a, h->stream)) { h->err = "per-solve input kernels failed"; return YGPU_ERR_DEVICE; }
    }
    YG_CHECK(h, hipStreamSynchronize(h->stream));
    return YGPU_OK;
  }

  int upload_state(const ygpu_ssa_problem* probs) {
    int rc;
    if (diva) {
      if ((rc = up2(f.ubar_x, YG_FIELD(ux_bar).data(), 0, false))) return rc;
      if ((rc = up2(f.ubar_y, YG_FIELD(uy_bar).data(), 0, false))) return rc;
      if ((rc = up2(f.duxdz, YG_FIELD(duxdz).data(), 1, true))) return rc;
      if ((rc = up2(f.duydz, YG_FIELD(duydz).data(), 1, true))) return rc;
    }
    if ((rc = up2(f.ux, YG_FIELD(ux_b).data(), 0, false))) return rc;
    if ((rc = up2(f.uy, YG_FIELD(uy_b).data(), 0, false))) return rc;
    if ((rc = up2(f.taub_x, YG_FIELD(taub_acx).data(), 0, true))) return rc;
    if ((rc = up2(f.taub_y, YG_FIELD(taub_acy).data(), 0, true))) return rc;
    if ((rc = up2(f.visc, YG_FIELD(visc_eff).data(), 1, false))) return rc;
    if ((rc = up2(f.beta, YG_FIELD(beta).data(), 0, true))) return rc;
    if ((rc = up2(f.beta_acx, YG_FIELD(beta_acx).data(), 0, true))) return rc;
    if ((rc = up2(f.beta_acy, YG_FIELD(beta_acy).data(), 0, true))) return rc;
    return YGPU_OK;
  }

  int download(ygpu_ssa_problem* probs) {
    int rc;
    if (diva) {
      if ((rc = down2(f.ubar_x, YG_FIELDW(ux_bar).data(), 0))) return rc;
      if ((rc = down2(f.ubar_y, YG_FIELDW(uy_bar).data(), 0))) return rc;
      if ((rc = down2(f.duxdz, YG_FIELDW(duxdz).data(), 1))) return rc;
      if ((rc = down2(f.duydz, YG_FIELDW(duydz).data(), 1))) return rc;
      if ((rc = down2(f.beta_eff, YG_FIELDW(beta_eff).data(), 0))) return rc;
      if ((rc = down2(dv.ux3, YG_FIELDW(ux).data(), 1))) return rc;
      if ((rc = down2(dv.uy3, YG_FIELDW(uy).data(), 1))) return rc;
    }
    if ((rc = down2(f.ux, YG_FIELDW(ux_b).data(), 0))) return rc;
    if ((rc = down2(f.uy, YG_FIELDW(uy_b).data(), 0))) return rc;
    if ((rc = down2(f.taub_x, YG_FIELDW(taub_acx).data(), 0))) return rc;
    if ((rc = down2(f.taub_y, YG_FIELDW(taub_acy).data(), 0))) return rc;
    if ((rc = down2(f.visc, YG_FIELDW(visc_eff).data(), 1))) return rc;
    if ((rc = down2(f.visc_int, YG_FIELDW(visc_eff_int).data(), 0))) return rc;
    if ((rc = down2(f.err_x, YG_FIELDW(ssa_err_acx).data(), 0))) return rc;
    if ((rc = down2(f.err_y, YG_FIELDW(ssa_err_acy).data(), 0))) return rc;
    if ((rc = down2(f.beta, YG_FIELDW(beta).data(), 0))) return rc;
    if ((rc = down2(f.beta_acx, YG_FIELDW(beta_acx).data(), 0))) return rc;
    if ((rc = down2(f.beta_acy, YG_FIELDW(beta_acy).data(), 0))) return rc;
    std::vector<int> it(nmem), li(nmem), lf(nmem), ln(nmem), dn(nmem);
    std::vector<double> rs(nmem);
    YG_CHECK(h, hipMemcpy(it.data(), d_iter_now, sizeof(int) * nmem, hipMemcpyDeviceToHost));
    YG_CHECK(h, hipMemcpy(li.data(), d_lin_iter, sizeof(int) * nmem, hipMemcpyDeviceToHost));
    YG_CHECK(h, hipMemcpy(lf.data(), d_lin_fail, sizeof(int) * nmem, hipMemcpyDeviceToHost));
    YG_CHECK(h, hipMemcpy(ln.data(), d_lim_n, sizeof(int) * nmem, hipMemcpyDeviceToHost));
    YG_CHECK(h, hipMemcpy(dn.data(), d_done, sizeof(int) * nmem, hipMemcpyDeviceToHost));
    YG_CHECK(h, hipMemcpy(rs.data(), d_resid, sizeof(double) * nmem, hipMemcpyDeviceToHost));
    for (int k = 0; k < nmem; ++k) {
      probs[k].ssa_iter_now = it[k];
      probs[k].ssa_lin_iter = li[k];
      probs[k].ssa_lin_fail = lf[k];
      probs[k].ssa_lim_n = ln[k];
      probs[k].picard_resid = rs[k];
      probs[k].status = dn[k] < 0 ? YGPU_ERR_NONFINITE : YGPU_OK;
    }
    return YGPU_OK;
  }

  // device copy of the initial INOUT state, so a resident batch can be
  // solved again from the same start without touching the host
  std::vector<std::pair<R*, R*>> backup;  // (live, saved)
  // every member's warm flag (keep_factors): 1 when its next solve continues
  // the previous one on the same cells, so its factors are still in place
  int set_warm(int v) {
    std::vector<int> w(nmem, v);
    YG_CHECK(h, hipMemcpyAsync(d_warm, w.data(), sizeof(int) * nmem, hipMemcpyHostToDevice, h->stream));
    YG_CHECK(h, hipStreamSynchronize(h->stream));
    return YGPU_OK;
  }

  int save_initial() {
    std::vector<std::pair<R*, long long>> st = {
        {f.ux, total_cells}, {f.uy, total_cells}, {f.visc, total_z}, {f.beta, total_cells},
        {f.beta_acx, total_cells}, {f.beta_acy, total_cells}, {f.taub_x, total_cells},
        {f.taub_y, total_cells}};
    if (diva) {
      st.push_back({f.ubar_x, total_cells});
      st.push_back({f.ubar_y, total_cells});
      st.push_back({f.duxdz, total_z});
      st.push_back({f.duydz, total_z});
    }
    for (auto& e : st) {
      R* saved;
      YG_CHECK(h, ar.alloc(&saved, e.second));
      YG_CHECK(h, hipMemcpyAsync(saved, e.first, sizeof(R) * e.second, hipMemcpyDeviceToDevice, h->stream));
      backup.push_back({e.first, saved});
      sizes.push_back(e.second);
    }
    YG_CHECK(h, hipStreamSynchronize(h->stream));
    return YGPU_OK;
  }
  std::vector<long long> sizes;
  int restore_initial() {
    for (size_t q = 0; q < backup.size(); ++q)
      YG_CHECK(h, hipMemcpyAsync(backup[q].first, backup[q].second, sizeof(R) * sizes[q],
                                 hipMemcpyDeviceToDevice, h->stream));
    return YGPU_OK;
  }

  // ---- one Picard iteration, and a graph of several ---------------------
  // Every kernel skips members whose done flag is set and each member counts
  // its own iterations, so iterations can run back to back without looking
  // at the flags in between: a graph of `chunk` iterations replayed until all
  // members are done gives the same result as the loop that checks after
  // each iteration.  ev0/ev1 (may be null) bracket the Krylov launch.
  void enqueue_iteration(const YgBatchGeom& G, const YgPicardState& ps, const YgKrylovArgs& ka,
                         hipEvent_t ev0, hipEvent_t ev1, hipStream_t st) {
    if (diva) yg_launch_diva_pre<R>(G, f, dv, fp, d_done, st);
    yg_launch_fields_pre<R>(G, f, fp, d_done, st);
    yg_launch_assemble<R>(G, f, fp, d_done, d_A64, kmode == YG_MODE_MIXED ? d_A32 : nullptr, d_b,
                          d_x, d_mask, st);
    if (ev0) (void)hipEventRecord(ev0, st);
    if (group > 1) yg_launch_krylov_group(ka, ngroups, gthreads, st);
    else yg_launch_krylov(ka, kmode, nmem, block, st);
    if (ev1) (void)hipEventRecord(ev1, st);
    yg_launch_epilogue<R>(G, f, fp, d_x, d_kit, d_kst, iter_max, ps, st);
  }

  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int gchunk = 0;
  bool graph_failed = false;

  void drop_graph() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    gexec = nullptr;
    graph = nullptr;
  }
  ~Batch() { drop_graph(); }

  // Captures `chunk` iterations on the handle's stream, once per batch and
  // chunk size; false if the runtime refused (the eager loop is used then).
  // Kernel statistics time each Krylov launch with HIP events, which a
  // captured graph does not support, so a run with statistics stays eager.
  bool ensure_graph(const YgBatchGeom& G, const YgPicardState& ps, const YgKrylovArgs& ka,
                    int chunk) {
    if (graph_failed) return false;
    if (gexec && gchunk == chunk) return true;
    drop_graph();
    yg_krylov_prepare();
    hipStream_t st = h->stream;
    if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      graph_failed = true;
      return false;
    }
    for (int k = 0; k < chunk; ++k) enqueue_iteration(G, ps, ka, nullptr, nullptr, st);
    const hipError_t ec = hipStreamEndCapture(st, &graph);
    const hipError_t el = hipGetLastError();
    if (ec != hipSuccess || el != hipSuccess || !graph ||
        hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0) != hipSuccess) {
      graph_failed = true;
      drop_graph();
      return false;
    }
    gchunk = chunk;
    return true;
  }

  // y = A x0 over every member with the stencil records of the last
  // assembled systems (x0: the last solution), reps timed launches
  int spmv(int fp32, int reps, double* y_host, double* ms, double* bytes) {
    if (fp32 && !d_A32) { h->err = "spmv: FP32 records exist in the mixed mode only"; return YGPU_ERR_ARG; }
    hipStream_t st = h->stream;
    const YgBatchGeom G = geom();
    const int wrapx = (fp.BC == YG_BC_PERIODIC || fp.BC == YG_BC_PERIODIC_X);
    const int wrapy = (fp.BC == YG_BC_PERIODIC || fp.BC == YG_BC_PERIODIC_Y ||
                       fp.BC == YG_BC_MISMIP3D || fp.BC == YG_BC_TROUGH);
    const long long n = 2 * total_cells;
    Arena tmp;
    double* y64 = nullptr;
    float *x32 = nullptr, *y32 = nullptr;
    if (fp32) {
      YG_CHECK(h, tmp.alloc(&x32, n));
      YG_CHECK(h, tmp.alloc(&y32, n));
      yg_launch_to_float(d_x, x32, n, st);
    } else {
      YG_CHECK(h, tmp.alloc(&y64, n));
    }
    auto launch = [&]() {
      if (fp32) yg_launch_spmv<float>(G, d_A32, x32, y32, wrapx, wrapy, st);
      else yg_launch_spmv<double>(G, d_A64, d_x, y64, wrapx, wrapy, st);
    };
    launch();   // untimed first launch
    hipEvent_t e0, e1;
    YG_CHECK(h, hipEventCreate(&e0));
    YG_CHECK(h, hipEventCreate(&e1));
    YG_CHECK(h, hipEventRecord(e0, st));
    for (int r = 0; r < reps; ++r) launch();
    YG_CHECK(h, hipEventRecord(e1, st));
    YG_CHECK(h, hipGetLastError());
    YG_CHECK(h, hipEventSynchronize(e1));
    float t = 0.f;
    YG_CHECK(h, hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ms) *ms = reps > 0 ? t / reps : 0.0;
    // 18 matrix values, one velocity pair read and one written per cell
    if (bytes) *bytes = (double)total_cells * (fp32 ? 22.0 * 4 : 22.0 * 8);
    if (y_host) {
      if (fp32) {
        std::vector<float> yf(n);
        YG_CHECK(h, hipMemcpy(yf.data(), y32, sizeof(float) * n, hipMemcpyDeviceToHost));
        for (long long k = 0; k < n; ++k) y_host[k] = yf[k];
      } else {
        YG_CHECK(h, hipMemcpy(y_host, y64, sizeof(double) * n, hipMemcpyDeviceToHost));
      }
    }
    return YGPU_OK;
  }

  // ---- neff_hook (velocity_diva.f90:99-107, 396-400) ----------------------
  // The caller's problems, when their hooks are to be called during run().
  const ygpu_ssa_problem* hook_probs = nullptr;
  bool has_hooks() const {
    if (!hook_probs || !diva) return false;
    for (int k = 0; k < nmem; ++k) if (hook_probs[k].neff_hook) return true;
    return false;
  }

  template <typename T>
  static bool c_bed_changed(T* c, const T* c_nm1, long long n, double conv) {
    // sum(abs(c_bed - c_bed_nm1)) > ssa_iter_conv * max(sum(abs(c_bed)), TOL_UNDERFLOW),
    // summed in the caller's real type in array order
    T d = T(0), a = T(0);
    for (long long q = 0; q < n; ++q) {
      d += std::fabs(c[q] - c_nm1[q]);
      a += std::fabs(c[q]);
    }
    return d > (T)conv * std::max(a, (T)1e-15);
  }

  // After an iteration of the members in `ran`: the basal velocity goes to
  // the callers' ux_b/uy_b, each hooked member's hook updates its c_bed, the
  // new c_bed goes to the device, and a member whose c_bed moved by more than
  // the Picard tolerance is not done unless it reached ssa_iter_max.
  int apply_hooks(const std::vector<int>& ran, std::vector<int>& done) {
    int rc;
    if ((rc = down2(f.ux, YG_FIELDW_OF(hook_probs, ux_b).data(), 0))) return rc;
    if ((rc = down2(f.uy, YG_FIELDW_OF(hook_probs, uy_b).data(), 0))) return rc;
    std::vector<int> it(nmem);
    YG_CHECK(h, hipMemcpy(it.data(), d_iter_now, sizeof(int) * nmem, hipMemcpyDeviceToHost));
    bool flags = false;
    for (int k : ran) {
      const ygpu_ssa_problem& P = hook_probs[k];
      if (!P.neff_hook) continue;
      const long long n = (long long)P.nx * P.ny;
      void* cb = const_cast<void*>(P.c_bed);
      std::vector<unsigned char> nm1((size_t)n * P.real_bytes);
      memcpy(nm1.data(), cb, nm1.size());
      P.neff_hook(P.hook_ctx, P.nx, P.ny, P.real_bytes, cb, P.ux_b, P.uy_b);
      const bool moved = P.real_bytes == 4
                             ? c_bed_changed((float*)cb, (const float*)nm1.data(), n, fp.iter_conv)
                             : c_bed_changed((double*)cb, (const double*)nm1.data(), n, fp.iter_conv);
      if (moved && done[k] == 1 && it[k] < iter_max) { done[k] = 0; flags = true; }
    }
    if ((rc = up2((R*)f.c_bed, YG_FIELD_OF(hook_probs, c_bed).data(), 0, false))) return rc;
    if (flags)
      YG_CHECK(h, hipMemcpy(d_done, done.data(), sizeof(int) * nmem, hipMemcpyHostToDevice));
    return YGPU_OK;
  }

  // the Picard loop (velocity_ssa.f90:172-334 / velocity_diva.f90:195-404)
  int run(double* seconds) {
    hipStream_t st = h->stream;
    int rc;
    YG_CHECK(h, hipMemsetAsync(d_done, 0, sizeof(int) * nmem, st));
    YG_CHECK(h, hipMemsetAsync(d_lin_iter, 0, sizeof(int) * nmem, st));
    YG_CHECK(h, hipMemsetAsync(d_lin_fail, 0, sizeof(int) * nmem, st));
    YG_CHECK(h, hipMemsetAsync(d_iter_now, 0, sizeof(int) * nmem, st));
    YG_CHECK(h, hipMemsetAsync(d_resid, 0, sizeof(double) * nmem, st));
    if (f.asm_err) YG_CHECK(h, hipMemsetAsync(f.asm_err, 0, sizeof(int) * nmem, st));
    hipEvent_t e0, e1;
    YG_CHECK(h, hipEventCreate(&e0));
    YG_CHECK(h, hipEventCreate(&e1));
    YG_CHECK(h, hipEventRecord(e0, st));
    const YgBatchGeom G = geom();
    yg_launch_init<R>(G, f, fp, nullptr, st);
    YgPicardState ps{d_iter_now, d_done, d_resid, d_lin_iter, d_lin_fail, d_lim_n,
                     d_epart, d_eticket, YG_EPI_BLOCKS};
    YgKrylovArgs ka{};
    ka.members = d_mem; ka.soff = d_soff; ka.mask = d_mask;
    ka.A64 = d_A64; ka.A32 = d_A32; ka.F64 = d_F64; ka.F32 = d_F32; ka.F16 = d_F16;
    ka.b = d_b; ka.x = d_x;
    ka.W64 = d_W64; ka.W32 = d_W32; ka.total_cells = total_cells;
    ka.wrapx = (fp.BC == YG_BC_PERIODIC || fp.BC == YG_BC_PERIODIC_X);
    ka.wrapy = (fp.BC == YG_BC_PERIODIC || fp.BC == YG_BC_PERIODIC_Y || fp.BC == YG_BC_MISMIP3D ||
                fp.BC == YG_BC_TROUGH);
    ka.skip = d_done; ka.out_iters = d_kit; ka.out_status = d_kst; ka.out_resid = d_kres;
    ka.out_sweeps = d_ksw;
    ka.picard_resid = d_resid;
    ka.warm = d_warm; ka.iter_now = d_iter_now;
    ka.plain64 = kmode == YG_MODE_FP64;
    ka.opts = opts.k;
    ka.group = group; ka.nmembers = nmem; ka.gsync = d_gsync; ka.group_fail = d_gfail;
    if (vargroup) {
      ka.gtab = d_gtab; ka.gmoff = d_gmoff; ka.gmem = d_gmem;
      ka.nwg = gnwg; ka.ngrp = ngroups;
    }
    if (d_gfail) YG_CHECK(h, hipMemsetAsync(d_gfail, 0, sizeof(int), st));
    long long* d_tm = nullptr;
    const bool phase = getenv("YGPU_PHASE_TIMERS") != nullptr;
    if (phase) {
      YG_CHECK(h, hipMalloc(&d_tm, sizeof(long long) * 8 * nmem));
      YG_CHECK(h, hipMemsetAsync(d_tm, 0, sizeof(long long) * 8 * nmem, st));
      ka.timers = d_tm;
    }
    std::vector<int> done(nmem), kit(nmem), ksw(nmem);
    hipEvent_t k0 = nullptr, k1 = nullptr;
    if (h->stats_on) {
      YG_CHECK(h, hipEventCreate(&k0));
      YG_CHECK(h, hipEventCreate(&k1));
    }
    const bool hooked = has_hooks();
    const int chunk = (phase || h->stats_on || hooked) ? 0 : opts.k.graph_chunk;
    const bool graphed = chunk > 0 && ensure_graph(G, ps, ka, chunk);
    if (graphed) {
      // replays of the captured chunk, one flag read per replay
      for (int base = 0; base < iter_max; base += chunk) {
        YG_CHECK(h, hipGraphLaunch(gexec, st));
        YG_CHECK(h, hipMemcpyAsync(done.data(), d_done, sizeof(int) * nmem, hipMemcpyDeviceToHost, st));
        YG_CHECK(h, hipStreamSynchronize(st));
        bool all = true;
        for (int k = 0; k < nmem; ++k) all &= done[k] != 0;
        if (all) break;
      }
    }
    for (int iter = 1; !graphed && iter <= iter_max; ++iter) {
      std::vector<int> active;
      if (h->stats_on || hooked)
        for (int k = 0; k < nmem; ++k) if (!done[k]) active.push_back(k);
      enqueue_iteration(G, ps, ka, k0, k1, st);
      YG_CHECK(h, hipGetLastError());
      YG_CHECK(h, hipMemcpyAsync(done.data(), d_done, sizeof(int) * nmem, hipMemcpyDeviceToHost, st));
      if (h->stats_on) {
        YG_CHECK(h, hipMemcpyAsync(kit.data(), d_kit, sizeof(int) * nmem, hipMemcpyDeviceToHost, st));
        YG_CHECK(h, hipMemcpyAsync(ksw.data(), d_ksw, sizeof(int) * nmem, hipMemcpyDeviceToHost, st));
      }
      YG_CHECK(h, hipStreamSynchronize(st));
      if (h->stats_on) {
        float kms = 0.f;
        YG_CHECK(h, hipEventElapsedTime(&kms, k0, k1));
        h->stats.krylov_launches += 1;
        h->stats.krylov_group = group;
        h->stats.krylov_block = group > 1 ? gthreads : block;
        h->stats.krylov_ms += kms;
        const bool mixed = kmode == YG_MODE_MIXED;
        for (int k : active) {
          const double C = (double)mem[k].nx * mem[k].ny;
          const int it = kit[k];
          // bytes per factor value: BF16 L + U cells 16 + 24 B, FP32 L + BF16 U 32 + 24 B
          const double pF = opts.k.factor16 == 1 ? 40.0 / 18.0 : opts.k.factor16 == 2 ? 56.0 / 18.0 : 4.0;
          h->stats.krylov_bytes += mixed ? krylov_bytes(C, it, 4, pF, 4, opts.k.precon == 2) +
                                               refinement_bytes(C, ksw[k])
                                         : krylov_bytes(C, it, 8, 8, 8, opts.k.precon == 2);
          h->stats.refine_sweeps += mixed ? ksw[k] : 0;
          h->stats.krylov_iters += it;
          h->stats.unknown_iters += 2.0 * C * it;
          h->stats.picard_iters += 1;
        }
      }
      if (hooked && (rc = apply_hooks(active, done))) return rc;
      bool all = true;
      for (int k = 0; k < nmem; ++k) all &= done[k] != 0;
      if (all) break;
    }
    if (k0) (void)hipEventDestroy(k0);
    if (k1) (void)hipEventDestroy(k1);
    if (phase) {
      std::vector<long long> tmh(8 * nmem);
      YG_CHECK(h, hipMemcpy(tmh.data(), d_tm, sizeof(long long) * 8 * nmem, hipMemcpyDeviceToHost));
      (void)hipFree(d_tm);
      int rate = 0;
      (void)hipDeviceGetAttribute(&rate, hipDeviceAttributeWallClockRate, h->device);
      const double us = rate > 0 ? 1e3 / rate : 1e-2;  // rate in kHz
      fprintf(stderr, "ygpu phases member0 (ms): factor %.3f sweepL %.3f sweepU %.3f spmv %.3f "
              "vectors %.3f group-barrier waits %.3f total %.3f\n",
              tmh[0] * us * 1e-3, tmh[1] * us * 1e-3, tmh[2] * us * 1e-3, tmh[3] * us * 1e-3,
              tmh[4] * us * 1e-3, tmh[5] * us * 1e-3, tmh[7] * us * 1e-3);
      double tsum = 0.0, tmax = 0.0;
      for (int k = 0; k < nmem; ++k) {
        tsum += tmh[8 * k + 7];
        tmax = std::max(tmax, (double)tmh[8 * k + 7]);
      }
      fprintf(stderr, "ygpu phases: Krylov time per member (ms) mean %.3f max %.3f over %d members\n",
              tsum / nmem * us * 1e-3, tmax * us * 1e-3, nmem);
      if (getenv("YGPU_PHASE_TIMERS")[0] == '2') {
        fprintf(stderr, "ygpu phases: per member (ms):");
        for (int k = 0; k < nmem; ++k) fprintf(stderr, " %.2f", tmh[8 * k + 7] * us * 1e-3);
        fprintf(stderr, "\n");
      }
    }
    if (d_gfail) {
      int gf = 0;
      YG_CHECK(h, hipMemcpyAsync(&gf, d_gfail, sizeof(int), hipMemcpyDeviceToHost, st));
      YG_CHECK(h, hipStreamSynchronize(st));
      if (gf) {
        h->err = "member group: a workgroup waited more than 2 s for its group "
                 "(the group's workgroups were not resident together)";
        return YGPU_ERR_DEVICE;
      }
    }
    if (f.asm_err) {
      std::vector<int> bad(nmem);
      YG_CHECK(h, hipMemcpyAsync(bad.data(), f.asm_err, sizeof(int) * nmem, hipMemcpyDeviceToHost, st));
      YG_CHECK(h, hipStreamSynchronize(st));
      for (int k = 0; k < nmem; ++k)
        if (bad[k]) {
          h->err = "energy assembly: domain " + std::to_string(k) +
                   " has a row outside the 9-point stencil";
          return YGPU_ERR_UNSUPPORTED;
        }
    }
    yg_launch_finish<R>(G, f, fp, ps, nullptr, st);
    if (diva) yg_launch_diva_finish<R>(G, f, dv, fp, nullptr, st);
    YG_CHECK(h, hipGetLastError());
    YG_CHECK(h, hipEventRecord(e1, st));
    YG_CHECK(h, hipEventSynchronize(e1));
    float ms = 0.f;
    YG_CHECK(h, hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (seconds) *seconds = ms * 1e-3;
    return YGPU_OK;
  }
};

}  // namespace yg
