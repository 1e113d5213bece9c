// ygpu_api.hip -- host side of libygpu.so: option parsing, the one-shot,
// batch and session entry points of the C-ABI and the CSR solve.
#include "ygpu_host.h"

namespace yg {

// Option strings are read in place, one (key, value) pair of blank-separated
// tokens at a time, with no heap allocation on the success path (the
// no-allocation entry point parses with this too).
namespace {
struct Tok {
  const char* p;
  int n;
  bool is(const char* w) const { return (int)strlen(w) == n && strncmp(p, w, n) == 0; }
  int to_int() const { char b[32]; const int m = std::min(n, 31); memcpy(b, p, m); b[m] = 0; return atoi(b); }
  double to_double() const { char b[64]; const int m = std::min(n, 63); memcpy(b, p, m); b[m] = 0; return atof(b); }
};
bool next_tok(const char*& s, Tok& t) {
  while (*s == ' ' || *s == '\t' || *s == '\n') ++s;
  if (!*s) return false;
  t.p = s;
  while (*s && *s != ' ' && *s != '\t' && *s != '\n') ++s;
  t.n = (int)(s - t.p);
  return true;
}
}  // namespace

int parse_opts(const char* s, Opts& o, std::string& err) {
  o.k.solver = 4; o.k.precon = 0; o.k.maxiter = 1000; o.k.initx_zeros = 1; o.k.conv_cond = 0;
  o.k.ir_max = 60; o.k.tol = 1e-12; o.k.inner_tol = 1e-3; o.k.ilu_blocks = -1; o.k.ir_target = 0.0;
  o.k.ilu_cols = 1;
  o.k.factor16 = 0; o.k.graph_chunk = 4;
  o.k.ilu_rev = 0;
  o.k.ir_delta = 1e-3;   // reliable updates (-ygpu_ir restart: 0)
  o.k.ir_loose = -1.0;   // automatic: 10 ssa_iter_conv where a Picard loop runs
  o.k.loose_inner = 1e-2;
  o.k.keep_factors = 0;  // 1: warm-started solves begin on their previous solve's factors
  o.substituted = nullptr;
  o.fields32 = 1;   // the mixed mode's field kernels compute in FP32 unless -ygpu_fields fp64
  o.pattern = 0;
  if (!s) return 0;
  Tok k, v;
  while (next_tok(s, k)) {
    if (!next_tok(s, v)) { err = "option string: missing value"; return YGPU_ERR_ARG; }
    if (k.is("-i")) {
      if (v.is("cg") || v.is("1")) o.k.solver = 1;
      else if (v.is("bicgstab") || v.is("4")) o.k.solver = 4;
      // The reference's namelists also name LIS's MINRES and BiCGSafe
      // (residual-formulation defaults, input/yelmo_defaults.nml:160,
      // par/*.nml).  The residual matrix is nonsymmetric with a diagonal of
      // both signs, where MINRES's SPD-preconditioner recurrences do not
      // apply; both solve the same system to the same tolerance with
      // BiCGSTAB here, and the call leaves a note (ygpu_last_note) saying so.
      else if (v.is("minres")) { o.k.solver = 4; o.substituted = "-i minres solved with BiCGSTAB"; }
      else if (v.is("bicgsafe")) { o.k.solver = 4; o.substituted = "-i bicgsafe solved with BiCGSTAB"; }
      else { err = "unsupported Krylov solver: " + std::string(v.p, v.n); return YGPU_ERR_UNSUPPORTED; }
    } else if (k.is("-p")) {
      if (v.is("none") || v.is("0")) o.k.precon = 0;
      else if (v.is("jacobi") || v.is("1")) o.k.precon = 1;
      else if (v.is("ilu") || v.is("2")) o.k.precon = 2;
      else { err = "unsupported preconditioner: " + std::string(v.p, v.n); return YGPU_ERR_UNSUPPORTED; }
    } else if (k.is("-maxiter")) o.k.maxiter = v.to_int();
    else if (k.is("-tol")) o.k.tol = v.to_double();
    else if (k.is("-initx_zeros")) o.k.initx_zeros = (v.is("true") || v.is("1"));
    else if (k.is("-conv_cond")) o.k.conv_cond = v.to_int();
    else if (k.is("-ir_inner_tol")) o.k.inner_tol = v.to_double();
    else if (k.is("-ygpu_ir_tol")) o.k.ir_target = v.to_double();
    else if (k.is("-ir_max")) o.k.ir_max = v.to_int();
    else if (k.is("-ygpu_ir")) {
      if (v.is("reliable")) { if (o.k.ir_delta <= 0.0) o.k.ir_delta = 1e-3; }
      else if (v.is("restart")) o.k.ir_delta = 0.0;
      else { err = "-ygpu_ir takes reliable or restart"; return YGPU_ERR_ARG; }
    }
    else if (k.is("-ygpu_ir_loose")) {
      o.k.ir_loose = v.to_double();
      if (!(o.k.ir_loose >= 0.0)) { err = "-ygpu_ir_loose takes a value >= 0 (0: off)"; return YGPU_ERR_ARG; }
    }
    else if (k.is("-ygpu_ir_loose_inner")) {
      o.k.loose_inner = v.to_double();
      if (!(o.k.loose_inner > 0.0 && o.k.loose_inner < 1.0)) {
        err = "-ygpu_ir_loose_inner takes a value in (0, 1)";
        return YGPU_ERR_ARG;
      }
    }
    else if (k.is("-ygpu_ir_delta")) {
      o.k.ir_delta = v.to_double();
      if (!(o.k.ir_delta > 0.0 && o.k.ir_delta < 1.0)) {
        err = "-ygpu_ir_delta takes a value in (0, 1)";
        return YGPU_ERR_ARG;
      }
    }
    else if (k.is("-ygpu_keep_factors")) {
      if (v.is("1") || v.is("true")) o.k.keep_factors = 1;
      else if (v.is("0") || v.is("false")) o.k.keep_factors = 0;
      else { err = "-ygpu_keep_factors takes 1 or 0"; return YGPU_ERR_ARG; }
    }
    else if (k.is("-ygpu_ilu_blocks")) o.k.ilu_blocks = v.to_int();
    else if (k.is("-ygpu_ilu_cols")) o.k.ilu_cols = std::max(1, v.to_int());
    else if (k.is("-ygpu_ilu_order")) {
      if (v.is("natural")) o.k.ilu_rev = 0;
      else if (v.is("reverse")) o.k.ilu_rev = 1;
      else { err = "-ygpu_ilu_order takes natural or reverse"; return YGPU_ERR_ARG; }
    }
    else if (k.is("-ygpu_graph")) o.k.graph_chunk = std::max(0, v.to_int());
    else if (k.is("-ygpu_group")) o.group = v.to_int();
    else if (k.is("-ygpu_group_threads")) {
      o.group_threads = v.to_int();
      if (o.group_threads != 256 && o.group_threads != 512) {
        err = "-ygpu_group_threads takes 256 or 512";
        return YGPU_ERR_ARG;
      }
    }
    else if (k.is("-ygpu_factor")) {
      if (v.is("bf16")) o.k.factor16 = 1;
      else if (v.is("bf16u")) o.k.factor16 = 2;
      else if (v.is("fp32")) o.k.factor16 = 0;
      else { err = "unknown factor storage: " + std::string(v.p, v.n); return YGPU_ERR_ARG; }
    }
    else if (k.is("-ygpu_ny")) o.ny_hint = v.to_int();
    else if (k.is("-ygpu_inputs")) {
      if (v.is("device")) o.dev_inputs = 1;
      else if (v.is("host")) o.dev_inputs = 0;
      else { err = "-ygpu_inputs takes device or host"; return YGPU_ERR_ARG; }
    }
    else if (k.is("-ygpu_taud_lim")) o.taud_lim = v.to_double();
    else if (k.is("-ygpu_grad_lim")) o.grad_lim = v.to_double();
    else if (k.is("-ygpu_slope_bg_x")) o.slope_x = v.to_double();
    else if (k.is("-ygpu_slope_bg_y")) o.slope_y = v.to_double();
    else if (k.is("-ygpu_lat_bc")) {
      if (v.is("none")) o.lat_bc = 0;
      else if (v.is("floating") || v.is("float")) o.lat_bc = 1;
      else if (v.is("marine")) o.lat_bc = 2;
      else if (v.is("all")) o.lat_bc = 3;
      else { err = "-ygpu_lat_bc takes none, floating, marine or all"; return YGPU_ERR_ARG; }
    }
    else if (k.is("-ygpu_pattern")) {
      if (v.is("auto")) o.pattern = 0;
      else if (v.is("scalar")) o.pattern = 1;
      else if (v.is("ssa")) o.pattern = 2;
      else { err = "unknown pattern: " + std::string(v.p, v.n); return YGPU_ERR_ARG; }
    }
    else if (k.is("-ygpu_fields")) {
      if (v.is("fp32")) o.fields32 = 1;
      else if (v.is("fp64")) o.fields32 = 0;
      else { err = "unknown field precision: " + std::string(v.p, v.n); return YGPU_ERR_ARG; }
    }
    else if (k.is("-ygpu_precision")) {
      if (v.is("fp64")) o.prec = 0;
      else if (v.is("mixed")) o.prec = 2;
      else { err = "unknown precision: " + std::string(v.p, v.n); return YGPU_ERR_ARG; }
    } else if (k.is("-omp_num_threads") || k.is("-print") || k.is("-scale") || k.is("-f")) {
      // accepted; no effect here
    } else { err = "unknown option: " + std::string(k.p, k.n); return YGPU_ERR_ARG; }
  }
  if (o.k.maxiter < 0) { err = "negative -maxiter"; return YGPU_ERR_ARG; }
  return 0;
}

// Algorithmic HBM bytes of the persistent Krylov kernel for one member of C
// cells that ran `iters` BiCGSTAB iterations with a factorization: matrix
// slots pA, factors pF, vectors pv bytes per value.  Per iteration: two
// stencil SpMVs (18 slot values + 8 B of mask/coordinates per cell), two
// ILU(0) applications (18 factor values per cell), 27 vector passes of two
// values per cell (the fused updates and dot products).
double krylov_bytes(double C, int iters, int pA, double pF, int pv, bool ilu) {
  const double setup = 18.0 * C * pA + 8.0 * C + 4.0 * 2.0 * C * pv;
  const double fact = ilu ? (18.0 * C * pA + 4.0 * C + 18.0 * C * pF) : 0.0;
  const double per_it = 36.0 * C * pA + 16.0 * C + (ilu ? 36.0 * C * pF : 4.0 * C * pF) +
                        54.0 * C * pv;
  return setup + fact + iters * per_it;
}

// Per FP64 residual pass: the 18 FP64 slot values + 8 B of mask/coordinates
// per cell, x and b read, r and A x written (4 vectors of 2 FP64 values per
// cell).  Per correction: r read and the FP32 scaled residual and zeroed
// correction written, then the correction read and x read and written.
double refinement_bytes(double C, int nres) {
  const double res_pass = 18.0 * C * 8 + 8.0 * C + 4.0 * 2.0 * C * 8;
  const double corr = 2.0 * C * 8 + 2.0 * 2.0 * C * 4 + 2.0 * C * 4 + 2.0 * 2.0 * C * 8;
  return nres * res_pass + std::max(0, nres - 1) * corr;
}

// skewed step table of an nx x ny member: soff[t] = cells before step t
void step_table(int nx, int ny, std::vector<int>& soff, int& T) {
  T = ny + 2 * (nx - 1);
  soff.assign(T + 1, 0);
  for (int t = 0; t < T; ++t) {
    const int a = t - (ny - 1);
    const int glo = a <= 0 ? 0 : (a + 1) / 2;
    const int ghi = std::min(nx - 1, t / 2);
    soff[t + 1] = soff[t] + std::max(0, ghi - glo + 1);
  }
}


bool fields32(const ygpu_ssa_params* par) {
  if (par->precision == YGPU_PREC_REF) return true;
  if (par->precision != YGPU_PREC_MIXED) return false;
  Opts o;
  std::string e;
  return parse_opts(par->ssa_lis_opt, o, e) == 0 && o.fields32;
}

int bc_code(int bnd) {
  switch (bnd) {
    case YGPU_BND_ZEROS: return YG_BC_ZEROS;
    case YGPU_BND_INFINITE: case YGPU_BND_MASK: return YG_BC_INFINITE;
    case YGPU_BND_MISMIP3D: return YG_BC_MISMIP3D;
    case YGPU_BND_TROUGH: return YG_BC_TROUGH;
    case YGPU_BND_PERIODIC: return YG_BC_PERIODIC;
    case YGPU_BND_PERIODIC_X: return YG_BC_PERIODIC_X;
    case YGPU_BND_PERIODIC_Y: return YG_BC_PERIODIC_Y;
  }
  return -1;
}

}  // namespace yg

using namespace yg;

namespace {

template <typename R>
int run_batch(ygpu_handle* h, const ygpu_ssa_params* par, int np, ygpu_ssa_problem* probs, int diva) {
  Batch<R> B;
  int rc = B.setup(h, par, np, probs, diva);
  if (rc) return rc;
  if ((rc = B.upload_inputs(probs))) return rc;
  if ((rc = B.upload_state(probs))) return rc;
  double secs = 0.0;
  B.hook_probs = probs;
  if ((rc = B.run(&secs))) return rc;
  if ((rc = B.download(probs))) return rc;
  if (diva) {
    // interior velocity of each level: 3D velocity minus basal velocity
    for (int k = 0; k < np; ++k) {
      ygpu_ssa_problem& P = probs[k];
      if (!P.ux_i || !P.uy_i || !P.ux || !P.uy || !P.ux_b || !P.uy_b) continue;
      const long long nc = (long long)P.nx * P.ny;
      for (int z = 0; z < P.nz; ++z)
        for (long long c = 0; c < nc; ++c) {
          if (P.real_bytes == 4) {
            ((float*)P.ux_i)[c + z * nc] = ((float*)P.ux)[c + z * nc] - ((float*)P.ux_b)[c];
            ((float*)P.uy_i)[c + z * nc] = ((float*)P.uy)[c + z * nc] - ((float*)P.uy_b)[c];
          } else {
            ((double*)P.ux_i)[c + z * nc] = ((double*)P.ux)[c + z * nc] - ((double*)P.ux_b)[c];
            ((double*)P.uy_i)[c + z * nc] = ((double*)P.uy)[c + z * nc] - ((double*)P.uy_b)[c];
          }
        }
    }
  }
  for (int k = 0; k < np; ++k) probs[k].solve_seconds = secs;
  return YGPU_OK;
}

// CSR (1-based, linear_solver_class) -> stencil records in the row order.
// The grid's ny comes from the widest coupling (uy(i+1,j) is 2 ny + 1
// columns right of ux(i,j)) unless ny_hint is given.  need_diag: every row
// must hold its diagonal (a preconditioner divides by it).
int csr_records(ygpu_handle* h, int n, const int32_t* a_ptr1, const int32_t* a_idx1,
                const double* a_val, int ny_hint, bool need_diag, int& nx, int& ny,
                std::vector<double>& A, std::vector<uint32_t>& mask) {
  const int cells = n / 2;
  ny = ny_hint;
  if (ny <= 0) {
    long long maxd = 0;
    for (int r = 0; r < n; ++r)
      for (int k = a_ptr1[r] - 1; k < a_ptr1[r + 1] - 1; ++k) maxd = std::max<long long>(maxd, (long long)(a_idx1[k] - 1) - r);
    ny = maxd >= 2 ? (int)(maxd / 2) : cells;
  }
  if (ny <= 0 || cells % ny) { h->err = "cannot infer the grid of this CSR pattern (give -ygpu_ny)"; return YGPU_ERR_UNSUPPORTED; }
  nx = cells / ny;
  if (nx > YG_MAX_LINES || ny > 65535) { h->err = "grid too large"; return YGPU_ERR_UNSUPPORTED; }
  A.assign((size_t)YG_NA * cells, 0.0);
  mask.assign(cells, 0u);
  static const int SU[2][9][3] = {
      {{0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {1, 0, 1}, {1, -1, 1}, {0, -1, 1}},
      {{0, 0, 1}, {0, 1, 1}, {0, -1, 1}, {1, 0, 1}, {-1, 0, 1}, {0, 0, 0}, {0, 1, 0}, {-1, 1, 0}, {-1, 0, 0}}};
  for (int r = 0; r < n; ++r) {
    const int cell = r / 2, comp = r & 1;
    const int i = cell / ny, j = cell % ny;
    const int q = cell;   // vectors and the stencil matrix keep the row order
    for (int k = a_ptr1[r] - 1; k < a_ptr1[r + 1] - 1; ++k) {
      const int col = a_idx1[k] - 1;
      if (col < 0 || col >= n) { h->err = "column index out of range"; return YGPU_ERR_ARG; }
      const int c2 = col / 2, comp2 = col & 1;
      const int di = c2 / ny - i, dj = c2 % ny - j;
      int s = -1;
      for (int t = 0; t < 9; ++t)
        if (SU[comp][t][0] == di && SU[comp][t][1] == dj && SU[comp][t][2] == comp2) { s = t; break; }
      if (s < 0) { h->err = "entry outside the C-grid SSA stencil"; return YGPU_ERR_UNSUPPORTED; }
      A[(size_t)YG_NA * q + 2 * s + comp] += a_val[k];
      mask[q] |= 1u << (16 * comp + s);
    }
    if (need_diag && !((mask[q] >> (16 * comp)) & 1u)) {
      h->err = "row without a diagonal entry";
      return YGPU_ERR_UNSUPPORTED;
    }
  }
  return YGPU_OK;
}

int run_any(ygpu_handle* h, const ygpu_ssa_params* par, int np, ygpu_ssa_problem* probs, int diva) {
  if (!h || !par || np < 1 || !probs) return YGPU_ERR_ARG;
  h->err.clear(); h->note.clear();
  for (int k = 0; k < np; ++k)
    if (probs[k].neff_hook && !diva) {
      h->err = "neff_hook is an argument of calc_velocity_diva only";
      return YGPU_ERR_ARG;
    }
  if (hipSetDevice(h->device) != hipSuccess) { h->err = "hipSetDevice failed"; return YGPU_ERR_DEVICE; }
  if (par->precision != YGPU_PREC_FP64 && par->precision != YGPU_PREC_MIXED &&
      par->precision != YGPU_PREC_REF) {
    h->err = "unknown precision";
    return YGPU_ERR_ARG;
  }
  // FP32 fields in the reference mode (the reference's wp = sp) and the
  // mixed mode with -ygpu_fields fp32; FP64 otherwise
  if (!fields32(par)) return run_batch<double>(h, par, np, probs, diva);
  return run_batch<float>(h, par, np, probs, diva);
}

}  // namespace

// ===========================================================================
// C-ABI
extern "C" {

int ygpu_create(int32_t device, ygpu_handle** out) {
  if (!out) return YGPU_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return YGPU_ERR_DEVICE;
  if (device < 0 || device >= n) return YGPU_ERR_ARG;
  if (hipSetDevice(device) != hipSuccess) return YGPU_ERR_DEVICE;
  // the no-allocation entry point's scratch address on this device
  if (yg_noalloc_init(device) != 0) return YGPU_ERR_DEVICE;
  ygpu_handle* h = new ygpu_handle();
  h->device = device;
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
    delete h;
    return YGPU_ERR_DEVICE;
  }
  *out = h;
  return YGPU_OK;
}

int ygpu_destroy(ygpu_handle* h) {
  if (!h) return YGPU_ERR_ARG;
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return YGPU_OK;
}

const char* ygpu_last_error(const ygpu_handle* h) { return h ? h->err.c_str() : "null handle"; }

const char* ygpu_last_note(const ygpu_handle* h) { return h ? h->note.c_str() : ""; }

void ygpu_default_params(ygpu_ssa_params* p) {
  if (!p) return;
  memset(p, 0, sizeof(*p));
  p->ssa_solver = 1;   // "energy" (input/yelmo_defaults.nml:159)
  p->boundaries = YGPU_BND_ZEROS;
  p->visc_method = 1;
  p->beta_method = 1;
  p->beta_gl_scale = 0;
  p->beta_gl_stag = 1;
  p->ssa_vel_lim_method = 1;
  p->ssa_iter_max = 20;
  p->ssa_write_log = 0;
  p->precision = YGPU_PREC_REF;
  p->visc_const = 1e7; p->beta_const = 1e3; p->beta_q = 1.0; p->beta_u0 = 100.0;
  p->beta_gl_f = 1.0; p->H_grnd_lim = 500.0; p->beta_min = 100.0; p->eps_0 = 1e-6;
  p->ssa_vel_lim_tau = 1e5; p->ssa_vel_max = 1e4; p->ssa_iter_rel = 0.7; p->ssa_iter_conv = 1e-2;
  p->rho_ice = 910.0; p->rho_sw = 1028.0; p->g = 9.81;
  p->no_slip = 0;
  p->block = 0;
  // ssa_lis_opt_energy (input/yelmo_defaults.nml:161)
  strncpy(p->ssa_lis_opt, "-i cg -p jacobi -maxiter 200 -tol 1.0e-2 -initx_zeros false",
          sizeof(p->ssa_lis_opt) - 1);
}

int ygpu_velocity_ssa_batch(ygpu_handle* h, const ygpu_ssa_params* par, int32_t nprob,
                            ygpu_ssa_problem* probs) {
  return run_any(h, par, nprob, probs, 0);
}

int ygpu_velocity_diva_batch(ygpu_handle* h, const ygpu_ssa_params* par, int32_t nprob,
                             ygpu_ssa_problem* probs) {
  return run_any(h, par, nprob, probs, 1);
}

int ygpu_velocity_ssa(ygpu_handle* h, const ygpu_ssa_params* par, int32_t nx, int32_t ny,
                      int32_t nz, float* ux_b, float* uy_b, float* taub_acx, float* taub_acy,
                      float* visc_eff, float* visc_eff_int, int32_t* ssa_mask_acx,
                      int32_t* ssa_mask_acy, float* ssa_err_acx, float* ssa_err_acy,
                      int32_t* ssa_iter_now, int32_t* ssa_lin_iter, int32_t* ssa_lin_fail,
                      int32_t* ssa_lim_n, float* beta, float* beta_acx, float* beta_acy,
                      const float* c_bed, const float* f_slide, const float* taud_acx,
                      const float* taud_acy, const float* taul_int_acx,
                      const float* taul_int_acy, const float* H_ice, const float* f_ice,
                      const float* H_grnd, const float* f_grnd, const float* f_grnd_acx,
                      const float* f_grnd_acy, const float* ATT, const float* zeta_aa,
                      const float* z_sl, const float* z_bed, const float* z_srf, float dx,
                      float dy, float n_glen) {
  ygpu_ssa_problem P;
  memset(&P, 0, sizeof(P));
  P.nx = nx; P.ny = ny; P.nz = nz; P.real_bytes = 4;
  P.dx = dx; P.dy = dy; P.n_glen = n_glen;
  P.ux_b = ux_b; P.uy_b = uy_b; P.taub_acx = taub_acx; P.taub_acy = taub_acy;
  P.visc_eff = visc_eff; P.visc_eff_int = visc_eff_int;
  P.ssa_mask_acx = ssa_mask_acx; P.ssa_mask_acy = ssa_mask_acy;
  P.ssa_err_acx = ssa_err_acx; P.ssa_err_acy = ssa_err_acy;
  P.beta = beta; P.beta_acx = beta_acx; P.beta_acy = beta_acy;
  P.c_bed = c_bed; P.f_slide = f_slide; P.taud_acx = taud_acx; P.taud_acy = taud_acy;
  P.taul_int_acx = taul_int_acx; P.taul_int_acy = taul_int_acy; P.H_ice = H_ice; P.f_ice = f_ice;
  P.H_grnd = H_grnd; P.f_grnd = f_grnd; P.f_grnd_acx = f_grnd_acx; P.f_grnd_acy = f_grnd_acy;
  P.ATT = ATT; P.zeta_aa = zeta_aa; P.z_sl = z_sl; P.z_bed = z_bed; P.z_srf = z_srf;
  int rc = run_any(h, par, 1, &P, 0);
  if (rc) return rc;
  if (ssa_iter_now) *ssa_iter_now = P.ssa_iter_now;
  if (ssa_lin_iter) *ssa_lin_iter = P.ssa_lin_iter;
  if (ssa_lin_fail) *ssa_lin_fail = P.ssa_lin_fail;
  if (ssa_lim_n) *ssa_lim_n = P.ssa_lim_n;
  return P.status;
}

int ygpu_velocity_diva(ygpu_handle* h, const ygpu_ssa_params* par, int32_t nx, int32_t ny,
                       int32_t nz, float* ux, float* uy, float* ux_bar, float* uy_bar, float* ux_b,
                       float* uy_b, float* ux_i, float* uy_i, float* taub_acx, float* taub_acy,
                       float* beta, float* beta_acx, float* beta_acy, float* beta_eff,
                       float* visc_eff, float* visc_eff_int, float* duxdz, float* duydz,
                       int32_t* ssa_mask_acx, int32_t* ssa_mask_acy, float* ssa_err_acx,
                       float* ssa_err_acy, int32_t* ssa_iter_now, int32_t* ssa_lin_iter,
                       int32_t* ssa_lin_fail, int32_t* ssa_lim_n, float* c_bed,
                       const float* f_slide, const float* taud_acx, const float* taud_acy,
                       const float* taul_int_acx, const float* taul_int_acy, const float* H_ice,
                       const float* f_ice, const float* H_grnd, const float* f_grnd,
                       const float* f_grnd_acx, const float* f_grnd_acy, const float* ATT,
                       const float* zeta_aa, const float* z_sl, const float* z_bed,
                       const float* z_srf, float dx, float dy, float n_glen,
                       ygpu_neff_hook neff_hook, void* hook_ctx) {
  ygpu_ssa_problem P;
  memset(&P, 0, sizeof(P));
  P.nx = nx; P.ny = ny; P.nz = nz; P.real_bytes = 4;
  P.dx = dx; P.dy = dy; P.n_glen = n_glen;
  P.ux = ux; P.uy = uy; P.ux_bar = ux_bar; P.uy_bar = uy_bar; P.ux_b = ux_b; P.uy_b = uy_b;
  P.ux_i = ux_i; P.uy_i = uy_i; P.taub_acx = taub_acx; P.taub_acy = taub_acy;
  P.beta = beta; P.beta_acx = beta_acx; P.beta_acy = beta_acy; P.beta_eff = beta_eff;
  P.visc_eff = visc_eff; P.visc_eff_int = visc_eff_int; P.duxdz = duxdz; P.duydz = duydz;
  P.ssa_mask_acx = ssa_mask_acx; P.ssa_mask_acy = ssa_mask_acy;
  P.ssa_err_acx = ssa_err_acx; P.ssa_err_acy = ssa_err_acy;
  P.c_bed = c_bed; P.f_slide = f_slide; P.taud_acx = taud_acx; P.taud_acy = taud_acy;
  P.taul_int_acx = taul_int_acx; P.taul_int_acy = taul_int_acy; P.H_ice = H_ice; P.f_ice = f_ice;
  P.H_grnd = H_grnd; P.f_grnd = f_grnd; P.f_grnd_acx = f_grnd_acx; P.f_grnd_acy = f_grnd_acy;
  P.ATT = ATT; P.zeta_aa = zeta_aa; P.z_sl = z_sl; P.z_bed = z_bed; P.z_srf = z_srf;
  P.neff_hook = neff_hook; P.hook_ctx = hook_ctx;
  int rc = run_any(h, par, 1, &P, 1);
  if (rc) return rc;
  if (ssa_iter_now) *ssa_iter_now = P.ssa_iter_now;
  if (ssa_lin_iter) *ssa_lin_iter = P.ssa_lin_iter;
  if (ssa_lin_fail) *ssa_lin_fail = P.ssa_lin_fail;
  if (ssa_lim_n) *ssa_lim_n = P.ssa_lim_n;
  return P.status;
}

// ---- inner boundary -------------------------------------------------------
int ygpu_csr_solve(ygpu_handle* h, int32_t n, int32_t nnz, const int32_t* a_ptr1,
                   const int32_t* a_idx1, const double* a_val, const double* b, double* x_inout,
                   const char* opts, int32_t* iters, int32_t* status, double* rel_resid,
                   double* seconds) {
  if (!h) return YGPU_ERR_ARG;
  h->err.clear(); h->note.clear();
  if (n <= 0 || !a_ptr1 || !a_idx1 || !a_val || !b || !x_inout || !iters || !status ||
      !rel_resid) {
    h->err = "invalid arguments";
    return YGPU_ERR_ARG;
  }
  if (a_ptr1[0] != 1 || a_ptr1[n] - 1 != nnz || nnz < 0) {
    h->err = "a_ptr must be 1-based with a_ptr(n+1)-1 == nnz";
    return YGPU_ERR_ARG;
  }
  int maxrow = 0;
  for (int r = 0; r < n; ++r) {
    if (a_ptr1[r + 1] < a_ptr1[r]) { h->err = "a_ptr must be non-decreasing"; return YGPU_ERR_ARG; }
    maxrow = std::max(maxrow, a_ptr1[r + 1] - a_ptr1[r]);
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
  // Scalar five-point systems (the implicit advection of
  // solver_advection.f90, at most five entries per row) go to their own
  // solver; the C-grid SSA stencil has rows of up to nine entries over two
  // unknowns per cell.
  if (o.pattern == 1 || (o.pattern == 0 && maxrow <= 5))
    return scalar_csr_solve(h, n, nnz, a_ptr1, a_idx1, a_val, b, x_inout, o, iters, status,
                            rel_resid, seconds);
  if (n & 1) {
    h->err = "n must be even for the C-grid SSA stencil (two unknowns per cell)";
    return YGPU_ERR_ARG;
  }
  const int cells = n / 2;
  int nx = 0, ny = 0;
  std::vector<double> A;
  std::vector<uint32_t> mask;
  if ((rc = csr_records(h, n, a_ptr1, a_idx1, a_val, o.ny_hint, o.k.precon != 0, nx, ny, A, mask)))
    return rc;
  std::vector<int> soff;
  int T;
  step_table(nx, ny, soff, T);
  if (T > YG_MAX_STEPS) { h->err = "grid too large for the step table"; return YGPU_ERR_UNSUPPORTED; }
  std::vector<double> bs(b, b + n), xs(x_inout, x_inout + n);
  const int kmode = o.prec == 2 ? YG_MODE_MIXED : YG_MODE_FP64;
  if (hipSetDevice(h->device) != hipSuccess) { h->err = "hipSetDevice failed"; return YGPU_ERR_DEVICE; }
  Arena ar;
  YgMember M{};
  M.nx = nx; M.ny = ny; M.nz = 1; M.T = T; M.coff = 0; M.toff = 0; M.dx = M.dy = 1.0;
  M.woff = 0;
  M.vlen = (int)(((2LL * cells + 3) / 4) * 4);
  M.maxcnt = 0;
  for (int t = 0; t < T; ++t) M.maxcnt = std::max(M.maxcnt, soff[t + 1] - soff[t]);
  if (M.maxcnt > YG_MAX_STEP_CELLS) {
    h->err = "a wavefront step holds more cells than the LDS stage";
    return YGPU_ERR_UNSUPPORTED;
  }
  YgKrylovArgs ka{};
  YgMember* d_mem; int* d_soff; uint32_t* d_mask; double *d_A64, *d_b, *d_x, *d_F64, *d_W64;
  float *d_A32 = nullptr, *d_F32 = nullptr, *d_W32 = nullptr;
  int *d_it, *d_st; double* d_res;
  YG_CHECK(h, ar.alloc(&d_mem, 1)); YG_CHECK(h, ar.alloc(&d_soff, soff.size()));
  YG_CHECK(h, ar.alloc(&d_mask, cells));
  YG_CHECK(h, ar.alloc(&d_A64, A.size())); YG_CHECK(h, ar.alloc(&d_b, n)); YG_CHECK(h, ar.alloc(&d_x, n));
  YG_CHECK(h, ar.alloc(&d_F64, (size_t)YG_NF * cells)); YG_CHECK(h, ar.alloc(&d_W64, (size_t)YG_NW64 * M.vlen));
  YG_CHECK(h, ar.alloc(&d_it, 1)); YG_CHECK(h, ar.alloc(&d_st, 1)); YG_CHECK(h, ar.alloc(&d_res, 1));
  hipStream_t st = h->stream;
  YG_CHECK(h, hipMemcpyAsync(d_mem, &M, sizeof(M), hipMemcpyHostToDevice, st));
  YG_CHECK(h, hipMemcpyAsync(d_soff, soff.data(), sizeof(int) * soff.size(), hipMemcpyHostToDevice, st));
  YG_CHECK(h, hipMemcpyAsync(d_mask, mask.data(), sizeof(uint32_t) * cells, hipMemcpyHostToDevice, st));
  YG_CHECK(h, hipMemcpyAsync(d_A64, A.data(), sizeof(double) * A.size(), hipMemcpyHostToDevice, st));
  YG_CHECK(h, hipMemcpyAsync(d_b, bs.data(), sizeof(double) * n, hipMemcpyHostToDevice, st));
  YG_CHECK(h, hipMemcpyAsync(d_x, xs.data(), sizeof(double) * n, hipMemcpyHostToDevice, st));
  std::vector<float> A32;
  if (kmode == YG_MODE_MIXED) {
    A32.resize(A.size());
    for (size_t q = 0; q < A.size(); ++q) A32[q] = (float)A[q];
    YG_CHECK(h, ar.alloc(&d_A32, A32.size())); YG_CHECK(h, ar.alloc(&d_F32, (size_t)YG_NF * cells));
    YG_CHECK(h, ar.alloc(&d_W32, (size_t)YG_NW32 * M.vlen));
    if (o.k.factor16) YG_CHECK(h, ar.alloc(&ka.F16, (size_t)YG_NF16 * cells));
    YG_CHECK(h, hipMemcpyAsync(d_A32, A32.data(), sizeof(float) * A32.size(), hipMemcpyHostToDevice, st));
  }
  ka.members = d_mem; ka.soff = d_soff; ka.mask = d_mask;
  ka.A64 = d_A64; ka.A32 = d_A32; ka.F64 = d_F64; ka.F32 = d_F32; ka.b = d_b; ka.x = d_x;
  ka.W64 = d_W64; ka.W32 = d_W32; ka.total_cells = cells; ka.wrapx = 0; ka.wrapy = 0;
  ka.skip = nullptr; ka.out_iters = d_it; ka.out_status = d_st; ka.out_resid = d_res;
  ka.opts = o.k;
  hipEvent_t e0, e1;
  YG_CHECK(h, hipEventCreate(&e0)); YG_CHECK(h, hipEventCreate(&e1));
  YG_CHECK(h, hipEventRecord(e0, st));
  const int block = nx <= 512 ? 512 : std::min(1024, ((nx + 63) / 64) * 64);
  yg_launch_krylov(ka, kmode, 1, block, st);
  YG_CHECK(h, hipGetLastError());
  YG_CHECK(h, hipEventRecord(e1, st));
  YG_CHECK(h, hipEventSynchronize(e1));
  float ms = 0.f;
  YG_CHECK(h, hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0); (void)hipEventDestroy(e1);
  if (seconds) *seconds = ms * 1e-3;
  YG_CHECK(h, hipMemcpy(xs.data(), d_x, sizeof(double) * n, hipMemcpyDeviceToHost));
  int it = 0, stt = 0; double res = 0.0;
  YG_CHECK(h, hipMemcpy(&it, d_it, sizeof(int), hipMemcpyDeviceToHost));
  YG_CHECK(h, hipMemcpy(&stt, d_st, sizeof(int), hipMemcpyDeviceToHost));
  YG_CHECK(h, hipMemcpy(&res, d_res, sizeof(double), hipMemcpyDeviceToHost));
  for (int r = 0; r < n; ++r) x_inout[r] = xs[r];
  *iters = it; *status = stt; *rel_resid = res;
  return YGPU_OK;
}

}  // extern "C"

int ygpu_stats_enable(ygpu_handle* h, int32_t on) {
  if (!h) return YGPU_ERR_ARG;
  h->stats_on = on != 0;
  h->stats = ygpu_stats{};
  return YGPU_OK;
}

int ygpu_stats_get(ygpu_handle* h, ygpu_stats* out) {
  if (!h || !out) return YGPU_ERR_ARG;
  *out = h->stats;
  h->stats = ygpu_stats{};
  return YGPU_OK;
}

// ---- resident batches --------------------------------------------------------
struct YgBatchBase {
  virtual ~YgBatchBase() {}
  virtual int run(int reset, double* secs) = 0;
  virtual int download(ygpu_ssa_problem* probs) = 0;
  virtual int spmv(int fp32, int reps, double* y, double* ms, double* bytes) = 0;
  int64_t bytes = 0;
};

template <typename R>
struct YgBatchT : YgBatchBase {
  Batch<R> B;
  int run(int reset, double* secs) override {
    int rc;
    if (reset && (rc = B.restore_initial())) return rc;
    return B.run(secs);
  }
  int download(ygpu_ssa_problem* probs) override { return B.download(probs); }
  int spmv(int fp32, int reps, double* y, double* ms, double* bytes) override {
    return B.spmv(fp32, reps, y, ms, bytes);
  }
};

struct ygpu_batch {
  ygpu_handle* h;
  YgBatchBase* b;
};

// ---- sessions and admission ------------------------------------------------
struct YgSessionBase {
  virtual ~YgSessionBase() {}
  virtual int solve(ygpu_ssa_problem* P) = 0;
  int64_t bytes = 0;
};

template <typename R>
struct YgSessionT : YgSessionBase {
  Batch<R> B;
  bool first = true;
  int solve(ygpu_ssa_problem* P) override {
    int rc;
    if ((rc = B.upload_inputs(P))) return rc;
    // the device-resident velocity, viscosity, friction and (mixed mode)
    // ILU(0) factors are the warm start
    if (first && (rc = B.upload_state(P))) return rc;
    if ((rc = B.set_warm(first ? 0 : 1))) return rc;
    double secs = 0.0;
    if ((rc = B.run(&secs))) return rc;
    if ((rc = B.download(P))) return rc;
    P->solve_seconds = secs;
    first = false;
    return YGPU_OK;
  }
};

struct ygpu_session {
  ygpu_handle* h;
  YgSessionBase* s;
};

extern "C" {

int ygpu_session_create(ygpu_handle* h, const ygpu_ssa_params* par, const ygpu_ssa_problem* prob,
                        ygpu_session** out, int64_t* device_bytes) {
  if (!h || !par || !prob || !out) return YGPU_ERR_ARG;
  *out = nullptr;
  h->err.clear(); h->note.clear();
  if (prob->neff_hook) { h->err = "neff_hook: one-shot DIVA entry points only"; return YGPU_ERR_ARG; }
  if (hipSetDevice(h->device) != hipSuccess) { h->err = "hipSetDevice failed"; return YGPU_ERR_DEVICE; }
  YgSessionBase* sb = nullptr;
  int rc;
  if (!fields32(par)) {
    auto* t = new YgSessionT<double>();
    rc = t->B.setup(h, par, 1, prob, 0);
    t->bytes = (int64_t)t->B.ar.bytes;
    sb = t;
  } else {
    auto* t = new YgSessionT<float>();
    rc = t->B.setup(h, par, 1, prob, 0);
    t->bytes = (int64_t)t->B.ar.bytes;
    sb = t;
  }
  if (rc) { delete sb; return rc; }
  if (device_bytes) *device_bytes = sb->bytes;
  *out = new ygpu_session{h, sb};
  return YGPU_OK;
}

int ygpu_session_solve(ygpu_session* s, ygpu_ssa_problem* prob) {
  if (!s || !prob) return YGPU_ERR_ARG;
  s->h->err.clear();
  if (hipSetDevice(s->h->device) != hipSuccess) { s->h->err = "hipSetDevice failed"; return YGPU_ERR_DEVICE; }
  return s->s->solve(prob);
}

int ygpu_batch_create(ygpu_handle* h, const ygpu_ssa_params* par, int32_t nprob,
                      const ygpu_ssa_problem* probs, int32_t diva, ygpu_batch** out,
                      int64_t* device_bytes) {
  if (!h || !par || nprob < 1 || !probs || !out) return YGPU_ERR_ARG;
  *out = nullptr;
  h->err.clear(); h->note.clear();
  for (int k = 0; k < nprob; ++k)
    if (probs[k].neff_hook) { h->err = "neff_hook: one-shot DIVA entry points only"; return YGPU_ERR_ARG; }
  if (hipSetDevice(h->device) != hipSuccess) { h->err = "hipSetDevice failed"; return YGPU_ERR_DEVICE; }
  YgBatchBase* bb = nullptr;
  int rc;
  auto make = [&](auto* t) {
    int r = t->B.setup(h, par, nprob, probs, diva ? 1 : 0);
    if (!r) r = t->B.upload_inputs(probs);
    if (!r) r = t->B.upload_state(probs);
    if (!r) r = t->B.save_initial();
    t->bytes = (int64_t)t->B.ar.bytes;
    bb = t;
    return r;
  };
  if (!fields32(par)) rc = make(new YgBatchT<double>());
  else rc = make(new YgBatchT<float>());
  if (rc) { delete bb; return rc; }
  if (device_bytes) *device_bytes = bb->bytes;
  *out = new ygpu_batch{h, bb};
  return YGPU_OK;
}

int ygpu_batch_run(ygpu_batch* b, int32_t reset, double* seconds) {
  if (!b) return YGPU_ERR_ARG;
  b->h->err.clear();
  if (hipSetDevice(b->h->device) != hipSuccess) { b->h->err = "hipSetDevice failed"; return YGPU_ERR_DEVICE; }
  return b->b->run(reset, seconds);
}

int ygpu_batch_download(ygpu_batch* b, ygpu_ssa_problem* probs) {
  if (!b || !probs) return YGPU_ERR_ARG;
  return b->b->download(probs);
}

int ygpu_batch_spmv(ygpu_batch* b, int32_t fp32, int32_t reps, double* y, double* ms_per_launch,
                    double* bytes_per_launch) {
  if (!b || reps < 0) return YGPU_ERR_ARG;
  b->h->err.clear();
  if (hipSetDevice(b->h->device) != hipSuccess) { b->h->err = "hipSetDevice failed"; return YGPU_ERR_DEVICE; }
  return b->b->spmv(fp32 ? 1 : 0, reps, y, ms_per_launch, bytes_per_launch);
}

int ygpu_csr_spmv(ygpu_handle* h, int32_t n, int32_t nnz, const int32_t* a_ptr1,
                  const int32_t* a_idx1, const double* a_val, const double* x, double* y,
                  const char* opts) {
  if (!h) return YGPU_ERR_ARG;
  h->err.clear(); h->note.clear();
  if (n <= 0 || (n & 1) || !a_ptr1 || !a_idx1 || !a_val || !x || !y) {
    h->err = "invalid arguments (n must be even: two unknowns per cell)";
    return YGPU_ERR_ARG;
  }
  if (a_ptr1[0] != 1 || a_ptr1[n] - 1 != nnz || nnz < 0) {
    h->err = "a_ptr must be 1-based with a_ptr(n+1)-1 == nnz";
    return YGPU_ERR_ARG;
  }
  Opts o;
  std::string e;
  int rc = parse_opts(opts, o, e);
  if (rc) { h->err = e; return rc; }
  int nx = 0, ny = 0;
  std::vector<double> A;
  std::vector<uint32_t> mask;
  if ((rc = csr_records(h, n, a_ptr1, a_idx1, a_val, o.ny_hint, false, nx, ny, A, mask))) return rc;
  if (hipSetDevice(h->device) != hipSuccess) { h->err = "hipSetDevice failed"; return YGPU_ERR_DEVICE; }
  Arena ar;
  YgMember M{};
  M.nx = nx; M.ny = ny; M.nz = 1; M.dx = M.dy = 1.0;
  YgMember* d_mem; long long* d_coff; double *d_A, *d_x, *d_y;
  YG_CHECK(h, ar.alloc(&d_mem, 1));
  YG_CHECK(h, ar.alloc(&d_coff, 1));
  YG_CHECK(h, ar.alloc(&d_A, A.size()));
  YG_CHECK(h, ar.alloc(&d_x, n));
  YG_CHECK(h, ar.alloc(&d_y, n));
  hipStream_t st = h->stream;
  const long long zero = 0;
  YG_CHECK(h, hipMemcpyAsync(d_mem, &M, sizeof(M), hipMemcpyHostToDevice, st));
  YG_CHECK(h, hipMemcpyAsync(d_coff, &zero, sizeof(zero), hipMemcpyHostToDevice, st));
  YG_CHECK(h, hipMemcpyAsync(d_A, A.data(), sizeof(double) * A.size(), hipMemcpyHostToDevice, st));
  YG_CHECK(h, hipMemcpyAsync(d_x, x, sizeof(double) * n, hipMemcpyHostToDevice, st));
  YgBatchGeom G{};
  G.members = d_mem; G.coffs = d_coff; G.zoffs = d_coff; G.soff = nullptr; G.nmem = 1;
  G.total_cells = n / 2; G.seg = d_coff; G.segm = nullptr; G.nseg = 1;
  const long long coff0 = 0;
  const std::vector<int> ch = yg_chunk_members(&coff0, 1, n / 2);
  int* d_ch = nullptr;
  YG_CHECK(h, ar.alloc(&d_ch, ch.size()));
  YG_CHECK(h, hipMemcpyAsync(d_ch, ch.data(), sizeof(int) * ch.size(), hipMemcpyHostToDevice, st));
  G.chunk_seg = d_ch;
  // the CSR boundary is non-periodic: a wrapped coupling is outside the stencil above
  yg_launch_spmv<double>(G, d_A, d_x, d_y, 0, 0, st);
  YG_CHECK(h, hipGetLastError());
  YG_CHECK(h, hipMemcpyAsync(y, d_y, sizeof(double) * n, hipMemcpyDeviceToHost, st));
  YG_CHECK(h, hipStreamSynchronize(st));
  return YGPU_OK;
}

int ygpu_batch_destroy(ygpu_batch* b) {
  if (!b) return YGPU_ERR_ARG;
  delete b->b;
  delete b;
  return YGPU_OK;
}

int ygpu_session_destroy(ygpu_session* s) {
  if (!s) return YGPU_ERR_ARG;
  delete s->s;
  delete s;
  return YGPU_OK;
}

int64_t ygpu_footprint_bytes(const ygpu_ssa_params* par, int32_t nx, int32_t ny, int32_t nz) {
  if (!par) return -1;
  const int64_t C = (int64_t)nx * ny, Z = C * nz;
  const int64_t rb = fields32(par) ? 4 : 8;
  int64_t b = C * (4 + 4 + YG_NA * 8 + 2 * 8 + 2 * 8 + YG_NW64 * 2 * 8 + 8);
  if (par->precision == YGPU_PREC_MIXED) b += C * (YG_NA * 4 + YG_NF * 4 + YG_NW32 * 2 * 4);
  else b += C * YG_NF * 8;
  // 2D fields, ATT and the viscosity (3D), ATT^(-1/n) at the quadrature nodes
  // (visc_method 1), and 2 3D fields of margin
  b += rb * (C * 36 + Z * 2 + (par->visc_method == 1 ? 4 * Z : 0) + 2 * Z);
  if (par->ssa_solver == 1) b += C * 2 * 4;   // copy_from of the energy form
  return b;
}

}  // extern "C"
