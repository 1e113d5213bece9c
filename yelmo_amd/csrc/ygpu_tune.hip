// ygpu_tune.hip -- the autotuner of the mixed-precision Krylov schedule.
//
// For one domain, the persistent Krylov kernel's schedule is the number of
// ILU(0) line blocks (one wave sweeps each block; more, shorter blocks mean
// shorter dependency chains per wave but a weaker block-Jacobi
// preconditioner and so more Krylov iterations) and the threads per member
// block (512 or 1024, which bounds the block count).  With G workgroups per
// member (member groups), the blocks are split over the G workgroups (a
// multiple of G of them, 512 threads each).  ygpu_autotune times the first
// Picard solve of the caller's own domain (its first ten iterations: the
// solve from rest and warm-started ones) under each candidate schedule on
// the device, for one workgroup per member and for
// each group size a batch of such domains may choose, and records the
// fastest per (grid size, G) in the handle; batches, sessions and servers
// created on the handle afterwards give every member of that size the
// schedule of the group size they run with.  The timed batch holds copies
// of the domain that fill the CUs (up to 64, within about 2 GB), so that
// the member shares the device's memory system as in a full batch: a
// schedule timed alone on an idle device favours more, shorter line blocks
// than a full device does.
#include "ygpu_host.h"

using namespace yg;

namespace {

struct Candidate {
  int nb, block, ncol = 1;
  double ms = 0.0;
  int iters = 0, status = -1;
};

std::vector<Candidate> candidates(int nx, int G) {
  std::vector<Candidate> out;
  const int nb0 = (nx + 63) / 64;
  if (G > 1) {
    // multiples of G, at least one block per 64 lines, at most one per wave
    const int nbmax = std::min(G * (YG_GROUP_THREADS / 64), nx);
    const int first = ((nb0 + G - 1) / G) * G;
    for (int nb = first, k = 0; nb <= nbmax && k < 4; nb += G, ++k)
      out.push_back(Candidate{nb, YG_GROUP_THREADS});
    return out;
  }
  for (int block : {512, 1024}) {
    if (block < nb0 * 64) continue;
    const int nbmax = std::min(block / 64, nx);
    std::vector<int> nbs;
    for (int f : {2, 3, 4, 6, 8})   // nb0 * f / 2: 1x, 1.5x, 2x, 3x, 4x
      nbs.push_back(std::min(nbmax, std::max(nb0, nb0 * f / 2)));
    std::sort(nbs.begin(), nbs.end());
    nbs.erase(std::unique(nbs.begin(), nbs.end()), nbs.end());
    for (int nb : nbs) out.push_back(Candidate{nb, block});
  }
  return out;
}

constexpr int kTunePicard = 10;   // Picard iterations of a timed solve

// Times the candidates of one group size G on a batch of `copies` copies of
// the domain (so that the member shares the device as in a full batch): the
// batch's Picard solve (at most kTunePicard iterations) from the same start,
// per candidate; records the fastest; returns it in *win (status -1: none
// converged).
template <typename R>
int tune_group(ygpu_handle* h, const ygpu_ssa_params* par, const ygpu_ssa_problem* prob, int diva,
               int reps, int G, int copies, Candidate* win) {
  ygpu_ssa_params p = *par;
  p.block = 0;
  // the first Picard iterations: the solve from rest and warm-started ones
  p.ssa_iter_max = std::min(p.ssa_iter_max, kTunePicard);
  const std::string opt = std::string(par->ssa_lis_opt) + " -ygpu_group " + std::to_string(G);
  if (opt.size() >= sizeof(p.ssa_lis_opt)) {
    h->err = "autotune: option string too long";
    return YGPU_ERR_ARG;
  }
  std::snprintf(p.ssa_lis_opt, sizeof(p.ssa_lis_opt), "%s", opt.c_str());
  std::vector<ygpu_ssa_problem> probs(copies, *prob);
  Batch<R> B;
  int rc = B.setup(h, &p, copies, probs.data(), diva);
  if (rc) return rc;
  if (B.opts.k.precon != 2) {
    h->err = "autotune: the schedule is that of the ILU(0) preconditioner (-p ilu)";
    return YGPU_ERR_ARG;
  }
  if (B.group != G) { win->status = -1; return YGPU_OK; }   // G does not apply to this domain
  if ((rc = B.upload_inputs(probs.data()))) return rc;
  if ((rc = B.upload_state(probs.data()))) return rc;
  if ((rc = B.save_initial())) return rc;
  std::vector<Candidate> cand = candidates(prob->nx, G);
  bool warm = false;
  for (Candidate& c : cand) {
    for (YgMember& M : B.mem) { M.nb = c.nb; M.ncol = c.ncol; }
    YG_CHECK(h, hipMemcpyAsync(B.d_mem, B.mem.data(), sizeof(YgMember) * copies,
                               hipMemcpyHostToDevice, h->stream));
    if (G == 1 && B.block != c.block) {   // a captured graph holds the block size
      B.block = c.block;
      B.drop_graph();
    }
    double best = 1e300;
    for (int r = 0; r < reps + (warm ? 0 : 1); ++r) {
      // the whole Picard solve from the same start (its later, warm-started
      // linear solves favour other schedules than the first one alone)
      if ((rc = B.restore_initial())) return rc;
      double secs = 0.0;
      if ((rc = B.run(&secs))) return rc;
      if (warm) best = std::min(best, 1e3 * secs);
      warm = true;
    }
    int lin_it = 0, lin_fail = 0;
    YG_CHECK(h, hipMemcpy(&lin_it, B.d_lin_iter, sizeof(int), hipMemcpyDeviceToHost));
    YG_CHECK(h, hipMemcpy(&lin_fail, B.d_lin_fail, sizeof(int), hipMemcpyDeviceToHost));
    c.iters = lin_it;
    c.status = lin_fail == 0 ? 0 : 4;
    c.ms = best;
  }
  win->status = -1;
  for (const Candidate& c : cand)
    if (c.status == 0 && (win->status != 0 || c.ms < win->ms)) *win = c;
  if (win->status == 0) h->sched[{prob->nx, prob->ny, G}] = {win->nb, win->block, win->ncol};
  return YGPU_OK;
}

int tune(ygpu_handle* h, const ygpu_ssa_params* par, const ygpu_ssa_problem* prob, int diva,
         int reps, int32_t* nb_out, int32_t* block_out, double* ms_out) {
  if (par->precision != YGPU_PREC_MIXED) {
    h->err = "autotune: the schedule belongs to the mixed-precision path";
    return YGPU_ERR_ARG;
  }
  int ncu = 0;
  YG_CHECK(h, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, h->device));
  const long long cells = (long long)prob->nx * prob->ny;
  // copies that fill the device's CUs, within about 2 GB of device memory
  const long long by_mem = std::max(1LL, (2LL << 30) / (cells * 1400));
  Opts o;
  std::string e;
  int rc0 = parse_opts(par->ssa_lis_opt, o, e);
  if (rc0) { h->err = e; return rc0; }
  const bool group_ok = o.k.solver == 4 && o.k.precon == 2 && o.k.ilu_blocks != 0;
  Candidate win1{};
  win1.status = -1;
  for (int G : {1, 2, 4, 8, 16}) {
    // the group sizes pick_group may choose for this domain
    if (G > 1 && (!group_ok || cells / G < 4096 || 8 * G > ncu)) continue;
    const int copies = (int)std::min<long long>({(long long)ncu / G, by_mem, 64LL});
    Candidate w{};
    const int rc = fields32(par)
                       ? tune_group<float>(h, par, prob, diva, reps, G, std::max(1, copies), &w)
                       : tune_group<double>(h, par, prob, diva, reps, G, std::max(1, copies), &w);
    if (rc) return rc;
    if (G == 1) win1 = w;
  }
  if (win1.status != 0) {
    h->err = "autotune: no schedule converged on this domain";
    return YGPU_ERR_UNSUPPORTED;
  }
  if (nb_out) *nb_out = win1.nb;
  if (block_out) *block_out = win1.block;
  if (ms_out) *ms_out = win1.ms;
  return YGPU_OK;
}

}  // namespace

extern "C" {

int ygpu_autotune(ygpu_handle* h, const ygpu_ssa_params* par, const ygpu_ssa_problem* prob,
                  int32_t diva, int32_t reps, int32_t* ilu_blocks, int32_t* block,
                  double* solve_ms) {
  if (!h || !par || !prob) return YGPU_ERR_ARG;
  h->err.clear(); h->note.clear();
  if (hipSetDevice(h->device) != hipSuccess) { h->err = "hipSetDevice failed"; return YGPU_ERR_DEVICE; }
  return tune(h, par, prob, diva ? 1 : 0, reps < 1 ? 1 : reps, ilu_blocks, block, solve_ms);
}

// One schedule for every group size (the member-group kernel rounds the
// block count up to a multiple of its group size).
int ygpu_set_schedule(ygpu_handle* h, int32_t nx, int32_t ny, int32_t ilu_blocks, int32_t block) {
  if (!h || nx < 3 || ny < 3) return YGPU_ERR_ARG;
  if (block != 0 && block != 512 && block != 1024) return YGPU_ERR_ARG;
  for (int G : {1, 2, 4, 8, 16}) {
    if (ilu_blocks <= 0) h->sched.erase({nx, ny, G});
    else h->sched[{nx, ny, G}] = {ilu_blocks, G > 1 ? YG_GROUP_THREADS : block, 0};
  }
  return YGPU_OK;
}

}  // extern "C"
