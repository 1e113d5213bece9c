// ygpu_serve.hip -- continuous batching of served solve sequences.
//
// A server owns a fixed table of member slots and one device arena of cells
// whose size is set by an HBM budget.  A request is one solve of a
// sequence; it is admitted into a slot as soon as a slot and a large enough
// cell range are free, and from then on advances one Picard iteration per
// step in the same launches as every other admitted solve, so requests that
// arrive while others run overlap on the device.  A finished solve's
// outputs are downloaded at once; its slot keeps the velocity, viscosity and
// friction resident as the warm start of the sequence's next solve, until
// the solve submitted with YGPU_SEQ_LAST finishes.
//
// Per-member state is what makes this work with the batch kernels: every
// kernel skips members whose done flag is set, each member counts its own
// Picard iterations, and per-cell kernels find their member through the
// sorted list of occupied cell ranges, so slots can sit anywhere in the
// arena.  A member's arithmetic depends only on its own data and the block
// size, which the server fixes, so a solve's result is the same whatever
// else shares the device.
#include <deque>
#include <map>
#include <memory>
#include <unordered_map>

#include "ygpu_host.h"

using namespace yg;

namespace {

// Cell ranges are handed out in whole granules: 64 cells keep the work
// vector offset (2 values per cell) a multiple of 4 and the records aligned.
constexpr long long kGranule = 64;
// A new sequence that cannot be placed may be overtaken by smaller ones for
// this many steps; after that, admission of new sequences waits for it.
constexpr int kMaxBypassSteps = 8;

// first-fit allocator of cell ranges; released ranges merge with neighbours
struct CellRanges {
  std::map<long long, long long> free_;  // first cell -> length
  long long cap = 0, used = 0;
  void init(long long c) {
    free_.clear();
    if (c > 0) free_[0] = c;
    cap = c;
    used = 0;
  }
  long long take(long long n) {
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < n) continue;
      const long long s = it->first, len = it->second;
      free_.erase(it);
      if (len > n) free_[s + n] = len - n;
      used += n;
      return s;
    }
    return -1;
  }
  void give(long long s, long long n) {
    used -= n;
    auto it = free_.emplace(s, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
};

long long granules(long long cells) { return (cells + kGranule - 1) / kGranule * kGranule; }

// host array of float or double -> R
template <typename R>
void pack(R* dst, const void* src, long long n, int rb, R fill) {
  if (!src) {
    for (long long q = 0; q < n; ++q) dst[q] = fill;
  } else if (rb == 4) {
    const float* p = (const float*)src;
    for (long long q = 0; q < n; ++q) dst[q] = (R)p[q];
  } else {
    const double* p = (const double*)src;
    for (long long q = 0; q < n; ++q) dst[q] = (R)p[q];
  }
}

template <typename R>
void unpack(void* dst, const R* src, long long n, int rb) {
  if (!dst) return;
  if (rb == 4) {
    float* p = (float*)dst;
    for (long long q = 0; q < n; ++q) p[q] = (float)src[q];
  } else {
    double* p = (double*)dst;
    for (long long q = 0; q < n; ++q) p[q] = (double)src[q];
  }
}

// Fields that cross the boundary, in one table per direction so upload,
// download and the request buffers agree on the order.
using InPtr = const void* ygpu_ssa_problem::*;
using OutPtr = void* ygpu_ssa_problem::*;

const InPtr kIn2D[] = {&ygpu_ssa_problem::c_bed,        &ygpu_ssa_problem::f_slide,
                       &ygpu_ssa_problem::taud_acx,     &ygpu_ssa_problem::taud_acy,
                       &ygpu_ssa_problem::taul_int_acx, &ygpu_ssa_problem::taul_int_acy,
                       &ygpu_ssa_problem::H_ice,        &ygpu_ssa_problem::f_ice,
                       &ygpu_ssa_problem::H_grnd,       &ygpu_ssa_problem::f_grnd,
                       &ygpu_ssa_problem::z_bed,        &ygpu_ssa_problem::z_sl};
constexpr int kNIn2D = sizeof(kIn2D) / sizeof(kIn2D[0]);

// INOUT state taken from the first solve of a sequence: (field, may be null)
struct StateRef { OutPtr p; bool may_null; };
const StateRef kState2D[] = {{&ygpu_ssa_problem::ux_b, false},     {&ygpu_ssa_problem::uy_b, false},
                             {&ygpu_ssa_problem::taub_acx, true},  {&ygpu_ssa_problem::taub_acy, true},
                             {&ygpu_ssa_problem::beta, true},      {&ygpu_ssa_problem::beta_acx, true},
                             {&ygpu_ssa_problem::beta_acy, true}};
const StateRef kState2DDiva[] = {{&ygpu_ssa_problem::ux_bar, false}, {&ygpu_ssa_problem::uy_bar, false}};
const StateRef kState3DDiva[] = {{&ygpu_ssa_problem::duxdz, true}, {&ygpu_ssa_problem::duydz, true}};

const OutPtr kOut2D[] = {&ygpu_ssa_problem::ux_b,        &ygpu_ssa_problem::uy_b,
                         &ygpu_ssa_problem::taub_acx,    &ygpu_ssa_problem::taub_acy,
                         &ygpu_ssa_problem::visc_eff_int, &ygpu_ssa_problem::ssa_err_acx,
                         &ygpu_ssa_problem::ssa_err_acy, &ygpu_ssa_problem::beta,
                         &ygpu_ssa_problem::beta_acx,    &ygpu_ssa_problem::beta_acy};
constexpr int kNOut2D = sizeof(kOut2D) / sizeof(kOut2D[0]);
const OutPtr kOut2DDiva[] = {&ygpu_ssa_problem::ux_bar, &ygpu_ssa_problem::uy_bar,
                             &ygpu_ssa_problem::beta_eff};
const OutPtr kOut3DDiva[] = {&ygpu_ssa_problem::duxdz, &ygpu_ssa_problem::duydz,
                             &ygpu_ssa_problem::ux, &ygpu_ssa_problem::uy};

// one submitted solve: an owned copy of its inputs, later its results
template <typename R>
struct Req {
  int64_t ticket = 0, seq = 0;
  int flags = 0;
  int nx = 0, ny = 0, nz = 0;
  double dx = 0, dy = 0;
  long long nc = 0;
  bool first = false;          // first solve of its sequence: carries the state
  std::vector<R> in2d, att, zeta, st2d, st3d;
  std::vector<int32_t> masks;
  int waited = 0;              // steps spent unplaced (new sequences)
  std::chrono::steady_clock::time_point t_admit;
  // results
  std::vector<R> out2d, out3d;
  int iter_now = 0, lin_iter = 0, lin_fail = 0, lim_n = 0, status = YGPU_OK;
  int asm_bad = 0;             // energy assembly met a row outside the stencil
  double resid = 0.0, seconds = 0.0;
};

enum SlotState { kFree = 0, kRunning = 1, kResident = 2 };

struct Slot {
  int state = kFree;
  int64_t seq = 0;
  int64_t ticket = 0;
  long long coff = 0, ncell = 0;  // cell range of the arena
  int nx = 0, ny = 0, nz = 0;
  // host copies whose addresses feed asynchronous copies: kept alive here
  YgMember M{};
  long long zoff = 0;
  std::vector<int> st;
};

}  // namespace

struct YgServerBase {
  virtual ~YgServerBase() {}
  virtual int submit(const ygpu_ssa_problem* p, int64_t seq, int flags, int64_t* ticket) = 0;
  virtual int step(int32_t* n_running, int32_t* n_queued) = 0;
  virtual int poll(int64_t* ticket) = 0;
  virtual int fetch(int64_t ticket, ygpu_ssa_problem* out) = 0;
  virtual int64_t device_bytes() const = 0;
};

namespace {

template <typename R>
struct Server : YgServerBase {
  Batch<R> B;
  ygpu_handle* h = nullptr;
  int S = 0, nzmax = 0, maxside = 0;
  double n_glen = 0.0;
  bool have_glen = false;
  CellRanges cells;
  std::vector<Slot> slots;
  std::deque<std::unique_ptr<Req<R>>> pending;
  std::unordered_map<int64_t, std::unique_ptr<Req<R>>> running;  // ticket -> solve in a slot
  std::unordered_map<int64_t, std::unique_ptr<Req<R>>> results;  // finished, not fetched
  std::deque<int64_t> completed;
  // sequences: slot holding the state (-1 while none), dims, whether it ended
  struct Seq { int slot = -1; int nx = 0, ny = 0, nz = 0; bool ended = false; };
  std::unordered_map<int64_t, Seq> seqs;
  int64_t next_ticket = 1;
  // device per-slot selection flags and occupied-range list; host mirrors
  int* d_sel = nullptr;
  long long* d_seg = nullptr;
  int* d_segm = nullptr;
  std::vector<int> h_done, h_sel, h_zero_i, h_idle;
  std::vector<long long> h_seg;
  std::vector<int> h_segm;
  std::vector<double> h_zero_d;
  int nseg = 0;
  long long hi_cell = 0;

  int64_t device_bytes() const override { return (int64_t)B.ar.bytes; }

  // device bytes per arena cell at this precision and nz
  static double bytes_per_cell(const Batch<R>& b, int nz) {
    double c = 4.0 + YG_NA * 8.0 + 2 * 8.0 + 2 * 8.0 + YG_NW64 * 2 * 8.0 + 8.0;
    if (b.kmode == YG_MODE_MIXED) c += YG_NA * 4.0 + YG_NF * 4.0 + YG_NW32 * 2 * 4.0;
    else c += YG_NF * 8.0;
    if (b.fp.energy) c += 2 * 4.0;
    c += sizeof(R) * (34.0 + 3.0 * nz + (b.diva ? 8.0 + 4.0 * nz : 0.0));
    return c;
  }

  int create(ygpu_handle* hh, const ygpu_ssa_params* par, int diva, int nslots, int max_side,
             int nz_max, int64_t budget) {
    h = hh;
    int rc = B.setup_params(hh, par, diva);
    if (rc) return rc;
    if (nslots < 1 || max_side < 3 || max_side > YG_MAX_LINES || nz_max < 2 || budget <= 0) {
      h->err = "server: bad slot count, side, nz or budget";
      return YGPU_ERR_ARG;
    }
    S = nslots; nzmax = nz_max; maxside = max_side;
    const double bpc = bytes_per_cell(B, nz_max);
    long long cap = (long long)((double)budget / bpc) / kGranule * kGranule;
    // no more cells than the slots can occupy at once
    cap = std::min(cap, (long long)S * granules((long long)max_side * max_side));
    if (cap < granules((long long)max_side * max_side)) {
      h->err = "server: the HBM budget does not hold one domain of the largest side";
      return YGPU_ERR_ARG;
    }
    cells.init(cap);
    B.nmem = S;
    B.mem.assign(S, YgMember{});
    B.coffs.assign(S + 1, 0);
    B.zoffs.assign(S + 1, 0);
    B.soff.assign((size_t)S * (YG_MAX_STEPS + 1), 0);
    B.total_cells = cap;
    B.total_z = cap * nz_max;
    B.total_v = 2 * cap;
    B.zstride = nz_max;
    B.block = Batch<R>::pick_block(par, max_side);
    if ((rc = B.allocate())) return rc;
    YG_CHECK(h, B.ar.alloc(&d_sel, S));
    YG_CHECK(h, B.ar.alloc(&d_seg, S));
    YG_CHECK(h, B.ar.alloc(&d_segm, S));
    slots.assign(S, Slot{});
    h_done.assign(S, 3);
    h_sel.assign(S, 1);
    h_idle.assign(S, 3);
    h_zero_i.assign(S, 0);
    h_zero_d.assign(S, 0.0);
    // every slot starts idle: skipped by every kernel
    YG_CHECK(h, hipMemcpyAsync(B.d_done, h_idle.data(), sizeof(int) * S, hipMemcpyHostToDevice, h->stream));
    YG_CHECK(h, hipStreamSynchronize(h->stream));
    return YGPU_OK;
  }

  int submit(const ygpu_ssa_problem* p, int64_t seq, int flags, int64_t* ticket) override {
    if (!p || !ticket) return YGPU_ERR_ARG;
    auto it = seqs.find(seq);
    const bool cont = (flags & YGPU_SEQ_CONTINUE) != 0;
    if (cont && (it == seqs.end() || it->second.ended)) {
      h->err = "submit: sequence " + std::to_string(seq) + " has no state to continue from";
      return YGPU_ERR_ARG;
    }
    if (!cont && it != seqs.end()) {
      h->err = "submit: sequence id " + std::to_string(seq) + " is in use until its last solve finishes";
      return YGPU_ERR_ARG;
    }
    if (p->nx > maxside || p->ny > maxside || p->nz > nzmax) {
      h->err = "submit: domain larger than the server's maximum side or nz";
      return YGPU_ERR_ARG;
    }
    if (cont && (p->nx != it->second.nx || p->ny != it->second.ny || p->nz != it->second.nz)) {
      h->err = "submit: a sequence keeps its grid";
      return YGPU_ERR_ARG;
    }
    if (have_glen && p->n_glen != n_glen) {
      h->err = "submit: every request of a server shares n_glen";
      return YGPU_ERR_ARG;
    }
    YgMember M;
    std::vector<int> st;
    int rc = B.describe(*p, 0, M, st);
    if (rc) return rc;
    if (!p->ssa_mask_acx || !p->ssa_mask_acy || !p->zeta_aa || !p->ATT) {
      h->err = "submit: missing masks, zeta_aa or ATT";
      return YGPU_ERR_ARG;
    }
    for (InPtr f : kIn2D)
      if (!(p->*f)) { h->err = "submit: missing input array"; return YGPU_ERR_ARG; }
    auto r = std::make_unique<Req<R>>();
    r->seq = seq; r->flags = flags; r->nx = p->nx; r->ny = p->ny; r->nz = p->nz;
    r->dx = p->dx; r->dy = p->dy;
    const long long nc = (long long)p->nx * p->ny;
    r->nc = nc;
    const int rb = p->real_bytes;
    r->in2d.resize((size_t)kNIn2D * nc);
    for (int q = 0; q < kNIn2D; ++q) pack<R>(&r->in2d[(size_t)q * nc], p->*kIn2D[q], nc, rb, R(0));
    r->att.resize((size_t)nc * p->nz);
    pack<R>(r->att.data(), p->ATT, nc * p->nz, rb, R(0));
    r->zeta.resize(p->nz);
    pack<R>(r->zeta.data(), p->zeta_aa, p->nz, rb, R(0));
    r->masks.resize(2 * nc);
    memcpy(r->masks.data(), p->ssa_mask_acx, sizeof(int32_t) * nc);
    memcpy(r->masks.data() + nc, p->ssa_mask_acy, sizeof(int32_t) * nc);
    r->first = !cont;
    if (r->first) {
      std::vector<StateRef> s2(std::begin(kState2D), std::end(kState2D));
      if (B.diva) s2.insert(s2.end(), std::begin(kState2DDiva), std::end(kState2DDiva));
      r->st2d.resize(s2.size() * nc);
      for (size_t q = 0; q < s2.size(); ++q) {
        const void* src = p->*(s2[q].p);
        if (!src && !s2[q].may_null) { h->err = "submit: missing state array"; return YGPU_ERR_ARG; }
        pack<R>(&r->st2d[q * nc], src, nc, rb, R(0));
      }
      const long long nzc = nc * p->nz;
      const size_t n3 = B.diva ? 3 : 1;
      r->st3d.resize(n3 * nzc);
      if (!p->visc_eff) { h->err = "submit: missing visc_eff"; return YGPU_ERR_ARG; }
      pack<R>(r->st3d.data(), p->visc_eff, nzc, rb, R(0));
      if (B.diva)
        for (int q = 0; q < 2; ++q) pack<R>(&r->st3d[(q + 1) * nzc], p->*kState3DDiva[q].p, nzc, rb, R(0));
    }
    if (!have_glen) { n_glen = p->n_glen; have_glen = true; B.fp.n_glen = n_glen; }
    r->ticket = next_ticket++;
    *ticket = r->ticket;
    Seq& sq = seqs[seq];
    if (!cont) { sq.nx = p->nx; sq.ny = p->ny; sq.nz = p->nz; }
    if (flags & YGPU_SEQ_LAST) sq.ended = true;
    pending.push_back(std::move(r));
    return YGPU_OK;
  }

  // sorted list of occupied cell ranges -> device
  int publish_ranges() {
    std::vector<std::pair<long long, int>> occ;
    hi_cell = 0;
    for (int s = 0; s < S; ++s)
      if (slots[s].state != kFree) {
        occ.push_back({slots[s].coff, s});
        hi_cell = std::max(hi_cell, slots[s].coff + slots[s].ncell);
      }
    std::sort(occ.begin(), occ.end());
    nseg = (int)occ.size();
    h_seg.resize(std::max(nseg, 1));
    h_segm.resize(std::max(nseg, 1));
    for (int k = 0; k < nseg; ++k) { h_seg[k] = occ[k].first; h_segm[k] = occ[k].second; }
    if (nseg) {
      YG_CHECK(h, hipMemcpyAsync(d_seg, h_seg.data(), sizeof(long long) * nseg, hipMemcpyHostToDevice, h->stream));
      YG_CHECK(h, hipMemcpyAsync(d_segm, h_segm.data(), sizeof(int) * nseg, hipMemcpyHostToDevice, h->stream));
    }
    return YGPU_OK;
  }

  YgBatchGeom geom() const {
    YgBatchGeom G = B.geom();
    G.total_cells = hi_cell;
    G.seg = d_seg; G.segm = d_segm; G.nseg = nseg;
    return G;
  }

  // copies one admitted solve into its slot (asynchronous on the stream)
  int load(int s, Req<R>& r) {
    Slot& sl = slots[s];
    const long long c0 = sl.coff, nc = r.nc;
    hipStream_t st = h->stream;
    R* in_dev[kNIn2D] = {(R*)B.f.c_bed, (R*)B.f.f_slide, (R*)B.f.taud_acx,
                         (R*)B.f.taud_acy, (R*)B.f.taul_acx, (R*)B.f.taul_acy, (R*)B.f.H_ice,
                         (R*)B.f.f_ice, (R*)B.f.H_grnd, (R*)B.f.f_grnd, (R*)B.f.z_bed,
                         (R*)B.f.z_sl};
    for (int q = 0; q < kNIn2D; ++q)
      YG_CHECK(h, hipMemcpyAsync(in_dev[q] + c0, &r.in2d[(size_t)q * nc], sizeof(R) * nc,
                                 hipMemcpyHostToDevice, st));
    YG_CHECK(h, hipMemcpyAsync((R*)B.f.ATT + sl.zoff, r.att.data(), sizeof(R) * r.att.size(),
                               hipMemcpyHostToDevice, st));
    YG_CHECK(h, hipMemcpyAsync((R*)B.f.zeta + (size_t)s * B.zstride, r.zeta.data(),
                               sizeof(R) * r.zeta.size(), hipMemcpyHostToDevice, st));
    YG_CHECK(h, hipMemcpyAsync((int32_t*)B.f.mx + c0, r.masks.data(), sizeof(int32_t) * nc,
                               hipMemcpyHostToDevice, st));
    YG_CHECK(h, hipMemcpyAsync((int32_t*)B.f.my + c0, r.masks.data() + nc, sizeof(int32_t) * nc,
                               hipMemcpyHostToDevice, st));
    if (r.first) {
      std::vector<R*> s2 = {B.f.ux, B.f.uy, B.f.taub_x, B.f.taub_y, B.f.beta, B.f.beta_acx, B.f.beta_acy};
      if (B.diva) { s2.push_back(B.f.ubar_x); s2.push_back(B.f.ubar_y); }
      for (size_t q = 0; q < s2.size(); ++q)
        YG_CHECK(h, hipMemcpyAsync(s2[q] + c0, &r.st2d[q * nc], sizeof(R) * nc, hipMemcpyHostToDevice, st));
      const long long nzc = nc * r.nz;
      std::vector<R*> s3 = {B.f.visc};
      if (B.diva) { s3.push_back(B.f.duxdz); s3.push_back(B.f.duydz); }
      for (size_t q = 0; q < s3.size(); ++q)
        YG_CHECK(h, hipMemcpyAsync(s3[q] + sl.zoff, &r.st3d[q * nzc], sizeof(R) * nzc,
                                   hipMemcpyHostToDevice, st));
    }
    // member record, offsets and step table of the slot
    YG_CHECK(h, hipMemcpyAsync(B.d_mem + s, &sl.M, sizeof(YgMember), hipMemcpyHostToDevice, st));
    YG_CHECK(h, hipMemcpyAsync(B.d_coffs + s, &sl.coff, sizeof(long long), hipMemcpyHostToDevice, st));
    YG_CHECK(h, hipMemcpyAsync(B.d_zoffs + s, &sl.zoff, sizeof(long long), hipMemcpyHostToDevice, st));
    YG_CHECK(h, hipMemcpyAsync(B.d_soff + sl.M.toff, sl.st.data(), sizeof(int) * sl.st.size(),
                               hipMemcpyHostToDevice, st));
    // fresh Picard counters
    YG_CHECK(h, hipMemcpyAsync(B.d_done + s, &h_zero_i[s], sizeof(int), hipMemcpyHostToDevice, st));
    YG_CHECK(h, hipMemcpyAsync(B.d_iter_now + s, &h_zero_i[s], sizeof(int), hipMemcpyHostToDevice, st));
    YG_CHECK(h, hipMemcpyAsync(B.d_lin_iter + s, &h_zero_i[s], sizeof(int), hipMemcpyHostToDevice, st));
    YG_CHECK(h, hipMemcpyAsync(B.d_lin_fail + s, &h_zero_i[s], sizeof(int), hipMemcpyHostToDevice, st));
    YG_CHECK(h, hipMemcpyAsync(B.d_resid + s, &h_zero_d[s], sizeof(double), hipMemcpyHostToDevice, st));
    if (B.f.asm_err)
      YG_CHECK(h, hipMemcpyAsync(B.f.asm_err + s, &h_zero_i[s], sizeof(int), hipMemcpyHostToDevice, st));
    return YGPU_OK;
  }

  // admits queued solves into slots; returns the number admitted
  int admit(int* n_admitted) {
    *n_admitted = 0;
    bool blocked_new = false;
    std::vector<int> admitted;
    for (auto it = pending.begin(); it != pending.end();) {
      Req<R>& r = **it;
      Seq& sq = seqs[r.seq];
      int s = -1;
      if (!r.first) {
        // continuation: its sequence's slot must hold the previous result
        if (sq.slot < 0 || slots[sq.slot].state != kResident) { ++it; continue; }
        s = sq.slot;
      } else {
        if (blocked_new || sq.slot >= 0) { ++it; continue; }
        int fs = -1;
        for (int k = 0; k < S; ++k) if (slots[k].state == kFree) { fs = k; break; }
        const long long need = granules(r.nc);
        const long long c0 = fs >= 0 ? cells.take(need) : -1;
        if (fs < 0 || c0 < 0) {
          if (++r.waited > kMaxBypassSteps) blocked_new = true;
          ++it;
          continue;
        }
        s = fs;
        Slot& sl = slots[s];
        sl.coff = c0; sl.ncell = need; sl.seq = r.seq;
        sl.nx = r.nx; sl.ny = r.ny; sl.nz = r.nz;
        sl.zoff = c0 * nzmax;
        ygpu_ssa_problem P{};
        P.nx = r.nx; P.ny = r.ny; P.nz = r.nz; P.real_bytes = 8; P.dx = r.dx; P.dy = r.dy;
        int rc = B.describe(P, s, sl.M, sl.st);
        if (rc) return rc;
        sl.M.coff = c0;
        sl.M.toff = (long long)s * (YG_MAX_STEPS + 1);
        sl.M.woff = 2 * c0;
        sq.slot = s;
      }
      Slot& sl = slots[s];
      sl.state = kRunning;
      sl.ticket = r.ticket;
      int rc = load(s, r);
      if (rc) return rc;
      r.t_admit = std::chrono::steady_clock::now();
      admitted.push_back(s);
      running[r.ticket] = std::move(*it);
      it = pending.erase(it);
    }
    *n_admitted = (int)admitted.size();
    if (admitted.empty()) return YGPU_OK;
    int rc = publish_ranges();
    if (rc) return rc;
    // set_inactive_margins and the error fields of the admitted slots only
    std::fill(h_sel.begin(), h_sel.end(), 1);
    for (int s : admitted) h_sel[s] = 0;
    YG_CHECK(h, hipMemcpyAsync(d_sel, h_sel.data(), sizeof(int) * S, hipMemcpyHostToDevice, h->stream));
    yg_launch_init<R>(geom(), B.f, B.fp, d_sel, h->stream);
    YG_CHECK(h, hipGetLastError());
    return YGPU_OK;
  }

  // downloads the results of a finished slot into its request
  int collect(int s, Req<R>& r) {
    Slot& sl = slots[s];
    const long long c0 = sl.coff, nc = r.nc, nzc = nc * r.nz;
    hipStream_t st = h->stream;
    std::vector<const R*> o2 = {B.f.ux, B.f.uy, B.f.taub_x, B.f.taub_y, B.f.visc_int, B.f.err_x,
                                B.f.err_y, B.f.beta, B.f.beta_acx, B.f.beta_acy};
    if (B.diva) { o2.push_back(B.f.ubar_x); o2.push_back(B.f.ubar_y); o2.push_back(B.f.beta_eff); }
    std::vector<const R*> o3 = {B.f.visc};
    if (B.diva) { o3.push_back(B.f.duxdz); o3.push_back(B.f.duydz); o3.push_back(B.dv.ux3); o3.push_back(B.dv.uy3); }
    r.out2d.resize(o2.size() * nc);
    r.out3d.resize(o3.size() * nzc);
    for (size_t q = 0; q < o2.size(); ++q)
      YG_CHECK(h, hipMemcpyAsync(&r.out2d[q * nc], o2[q] + c0, sizeof(R) * nc, hipMemcpyDeviceToHost, st));
    for (size_t q = 0; q < o3.size(); ++q)
      YG_CHECK(h, hipMemcpyAsync(&r.out3d[q * nzc], o3[q] + sl.zoff, sizeof(R) * nzc, hipMemcpyDeviceToHost, st));
    YG_CHECK(h, hipMemcpyAsync(&r.iter_now, B.d_iter_now + s, sizeof(int), hipMemcpyDeviceToHost, st));
    YG_CHECK(h, hipMemcpyAsync(&r.lin_iter, B.d_lin_iter + s, sizeof(int), hipMemcpyDeviceToHost, st));
    YG_CHECK(h, hipMemcpyAsync(&r.lin_fail, B.d_lin_fail + s, sizeof(int), hipMemcpyDeviceToHost, st));
    YG_CHECK(h, hipMemcpyAsync(&r.lim_n, B.d_lim_n + s, sizeof(int), hipMemcpyDeviceToHost, st));
    YG_CHECK(h, hipMemcpyAsync(&r.resid, B.d_resid + s, sizeof(double), hipMemcpyDeviceToHost, st));
    r.asm_bad = 0;
    if (B.f.asm_err)
      YG_CHECK(h, hipMemcpyAsync(&r.asm_bad, B.f.asm_err + s, sizeof(int), hipMemcpyDeviceToHost, st));
    return YGPU_OK;
  }

  int step(int32_t* n_running, int32_t* n_queued) override {
    int rc, nadm = 0;
    if ((rc = admit(&nadm))) return rc;
    int nrun = 0;
    for (int s = 0; s < S; ++s) nrun += slots[s].state == kRunning;
    if (nrun > 0) {
      hipStream_t st = h->stream;
      const YgBatchGeom G = geom();
      YgPicardState ps{B.d_iter_now, B.d_done, B.d_resid, B.d_lin_iter, B.d_lin_fail, B.d_lim_n};
      YgKrylovArgs ka{};
      ka.members = B.d_mem; ka.soff = B.d_soff; ka.mask = B.d_mask;
      ka.A64 = B.d_A64; ka.A32 = B.d_A32; ka.F64 = B.d_F64; ka.F32 = B.d_F32;
      ka.b = B.d_b; ka.x = B.d_x; ka.W64 = B.d_W64; ka.W32 = B.d_W32;
      ka.total_cells = hi_cell;
      ka.wrapx = (B.fp.BC == YG_BC_PERIODIC || B.fp.BC == YG_BC_PERIODIC_X);
      ka.wrapy = (B.fp.BC == YG_BC_PERIODIC || B.fp.BC == YG_BC_PERIODIC_Y ||
                  B.fp.BC == YG_BC_MISMIP3D || B.fp.BC == YG_BC_TROUGH);
      ka.skip = B.d_done; ka.out_iters = B.d_kit; ka.out_status = B.d_kst; ka.out_resid = B.d_kres;
      ka.opts = B.opts.k;
      if (B.diva) yg_launch_diva_pre<R>(G, B.f, B.dv, B.fp, B.d_done, st);
      yg_launch_fields_pre<R>(G, B.f, B.fp, B.d_done, st);
      yg_launch_assemble<R>(G, B.f, B.fp, B.d_done, B.d_A64,
                            B.kmode == YG_MODE_MIXED ? B.d_A32 : nullptr, B.d_b, B.d_x, B.d_mask, st);
      yg_launch_krylov(ka, B.kmode, S, B.block, st);
      yg_launch_epilogue<R>(G, B.f, B.fp, B.d_x, B.d_kit, B.d_kst, B.iter_max, ps, st);
      YG_CHECK(h, hipGetLastError());
      YG_CHECK(h, hipMemcpyAsync(h_done.data(), B.d_done, sizeof(int) * S, hipMemcpyDeviceToHost, st));
      YG_CHECK(h, hipStreamSynchronize(st));
      std::vector<int> fin;
      for (int s = 0; s < S; ++s)
        if (slots[s].state == kRunning && h_done[s] != 0) fin.push_back(s);
      if (!fin.empty()) {
        std::fill(h_sel.begin(), h_sel.end(), 1);
        for (int s : fin) h_sel[s] = 0;
        YG_CHECK(h, hipMemcpyAsync(d_sel, h_sel.data(), sizeof(int) * S, hipMemcpyHostToDevice, st));
        yg_launch_finish<R>(G, B.f, B.fp, ps, d_sel, st);
        if (B.diva) yg_launch_diva_finish<R>(G, B.f, B.dv, B.fp, d_sel, st);
        YG_CHECK(h, hipGetLastError());
        for (int s : fin)
          if ((rc = collect(s, *running[slots[s].ticket]))) return rc;
        YG_CHECK(h, hipStreamSynchronize(st));
        const auto now = std::chrono::steady_clock::now();
        for (int s : fin) {
          Slot& sl = slots[s];
          auto node = running.find(sl.ticket);
          std::unique_ptr<Req<R>> r = std::move(node->second);
          running.erase(node);
          r->status = r->asm_bad ? YGPU_ERR_UNSUPPORTED
                                 : h_done[s] < 0 ? YGPU_ERR_NONFINITE : YGPU_OK;
          r->seconds = std::chrono::duration<double>(now - r->t_admit).count();
          if (r->flags & YGPU_SEQ_LAST) {
            cells.give(sl.coff, sl.ncell);
            sl = Slot{};
            seqs.erase(r->seq);
          } else {
            sl.state = kResident;
          }
          // finished members stay skipped (their done flag is set)
          completed.push_back(r->ticket);
          results[r->ticket] = std::move(r);
        }
        if ((rc = publish_ranges())) return rc;
      }
    }
    if (n_running) {
      int k = 0;
      for (int s = 0; s < S; ++s) k += slots[s].state == kRunning;
      *n_running = k;
    }
    if (n_queued) *n_queued = (int)pending.size();
    return YGPU_OK;
  }

  int poll(int64_t* ticket) override {
    if (!ticket) return YGPU_ERR_ARG;
    if (completed.empty()) return 0;
    *ticket = completed.front();
    completed.pop_front();
    return 1;
  }

  int fetch(int64_t ticket, ygpu_ssa_problem* out) override {
    auto it = results.find(ticket);
    if (it == results.end()) { h->err = "fetch: unknown or unfinished ticket"; return YGPU_ERR_ARG; }
    Req<R>& r = *it->second;
    if (!out || out->nx != r.nx || out->ny != r.ny || out->nz != r.nz ||
        (out->real_bytes != 4 && out->real_bytes != 8)) {
      h->err = "fetch: output arrays do not match the request's grid";
      return YGPU_ERR_ARG;
    }
    const int rb = out->real_bytes;
    const long long nc = r.nc, nzc = nc * r.nz;
    for (int q = 0; q < kNOut2D; ++q) unpack<R>(out->*kOut2D[q], &r.out2d[(size_t)q * nc], nc, rb);
    unpack<R>(out->visc_eff, r.out3d.data(), nzc, rb);
    if (B.diva) {
      for (int q = 0; q < 3; ++q)
        unpack<R>(out->*kOut2DDiva[q], &r.out2d[(size_t)(kNOut2D + q) * nc], nc, rb);
      for (int q = 0; q < 4; ++q)
        unpack<R>(out->*kOut3DDiva[q], &r.out3d[(size_t)(q + 1) * nzc], nzc, rb);
      // interior velocity: 3D velocity minus basal velocity (velocity_diva.f90:419-423)
      if (out->ux_i && out->uy_i) {
        const R* ux3 = &r.out3d[3 * nzc];
        const R* uy3 = &r.out3d[4 * nzc];
        std::vector<R> ui(nzc);
        for (long long c = 0; c < nzc; ++c) ui[c] = ux3[c] - r.out2d[c % nc];
        unpack<R>(out->ux_i, ui.data(), nzc, rb);
        for (long long c = 0; c < nzc; ++c) ui[c] = uy3[c] - r.out2d[nc + c % nc];
        unpack<R>(out->uy_i, ui.data(), nzc, rb);
      }
    }
    out->ssa_iter_now = r.iter_now;
    out->ssa_lin_iter = r.lin_iter;
    out->ssa_lin_fail = r.lin_fail;
    out->ssa_lim_n = r.lim_n;
    out->picard_resid = r.resid;
    out->status = r.status;
    out->solve_seconds = r.seconds;
    results.erase(it);
    return YGPU_OK;
  }
};

}  // namespace

struct ygpu_server {
  ygpu_handle* h;
  YgServerBase* s;
};

extern "C" {

int ygpu_server_create(ygpu_handle* h, const ygpu_ssa_params* par, int32_t diva, int32_t max_slots,
                       int32_t max_side, int32_t nz_max, int64_t hbm_budget_bytes,
                       ygpu_server** out, int64_t* device_bytes) {
  if (!h || !par || !out) return YGPU_ERR_ARG;
  *out = nullptr;
  h->err.clear();
  if (hipSetDevice(h->device) != hipSuccess) { h->err = "hipSetDevice failed"; return YGPU_ERR_DEVICE; }
  YgServerBase* sb = nullptr;
  int rc;
  if (par->precision != YGPU_PREC_REF) {
    auto* t = new Server<double>();
    sb = t;
    rc = t->create(h, par, diva ? 1 : 0, max_slots, max_side, nz_max, hbm_budget_bytes);
  } else {
    auto* t = new Server<float>();
    sb = t;
    rc = t->create(h, par, diva ? 1 : 0, max_slots, max_side, nz_max, hbm_budget_bytes);
  }
  if (rc) { delete sb; return rc; }
  if (device_bytes) *device_bytes = sb->device_bytes();
  *out = new ygpu_server{h, sb};
  return YGPU_OK;
}

int ygpu_submit(ygpu_server* s, const ygpu_ssa_problem* prob, int64_t seq, int32_t flags,
                int64_t* ticket) {
  if (!s) return YGPU_ERR_ARG;
  s->h->err.clear();
  return s->s->submit(prob, seq, flags, ticket);
}

int ygpu_server_step(ygpu_server* s, int32_t* n_running, int32_t* n_queued) {
  if (!s) return YGPU_ERR_ARG;
  s->h->err.clear();
  if (hipSetDevice(s->h->device) != hipSuccess) { s->h->err = "hipSetDevice failed"; return YGPU_ERR_DEVICE; }
  return s->s->step(n_running, n_queued);
}

int ygpu_poll(ygpu_server* s, int64_t* ticket) {
  if (!s) return YGPU_ERR_ARG;
  return s->s->poll(ticket);
}

int ygpu_fetch(ygpu_server* s, int64_t ticket, ygpu_ssa_problem* out) {
  if (!s) return YGPU_ERR_ARG;
  s->h->err.clear();
  return s->s->fetch(ticket, out);
}

int ygpu_server_destroy(ygpu_server* s) {
  if (!s) return YGPU_ERR_ARG;
  if (hipSetDevice(s->h->device) == hipSuccess) (void)hipStreamSynchronize(s->h->stream);
  delete s->s;
  delete s;
  return YGPU_OK;
}

}  // extern "C"
