// ygpu_serve.hip -- continuous batching of served solve sequences.
//
// A server owns member slots and one device arena of cells whose size is
// set by an HBM budget, split into lanes by domain size.  Each lane has its
// own slots, its own part of the arena and its own HIP stream.  A request is
// one solve of a sequence; it is admitted into a slot of the smallest lane
// that takes its grid (or a larger one with room) as soon as a slot and a
// large enough cell range are free.  From then on it advances one Picard
// iteration per lane step, in the same launches as every other admitted
// solve of its lane, so requests that arrive while others run overlap on the
// device.  Lanes step independently: a lane of small domains runs many
// Picard iterations while a lane of large domains runs one, and their
// kernels share the CUs.  A finished solve's outputs are downloaded at
// once; its slot keeps the velocity, viscosity and friction resident as the
// warm start of the sequence's next solve, until the solve submitted with
// YGPU_SEQ_LAST finishes.
//
// Per-member state is what makes this work with the batch kernels: every
// kernel skips members whose done flag is set, each member counts its own
// Picard iterations, and per-cell kernels find their member through the
// sorted list of occupied cell ranges, so slots can sit anywhere in a lane's
// arena.  A member's arithmetic depends only on its own data and the lane's
// block size, so a solve's result is the same whatever else shares the
// device.
#include <deque>
#include <map>
#include <thread>
#include <memory>
#include <unordered_map>

#include "ygpu_host.h"

using namespace yg;

namespace {

// Cell ranges are handed out in whole granules: 64 cells keep the work
// vector offset (2 values per cell) a multiple of 4 and the records aligned.
constexpr long long kGranule = 64;
// A new sequence that cannot be placed may be overtaken by later ones for
// this many lane steps; after that, no later sequence takes a lane it could
// use until it is placed.
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
                       &ygpu_ssa_problem::z_bed,        &ygpu_ssa_problem::z_sl,
                       &ygpu_ssa_problem::f_grnd_acx,   &ygpu_ssa_problem::f_grnd_acy};
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
  int lane = 0;
  int64_t seq = 0;
  int64_t ticket = 0;
  long long coff = 0, ncell = 0;  // cell range of the arena
  int nx = 0, ny = 0, nz = 0;
  // host copies whose addresses feed asynchronous copies: kept alive here
  YgMember M{};
  long long zoff = 0;
  std::vector<int> st;
};

// One size class: slots [s0, s0 + ns), cells [c0, c0 + cap), its stream.
struct Lane {
  int s0 = 0, ns = 0, side = 0, block = 512;
  long long c0 = 0, cap = 0;
  CellRanges cells;          // offsets relative to c0
  hipStream_t stream = nullptr;
  hipEvent_t done_ev = nullptr;
  int* h_done = nullptr;     // pinned: done flags of the lane's slots
  bool inflight = false;
  int nseg = 0;
  long long hi_cell = 0;     // one past the last occupied cell
  std::vector<long long> h_seg;
  std::vector<int> h_segm, h_sel;
  // member groups (the largest lane only): workgroups per member, groups,
  // their synchronization words and the wait-timeout flag (device, pinned)
  int group = 1, ngroups = 0;
  int *d_gsync = nullptr, *d_gfail = nullptr, *h_gfail = nullptr;
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
  std::vector<Lane> lanes;
  int gthreads = YG_GROUP_THREADS;   // threads per member-group workgroup (setup_groups)
  std::vector<Slot> slots;
  std::deque<std::unique_ptr<Req<R>>> pending;
  std::unordered_map<int64_t, std::unique_ptr<Req<R>>> running;  // ticket -> solve in a slot
  std::unordered_map<int64_t, std::unique_ptr<Req<R>>> results;  // finished, not fetched
  std::deque<int64_t> completed;
  // sequences: slot holding the state (-1 while none), dims, whether it ended
  struct Seq { int slot = -1; int nx = 0, ny = 0, nz = 0; bool ended = false; };
  std::unordered_map<int64_t, Seq> seqs;
  int64_t next_ticket = 1;
  // device per-slot selection flags and occupied-range lists (per lane)
  int* d_sel = nullptr;
  long long* d_seg = nullptr;
  int* d_segm = nullptr;
  std::vector<int> h_zero_i, h_idle, h_warm;
  std::vector<double> h_zero_d;

  int64_t device_bytes() const override { return (int64_t)B.ar.bytes; }

  ~Server() override {
    for (Lane& L : lanes) {
      if (L.stream) (void)hipStreamSynchronize(L.stream);
      if (L.done_ev) (void)hipEventDestroy(L.done_ev);
      if (L.stream) (void)hipStreamDestroy(L.stream);
      if (L.h_done) (void)hipHostFree(L.h_done);
      if (L.h_gfail) (void)hipHostFree(L.h_gfail);
    }
  }

  // device bytes per arena cell at this precision and nz
  static double bytes_per_cell(const Batch<R>& b, int nz) {
    double c = 4.0 + YG_NA * 8.0 + 2 * 8.0 + 2 * 8.0 + YG_NW64 * 2 * 8.0 + 8.0;
    if (b.kmode == YG_MODE_MIXED) c += YG_NA * 4.0 + YG_NF * 4.0 + YG_NW32 * 2 * 4.0;
    else c += YG_NF * 8.0;
    if (b.fp.energy) c += 2 * 4.0;
    if (b.kmode == YG_MODE_MIXED && b.opts.k.factor16) c += 2.0 * YG_NF16;
    c += sizeof(R) * (34.0 + 2.0 * nz + (b.diva ? 8.0 + 4.0 * nz : 0.0) +
                      (b.fp.visc_method == 1 ? 4.0 * nz : 0.0));
    return c;
  }

  int create(ygpu_handle* hh, const ygpu_ssa_params* par, int diva, int nlanes,
             const int32_t* lane_side, const int32_t* lane_slots, int nz_max, int64_t budget) {
    h = hh;
    int rc = B.setup_params(hh, par, diva);
    if (rc) return rc;
    if (B.opts.dev_inputs) {
      h->err = "-ygpu_inputs device: batches, sessions and one-shot calls only";
      return YGPU_ERR_UNSUPPORTED;
    }
    if (nlanes < 1 || nlanes > 8 || !lane_side || !lane_slots || nz_max < 2 || budget <= 0) {
      h->err = "server: bad lane list, nz or budget";
      return YGPU_ERR_ARG;
    }
    nzmax = nz_max;
    S = 0;
    double want = 0.0;  // cells the lanes need to fill every slot at full size
    for (int l = 0; l < nlanes; ++l) {
      if (lane_slots[l] < 1 || lane_side[l] < 3 || lane_side[l] > YG_MAX_LINES ||
          (l > 0 && lane_side[l] <= lane_side[l - 1])) {
        h->err = "server: lane sides must increase, each lane needs a slot";
        return YGPU_ERR_ARG;
      }
      S += lane_slots[l];
      want += (double)lane_slots[l] * granules((long long)lane_side[l] * lane_side[l]);
    }
    maxside = lane_side[nlanes - 1];
    const double bpc = bytes_per_cell(B, nz_max);
    const double have = (double)budget / bpc;   // cells the budget holds
    lanes.assign(nlanes, Lane{});
    long long c0 = 0;
    int s0 = 0;
    for (int l = 0; l < nlanes; ++l) {
      Lane& L = lanes[l];
      const long long full = (long long)lane_slots[l] * granules((long long)lane_side[l] * lane_side[l]);
      long long cap = want <= have ? full : (long long)(have * (double)full / want);
      cap = cap / kGranule * kGranule;
      if (cap < granules((long long)lane_side[l] * lane_side[l])) {
        h->err = "server: the HBM budget does not hold one domain of lane " + std::to_string(l);
        return YGPU_ERR_ARG;
      }
      L.s0 = s0; L.ns = lane_slots[l]; L.side = lane_side[l];
      L.c0 = c0; L.cap = cap;
      L.cells.init(cap);
      L.block = Batch<R>::pick_block(par, lane_side[l]);
      L.h_seg.assign(L.ns, 0);
      L.h_segm.assign(L.ns, 0);
      L.h_sel.assign(L.ns, 1);
      s0 += L.ns;
      c0 += cap;
    }
    B.nmem = S;
    B.mem.assign(S, YgMember{});
    B.coffs.assign(S + 1, 0);
    B.zoffs.assign(S + 1, 0);
    B.soff.assign((size_t)S * (YG_MAX_STEPS + 1), 0);
    B.total_cells = c0;
    B.total_z = c0 * nz_max;
    B.total_v = 2 * c0;
    B.zstride = nz_max;
    B.block = lanes.back().block;
    if ((rc = B.allocate())) return rc;
    YG_CHECK(h, B.ar.alloc(&d_sel, S));
    YG_CHECK(h, B.ar.alloc(&d_seg, S));
    YG_CHECK(h, B.ar.alloc(&d_segm, S));
    for (Lane& L : lanes) {
      YG_CHECK(h, hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking));
      YG_CHECK(h, hipEventCreateWithFlags(&L.done_ev, hipEventDisableTiming));
      YG_CHECK(h, hipHostMalloc((void**)&L.h_done, sizeof(int) * L.ns, hipHostMallocDefault));
    }
    if ((rc = setup_groups())) return rc;
    slots.assign(S, Slot{});
    for (Lane& L : lanes)
      for (int k = 0; k < L.ns; ++k) slots[L.s0 + k].lane = (int)(&L - lanes.data());
    h_idle.assign(S, 3);
    h_zero_i.assign(S, 0);
    h_warm.assign(S, 0);
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
      h->err = "submit: domain larger than the server's largest lane or nz";
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
    YgMember M{};
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

  // sorted list of a lane's occupied cell ranges -> device
  int publish_ranges(Lane& L) {
    std::vector<std::pair<long long, int>> occ;
    L.hi_cell = L.c0;
    for (int k = 0; k < L.ns; ++k) {
      const Slot& sl = slots[L.s0 + k];
      if (sl.state == kFree) continue;
      occ.push_back({sl.coff, k});
      L.hi_cell = std::max(L.hi_cell, sl.coff + sl.ncell);
    }
    std::sort(occ.begin(), occ.end());
    L.nseg = (int)occ.size();
    for (int k = 0; k < L.nseg; ++k) { L.h_seg[k] = occ[k].first; L.h_segm[k] = occ[k].second; }
    if (L.nseg) {
      YG_CHECK(h, hipMemcpyAsync(d_seg + L.s0, L.h_seg.data(), sizeof(long long) * L.nseg,
                                 hipMemcpyHostToDevice, L.stream));
      YG_CHECK(h, hipMemcpyAsync(d_segm + L.s0, L.h_segm.data(), sizeof(int) * L.nseg,
                                 hipMemcpyHostToDevice, L.stream));
    }
    return YGPU_OK;
  }

  // the lane's slots as a batch: member arrays offset by the lane's first slot
  YgBatchGeom geom(const Lane& L) const {
    YgBatchGeom G = B.geom();
    G.members = B.d_mem + L.s0;
    G.coffs = B.d_coffs + L.s0;
    G.zoffs = B.d_zoffs + L.s0;
    G.nmem = L.ns;
    G.total_cells = L.hi_cell;
    G.seg = d_seg + L.s0; G.segm = d_segm + L.s0; G.nseg = L.nseg;
    G.chunk_seg = nullptr;   // ranges change with admissions: searched
    return G;
  }
  YgFields<R> fields(const Lane& L) const {
    YgFields<R> f = B.f;
    f.zeta += (size_t)L.s0 * B.zstride;
    if (f.asm_err) f.asm_err += L.s0;
    return f;
  }
  YgPicardState pstate(const Lane& L) const {
    return YgPicardState{B.d_iter_now + L.s0, B.d_done + L.s0, B.d_resid + L.s0,
                         B.d_lin_iter + L.s0, B.d_lin_fail + L.s0, B.d_lim_n + L.s0,
                         B.d_epart + (size_t)L.s0 * YG_EPI_BLOCKS * 4, B.d_eticket + L.s0,
                         YG_EPI_BLOCKS};
  }

  // copies one admitted solve into its slot (asynchronous on the lane's stream)
  int load(int s, Req<R>& r) {
    Slot& sl = slots[s];
    const long long c0 = sl.coff, nc = r.nc;
    hipStream_t st = lanes[sl.lane].stream;
    R* in_dev[kNIn2D] = {(R*)B.f.c_bed, (R*)B.f.f_slide, (R*)B.f.taud_acx,
                         (R*)B.f.taud_acy, (R*)B.f.taul_acx, (R*)B.f.taul_acy, (R*)B.f.H_ice,
                         (R*)B.f.f_ice, (R*)B.f.H_grnd, (R*)B.f.f_grnd, (R*)B.f.z_bed,
                         (R*)B.f.z_sl, (R*)B.f.f_grnd_acx, (R*)B.f.f_grnd_acy};
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
    // a continuation finds its previous solve's factors in the slot
    h_warm[s] = r.first ? 0 : 1;
    YG_CHECK(h, hipMemcpyAsync(B.d_warm + s, &h_warm[s], sizeof(int), hipMemcpyHostToDevice, st));
    YG_CHECK(h, hipMemcpyAsync(B.d_lin_iter + s, &h_zero_i[s], sizeof(int), hipMemcpyHostToDevice, st));
    YG_CHECK(h, hipMemcpyAsync(B.d_lin_fail + s, &h_zero_i[s], sizeof(int), hipMemcpyHostToDevice, st));
    YG_CHECK(h, hipMemcpyAsync(B.d_resid + s, &h_zero_d[s], sizeof(double), hipMemcpyHostToDevice, st));
    if (B.f.asm_err)
      YG_CHECK(h, hipMemcpyAsync(B.f.asm_err + s, &h_zero_i[s], sizeof(int), hipMemcpyHostToDevice, st));
    return YGPU_OK;
  }

  // lanes a grid may use, smallest first
  int first_lane(int nx, int ny) const {
    const int side = std::max(nx, ny);
    for (size_t l = 0; l < lanes.size(); ++l)
      if (lanes[l].side >= side) return (int)l;
    return -1;
  }

  // admits queued solves into slots of idle lanes; count_wait: a lane step
  // ended since the last call (the unit of a request's waiting time)
  int admit(bool count_wait) {
    int block_from = (int)lanes.size();   // lanes held for a long-waiting request
    std::vector<std::vector<int>> admitted(lanes.size());
    for (auto it = pending.begin(); it != pending.end();) {
      Req<R>& r = **it;
      Seq& sq = seqs[r.seq];
      int s = -1;
      if (!r.first) {
        // continuation: its sequence's slot must hold the previous result,
        // and its lane must be between steps
        if (sq.slot < 0 || slots[sq.slot].state != kResident ||
            lanes[slots[sq.slot].lane].inflight) { ++it; continue; }
        s = sq.slot;
      } else {
        if (sq.slot >= 0) { ++it; continue; }
        const long long need = granules(r.nc);
        const int l0 = first_lane(r.nx, r.ny);
        for (int l = l0; l >= 0 && l < block_from && s < 0; ++l) {
          Lane& L = lanes[l];
          if (L.inflight) continue;
          int fs = -1;
          for (int k = 0; k < L.ns; ++k)
            if (slots[L.s0 + k].state == kFree) { fs = L.s0 + k; break; }
          if (fs < 0) continue;
          const long long off = L.cells.take(need);
          if (off < 0) continue;
          s = fs;
          Slot& sl = slots[s];
          sl.coff = L.c0 + off; sl.ncell = need; sl.seq = r.seq;
          sl.nx = r.nx; sl.ny = r.ny; sl.nz = r.nz;
          sl.zoff = sl.coff * nzmax;
          ygpu_ssa_problem P{};
          P.nx = r.nx; P.ny = r.ny; P.nz = r.nz; P.real_bytes = 8; P.dx = r.dx; P.dy = r.dy;
          B.group = L.group;   // the schedule of the lane's group size
          int rc = B.describe(P, s, sl.M, sl.st);
          B.group = 1;
          if (rc) return rc;
          sl.M.coff = sl.coff;
          sl.M.toff = (long long)s * (YG_MAX_STEPS + 1);
          sl.M.woff = 2 * sl.coff;
          sq.slot = s;
        }
        if (s < 0) {
          if (count_wait) ++r.waited;
          if (r.waited > kMaxBypassSteps) block_from = std::min(block_from, l0);
          ++it;
          continue;
        }
      }
      Slot& sl = slots[s];
      sl.state = kRunning;
      sl.ticket = r.ticket;
      int rc = load(s, r);
      if (rc) return rc;
      r.t_admit = std::chrono::steady_clock::now();
      admitted[sl.lane].push_back(s);
      running[r.ticket] = std::move(*it);
      it = pending.erase(it);
    }
    for (size_t l = 0; l < lanes.size(); ++l) {
      if (admitted[l].empty()) continue;
      Lane& L = lanes[l];
      int rc = publish_ranges(L);
      if (rc) return rc;
      // set_inactive_margins and the error fields of the admitted slots only
      std::fill(L.h_sel.begin(), L.h_sel.end(), 1);
      for (int s : admitted[l]) L.h_sel[s - L.s0] = 0;
      YG_CHECK(h, hipMemcpyAsync(d_sel + L.s0, L.h_sel.data(), sizeof(int) * L.ns,
                                 hipMemcpyHostToDevice, L.stream));
      yg_launch_init<R>(geom(L), fields(L), B.fp, d_sel + L.s0, L.stream);
      YG_CHECK(h, hipGetLastError());
      // the lane waits for its host copies before the next admission reuses them
      YG_CHECK(h, hipStreamSynchronize(L.stream));
    }
    return YGPU_OK;
  }

  // Member groups (mixed precision, BiCGSTAB, line-block ILU(0)): G
  // workgroups per member, every member of a lane with its lane's G.
  // -ygpu_group G sets G for the largest lane.  By default each lane takes G
  // from the batch rule for one domain of its largest size, largest lane
  // first, as long as the Krylov launches of all lanes (slots rounded up to 8,
  // times G workgroups, each holding a CU) fit the CUs at once.  Then every
  // workgroup of every group can be resident together, so no group waits for
  // a CU that another group holds, and every other kernel ends without
  // waiting for anything.
  int setup_groups() {
    const int want = B.opts.group;
    const bool fits = Batch<R>::group_solver_fits(B.kmode, B.opts.k);
    if (want == 1 || !fits) return YGPU_OK;
    if (want != 0 && want != 2 && want != 4 && want != 8) {
      h->err = "-ygpu_group takes 0, 1, 2, 4 or 8";
      return YGPU_ERR_ARG;
    }
    int ncu = 0;
    YG_CHECK(h, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, h->device));
    // threads per group workgroup: what a batch of one such member takes
    // (256 unless -ygpu_group_threads says otherwise), so a served sequence
    // gives the bits of the same sequence solved alone in a session; the CU
    // budget below still counts every workgroup as holding a whole CU
    gthreads = B.opts.group_threads ? B.opts.group_threads : YG_GROUP_THREADS;
    auto groups8 = [](const Lane& L) { return ((L.ns + 7) / 8) * 8; };
    auto enable = [&](Lane& L, int G) -> int {
      L.group = G;
      L.ngroups = groups8(L);
      YG_CHECK(h, B.ar.alloc(&L.d_gsync, (size_t)L.ngroups * (YG_GSYNC_BYTES / 4)));
      YG_CHECK(h, B.ar.alloc(&L.d_gfail, 1));
      YG_CHECK(h, hipMemsetAsync(L.d_gfail, 0, sizeof(int), h->stream));
      YG_CHECK(h, hipHostMalloc((void**)&L.h_gfail, sizeof(int), hipHostMallocDefault));
      *L.h_gfail = 0;
      return YGPU_OK;
    };
    auto fits_lines = [&](const Lane& L, int G) { return L.side <= 64 * G * (gthreads / 64); };
    if (want != 0) {
      Lane& L = lanes.back();
      int G = want;
      while (G > 1 && groups8(L) * G > ncu) G /= 2;
      if (G >= 2 && fits_lines(L, G)) {
        int rc = enable(L, G);
        if (rc) return rc;
      }
    } else {
      int used = 0;   // Krylov workgroups of all lanes per launch round
      for (const Lane& L : lanes) used += groups8(L);
      for (auto it = lanes.rbegin(); it != lanes.rend(); ++it) {
        Lane& L = *it;
        const long long cells = (long long)L.side * L.side;
        int G = 1;
        for (int g : {8, 4, 2})
          if (8 * g <= ncu && cells / g >= 4096) { G = g; break; }
        while (G > 1 && used + groups8(L) * (G - 1) > ncu) G /= 2;
        if (G < 2 || !fits_lines(L, G)) continue;
        used += groups8(L) * (G - 1);
        int rc = enable(L, G);
        if (rc) return rc;
      }
    }
    YG_CHECK(h, hipStreamSynchronize(h->stream));
    return YGPU_OK;
  }

  // one Picard iteration of every running solve of the lane, then an
  // asynchronous read of its done flags
  int launch(Lane& L) {
    hipStream_t st = L.stream;
    const YgBatchGeom G = geom(L);
    const YgFields<R> f = fields(L);
    const YgPicardState ps = pstate(L);
    YgKrylovArgs ka{};
    ka.members = B.d_mem + L.s0; ka.soff = B.d_soff; ka.mask = B.d_mask;
    ka.A64 = B.d_A64; ka.A32 = B.d_A32; ka.F64 = B.d_F64; ka.F32 = B.d_F32; ka.F16 = B.d_F16;
    ka.b = B.d_b; ka.x = B.d_x; ka.W64 = B.d_W64; ka.W32 = B.d_W32;
    ka.total_cells = L.hi_cell;
    ka.wrapx = (B.fp.BC == YG_BC_PERIODIC || B.fp.BC == YG_BC_PERIODIC_X);
    ka.wrapy = (B.fp.BC == YG_BC_PERIODIC || B.fp.BC == YG_BC_PERIODIC_Y ||
                B.fp.BC == YG_BC_MISMIP3D || B.fp.BC == YG_BC_TROUGH);
    ka.skip = ps.done; ka.out_iters = B.d_kit + L.s0; ka.out_status = B.d_kst + L.s0;
    ka.out_resid = B.d_kres + L.s0;
    ka.picard_resid = ps.resid;
    ka.warm = B.d_warm + L.s0; ka.iter_now = ps.iter_now;
    ka.plain64 = B.kmode == YG_MODE_FP64;
    ka.opts = B.opts.k;
    if (B.diva) yg_launch_diva_pre<R>(G, f, B.dv, B.fp, ps.done, st);
    yg_launch_fields_pre<R>(G, f, B.fp, ps.done, st);
    yg_launch_assemble<R>(G, f, B.fp, ps.done, B.d_A64,
                          B.kmode == YG_MODE_MIXED ? B.d_A32 : nullptr, B.d_b, B.d_x, B.d_mask, st);
    if (L.group > 1) {
      ka.group = L.group; ka.nmembers = L.ns; ka.gsync = L.d_gsync; ka.group_fail = L.d_gfail;
      yg_launch_krylov_group(ka, L.ngroups, gthreads, st);
    } else {
      yg_launch_krylov(ka, B.kmode, L.ns, L.block, st);
    }
    yg_launch_epilogue<R>(G, f, B.fp, B.d_x, B.d_kit + L.s0, B.d_kst + L.s0, B.iter_max, ps, st);
    YG_CHECK(h, hipGetLastError());
    YG_CHECK(h, hipMemcpyAsync(L.h_done, ps.done, sizeof(int) * L.ns, hipMemcpyDeviceToHost, st));
    if (L.h_gfail)
      YG_CHECK(h, hipMemcpyAsync(L.h_gfail, L.d_gfail, sizeof(int), hipMemcpyDeviceToHost, st));
    YG_CHECK(h, hipEventRecord(L.done_ev, st));
    L.inflight = true;
    return YGPU_OK;
  }

  // downloads the results of a finished slot into its request
  int collect(int s, Req<R>& r, hipStream_t st) {
    Slot& sl = slots[s];
    const long long c0 = sl.coff, nc = r.nc, nzc = nc * r.nz;
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

  // a lane's step has ended: finish and download the solves that completed
  int complete(Lane& L) {
    L.inflight = false;
    if (L.h_gfail && *L.h_gfail) {
      h->err = "member group: a workgroup waited more than 2 s for its group";
      return YGPU_ERR_DEVICE;
    }
    std::vector<int> fin;
    for (int k = 0; k < L.ns; ++k)
      if (slots[L.s0 + k].state == kRunning && L.h_done[k] != 0) fin.push_back(L.s0 + k);
    if (fin.empty()) return YGPU_OK;
    hipStream_t st = L.stream;
    std::fill(L.h_sel.begin(), L.h_sel.end(), 1);
    for (int s : fin) L.h_sel[s - L.s0] = 0;
    YG_CHECK(h, hipMemcpyAsync(d_sel + L.s0, L.h_sel.data(), sizeof(int) * L.ns, hipMemcpyHostToDevice, st));
    const YgBatchGeom G = geom(L);
    const YgFields<R> f = fields(L);
    yg_launch_finish<R>(G, f, B.fp, pstate(L), d_sel + L.s0, st);
    if (B.diva) yg_launch_diva_finish<R>(G, f, B.dv, B.fp, d_sel + L.s0, st);
    YG_CHECK(h, hipGetLastError());
    int rc;
    for (int s : fin)
      if ((rc = collect(s, *running[slots[s].ticket], st))) return rc;
    YG_CHECK(h, hipStreamSynchronize(st));
    const auto now = std::chrono::steady_clock::now();
    bool freed = false;
    for (int s : fin) {
      Slot& sl = slots[s];
      auto node = running.find(sl.ticket);
      std::unique_ptr<Req<R>> r = std::move(node->second);
      running.erase(node);
      r->status = r->asm_bad ? YGPU_ERR_UNSUPPORTED
                             : L.h_done[s - L.s0] < 0 ? YGPU_ERR_NONFINITE : YGPU_OK;
      r->seconds = std::chrono::duration<double>(now - r->t_admit).count();
      if (r->flags & YGPU_SEQ_LAST) {
        L.cells.give(sl.coff - L.c0, sl.ncell);
        const int lane = sl.lane;
        sl = Slot{};
        sl.lane = lane;
        seqs.erase(r->seq);
        freed = true;
      } else {
        sl.state = kResident;
      }
      // finished members stay skipped (their done flag is set)
      completed.push_back(r->ticket);
      results[r->ticket] = std::move(r);
    }
    if (freed && (rc = publish_ranges(L))) return rc;
    return YGPU_OK;
  }

  bool lane_has_work(const Lane& L) const {
    for (int k = 0; k < L.ns; ++k)
      if (slots[L.s0 + k].state == kRunning) return true;
    return false;
  }

  // Advances every lane by at least nothing and returns once one lane has
  // finished a Picard iteration (or when nothing runs).
  int step(int32_t* n_running, int32_t* n_queued) override {
    int rc;
    bool progressed = false;
    for (int spin = 0;; ++spin) {
      bool ended = spin == 0;
      for (Lane& L : lanes) {
        if (!L.inflight) continue;
        const hipError_t q = hipEventQuery(L.done_ev);
        if (q == hipErrorNotReady) continue;
        YG_CHECK(h, q);
        if ((rc = complete(L))) return rc;
        progressed = ended = true;
      }
      if ((rc = admit(ended))) return rc;
      bool any_inflight = false;
      for (Lane& L : lanes) {
        if (!L.inflight && lane_has_work(L) && (rc = launch(L))) return rc;
        any_inflight |= L.inflight;
      }
      if (progressed || !any_inflight) break;
      // a short back-off: the lanes' kernels run for milliseconds
      std::this_thread::sleep_for(std::chrono::microseconds(spin < 100 ? 20 : 200));
    }
    if (n_running) {
      int k = 0;
      for (const Slot& sl : slots) k += sl.state == kRunning;
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

int ygpu_server_create(ygpu_handle* h, const ygpu_ssa_params* par, int32_t diva, int32_t nlanes,
                       const int32_t* lane_max_side, const int32_t* lane_slots, int32_t nz_max,
                       int64_t hbm_budget_bytes, ygpu_server** out, int64_t* device_bytes) {
  if (!h || !par || !out) return YGPU_ERR_ARG;
  *out = nullptr;
  h->err.clear(); h->note.clear();
  if (hipSetDevice(h->device) != hipSuccess) { h->err = "hipSetDevice failed"; return YGPU_ERR_DEVICE; }
  YgServerBase* sb = nullptr;
  int rc;
  if (!fields32(par)) {
    auto* t = new Server<double>();
    sb = t;
    rc = t->create(h, par, diva ? 1 : 0, nlanes, lane_max_side, lane_slots, nz_max, hbm_budget_bytes);
  } else {
    auto* t = new Server<float>();
    sb = t;
    rc = t->create(h, par, diva ? 1 : 0, nlanes, lane_max_side, lane_slots, nz_max, hbm_budget_bytes);
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
  if (prob && prob->neff_hook) { s->h->err = "neff_hook: one-shot DIVA entry points only"; return YGPU_ERR_ARG; }
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
