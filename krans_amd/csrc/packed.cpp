// Packed cyclic Lanczos: the Golub-Kahan cycles of several small solves in
// shared, segmented launches.
//
// A small solve spends its Golub-Kahan cycle waiting on kernel dispatch
// rather than on HBM: every step is a chain of dependent launches over short
// vectors. A pack stacks the Krylov storage of its members in one set of
// arrays (each member's rows padded to kPackRows, the row block of the V^T w
// and update kernels, so no block straddles two members) and builds one
// cycle for all of them with one chain of launches: the block-diagonal
// product with a per-member beta term, V^T w partials over the stacked
// columns, per-member finishing sums, updates with per-member coefficients,
// per-member norms and normalised copies. Locked-basis columns beyond a
// member's own count are zero in its rows, and the middle of each CGS2 pair
// runs as an update followed by a projection, so a member's arithmetic reads
// nothing but its own rows: its result is bitwise the same alone in a pack
// and in any mix.
//
// Everything else of a solve (the beta_j replay, the projected problem, the
// candidates, the projector polish, restarts, fallback) is the member's own
// CyclicDriver, run on the member's thread and stream exactly as in a solo
// solve (CyclicLanczosSolver.cpp:291-616). Rounds are opportunistic: once a
// member is ready the leader gathers whoever else becomes ready within a
// short window (all remaining members at most), builds the cycle of that set
// on its own stream after their events, and releases each with an event of
// its own recorded after the cycle. A member busy with its own restart work
// joins a later round instead of holding the others back.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <limits>
#include <mutex>
#include <thread>

#include "cyclic_driver.hpp"
#include "packed.hpp"

namespace kh {
namespace {

inline int64_t pad_rows(int64_t x) {
  const int64_t r = std::max<int64_t>(x, 1);
  return (r + kPackRows - 1) / kPackRows * kPackRows;
}

inline std::size_t esz(Prec p) { return p == Prec::F64 ? 8 : 4; }

int64_t target_of(const CyclicParams &P) {
  return std::max<int64_t>(1, P.nullity_known >= 0 ? P.nullity_known : P.k);
}

// How long the leader waits, after the first member of a round is ready,
// for the others to join it.
constexpr auto kGatherWindow = std::chrono::microseconds(300);

// The members' rendezvous with the leader.
struct Rendezvous {
  std::mutex mu;
  std::condition_variable cv;
  int total = 0, settled = 0, left = 0;
  std::vector<int> ready;           // members waiting for a cycle
  std::vector<int64_t> ready_nb;    // their locked-basis sizes
  std::vector<hipEvent_t> ready_ev; // their start vectors' events
  // per member: set when its cycle is queued, with that cycle's counts and
  // its release event (recorded on the leader stream after the cycle)
  std::vector<char> go;
  std::vector<CyclicDriver::CycleQueued> out;
  std::vector<hipEvent_t> release;
  bool failed = false;
  std::string error;
};

struct PackCycle {
  Context &c;
  const DeviceCsr &A;
  PackLayout L;
  Prec kp;
  int64_t ncv, ycap, LM, LN, hs;
  void *vb, *x, *wx, *wy;
  double *H1, *H2, *np;

  void *ycol(int64_t j) const {
    return static_cast<char *>(vb) + (std::size_t)(ycap - 1 - j) * LN * esz(kp);
  }
  void *xcol(int64_t j) const {
    return static_cast<char *>(x) + (std::size_t)j * LM * esz(kp);
  }

  // One Golub-Kahan cycle of every active segment (the build of
  // CyclicLanczosSolver.cpp:387-421 in the form of CyclicDriver::queue_cycle
  // with CGS2), nbmax the largest locked basis among them.
  CyclicDriver::CycleQueued run(int nbmax) {
    int64_t nyc = 1, nxc = 0, nmv = 0, nmtv = 0;
    for (int64_t j = 0; j < ncv; ++j) {
      const bool from_y = nyc > nxc;
      const int side = from_y ? 0 : 1;
      void *w = from_y ? wx : wy;
      if (from_y) {
        pack_spmv(c, A, false, kp, L, ycol(nyc - 1), w, j == 0 ? nullptr : xcol(nxc - 1),
                  (int)(j - 1));
        ++nmv;
      } else {
        pack_spmv(c, A, true, kp, L, xcol(nxc - 1), w, ycol(nyc - 1), (int)(j - 1));
        ++nmtv;
      }
      // x side: X_0 .. X_{nxc-1}; y side: [Y_{nyc-1} .. Y_0 | locked basis],
      // the columns beyond a member's own basis zero in its rows
      const void *V = from_y ? x : ycol(nyc - 1);
      const int64_t ld = from_y ? LM : LN;
      const int s = (int)(from_y ? nxc : nyc + nbmax);
      if (s > 0) {
        pack_vtw(c, kp, L, side, V, ld, s, w, c.partials(1));
        pack_finish(c, L, side, c.partials(1), s, H1, hs);
        pack_vh_sub(c, kp, L, side, V, ld, s, H1, hs, w, nullptr);
        pack_vtw(c, kp, L, side, V, ld, s, w, c.partials(1));
        pack_finish(c, L, side, c.partials(1), s, H2, hs);
        pack_vh_sub(c, kp, L, side, V, ld, s, H2, hs, w, np);
      } else {
        pack_vh_sub(c, kp, L, side, w, ld, 0, H1, hs, w, np);  // the norm partials only
      }
      pack_norm(c, L, side, np, (int)j);
      if (nyc + nxc >= ncv)
        break;
      pack_scale_copy(c, kp, L, side, (int)j, w, from_y ? xcol(nxc) : ycol(nyc));
      if (from_y)
        ++nxc;
      else
        ++nyc;
    }
    return {nyc + nxc, nmv, nmtv};
  }
};

}  // namespace

bool packable(const CyclicParams &P) {
  return P.reorth == 2 && P.y0_host == nullptr && !P.return_left_space;
}

int64_t pack_ncv_eff(const CyclicParams &P) {
  const int64_t k = P.nullity_known >= 0 ? P.nullity_known : P.k;
  return std::max<int64_t>(std::max<int64_t>(2, P.ncv), k + 2);
}

double pack_footprint(const std::vector<const PackRequest *> &reqs) {
  if (reqs.empty())
    return 0.0;
  double nnz = 0.0;
  int64_t LM = 0, LN = 0, kmax = 1;
  for (const auto *r : reqs) {
    LM += pad_rows(r->m);
    LN += pad_rows(r->n);
    nnz += (double)r->rowptr[r->m];
    kmax = std::max(kmax, target_of(r->P));
  }
  const Prec kp = reqs.front()->P.krylov;
  const int64_t cap = pack_ncv_eff(reqs.front()->P) / 2 + 2;
  double b = 2.0 * (16.0 * nnz + 4.0 * (double)(LM + LN + 2)) + 8.0 * LM;  // A_pack + CSC twin
  b += (double)esz(kp) * (double)((cap + kmax) * LN + cap * LM + LM + LN);  // Krylov storage
  b += 8.0 * (double)(LM + LN) / kPackRows + 16.0 * (double)reqs.size() * (double)(cap + kmax);
  return b;
}

void solve_packed(Context &leader, const std::vector<Context *> &members,
                  const std::vector<PackRequest *> &all) {
  // a malformed matrix fails its own request before the pack is built
  std::vector<PackRequest *> reqs;
  for (PackRequest *q : all) {
    try {
      validate_csr(q->m, q->n, q->rowptr, q->col, q->val);
      reqs.push_back(q);
    } catch (const std::invalid_argument &ex) {
      q->status = 1;
      q->error = ex.what();
      if (q->on_done)
        q->on_done();
    }
  }
  const int P = (int)reqs.size();
  if (P == 0)
    return;
  if ((int)members.size() < P)
    throw std::invalid_argument("solve_packed: fewer member contexts than requests");
  const Prec kp = reqs[0]->P.krylov;
  const int64_t ncv = pack_ncv_eff(reqs[0]->P);
  for (const auto *r : reqs)
    if (r->P.krylov != kp || pack_ncv_eff(r->P) != ncv || !packable(r->P))
      throw std::invalid_argument("solve_packed: requests of one pack need CGS2, the same "
                                  "Krylov precision and the same ncv");

  // layout: member i's m-side rows at xoff[i], n-side rows at yoff[i]
  std::vector<int64_t> xoff((std::size_t)P), yoff((std::size_t)P);
  int64_t LM = 0, LN = 0, kmax = 1, nnz = 0;
  for (int i = 0; i < P; ++i) {
    const PackRequest &q = *reqs[(std::size_t)i];
    if (q.m < 0 || q.n < 0 || !q.rowptr)
      throw std::invalid_argument("solve_packed: bad matrix");
    xoff[(std::size_t)i] = LM;
    yoff[(std::size_t)i] = LN;
    LM += pad_rows(q.m);
    LN += pad_rows(q.n);
    nnz += q.rowptr[q.m];
    kmax = std::max(kmax, target_of(q.P));
  }
  const int64_t i32max = std::numeric_limits<int32_t>::max();
  if (LM > i32max || LN > i32max || nnz > i32max)
    throw std::invalid_argument("solve_packed: pack exceeds int32 indexing");

  // the block-diagonal pack matrix (columns shifted by each member's offset:
  // rows keep their ascending order), padding rows empty
  std::vector<int32_t> rp((std::size_t)LM + 1), ci;
  std::vector<double> va;
  ci.reserve((std::size_t)nnz);
  va.reserve((std::size_t)nnz);
  for (int i = 0; i < P; ++i) {
    const PackRequest &q = *reqs[(std::size_t)i];
    const int64_t rows = pad_rows(q.m);
    for (int64_t r = 0; r < rows; ++r) {
      rp[(std::size_t)(xoff[(std::size_t)i] + r)] = (int32_t)ci.size();
      if (r < q.m)
        for (int32_t p = q.rowptr[r]; p < q.rowptr[r + 1]; ++p) {
          ci.push_back(q.col[p] + (int32_t)yoff[(std::size_t)i]);
          va.push_back(q.val[p]);
        }
    }
  }
  rp[(std::size_t)LM] = (int32_t)ci.size();
  std::unique_ptr<DeviceCsr> Apack(csr_create(leader, LM, LN, rp.data(), ci.data(), va.data()));

  // packed Golub-Kahan storage, zeroed (the locked-basis columns must start
  // at zero in every member's rows)
  const int64_t ycap = CyclicDriver::ycap_for(ncv), xcap = ycap, hs = ycap + kmax;
  const std::size_t e = esz(kp);
  DBuf VB((std::size_t)(ycap + kmax) * LN * e), X((std::size_t)xcap * LM * e);
  DBuf WX((std::size_t)LM * e), WY((std::size_t)LN * e);
  for (DBuf *b : {&VB, &X, &WX, &WY})
    KH_CHECK(hipMemsetAsync(b->p, 0, b->bytes, leader.stream));
  // segment tables
  std::vector<PackSeg> segs((std::size_t)P);
  std::vector<int32_t> som((std::size_t)(LM / kPackRows)), son((std::size_t)(LN / kPackRows));
  for (int i = 0; i < P; ++i) {
    PackSeg &S = segs[(std::size_t)i];
    S.blk0[0] = xoff[(std::size_t)i] / kPackRows;
    S.nblk[0] = pad_rows(reqs[(std::size_t)i]->m) / kPackRows;
    S.blk0[1] = yoff[(std::size_t)i] / kPackRows;
    S.nblk[1] = pad_rows(reqs[(std::size_t)i]->n) / kPackRows;
    for (int64_t b = 0; b < S.nblk[0]; ++b)
      som[(std::size_t)(S.blk0[0] + b)] = i;
    for (int64_t b = 0; b < S.nblk[1]; ++b)
      son[(std::size_t)(S.blk0[1] + b)] = i;
  }
  DBuf dsegs((std::size_t)P * sizeof(PackSeg)), dsom(som.size() * 4), dson(son.size() * 4);
  KH_CHECK(hipMemcpyAsync(dsom.p, som.data(), dsom.bytes, hipMemcpyHostToDevice, leader.stream));
  KH_CHECK(hipMemcpyAsync(dson.p, son.data(), dson.bytes, hipMemcpyHostToDevice, leader.stream));
  DBuf H1((std::size_t)P * hs * 8), H2((std::size_t)P * hs * 8);
  DBuf NP((std::size_t)std::max(LM, LN) / kPackRows * 8);
  leader.sync();
  PackLayout L;
  L.segs = dsegs.as<PackSeg>();
  L.nseg = P;
  L.seg_of_blk[0] = dsom.as<int32_t>();
  L.seg_of_blk[1] = dson.as<int32_t>();
  L.nblk[0] = LM / kPackRows;
  L.nblk[1] = LN / kPackRows;
  PackCycle cyc{leader, *Apack, L, kp, ncv, ycap, LM, LN, hs, VB.p, X.p, WX.p, WY.p,
                H1.as<double>(), H2.as<double>(), NP.as<double>()};

  Rendezvous R;
  R.total = P;
  R.go.assign((std::size_t)P, 0);
  R.out.resize((std::size_t)P);
  R.release.assign((std::size_t)P, nullptr);
  struct Events {  // destroys the release events when the pack ends
    std::vector<hipEvent_t> &v;
    ~Events() {
      for (hipEvent_t e : v)
        if (e)
          (void)hipEventDestroy(e);
    }
  } release_guard{R.release};
  for (int i = 0; i < P; ++i)
    KH_CHECK(hipEventCreateWithFlags(&R.release[(std::size_t)i], hipEventDisableTiming));
  auto member = [&](int i) {
    PackRequest &q = *reqs[(std::size_t)i];
    Context &mc = *members[(std::size_t)i];
    bool settled = false, left = false;
    auto settle = [&](double *bnorm) {
      std::lock_guard<std::mutex> lk(R.mu);
      segs[(std::size_t)i].bnorm = bnorm;
      ++R.settled;
      settled = true;
      R.cv.notify_all();
    };
    auto leave = [&] {
      std::lock_guard<std::mutex> lk(R.mu);
      ++R.left;
      left = true;
      R.cv.notify_all();
    };
    hipEvent_t E = nullptr;
    const hipEvent_t F = R.release[(std::size_t)i];
    try {
      KH_CHECK(hipSetDevice(mc.device));
      std::unique_ptr<DeviceCsr> A(csr_create(mc, q.m, q.n, q.rowptr, q.col, q.val));
      RefStream rng(q.seed, (std::size_t)std::max<int64_t>(1 << 16, 6 * q.n));
      CsrOp op(*A, false);
      CyclicDriver d(mc, op, rng, q.P, q.stats);
      d.use_storage(static_cast<char *>(VB.p) + (std::size_t)yoff[(std::size_t)i] * e, LN,
                    static_cast<char *>(X.p) + (std::size_t)xoff[(std::size_t)i] * e, LM,
                    static_cast<char *>(WX.p) + (std::size_t)xoff[(std::size_t)i] * e,
                    static_cast<char *>(WY.p) + (std::size_t)yoff[(std::size_t)i] * e);
      DeviceBasis B;
      if (d.prepare()) {
        settle(nullptr);
        leave();
        B = d.take_basis();
      } else {
        settle(d.beta_slots());
        KH_CHECK(hipEventCreateWithFlags(&E, hipEventDisableTiming));
        int st = CyclicDriver::kStop;
        while (d.cycle_begin()) {
          KH_CHECK(hipEventRecord(E, mc.stream));
          CyclicDriver::CycleQueued cq{};
          {
            std::unique_lock<std::mutex> lk(R.mu);
            R.go[(std::size_t)i] = 0;
            R.ready.push_back(i);
            R.ready_nb.push_back(d.basis_count());
            R.ready_ev.push_back(E);
            R.cv.notify_all();
            R.cv.wait(lk, [&] { return R.go[(std::size_t)i] || R.failed; });
            if (R.failed)
              throw std::runtime_error("packed cycle failed: " + R.error);
            cq = R.out[(std::size_t)i];
          }
          KH_CHECK(hipStreamWaitEvent(mc.stream, F, 0));
          st = d.cycle_end(cq.queued, cq.nmv, cq.nmtv);
          if (st != CyclicDriver::kContinue)
            break;
          st = CyclicDriver::kStop;
        }
        leave();
        B = st == CyclicDriver::kDone ? d.finish() : d.after_loop();
      }
      q.count = B.count;
      q.basis.resize((std::size_t)(B.count * B.len));
      if (B.count)
        KH_CHECK(hipMemcpyAsync(q.basis.data(), B.buf.p, q.basis.size() * 8,
                                hipMemcpyDeviceToHost, mc.stream));
      mc.sync();
      q.status = 0;
    } catch (const std::invalid_argument &ex) {
      q.status = 1;
      q.error = ex.what();
    } catch (const DeviceOOM &ex) {
      q.status = 3;
      q.error = ex.what();
    } catch (const std::exception &ex) {
      q.status = 2;
      q.error = ex.what();
    }
    if (!settled)
      settle(nullptr);
    if (!left)
      leave();
    if (E) {
      (void)hipStreamSynchronize(mc.stream);
      (void)hipEventDestroy(E);
    }
    if (q.on_done)
      q.on_done();
  };

  std::vector<std::thread> threads;
  threads.reserve((std::size_t)P);
  int started = 0;
  try {
    for (; started < P; ++started)
      threads.emplace_back(member, started);
  } catch (const std::exception &ex) {
    // members that could not start fail here and count as settled and left
    std::lock_guard<std::mutex> lk(R.mu);
    for (int i = started; i < P; ++i) {
      PackRequest &q = *reqs[(std::size_t)i];
      q.status = 2;
      q.error = std::string("packed solve could not start: ") + ex.what();
      ++R.settled;
      ++R.left;
    }
    R.cv.notify_all();
  }
  for (int i = started; i < P; ++i)
    if (reqs[(std::size_t)i]->on_done)
      reqs[(std::size_t)i]->on_done();

  try {
    {
      std::unique_lock<std::mutex> lk(R.mu);
      R.cv.wait(lk, [&] { return R.settled == R.total; });
    }
    while (true) {
      std::vector<int> rd;
      std::vector<int64_t> nbs;
      std::vector<hipEvent_t> evs;
      {
        std::unique_lock<std::mutex> lk(R.mu);
        R.cv.wait(lk, [&] { return !R.ready.empty() || R.left == R.total; });
        if (R.ready.empty())
          break;  // every member has left
        R.cv.wait_for(lk, kGatherWindow,
                      [&] { return (int)R.ready.size() + R.left == R.total; });
        rd.swap(R.ready);
        nbs.swap(R.ready_nb);
        evs.swap(R.ready_ev);
      }
      int nbmax = 0;
      for (auto &S : segs)
        S.active = 0;
      for (std::size_t k = 0; k < rd.size(); ++k) {
        PackSeg &S = segs[(std::size_t)rd[k]];
        S.active = 1;
        S.nb = (int32_t)nbs[k];
        nbmax = std::max(nbmax, (int)nbs[k]);
      }
      leader.upload(dsegs.p, segs.data(), segs.size() * sizeof(PackSeg));
      for (hipEvent_t ev : evs)
        KH_CHECK(hipStreamWaitEvent(leader.stream, ev, 0));
      const CyclicDriver::CycleQueued cq = cyc.run(nbmax);
      for (int i : rd)
        KH_CHECK(hipEventRecord(R.release[(std::size_t)i], leader.stream));
      {
        std::lock_guard<std::mutex> lk(R.mu);
        for (int i : rd) {
          R.out[(std::size_t)i] = cq;
          R.go[(std::size_t)i] = 1;
        }
        R.cv.notify_all();
      }
    }
  } catch (const std::exception &ex) {
    std::lock_guard<std::mutex> lk(R.mu);
    R.failed = true;
    R.error = ex.what();
    R.cv.notify_all();
  }
  for (auto &t : threads)
    t.join();
  leader.sync();
}

}  // namespace kh
