// Continuous-batching null-space server.
//
// Requests (a CSR matrix, cyclic-Lanczos options, precision, seed) enter a
// FIFO queue. A request is admitted when a solver slot is free and its
// device footprint (matrix + Krylov state, cyclic_footprint) fits in the HBM
// budget beside the requests already running; requests that arrive while
// others are on the GPU start as soon as a slot and memory free up. Each
// slot owns its own context (device stream), so admitted solves run
// concurrently on the device; every solve is deterministic and independent
// of the others (own stream, buffers and random stream), so its result is the
// same alone and in the mix. Latencies (submit -> result) feed p50/p99.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "server.hpp"

namespace kh {

using Clock = std::chrono::steady_clock;

struct Server::Req {
  int64_t id = 0;
  int64_t m = 0, n = 0;
  std::vector<int32_t> rp, ci;
  std::vector<double> va;
  const DeviceCsr *shared = nullptr;  // resident matrix (not owned)
  CyclicParams P;
  unsigned seed = 1;
  double footprint = 0;
  Clock::time_point t_submit, t_start, t_end;
  bool done = false;
  int status = 0;
  std::string error;
  std::vector<double> basis;  // n x count
  int64_t count = 0;
  CyclicStats stats;
  int packed = 0;  // members of its pack (0: solved alone)
};

// Requests return the right null space only: the left space of
// find_nullspace_cyclic_lanczos_ex (CyclicLanczosSolver.cpp:618-660) is a
// second solve on A^T, submitted as its own request by the caller.
void Server::check_params(const CyclicParams &P) {
  if (P.return_left_space)
    throw std::invalid_argument(
        "server: return_left_space is not served; submit A^T as its own request");
}

Server::Server(int device, int slots, double hbm_budget)
    : device_(device), budget_(hbm_budget) {
  if (slots < 1 || slots > 64)
    throw std::invalid_argument("server: slots must be in [1, 64]");
  if (budget_ <= 0) {
    KH_CHECK(hipSetDevice(device));
    std::size_t fr = 0, tot = 0;
    KH_CHECK(hipMemGetInfo(&fr, &tot));
    budget_ = 0.85 * (double)fr;
  }
  t0_ = Clock::now();
  for (int i = 0; i < slots; ++i)
    workers_.emplace_back([this, i] { work(i); });
}

Server::~Server() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
  }
  cv_.notify_all();
  for (auto &t : workers_)
    t.join();
}

int64_t Server::submit(int64_t m, int64_t n, const int32_t *rowptr, const int32_t *col,
                       const double *val, const CyclicParams &P, unsigned seed) {
  if (m < 0 || n < 0 || !rowptr)
    throw std::invalid_argument("server: bad matrix");
  check_params(P);
  const int64_t nnz = rowptr[m];
  if (nnz < 0 || rowptr[0] != 0)
    throw std::invalid_argument("server: bad row pointers");
  if (nnz > 0 && (!col || !val))
    throw std::invalid_argument("server: null column indices or values");
  auto r = std::make_shared<Req>();
  r->m = m;
  r->n = n;
  r->rp.assign(rowptr, rowptr + m + 1);
  r->ci.assign(col, col + nnz);
  r->va.assign(val, val + nnz);
  r->P = P;
  r->P.y0_host = nullptr;
  r->seed = seed;
  const double mat = 2.0 * (16.0 * nnz + 4.0 * (m + n + 2)) + 8.0 * m;
  r->footprint = mat + cyclic_footprint(m, n, P);
  return enqueue(std::move(r));
}

int64_t Server::submit_resident(const DeviceCsr *A, const CyclicParams &P, unsigned seed) {
  if (!A)
    throw std::invalid_argument("server: null matrix");
  check_params(P);
  auto r = std::make_shared<Req>();
  r->m = A->m;
  r->n = A->n;
  r->shared = A;
  r->P = P;
  r->P.y0_host = nullptr;
  r->seed = seed;
  r->footprint = cyclic_footprint(A->m, A->n, P);
  return enqueue(std::move(r));
}

void Server::set_packing(int pack_max, int64_t pack_len) {
  if (pack_max < 1 || pack_max > 64)
    throw std::invalid_argument("server: pack_max must be in [1, 64]");
  if (pack_len < 0)
    throw std::invalid_argument("server: pack_len must be >= 0");
  std::lock_guard<std::mutex> lk(mu_);
  pack_max_ = pack_max;
  pack_len_ = pack_len;
}

bool Server::packs(const Req &r) const {
  return pack_max_ > 1 && !r.shared && r.m + r.n <= pack_len_ && packable(r.P);
}

void Server::complete(Req &r) {
  r.t_end = Clock::now();
  r.done = true;
  used_ -= r.footprint;
  --running_;
  latencies_.push_back(std::chrono::duration<double>(r.t_end - r.t_submit).count());
  vectors_ += r.count;
  last_end_ = r.t_end;
}

int64_t Server::enqueue(std::shared_ptr<Req> r) {
  if (r->footprint > budget_)
    throw std::invalid_argument("server: request footprint exceeds the HBM budget");
  std::lock_guard<std::mutex> lk(mu_);
  r->id = next_id_++;
  r->t_submit = Clock::now();
  all_[r->id] = r;
  queue_.push_back(r);
  cv_.notify_all();
  return r->id;
}

void Server::work(int slot) {
  std::unique_ptr<Context> ctx;
  try {
    ctx = std::make_unique<Context>(device_);
  } catch (const std::exception &e) {
    {
      std::lock_guard<std::mutex> lk(mu_);
      slot_errors_.push_back(e.what());
    }
    // a client may already wait for a request that no slot can now run
    cv_.notify_all();
    return;
  }
  (void)slot;
  // the member contexts of this slot's packs (own stream each), kept
  std::vector<std::unique_ptr<Context>> members;
  while (true) {
    std::vector<std::shared_ptr<Req>> batch;
    double pack_extra = 0.0;
    {
      std::unique_lock<std::mutex> lk(mu_);
      // FIFO admission: the head starts when its footprint fits
      cv_.wait(lk, [&] {
        return (stop_ && queue_.empty()) ||
               (!queue_.empty() && used_ + queue_.front()->footprint <= budget_);
      });
      if (queue_.empty())
        return;  // stopping
      auto r = queue_.front();
      queue_.pop_front();
      batch.push_back(r);
      if (packs(*r)) {
        // compatible small requests among the next queued ones join the pack
        // while the pack (members + packed storage) fits the budget
        std::vector<const PackRequest *> probe;
        auto as_pack = [](const Req &q) {
          PackRequest pr;
          pr.m = q.m, pr.n = q.n, pr.rowptr = q.rp.data(), pr.P = q.P;
          return pr;
        };
        std::vector<PackRequest> shapes{as_pack(*r)};
        double members_fp = r->footprint;
        const int64_t ncv = pack_ncv_eff(r->P);
        int scanned = 0;
        for (auto it = queue_.begin();
             it != queue_.end() && (int)batch.size() < pack_max_ && scanned < 64; ++scanned) {
          const Req &q = **it;
          if (!packs(q) || q.P.krylov != r->P.krylov || pack_ncv_eff(q.P) != ncv) {
            ++it;
            continue;
          }
          shapes.push_back(as_pack(q));
          probe.clear();
          for (const auto &sp : shapes)
            probe.push_back(&sp);
          const double extra = pack_footprint(probe);
          if (used_ + members_fp + q.footprint + extra > budget_) {
            shapes.pop_back();
            break;
          }
          members_fp += q.footprint;
          batch.push_back(*it);
          it = queue_.erase(it);
        }
        probe.clear();
        for (const auto &sp : shapes)
          probe.push_back(&sp);
        pack_extra = pack_footprint(probe);
      }
      const auto now = Clock::now();
      for (auto &q : batch) {
        used_ += q->footprint;
        q->t_start = now;
        q->packed = packs(*q) ? (int)batch.size() : 0;
        ++running_;
      }
      used_ += pack_extra;
    }
    if (batch.front()->packed > 0) {
      // one pack (possibly of one): members report as they finish
      std::vector<PackRequest> prs(batch.size());
      std::vector<PackRequest *> pp;
      std::vector<Context *> mctx;
      for (std::size_t i = 0; i < batch.size(); ++i) {
        Req &q = *batch[i];
        PackRequest &pr = prs[i];
        pr.m = q.m, pr.n = q.n, pr.rowptr = q.rp.data(), pr.col = q.ci.data();
        pr.val = q.va.data(), pr.P = q.P, pr.seed = q.seed;
        pr.on_done = [this, &pr, &q] {
          std::lock_guard<std::mutex> lk(mu_);
          q.status = pr.status;
          q.error = pr.error;
          q.count = pr.count;
          q.basis = std::move(pr.basis);
          q.stats = pr.stats;
          complete(q);
          cv_.notify_all();
        };
        pp.push_back(&pr);
      }
      std::string fail;
      try {
        while (members.size() < batch.size())
          members.push_back(std::make_unique<Context>(device_, std::size_t(1) << 20));
        for (std::size_t i = 0; i < batch.size(); ++i)
          mctx.push_back(members[i].get());
        solve_packed(*ctx, mctx, pp);
      } catch (const std::exception &e) {
        fail = e.what();  // the pack could not start: every member fails alone
      }
      {
        std::lock_guard<std::mutex> lk(mu_);
        if (!fail.empty())
          for (auto &q : batch)
            if (!q->done) {
              q->status = 2;
              q->error = "server: pack failed: " + fail;
              complete(*q);
            }
        used_ -= pack_extra;
      }
      cv_.notify_all();
      continue;
    }
    auto r = batch.front();
    try {
      std::unique_ptr<DeviceCsr> own;
      const DeviceCsr *A = r->shared;
      if (!A) {
        own.reset(csr_create(*ctx, r->m, r->n, r->rp.data(), r->ci.data(), r->va.data()));
        A = own.get();
      }
      RefStream rng(r->seed, (std::size_t)std::max<int64_t>(1 << 16, 6 * r->n));
      CsrOp op(*A, false);
      DeviceBasis B = cyclic_right(*ctx, op, rng, r->P, r->stats);
      r->count = B.count;
      r->basis.resize((std::size_t)(B.count * B.len));
      if (B.count)
        KH_CHECK(hipMemcpyAsync(r->basis.data(), B.buf.p, r->basis.size() * 8,
                                hipMemcpyDeviceToHost, ctx->stream));
      ctx->sync();
      r->status = 0;
    } catch (const std::invalid_argument &e) {
      r->status = 1;
      r->error = e.what();
    } catch (const DeviceOOM &e) {
      r->status = 3;
      r->error = e.what();
    } catch (const std::exception &e) {
      r->status = 2;
      r->error = e.what();
    }
    {
      std::lock_guard<std::mutex> lk(mu_);
      complete(*r);
    }
    cv_.notify_all();
  }
}

Server::Result Server::wait(int64_t id) {
  std::shared_ptr<Req> r;
  std::unique_lock<std::mutex> lk(mu_);
  auto it = all_.find(id);
  if (it == all_.end())
    throw std::invalid_argument("server: unknown request id");
  r = it->second;
  cv_.wait(lk, [&] { return r->done || (!slot_errors_.empty() && running_ == 0 &&
                                         workers_alive() == 0); });
  if (!r->done)
    throw std::runtime_error("server: no solver slot could start: " + slot_errors_.front());
  Result out;
  out.status = r->status;
  out.error = r->error;
  out.n = r->n;
  out.count = r->count;
  out.basis = std::move(r->basis);
  out.stats = r->stats;
  out.packed = r->packed;
  out.latency = std::chrono::duration<double>(r->t_end - r->t_submit).count();
  out.queued = std::chrono::duration<double>(r->t_start - r->t_submit).count();
  all_.erase(it);
  return out;
}

int Server::workers_alive() const {
  return (int)workers_.size() - (int)slot_errors_.size();
}

Server::Stats Server::stats() {
  std::lock_guard<std::mutex> lk(mu_);
  Stats s;
  s.done = (int64_t)latencies_.size();
  s.vectors = vectors_;
  std::vector<double> l = latencies_;
  std::sort(l.begin(), l.end());
  auto pct = [&](double q) {
    if (l.empty())
      return 0.0;
    const double pos = q * (double)(l.size() - 1);
    const std::size_t lo = (std::size_t)pos, hi = std::min(lo + 1, l.size() - 1);
    return l[lo] + (l[hi] - l[lo]) * (pos - (double)lo);
  };
  s.p50 = pct(0.5);
  s.p99 = pct(0.99);
  s.wall = std::chrono::duration<double>(last_end_ - t0_).count();
  s.running = running_;
  s.queued = (int64_t)queue_.size();
  s.used_bytes = used_;
  s.budget_bytes = budget_;
  return s;
}

void Server::reset_clock() {
  std::lock_guard<std::mutex> lk(mu_);
  latencies_.clear();
  vectors_ = 0;
  t0_ = Clock::now();
  last_end_ = t0_;
}

}  // namespace kh
