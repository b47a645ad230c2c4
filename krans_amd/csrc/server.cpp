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
  while (true) {
    std::shared_ptr<Req> r;
    {
      std::unique_lock<std::mutex> lk(mu_);
      // FIFO admission: the head starts when its footprint fits
      cv_.wait(lk, [&] {
        return (stop_ && queue_.empty()) ||
               (!queue_.empty() && used_ + queue_.front()->footprint <= budget_);
      });
      if (queue_.empty())
        return;  // stopping
      r = queue_.front();
      queue_.pop_front();
      used_ += r->footprint;
      r->t_start = Clock::now();
      ++running_;
    }
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
      r->t_end = Clock::now();
      r->done = true;
      used_ -= r->footprint;
      --running_;
      const double lat = std::chrono::duration<double>(r->t_end - r->t_submit).count();
      latencies_.push_back(lat);
      vectors_ += r->count;
      last_end_ = r->t_end;
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
