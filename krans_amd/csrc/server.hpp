// Continuous-batching null-space server (see server.cpp).
#pragma once
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "packed.hpp"
#include "solver.hpp"

namespace kh {

class Server {
public:
  // slots: concurrent solves (one device stream each); hbm_budget <= 0 means
  // 85% of the device memory free at start.
  Server(int device, int slots, double hbm_budget);
  ~Server();
  Server(const Server &) = delete;
  Server &operator=(const Server &) = delete;

  // Host CSR (copied); the matrix is uploaded by the slot that runs it.
  int64_t submit(int64_t m, int64_t n, const int32_t *rowptr, const int32_t *col,
                 const double *val, const CyclicParams &P, unsigned seed);
  // A matrix already resident on the device (must outlive the request).
  int64_t submit_resident(const DeviceCsr *A, const CyclicParams &P, unsigned seed);

  struct Result {
    int status = 0;  // 0 ok, 1 invalid argument, 2 runtime error, 3 device OOM
    std::string error;
    int64_t n = 0, count = 0;
    std::vector<double> basis;  // n x count, column-major
    CyclicStats stats;
    double latency = 0, queued = 0;
    int packed = 0;  // members of the pack its Golub-Kahan cycles ran in (0: solo)
  };
  Result wait(int64_t id);

  struct Stats {
    int64_t done = 0, vectors = 0, running = 0, queued = 0;
    double p50 = 0, p99 = 0, wall = 0, used_bytes = 0, budget_bytes = 0;
  };
  Stats stats();
  void reset_clock();
  // Packing of small requests: a request with m + n <= pack_len (host CSR,
  // CGS2) is solved in a pack with up to pack_max - 1 compatible queued
  // requests (same Krylov precision and ncv), alone in a pack of one
  // otherwise; pack_max <= 1 turns packing off (every request solo, the
  // default). Packs trade latency for launch sharing: a pack's cycles read
  // every member's rows, so they pay off when small requests dominate.
  void set_packing(int pack_max, int64_t pack_len);

private:
  struct Req;
  static void check_params(const CyclicParams &P);
  int64_t enqueue(std::shared_ptr<Req> r);
  void work(int slot);
  int workers_alive() const;
  bool packs(const Req &r) const;
  void complete(Req &r);  // bookkeeping of a finished request (lock held)
  int pack_max_ = 1;  // packing off until set_packing
  int64_t pack_len_ = int64_t(1) << 19;

  int device_;
  double budget_;
  double used_ = 0;
  int64_t next_id_ = 1, vectors_ = 0, running_ = 0;
  bool stop_ = false;
  std::mutex mu_;
  std::condition_variable cv_;
  std::deque<std::shared_ptr<Req>> queue_;
  std::map<int64_t, std::shared_ptr<Req>> all_;
  std::vector<double> latencies_;
  std::vector<std::string> slot_errors_;
  std::chrono::steady_clock::time_point t0_, last_end_;
  std::vector<std::thread> workers_;
};

}  // namespace kh
