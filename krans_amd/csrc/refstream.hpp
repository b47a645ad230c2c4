// The reference-compatible random stream (host only: no HIP headers, so the
// stream and its threads also build for CPU sanitizer runs).
#pragma once
#include <condition_variable>
#include <cstddef>
#include <cstdint>
#include <mutex>
#include <thread>
#include <vector>

namespace kh {

// The reference's random stream: std::mt19937(seed) feeding one
// std::normal_distribution<double> whose cached second draw persists across
// calls (StdVectorFactory.hpp:16-58, EigenVectorFactory.hpp:13-38). Vectors
// are generated on the host, in the reference's order, and uploaded. A
// background thread keeps the next values of the sequence ready (the
// sequence does not depend on how requests are split), so generation
// overlaps the device work.
class RefStream {
public:
  explicit RefStream(unsigned seed, std::size_t prefetch = std::size_t(1) << 22);
  ~RefStream();
  RefStream(const RefStream &) = delete;
  RefStream &operator=(const RefStream &) = delete;
  void fill(double *out, std::size_t n);

  // MT19937 (Matsumoto & Nishimura 1998) with std::mt19937's seeding: the
  // same 32-bit outputs, produced 624 at a time (one twist of the state,
  // then the tempering of every word).
  class Mt19937 {
  public:
    explicit Mt19937(uint32_t seed);
    void block(uint32_t *out);  // the next 624 outputs
  private:
    uint32_t mt_[624];
  };

private:
  void produce();
  Mt19937 gen_;
  std::vector<double> ring_;
  std::size_t cap_ = 0, head_ = 0, count_ = 0;
  bool stop_ = false;
  std::mutex mu_;
  std::condition_variable cv_;
  std::thread th_;
};

}  // namespace kh
