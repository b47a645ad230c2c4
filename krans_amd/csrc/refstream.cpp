// Prefetching implementation of the reference-compatible random stream.
#include <algorithm>

#include "solver.hpp"

namespace kh {

namespace {
constexpr std::size_t kChunk = std::size_t(1) << 15;
}  // namespace

RefStream::RefStream(unsigned seed, std::size_t prefetch)
    : gen_(seed), dist_(0.0, 1.0), ring_(std::max(prefetch, kChunk)), cap_(ring_.size()) {
  th_ = std::thread([this] { produce(); });
}

RefStream::~RefStream() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
  }
  cv_.notify_all();
  th_.join();
}

void RefStream::produce() {
  std::vector<double> tmp(kChunk);
  while (true) {
    std::size_t room;
    {
      std::unique_lock<std::mutex> lk(mu_);
      cv_.wait(lk, [&] { return stop_ || count_ < cap_; });
      if (stop_)
        return;
      room = std::min(kChunk, cap_ - count_);
    }
    // generate outside the lock: gen_/dist_ belong to this thread only
    for (std::size_t i = 0; i < room; ++i)
      tmp[i] = dist_(gen_);
    {
      std::lock_guard<std::mutex> lk(mu_);
      std::size_t tail = (head_ + count_) % cap_;
      for (std::size_t i = 0; i < room; ++i) {
        ring_[tail] = tmp[i];
        tail = (tail + 1 == cap_) ? 0 : tail + 1;
      }
      count_ += room;
    }
    cv_.notify_all();
  }
}

void RefStream::fill(double *out, std::size_t n) {
  std::size_t got = 0;
  while (got < n) {
    std::unique_lock<std::mutex> lk(mu_);
    cv_.wait(lk, [&] { return count_ > 0; });
    const std::size_t take = std::min(count_, n - got);
    const std::size_t first = std::min(take, cap_ - head_);
    std::copy(ring_.begin() + (long)head_, ring_.begin() + (long)(head_ + first), out + got);
    if (take > first)
      std::copy(ring_.begin(), ring_.begin() + (long)(take - first), out + got + first);
    head_ = (head_ + take) % cap_;
    count_ -= take;
    got += take;
    lk.unlock();
    cv_.notify_all();
  }
}

}  // namespace kh
