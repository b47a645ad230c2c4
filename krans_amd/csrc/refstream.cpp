// Prefetching implementation of the reference-compatible random stream: the
// values of std::normal_distribution<double>(0, 1) driven by
// std::mt19937(seed), as libstdc++ produces them, bit for bit.
//
// The generator is MT19937 with std::mt19937's seeding (checked output for
// output against std::mt19937), run a state block at a time.
// libstdc++'s normal_distribution (bits/random.tcc) is the Marsaglia polar
// method: each attempt draws two uniforms u = generate_canonical<double, 53>
// (two 32-bit outputs w0, w1: (w0 + w1 * 2^32) / 2^64, clamped below 1),
// forms x = 2u1 - 1, y = 2u2 - 1, r2 = x^2 + y^2, rejects r2 > 1 or r2 == 0,
// and otherwise returns y * m and then x * m with m = sqrt(-2 log(r2) / r2)
// (each result then goes through * stddev + mean, i.e. * 1.0 + 0.0). Every
// attempt consumes exactly four generator outputs, so attempt c reads outputs
// 4c..4c+3 whether it is accepted or not. The producer thread therefore draws
// the raw outputs sequentially (cheap) and converts a batch of attempts on
// worker threads, each keeping its accepted pairs in order; concatenating the
// workers' outputs in order reproduces the sequential stream exactly. The
// expressions below are those of libstdc++ term by term, without contraction.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <thread>

#include "refstream.hpp"

namespace kh {

namespace {
constexpr std::size_t kPairs = std::size_t(1) << 18;  // attempts per batch

inline double canonical53(uint32_t w0, uint32_t w1) {
#pragma clang fp contract(off)
  double sum = 0.0, tmp = 1.0;
  sum += (double)w0 * tmp;
  tmp *= 4294967296.0;
  sum += (double)w1 * tmp;
  tmp *= 4294967296.0;
  double r = sum / tmp;
  if (r >= 1.0)
    r = std::nextafter(1.0, 0.0);
  return r;
}

// Attempts [c0, c1) of the batch `w` (4 outputs each) into `out`, in order.
void convert(const uint32_t *w, std::size_t c0, std::size_t c1, std::vector<double> &out) {
#pragma clang fp contract(off)
  out.clear();
  for (std::size_t c = c0; c < c1; ++c) {
    const double x = 2.0 * canonical53(w[4 * c], w[4 * c + 1]) - 1.0;
    const double y = 2.0 * canonical53(w[4 * c + 2], w[4 * c + 3]) - 1.0;
    const double r2 = x * x + y * y;
    if (r2 > 1.0 || r2 == 0.0)
      continue;
    const double mult = std::sqrt(-2 * std::log(r2) / r2);
    const double a = y * mult, b = x * mult;
    out.push_back(a * 1.0 + 0.0);
    out.push_back(b * 1.0 + 0.0);
  }
}
}  // namespace

// std::mt19937's seeding: x_0 = seed, x_i = 1812433253 (x_{i-1} ^ (x_{i-1} >> 30)) + i.
RefStream::Mt19937::Mt19937(uint32_t seed) {
  mt_[0] = seed;
  for (uint32_t i = 1; i < 624; ++i)
    mt_[i] = 1812433253u * (mt_[i - 1] ^ (mt_[i - 1] >> 30)) + i;
}

// One twist of the whole state (the first 227 words read the old words
// 397 ahead, the rest the already twisted words 227 behind), then the
// tempering of every word: outputs 624k .. 624k + 623 of the generator.
void RefStream::Mt19937::block(uint32_t *out) {
  constexpr uint32_t kUpper = 0x80000000u, kLower = 0x7fffffffu, kA = 0x9908b0dfu;
  auto twist = [](uint32_t a, uint32_t b, uint32_t c) {
    const uint32_t y = (a & kUpper) | (b & kLower);
    return c ^ (y >> 1) ^ ((b & 1u) ? kA : 0u);
  };
  int i = 0;
  for (; i < 624 - 397; ++i)
    mt_[i] = twist(mt_[i], mt_[i + 1], mt_[i + 397]);
  for (; i < 623; ++i)
    mt_[i] = twist(mt_[i], mt_[i + 1], mt_[i + 397 - 624]);
  mt_[623] = twist(mt_[623], mt_[0], mt_[396]);
  for (int k = 0; k < 624; ++k) {
    uint32_t y = mt_[k];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    out[k] = y;
  }
}

RefStream::RefStream(unsigned seed, std::size_t prefetch)
    : gen_(seed), ring_(std::max<std::size_t>(prefetch, 4 * kPairs)), cap_(ring_.size()) {
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
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const std::size_t nt = std::min<std::size_t>(8, std::max<std::size_t>(1, hw / 2));
  // double buffering: the generator outputs of batch b+1 are drawn on this
  // thread while the workers convert batch b
  std::vector<uint32_t> words[2] = {std::vector<uint32_t>(4 * kPairs),
                                    std::vector<uint32_t>(4 * kPairs)};
  std::vector<std::vector<double>> outs(nt);
  // generator outputs come 624 at a time; `left` keeps the unused tail of
  // the last block for the next batch
  uint32_t blk[624];
  int left = 0;
  auto draw = [&](std::vector<uint32_t> &w) {
    std::size_t k = 0;
    const std::size_t n = w.size();
    while (k < n) {
      if (left == 0) {
        gen_.block(blk);
        left = 624;
      }
      const std::size_t take = std::min<std::size_t>((std::size_t)left, n - k);
      std::copy(blk + (624 - left), blk + (624 - left) + take, w.data() + k);
      k += take;
      left -= (int)take;
    }
  };
  auto push = [this](const std::vector<double> &o) {
    std::size_t done = 0;
    while (done < o.size()) {
      std::unique_lock<std::mutex> lk(mu_);
      cv_.wait(lk, [&] { return stop_ || count_ < cap_; });
      if (stop_)
        return false;
      const std::size_t room = std::min(cap_ - count_, o.size() - done);
      const std::size_t tail = (head_ + count_) % cap_;
      const std::size_t first = std::min(room, cap_ - tail);  // up to the ring's end
      std::copy(o.begin() + (long)done, o.begin() + (long)(done + first), ring_.begin() + (long)tail);
      std::copy(o.begin() + (long)(done + first), o.begin() + (long)(done + room), ring_.begin());
      count_ += room;
      done += room;
      lk.unlock();
      cv_.notify_all();
    }
    return true;
  };
  int cur = 0;
  draw(words[cur]);
  while (true) {
    {
      std::lock_guard<std::mutex> lk(mu_);
      if (stop_)
        return;
    }
    const uint32_t *w = words[cur].data();
    std::vector<std::thread> ths;
    for (std::size_t t = 0; t < nt; ++t) {
      const std::size_t c0 = kPairs * t / nt, c1 = kPairs * (t + 1) / nt;
      ths.emplace_back([&outs, w, t, c0, c1] { convert(w, c0, c1, outs[t]); });
    }
    draw(words[cur ^ 1]);
    for (auto &t : ths)
      t.join();
    for (const auto &o : outs)
      if (!push(o))
        return;
    cur ^= 1;
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
