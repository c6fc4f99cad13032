// Binomial checkpointing schedule (host).
#include "checkpoint.hpp"

#include <stdexcept>
#include <string>

namespace festung {

int64_t binomial_advances(int64_t n, int64_t s) {
  if (n <= 1) return 0;
  if (s <= 0) return n * (n - 1) / 2;
  // B(r) = C(s + 1 + r, s + 1), the number of steps that s + 1 snapshots
  // reverse with r forward sweeps
  unsigned __int128 B = 1;
  int64_t r = 0;
  while (B < static_cast<unsigned __int128>(n)) {
    ++r;
    B = B * static_cast<unsigned __int128>(s + 1 + r) / static_cast<unsigned __int128>(r);
  }
  // C(s + 1 + r, s + 2) = C(s + 1 + r, s + 1) r / (s + 2)
  const unsigned __int128 C2 = B * static_cast<unsigned __int128>(r) / static_cast<unsigned __int128>(s + 2);
  return r * n - static_cast<int64_t>(C2);
}

namespace {

// the first k in [1, n-1] minimising k + F(n - k, s - 1) + F(k, s); the
// function is convex in k, so the first k where it stops decreasing
int32_t best_split(int32_t n, int32_t s) {
  auto g = [&](int64_t k) { return k + binomial_advances(n - k, s - 1) + binomial_advances(k, s); };
  int32_t lo = 1, hi = n - 1;
  while (lo < hi) {
    const int32_t mid = lo + (hi - lo) / 2;
    if (g(mid + 1) - g(mid) >= 0) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

struct Builder {
  std::vector<CkAction> out;
  std::vector<int32_t> freeSlots;

  // reverse steps a .. b-1; state a is in buffer sa; s snapshot slots free
  void rev(int32_t a, int32_t b, int32_t sa, int32_t s) {
    const int32_t n = b - a;
    if (n == 1) {
      out.push_back({CkAction::REVERSE, a, 0, sa, 0});
      return;
    }
    if (s == 0) {
      for (int32_t j = b - 1; j >= a; --j) {
        if (j > a) out.push_back({CkAction::ADVANCE, a, j - a, sa, kRot});
        out.push_back({CkAction::REVERSE, j, 0, j > a ? kRot : sa, 0});
      }
      return;
    }
    const int32_t k = best_split(n, s);
    const int32_t slot = freeSlots.back();
    freeSlots.pop_back();
    out.push_back({CkAction::ADVANCE, a, k, sa, slot});
    rev(a + k, b, slot, s - 1);
    freeSlots.push_back(slot);
    rev(a, a + k, sa, s);
  }
};

}  // namespace

std::vector<CkAction> binomial_schedule(int32_t n, int32_t c) {
  if (n < 1 || c < 1) throw std::invalid_argument("binomial schedule: need n >= 1 steps and c >= 1 snapshots");
  if (c > n) c = n;
  Builder bld;
  for (int32_t i = c - 1; i >= 1; --i) bld.freeSlots.push_back(i);
  bld.rev(0, n, 0, c - 1);
  // simulate: which step's state each buffer holds
  std::vector<int64_t> held(c, -1);
  held[0] = 0;
  int64_t rot = -1, next = n - 1;
  auto holds = [&](int32_t buf) -> int64_t { return buf == kRot ? rot : held.at(buf); };
  for (const CkAction& x : bld.out) {
    if (holds(x.src) != x.from) throw std::logic_error("binomial schedule: source buffer holds another state");
    if (x.kind == CkAction::ADVANCE) {
      if (x.steps < 1 || x.from + x.steps >= n) throw std::logic_error("binomial schedule: bad advance");
      if (x.dst == kRot) rot = x.from + x.steps;
      else held.at(x.dst) = x.from + x.steps;
    } else {
      if (x.from != next) throw std::logic_error("binomial schedule: steps reversed out of order");
      --next;
    }
  }
  if (next != -1) throw std::logic_error("binomial schedule: not every step reversed");
  return bld.out;
}

}  // namespace festung
