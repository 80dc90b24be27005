// Host-side check of the planner (fyfft_amd/csrc/planner.cpp) built with
// AddressSanitizer and UndefinedBehaviorSanitizer by tests/test_sanitizers.py.
// Walks every length up to 2^17 and a sample up to 2^24 through the planner
// and checks its invariants; any sanitizer report or failed check ends the
// process with a non-zero status.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "planner.hpp"

using namespace fyfft;

static int failures = 0;
#define CHECK(c)                                                      \
  do {                                                                \
    if (!(c)) {                                                       \
      std::fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
      if (++failures > 20) std::exit(1);                              \
    }                                                                 \
  } while (0)

static void check_length(size_t n, const std::vector<char>& composite) {
  const int st = strategy_for(n);
  CHECK(st >= kDirect && st <= kDft);
  if (n < composite.size()) CHECK(is_prime(n) == (n >= 2 && !composite[n]));
  if (is_smooth7(n) && n >= 2) {
    const std::vector<int> r = radices_for(n);
    CHECK(!r.empty() && r.size() <= 16);
    size_t prod = 1;
    for (int x : r) {
      CHECK(x == 2 || x == 3 || x == 4 || x == 5 || x == 7 || x == 8 || x == 9 || x == 16);
      prod *= (size_t)x;
    }
    CHECK(prod == n);
    for (size_t elem : {8u, 16u}) {
      for (int opt = 0; opt < 4; ++opt) {
        LdsGeom g;
        if (lds_geometry(n, elem, 163840, &g, (opt & 1) != 0, (opt & 2) != 0)) {
          CHECK(g.threads >= 1 && g.threads <= 1024 && g.per_block >= 1);
          CHECK((size_t)g.ppt * g.threads >= n);
          CHECK(g.lds_bytes <= 163840);
          CHECK((size_t)g.threads * g.per_block <= 1024);
          if (g.small_radix) {  // the lean kernel: no radix-16 pass, at most 8 points per thread
            for (int x : r) CHECK(x != 16);
            CHECK((size_t)g.threads * (elem == 8 ? 8 : 4) >= n);
          }
        }
      }
      size_t a = 0, b = 0;
      if (four_step_split(n, 4096, &a, &b)) CHECK(a * b == n && a <= 4096 && b <= 4096 && a <= b);
      if (four_step_split_lds(n, elem, &a, &b)) CHECK(a * b == n && a * a <= n);
    }
  } else {
    CHECK(radices_for(n).empty());
  }
}

int main() {
  const size_t kSieve = 1 << 20;
  std::vector<char> composite(kSieve, 0);
  for (size_t i = 2; i * i < kSieve; ++i)
    if (!composite[i])
      for (size_t j = i * i; j < kSieve; j += i) composite[j] = 1;
  size_t count = 0;
  for (size_t n = 0; n <= (1u << 17); ++n, ++count) check_length(n, composite);
  for (size_t n = (1u << 17); n <= (1u << 24); n += 977, ++count) check_length(n, composite);
  // primitive roots: g^(p-1) = 1 and g^((p-1)/q) != 1 for every prime q | p-1
  size_t roots = 0;
  for (size_t p = 3; p < 300000; p += 2) {
    if (composite[p]) continue;
    const uint64_t g = primitive_root(p);
    CHECK(g >= 2 && g < p);
    CHECK(power_mod(g, p - 1, p) == 1);
    size_t m = p - 1;
    for (size_t q = 2; q * q <= m; ++q) {
      if (m % q) continue;
      CHECK(power_mod(g, (p - 1) / q, p) != 1);
      while (m % q == 0) m /= q;
    }
    if (m > 1) CHECK(power_mod(g, (p - 1) / m, p) != 1);
    ++roots;
  }
  // 128-bit products in power_mod: a prime near 2^40
  const uint64_t big = 1099511627791ull;
  CHECK(is_prime(big));
  CHECK(power_mod(3, big - 1, big) == 1);
  std::printf("planner check: %zu lengths, %zu primitive roots, %d failures\n", count, roots, failures);
  return failures ? 1 : 0;
}
