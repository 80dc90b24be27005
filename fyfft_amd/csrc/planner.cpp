// Host planner (see planner.hpp).
#include "planner.hpp"

#include <cmath>

#include <algorithm>

namespace fyfft {

// The reference's directly supported lengths (src/foye_FFT_radix_wrap.hpp:93-98).
static const size_t kRegistry[38] = {
    2, 3, 4, 5, 7, 8, 9, 16, 25, 27, 32, 49, 64, 81,
    125, 128, 243, 256, 343, 512, 625, 729, 1024,
    2048, 2187, 2401, 3125, 4096, 6561, 8192, 15625,
    16384, 16807, 19683, 32768, 59049, 65536, 78125};

bool in_reference_registry(size_t n) {
  return std::binary_search(kRegistry, kRegistry + 38, n);
}

bool is_smooth7(size_t n) {
  if (n == 0) return false;
  for (size_t p : {2u, 3u, 5u, 7u})
    while (n % p == 0) n /= p;
  return n == 1;
}

bool is_prime(size_t n) {
  if (n < 2) return false;
  if (n < 4) return true;
  if (n % 2 == 0 || n % 3 == 0) return false;
  for (size_t i = 5; i * i <= n; i += 6)
    if (n % i == 0 || n % (i + 2) == 0) return false;
  return true;
}

uint64_t power_mod(uint64_t base, uint64_t e, uint64_t mod) {
  unsigned __int128 r = 1, b = base % mod;
  while (e) {
    if (e & 1) r = (r * b) % mod;
    b = (b * b) % mod;
    e >>= 1;
  }
  return (uint64_t)r;
}

uint64_t primitive_root(uint64_t p) {
  if (p == 2) return 1;
  const uint64_t phi = p - 1;
  uint64_t m = phi;
  std::vector<uint64_t> f;
  for (uint64_t i = 2; i * i <= m; ++i)
    if (m % i == 0) {
      f.push_back(i);
      while (m % i == 0) m /= i;
    }
  if (m > 1) f.push_back(m);
  for (uint64_t g = 2; g < p; ++g) {
    bool ok = true;
    for (uint64_t q : f)
      if (power_mod(g, phi / q, p) == 1) { ok = false; break; }
    if (ok) return g;
  }
  return 0;
}

int strategy_for(size_t n) {
  if (in_reference_registry(n)) return kDirect;
  if (is_smooth7(n)) return kMix;
  if (is_prime(n) && is_smooth7(n - 1)) return kRader;
  return kBluestein;
}

std::vector<int> radices_for(size_t n) {
  std::vector<int> out;
  if (n < 2 || !is_smooth7(n)) return out;
  int e[8] = {0};
  for (int p : {2, 3, 5, 7})
    while (n % p == 0) { n /= p; ++e[p]; }
  const int a = e[2];
  if (a > 0) {
    const int np = (a + 3) / 4;
    const int base = a / np, extra = a % np;
    for (int i = 0; i < np; ++i) out.push_back(1 << (base + (i < extra ? 1 : 0)));
  }
  for (int i = 0; i < e[7]; ++i) out.push_back(7);
  for (int i = 0; i < e[5]; ++i) out.push_back(5);
  // threes in pairs as radix 9 (3 x 3 in registers): half the LDS passes
  for (int i = 0; i < e[3] / 2; ++i) out.push_back(9);
  if (e[3] % 2) out.push_back(3);
  return out;
}

std::vector<int> radices_composite(size_t n) {
  std::vector<int> out;
  if (n < 2 || !is_smooth7(n)) return out;
  int e[4] = {0, 0, 0, 0};
  const int pr[4] = {2, 3, 5, 7};
  for (int i = 0; i < 4; ++i)
    while (n % pr[i] == 0) { n /= pr[i]; ++e[i]; }
  static const int kRad[] = {25, 21, 20, 16, 15, 14, 12, 10, 9, 8, 7, 6, 5, 4, 3, 2};
  struct Ex { int r, x[4]; };
  std::vector<Ex> rad;
  for (int r : kRad) {
    Ex q{r, {0, 0, 0, 0}};
    int m = r;
    for (int i = 0; i < 4; ++i)
      while (m % pr[i] == 0) { m /= pr[i]; ++q.x[i]; }
    rad.push_back(q);
  }
  auto pass_cost = [](int r) {
    const int maxb = (16 + r - 1) / r;  // butterflies per thread at 16 points per thread
    const double used = (16.0 / r) / maxb;
    return 1.0 + 0.3 * (1.0 / used - 1.0);
  };
  // best[a][b][c][d]: cost and first radix for the exponent vector
  const int A = e[0] + 1, B = e[1] + 1, C = e[2] + 1, D = e[3] + 1;
  auto idx = [&](int a, int b, int c, int d) { return ((a * B + b) * C + c) * D + d; };
  std::vector<double> cost((size_t)A * B * C * D, 1e300);
  std::vector<int> first((size_t)A * B * C * D, 0);
  cost[0] = 0;
  for (int a = 0; a < A; ++a)
    for (int b = 0; b < B; ++b)
      for (int c = 0; c < C; ++c)
        for (int d = 0; d < D; ++d) {
          if (a + b + c + d == 0) continue;
          const size_t i = idx(a, b, c, d);
          for (const Ex& q : rad) {
            if (q.x[0] > a || q.x[1] > b || q.x[2] > c || q.x[3] > d) continue;
            const double v = cost[idx(a - q.x[0], b - q.x[1], c - q.x[2], d - q.x[3])] + pass_cost(q.r);
            if (v < cost[i] - 1e-12) {
              cost[i] = v;
              first[i] = q.r;
            }
          }
        }
  int a = e[0], b = e[1], c = e[2], d = e[3];
  while (a + b + c + d > 0) {
    const int r = first[idx(a, b, c, d)];
    out.push_back(r);
    for (const Ex& q : rad)
      if (q.r == r) a -= q.x[0], b -= q.x[1], c -= q.x[2], d -= q.x[3];
  }
  std::sort(out.begin(), out.end(), [](int x, int y) { return x > y; });
  return out;
}

size_t lds_capacity_elems(size_t elem_bytes, size_t lds_cap) { return lds_cap / elem_bytes; }

bool lds_geometry(size_t n, size_t elem_bytes, size_t lds_cap, LdsGeom* g, bool short_threads, bool small_radix) {
  if (n < 1 || n > (size_t)1 << 20) return false;
  LdsGeom r;
  r.n = (int)n;
  // points per thread: the generic kernels exist for F32 PPT 16/32 and
  // F64 PPT 8/16 (F64 keeps PPT low to stay within the VGPR budget)
  const int ppt_lo = elem_bytes >= 16 ? 8 : 16;
  r.ppt = (int)n <= ppt_lo * 1024 ? ppt_lo : 2 * ppt_lo;
  r.threads = (int)((n + r.ppt - 1) / r.ppt);
  if (r.ppt > ppt_lo) {
    // beyond ppt_lo * 1024 points: all 1024 threads (more waves to hide
    // latency with one workgroup per CU), ppt_lo < points per thread <= 2 * ppt_lo
    r.threads = 1024;
    r.ppt = (int)((n + 1023) / 1024);
  }
  if (short_threads && elem_bytes < 16 && r.threads < 32) {
    r.ppt = 8;
    r.threads = (int)((n + 7) / 8);
  }
  const bool fp64 = elem_bytes >= 16;
  if (small_radix && n >= (fp64 ? 64u : 128u) && n < (fp64 ? 400u : 2048u)) {
    const std::vector<int> rad = radices_for(n);
    if (std::find(rad.begin(), rad.end(), 16) == rad.end()) {
      r.ppt = fp64 ? 4 : 8;
      r.threads = (int)((n + r.ppt - 1) / r.ppt);
      r.small_radix = true;
    }
  }
  if (r.threads > 1024 || r.ppt > 2 * ppt_lo) return false;
  r.pad_shift = kLdsPadShift;
  size_t per = (size_t)((n - 1) + ((n - 1) >> r.pad_shift) + 1) * elem_bytes;
  if (per > lds_cap) {
    r.pad_shift = 30;  // no padding
    per = n * elem_bytes;
    if (per > lds_cap) return false;
  }
  int tpb = std::max(1, 256 / r.threads);
  tpb = std::min<size_t>(tpb, lds_cap / per);
  tpb = std::max(1, std::min(tpb, 1024 / r.threads));
  r.per_block = tpb;
  r.lds_bytes = per * (size_t)tpb;
  *g = r;
  return true;
}

bool four_step_split_lds(size_t n, size_t elem_bytes, size_t* n1, size_t* n2) {
  if (!is_smooth7(n)) return false;
  // column tiles of 16 sub-transforms within 40 KB of LDS (four or more
  // workgroups per CU) and rows that one workgroup holds
  const size_t max1 = 40960 / (16 * elem_bytes), max2 = elem_bytes <= 8 ? 16384 : 8192;
  size_t best = 0;
  for (size_t d = 2; d * d <= n && d <= max1; ++d)
    if (n % d == 0 && n / d <= max2) best = d;
  if (!best) return false;
  *n1 = best;
  *n2 = n / best;
  return true;
}

bool four_step_split_model(size_t n, size_t cap_elems, size_t elem_bytes, size_t* n1, size_t* n2) {
  if (!is_smooth7(n) || (n & (n - 1)) == 0) return false;
  auto pass_cost = [](size_t m) {
    double c = 0;
    for (int r : radices_for(m)) c += r == 3 ? 1.3 : r == 5 ? 1.15 : r == 7 ? 1.2 : r == 9 ? 1.1 : 1.0;
    return c;
  };
  double best = 1e300;
  for (size_t d = 2; d * d <= 64 * n && d <= cap_elems; ++d) {
    if (n % d) continue;
    const size_t e = n / d;
    if (e > cap_elems || d * 8 < e || e * 8 < d || !is_smooth7(d)) continue;
    const double c = pass_cost(d) + pass_cost(e) + ((d * elem_bytes) % 128 ? 0.5 : 0.0) +
                     ((e * elem_bytes) % 128 ? 0.5 : 0.0) + 0.05 * std::fabs(std::log((double)d / (double)e));
    if (c < best) {
      best = c;
      *n1 = d;
      *n2 = e;
    }
  }
  return best < 1e300;
}

bool four_step_split(size_t n, size_t cap_elems, size_t* n1, size_t* n2) {
  if (!is_smooth7(n)) return false;
  size_t best1 = 0;
  for (size_t d = 2; d * d <= n; ++d) {
    if (n % d) continue;
    const size_t e = n / d;
    if (d <= cap_elems && e <= cap_elems) best1 = d;  // largest d <= sqrt(n)
  }
  if (!best1) return false;
  *n1 = best1;
  *n2 = n / best1;
  return true;
}

}  // namespace fyfft
