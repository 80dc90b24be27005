// Driver for the drop-in header test: the reference's usage pattern
// (example.cpp:10-35: construct an invoker, dispatch forward and inverse on
// host buffers) through include/foye_FFT.hpp.
//   dropin_main <N> <in.bin> <fwd.bin> <inv.bin>
// Reads N interleaved complex<float> from in.bin, writes dispatch<false>
// and dispatch<true> outputs, then checks in-place dispatch and the error
// path, printing the strategy.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "foye_FFT.hpp"

static bool read_all(const char* p, std::vector<float>& v) {
  FILE* f = std::fopen(p, "rb");
  if (!f) return false;
  size_t got = std::fread(v.data(), sizeof(float), v.size(), f);
  std::fclose(f);
  return got == v.size();
}
static bool write_all(const char* p, const std::vector<float>& v) {
  FILE* f = std::fopen(p, "wb");
  if (!f) return false;
  size_t put = std::fwrite(v.data(), sizeof(float), v.size(), f);
  std::fclose(f);
  return put == v.size();
}

int main(int argc, char** argv) {
  if (argc != 5) { std::fprintf(stderr, "usage: %s N in fwd inv\n", argv[0]); return 2; }
  const std::size_t n = std::strtoull(argv[1], nullptr, 10);
  std::vector<float> in(2 * n), fwd(2 * n), inv(2 * n);
  if (!read_all(argv[2], in)) { std::fprintf(stderr, "cannot read input\n"); return 2; }
  using fy::fft::FFT1D_C2C_invoker;
  FFT1D_C2C_invoker invoker = n;  // non-explicit constructor, as in the reference
  invoker.dispatch<false>(in.data(), fwd.data());
  invoker.dispatch<true>(in.data(), inv.data());
  // in place must give the out-of-place result
  std::vector<float> buf = in;
  invoker.dispatch<false>(buf.data(), buf.data());
  if (std::memcmp(buf.data(), fwd.data(), buf.size() * sizeof(float)) != 0) {
    std::fprintf(stderr, "in-place result differs\n");
    return 1;
  }
  // move keeps the plan usable
  FFT1D_C2C_invoker moved = std::move(invoker);
  moved.dispatch<false>(in.data(), buf.data());
  if (std::memcmp(buf.data(), fwd.data(), buf.size() * sizeof(float)) != 0) return 1;
  // out-of-range length is an error, not undefined behaviour
  bool threw = false;
  try { FFT1D_C2C_invoker bad(1); } catch (const std::invalid_argument&) { threw = true; }
  if (!threw) { std::fprintf(stderr, "no exception for N=1\n"); return 1; }
  std::printf("strategy %d\n", (int)FFT1D_C2C_invoker::acquired_execution_strategy(n));
  // the reference's per-strategy invokers (foye_FFT.hpp:81-84), each used
  // through the base class where its is_processable accepts N: forward and
  // backward must agree with dispatch<false> / dispatch<true>
  auto close_to = [](const std::vector<float>& a, const std::vector<float>& b) {
    double num = 0, den = 0;
    for (size_t i = 0; i < a.size(); ++i) {
      num += (double)(a[i] - b[i]) * (a[i] - b[i]);
      den += (double)b[i] * b[i];
    }
    return num <= 1e-10 * den;  // rel-L2 <= 1e-5
  };
  int used = 0;
  auto agree = [&](fy::fft::basic_unsmooth_path_invoker& base, const char* name) {
    std::vector<float> f(2 * n), b(2 * n);
    base.forward(in.data(), f.data());
    base.backward(in.data(), b.data());
    if (!close_to(f, fwd) || !close_to(b, inv)) {
      std::fprintf(stderr, "%s invoker disagrees with dispatch\n", name);
      std::exit(1);
    }
    ++used;
  };
  if (fy::fft::FFT1D_mixed_radix_invoker::is_processable(n)) {
    fy::fft::FFT1D_mixed_radix_invoker m(n);
    agree(m, "mixed-radix");
  }
  if (fy::fft::FFT1D_C2C_rader_invoker::is_processable(n)) {
    fy::fft::FFT1D_C2C_rader_invoker r(n);
    agree(r, "rader");
  }
  if (fy::fft::FFT1D_bluestein_invoker::is_processable(n)) {
    fy::fft::FFT1D_bluestein_invoker b(n);
    agree(b, "bluestein");
  }
  fy::fft::FFT1D_C2C_dft_invoker d(n);
  agree(d, "dft");
  std::printf("invokers %d\n", used);
  if (!write_all(argv[3], fwd) || !write_all(argv[4], inv)) return 2;
  return 0;
}
