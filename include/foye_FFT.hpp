// foye_FFT.hpp -- drop-in replacement for kjear/fyFFT's public header.
//
// Same namespace, class, constructor, dispatch<invert> and strategy query as
// the reference (foye_FFT.hpp:79-185), implemented over the C ABI of
// libfyfft.so (include/fyfft.h).  Plain C++17; no HIP headers are needed to
// use it.  Link with -lfyfft.
//
// Semantics kept from the reference:
//   * FFT1D_C2C_invoker(N) plans a length-N complex<float> transform
//     (the constructor is not explicit, as in the reference);
//   * dispatch<false>(in, out) is the unnormalised forward transform,
//     dispatch<true>(in, out) the inverse scaled by 1/N; both take HOST
//     pointers to N interleaved (re, im) floats, run synchronously and allow
//     in == out;
//   * one invoker is not reentrant (its device staging buffer is reused);
//     distinct invokers may run concurrently;
//   * acquired_execution_strategy(N) is static and needs no GPU.
// Differences, by design:
//   * lengths are checked: N outside [min_input_length, FYFFT_MAX_LENGTH]
//     throws std::invalid_argument where the reference's __assume makes it
//     undefined (foye_FFT.hpp:103-104); N above 80000 is accepted;
//   * device failures throw std::runtime_error;
//   * acquired_execution_strategy reports THIS library's strategy: primes
//     whose N-1 is not 2/3/5/7-smooth report bluestein (the reference uses
//     a zero-padded Rader up to 39064 and a DFT table beyond), and DFT is
//     never reported;
//   * dispatch_device<invert>(d_in, d_out, stream) is the device-pointer
//     fast path: asynchronous on `stream` (a hipStream_t, nullptr = null
//     stream).
// The reference's public per-strategy invokers (foye_FFT.hpp:81-84) are
// here too, with its names, constructors, is_processable predicates and
// virtual forward/backward(const float*, float*) on host pointers
// (utility.hpp:6-11): mix_radix::FFT1D_mixed_radix_invoker,
// rader::FFT1D_C2C_rader_invoker, bluestein::FFT1D_bluestein_invoker and
// dft::FFT1D_C2C_dft_invoker, over fyfft_plan_create_strategy.  backward is
// the inverse scaled by 1/N, as in the reference.  A Rader invoker for a
// prime whose N-1 is not 2/3/5/7-smooth and the DFT invoker compute the same
// transform through the library's own strategy.
#ifndef FOYE_FFT_DROPIN_HPP_
#define FOYE_FFT_DROPIN_HPP_

#include <cstddef>
#include <stdexcept>
#include <string>
#include <utility>

#include "fyfft.h"

namespace fy::fft {

constexpr std::size_t min_input_length = 2;
constexpr std::size_t max_input_length = 80'000;  // the reference's limit

struct FFT1D_C2C_invoker {
  enum class FFT_execution_strategy { kernel_direct, mix_radix, rader, bluestein, DFT };

  FFT1D_C2C_invoker(std::size_t length) : N(length) {
    if (length < min_input_length || length > FYFFT_MAX_LENGTH)
      throw std::invalid_argument("FFT1D_C2C_invoker: length " + std::to_string(length) +
                                  " outside [2, " + std::to_string(FYFFT_MAX_LENGTH) + "]");
    check(fyfft_plan_create(&plan_, length, 1, FYFFT_F32), "plan");
  }
  ~FFT1D_C2C_invoker() { fyfft_plan_destroy(plan_); }
  FFT1D_C2C_invoker(const FFT1D_C2C_invoker&) = delete;
  FFT1D_C2C_invoker& operator=(const FFT1D_C2C_invoker&) = delete;
  FFT1D_C2C_invoker(FFT1D_C2C_invoker&& o) noexcept : plan_(std::exchange(o.plan_, nullptr)), N(o.N) {}
  FFT1D_C2C_invoker& operator=(FFT1D_C2C_invoker&& o) noexcept {
    if (this != &o) {
      fyfft_plan_destroy(plan_);
      plan_ = std::exchange(o.plan_, nullptr);
      N = o.N;
    }
    return *this;
  }

  // Host pointers, synchronous (foye_FFT.hpp:128-151).
  template <bool invert>
  void dispatch(const float* input, float* output) {
    check(fyfft_execute_host(plan_, input, output, invert ? FYFFT_INVERSE : FYFFT_FORWARD),
          "dispatch");
  }

  // Device pointers, asynchronous on `stream`.
  template <bool invert>
  void dispatch_device(const float* d_input, float* d_output, void* stream = nullptr) {
    check(fyfft_execute(plan_, d_input, d_output, invert ? FYFFT_INVERSE : FYFFT_FORWARD, stream),
          "dispatch_device");
  }

  static FFT_execution_strategy acquired_execution_strategy(std::size_t length) {
    int s = 0;
    if (fyfft_length_strategy(length, &s) != FYFFT_SUCCESS)
      throw std::invalid_argument("acquired_execution_strategy: invalid length");
    return static_cast<FFT_execution_strategy>(s);
  }

  std::size_t length() const { return N; }
  fyfft_plan* native_plan() const { return plan_; }

 private:
  static void check(fyfft_status s, const char* what) {
    if (s != FYFFT_SUCCESS)
      throw std::runtime_error(std::string("fyfft ") + what + ": " + fyfft_status_string(s) +
                               " " + fyfft_last_error());
  }
  fyfft_plan* plan_ = nullptr;
  std::size_t N;
};

// Base of the strategy invokers (utility.hpp:6-11).
struct basic_unsmooth_path_invoker {
  virtual ~basic_unsmooth_path_invoker() = default;
  virtual void forward(const float* input, float* output) {}
  virtual void backward(const float* input, float* output) {}
};

namespace detail {
// One length-N complex<float> plan of a fixed strategy on host pointers.
template <int Strategy>
struct strategy_invoker : basic_unsmooth_path_invoker {
  explicit strategy_invoker(std::size_t length) {
    if (length < 1 || length > FYFFT_MAX_LENGTH)
      throw std::invalid_argument("fyfft strategy invoker: length " + std::to_string(length) + " out of range");
    check(fyfft_plan_create_strategy(&plan_, length, 1, FYFFT_F32, Strategy), "plan");
  }
  ~strategy_invoker() override { fyfft_plan_destroy(plan_); }
  strategy_invoker(const strategy_invoker&) = delete;
  strategy_invoker& operator=(const strategy_invoker&) = delete;
  strategy_invoker(strategy_invoker&& o) noexcept : plan_(std::exchange(o.plan_, nullptr)) {}
  strategy_invoker& operator=(strategy_invoker&& o) noexcept {
    if (this != &o) {
      fyfft_plan_destroy(plan_);
      plan_ = std::exchange(o.plan_, nullptr);
    }
    return *this;
  }
  void forward(const float* input, float* output) override {
    check(fyfft_execute_host(plan_, input, output, FYFFT_FORWARD), "forward");
  }
  void backward(const float* input, float* output) override {
    check(fyfft_execute_host(plan_, input, output, FYFFT_INVERSE), "backward");
  }
  static bool is_processable(std::size_t length) {
    int ok = 0;
    return fyfft_strategy_processable(Strategy, length, &ok) == FYFFT_SUCCESS && ok != 0;
  }
  fyfft_plan* native_plan() const { return plan_; }

 private:
  static void check(fyfft_status s, const char* what) {
    if (s != FYFFT_SUCCESS)
      throw std::runtime_error(std::string("fyfft ") + what + ": " + fyfft_status_string(s) + " " +
                               fyfft_last_error());
  }
  fyfft_plan* plan_ = nullptr;
};
}  // namespace detail

namespace mix_radix {
struct FFT1D_mixed_radix_invoker : detail::strategy_invoker<FYFFT_MIX_RADIX> {
  FFT1D_mixed_radix_invoker(std::size_t n) : strategy_invoker(n) {}
};
}  // namespace mix_radix
namespace rader {
struct FFT1D_C2C_rader_invoker : detail::strategy_invoker<FYFFT_RADER> {
  FFT1D_C2C_rader_invoker(std::size_t length) : strategy_invoker(length) {}
};
}  // namespace rader
namespace bluestein {
struct FFT1D_bluestein_invoker : detail::strategy_invoker<FYFFT_BLUESTEIN> {
  static constexpr std::size_t min_length = 2;
  static constexpr std::size_t max_length = 39062;  // the reference's limit (bluestein.hpp:189)
  FFT1D_bluestein_invoker(std::size_t length) : strategy_invoker(length) {}
};
}  // namespace bluestein
namespace dft {
struct FFT1D_C2C_dft_invoker : detail::strategy_invoker<FYFFT_DFT> {
  FFT1D_C2C_dft_invoker(std::size_t length) : strategy_invoker(length) {}
};
}  // namespace dft

using fy::fft::mix_radix::FFT1D_mixed_radix_invoker;
using fy::fft::rader::FFT1D_C2C_rader_invoker;
using fy::fft::bluestein::FFT1D_bluestein_invoker;
using fy::fft::dft::FFT1D_C2C_dft_invoker;

}  // namespace fy::fft

#endif  // FOYE_FFT_DROPIN_HPP_
