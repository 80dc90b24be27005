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

}  // namespace fy::fft

#endif  // FOYE_FFT_DROPIN_HPP_
