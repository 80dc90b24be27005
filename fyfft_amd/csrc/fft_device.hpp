// Device-side building blocks of the MI355X FFT: complex helpers, radix
// 2/3/4/5/7/8/16 DFT codelets held in registers, and the Stockham
// (self-sorting) pass that moves one transform between LDS / HBM and
// registers.
//
// Pass formulation.  With N = R_0 * R_1 * ... * R_{P-1}, pass p has radix
// R = R_p and span Ls = R_0 * ... * R_{p-1} (Ls = 1 for pass 0).  Butterfly
// j in [0, N/R) with k = j mod Ls reads x[j + r*N/R] (r < R), multiplies
// input r by w_{Ls*R}^{k*r}, runs an R-point DFT and writes
// y[(j-k)*R + k + r*Ls].  After the last pass the data is in natural order,
// so there is no digit-reversal step (the reference's direct invokers
// permute first: radix2.hpp:308-425, radix4.hpp:505-553).  Pass 0 reads its
// inputs with stride N/R and the last pass writes with stride N/R, so both
// touch HBM fully coalesced.
//
// Twiddle table layout (one per plan, FP64-generated): the factors of pass
// p start at element Ls_p - 1 and hold w_{Ls*R}^{k*r} at
// [Ls - 1 + k*(R-1) + (r-1)], k < Ls, 1 <= r < R.  The slices telescope to
// N - 1 entries in total.  Forward uses the table, inverse/adjoint its
// conjugate.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fyfft {
namespace dev {

template <typename V> struct scalar_of;
template <> struct scalar_of<float2> { using type = float; };
template <> struct scalar_of<double2> { using type = double; };
template <typename V> using S = typename scalar_of<V>::type;

template <typename V> __device__ __forceinline__ V cadd(V a, V b) { return V{a.x + b.x, a.y + b.y}; }
template <typename V> __device__ __forceinline__ V csub(V a, V b) { return V{a.x - b.x, a.y - b.y}; }
template <typename V> __device__ __forceinline__ V cscale(V a, S<V> s) { return V{a.x * s, a.y * s}; }
template <typename V> __device__ __forceinline__ V cmul(V a, V b) {
  return V{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// a * conj(b)
template <typename V> __device__ __forceinline__ V cmulc(V a, V b) {
  return V{a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
// a * w when DIR < 0 (forward), a * conj(w) when DIR > 0
template <int DIR, typename V> __device__ __forceinline__ V twmul(V a, V w) {
  if constexpr (DIR < 0) return cmul(a, w); else return cmulc(a, w);
}
// a * (DIR * i): multiplication by exp(DIR * i*pi/2)
template <int DIR, typename V> __device__ __forceinline__ V rot90(V a) {
  if constexpr (DIR < 0) return V{a.y, -a.x}; else return V{-a.y, a.x};
}
// a * (c + DIR*i*s) for constants c, s.  In FP32 each constant is split
// into its rounded value and the rounding remainder (c = c_hi + c_lo), and
// the remainder's product is added first: the codelets' rotations by
// exp(i*pi*m/8) and exp(2*pi*i*m/9) act on most outputs of every pass, so
// the constants' own rounding would otherwise add a correlated error of up
// to 2e-8 per pass (N = 4096 forward: 1.36e-7 -> 1.19e-7 rel-L2 together
// with the depth-one twiddle products of apply_twiddles).
template <int DIR, typename V> __device__ __forceinline__ V rotc(V a, double c, double s) {
  const double sd = DIR < 0 ? -s : s;
  const S<V> cc = (S<V>)c, ss = (S<V>)sd;
  if constexpr (sizeof(S<V>) == 4) {
    const float cl = (float)(c - (double)cc), sl = (float)(sd - (double)ss);
    const float lre = a.x * cl - a.y * sl, lim = a.x * sl + a.y * cl;
    return V{__builtin_fmaf(a.x, cc, __builtin_fmaf(-a.y, ss, lre)),
             __builtin_fmaf(a.x, ss, __builtin_fmaf(a.y, cc, lim))};
  }
  return V{a.x * cc - a.y * ss, a.x * ss + a.y * cc};
}

// Streaming global accesses: every element is read once and written once, so
// loads and stores are non-temporal (they do not displace the twiddle tables
// from the caches).
typedef float f32x2_t __attribute__((ext_vector_type(2)));
typedef double f64x2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float2 gload(const float2* p) {
  const f32x2_t v = __builtin_nontemporal_load(reinterpret_cast<const f32x2_t*>(p));
  return float2{v.x, v.y};
}
__device__ __forceinline__ double2 gload(const double2* p) {
  const f64x2_t v = __builtin_nontemporal_load(reinterpret_cast<const f64x2_t*>(p));
  return double2{v.x, v.y};
}
__device__ __forceinline__ void gstore(float2* p, float2 v) {
  __builtin_nontemporal_store(f32x2_t{v.x, v.y}, reinterpret_cast<f32x2_t*>(p));
}
__device__ __forceinline__ void gstore(double2* p, double2 v) {
  __builtin_nontemporal_store(f64x2_t{v.x, v.y}, reinterpret_cast<f64x2_t*>(p));
}

// 16-bit storage element: two 16-bit floats (re, im) in 4 bytes (H = __bf16
// or _Float16); loads widen to FP32, stores round to nearest even.
template <typename H> struct cplx16 { H re, im; };

template <typename H> __device__ __forceinline__ float2 load16(const cplx16<H>* p) {
  const unsigned u = __builtin_nontemporal_load(reinterpret_cast<const unsigned*>(p));
  const cplx16<H> c = __builtin_bit_cast(cplx16<H>, u);
  return float2{(float)c.re, (float)c.im};
}
template <typename H> __device__ __forceinline__ void store16(cplx16<H>* p, float2 v) {
  const cplx16<H> c{(H)v.x, (H)v.y};
  __builtin_nontemporal_store(__builtin_bit_cast(unsigned, c), reinterpret_cast<unsigned*>(p));
}

// cos/sin(2*pi*j/R) for the odd prime radices, exact decimal expansions.
template <int R> __device__ __forceinline__ constexpr double cos_r(int j) {
  j %= R;
  if constexpr (R == 3) {
    return j == 0 ? 1.0 : -0.5;
  } else if constexpr (R == 5) {
    return j == 0 ? 1.0 : (j == 1 || j == 4) ? 0.30901699437494742410229341718281906
                                             : -0.80901699437494742410229341718281906;
  } else {
    return j == 0 ? 1.0
         : (j == 1 || j == 6) ? 0.62348980185873353052500488400423981
         : (j == 2 || j == 5) ? -0.22252093395631440428890256449679476
                              : -0.90096886790241912623610231950744505;
  }
}
template <int R> __device__ __forceinline__ constexpr double sin_r(int j) {
  j %= R;
  if constexpr (R == 3) {
    return j == 0 ? 0.0 : j == 1 ? 0.86602540378443864676372317075293618
                                 : -0.86602540378443864676372317075293618;
  } else if constexpr (R == 5) {
    return j == 0 ? 0.0
         : j == 1 ? 0.95105651629515357211643933337938214
         : j == 2 ? 0.58778525229247312916870595463907277
         : j == 3 ? -0.58778525229247312916870595463907277
                  : -0.95105651629515357211643933337938214;
  } else {
    return j == 0 ? 0.0
         : j == 1 ? 0.78183148246802980870844452667405775
         : j == 2 ? 0.97492791218182360701813168299393122
         : j == 3 ? 0.43388373911755812047576833284835875
         : j == 4 ? -0.43388373911755812047576833284835875
         : j == 5 ? -0.97492791218182360701813168299393122
                  : -0.78183148246802980870844452667405775;
  }
}

constexpr double kSqrtHalf = 0.70710678118654752440084436210484904;
constexpr double kCos8 = 0.92387953251128675612818318939678829;   // cos(pi/8)
constexpr double kSin8 = 0.38268343236508977172845998403039887;   // sin(pi/8)

// ---------------------------------------------------------------- codelets
// In-place R-point DFT of v[0..R) with sign DIR (-1 forward, +1 inverse),
// natural order in and out, no scaling.
template <int R, int DIR, typename V> struct Dft;

template <int DIR, typename V> struct Dft<2, DIR, V> {
  static __device__ __forceinline__ void run(V* v) {
    V a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <int DIR, typename V> struct Dft<4, DIR, V> {
  static __device__ __forceinline__ void run(V* v) {
    V s0 = cadd(v[0], v[2]), d0 = csub(v[0], v[2]);
    V s1 = cadd(v[1], v[3]), d1 = rot90<DIR>(csub(v[1], v[3]));
    v[0] = cadd(s0, s1);
    v[2] = csub(s0, s1);
    v[1] = cadd(d0, d1);
    v[3] = csub(d0, d1);
  }
};

template <int DIR, typename V> struct Dft<8, DIR, V> {
  static __device__ __forceinline__ void run(V* v) {
    V e[4] = {v[0], v[2], v[4], v[6]};
    V o[4] = {v[1], v[3], v[5], v[7]};
    Dft<4, DIR, V>::run(e);
    Dft<4, DIR, V>::run(o);
    o[1] = rotc<DIR>(o[1], kSqrtHalf, kSqrtHalf);
    o[2] = rot90<DIR>(o[2]);
    o[3] = rotc<DIR>(o[3], -kSqrtHalf, kSqrtHalf);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      v[k] = cadd(e[k], o[k]);
      v[k + 4] = csub(e[k], o[k]);
    }
  }
};

// w16^m multiply for m = n2*k1 with n2, k1 in 1..3
template <int DIR, typename V> __device__ __forceinline__ V w16(V a, int m) {
  switch (m) {
    case 1: return rotc<DIR>(a, kCos8, kSin8);
    case 2: return rotc<DIR>(a, kSqrtHalf, kSqrtHalf);
    case 3: return rotc<DIR>(a, kSin8, kCos8);
    case 4: return rot90<DIR>(a);
    case 6: return rotc<DIR>(a, -kSqrtHalf, kSqrtHalf);
    case 9: return rotc<DIR>(a, -kCos8, -kSin8);
    default: return a;
  }
}

// 16 = 4 x 4: x[4*n1 + n2]; DFT4 over n1, twiddle w16^(n2*k1), DFT4 over n2,
// output X[k1 + 4*k2].
template <int DIR, typename V> struct Dft<16, DIR, V> {
  static __device__ __forceinline__ void run(V* v) {
    V y[16];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
      V t[4] = {v[n2], v[4 + n2], v[8 + n2], v[12 + n2]};
      Dft<4, DIR, V>::run(t);
#pragma unroll
      for (int k1 = 0; k1 < 4; ++k1) y[n2 * 4 + k1] = (n2 && k1) ? w16<DIR>(t[k1], n2 * k1) : t[k1];
    }
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
      V t[4] = {y[k1], y[4 + k1], y[8 + k1], y[12 + k1]};
      Dft<4, DIR, V>::run(t);
#pragma unroll
      for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = t[k2];
    }
  }
};

// Odd prime radix by symmetric pairs:
//   X_k = x0 + sum_m (x_m + x_{R-m}) cos(2pi mk/R) + DIR*i * sum_m (x_m - x_{R-m}) sin(2pi mk/R)
template <int R, int DIR, typename V> struct DftOdd {
  static __device__ __forceinline__ void run(V* v) {
    constexpr int H = (R - 1) / 2;
    V s[H], d[H];
#pragma unroll
    for (int m = 1; m <= H; ++m) {
      s[m - 1] = cadd(v[m], v[R - m]);
      d[m - 1] = csub(v[m], v[R - m]);
    }
    const V x0 = v[0];
    V sum = x0;
#pragma unroll
    for (int m = 0; m < H; ++m) sum = cadd(sum, s[m]);
    v[0] = sum;
#pragma unroll
    for (int k = 1; k <= H; ++k) {
      V a = x0, b = V{0, 0};
#pragma unroll
      for (int m = 1; m <= H; ++m) {
        const S<V> c = (S<V>)cos_r<R>(m * k), sn = (S<V>)sin_r<R>(m * k);
        a = V{a.x + s[m - 1].x * c, a.y + s[m - 1].y * c};
        b = V{b.x + d[m - 1].x * sn, b.y + d[m - 1].y * sn};
      }
      const V ib = rot90<DIR>(b);
      v[k] = cadd(a, ib);
      v[R - k] = csub(a, ib);
    }
  }
};
template <int DIR, typename V> struct Dft<3, DIR, V> : DftOdd<3, DIR, V> {};

// cos/sin(2*pi*m/9) for m = 1, 2, 4
constexpr double kCos9[3] = {0.766044443118978035202392650555416674, 0.173648177666930348851716626769314796,
                             -0.939692620785908384054109277324731470};
constexpr double kSin9[3] = {0.642787609686539326322643409907263433, 0.984807753012208059366743024589523014,
                             0.342020143325668733044099614682259581};

// 9 = 3 x 3: x[3*n1 + n2]; DFT3 over n1, twiddle w9^(n2*k1), DFT3 over n2,
// output X[k1 + 3*k2].
template <int DIR, typename V> struct Dft<9, DIR, V> {
  static __device__ __forceinline__ void run(V* v) {
    V y[9];
#pragma unroll
    for (int n2 = 0; n2 < 3; ++n2) {
      V t[3] = {v[n2], v[3 + n2], v[6 + n2]};
      DftOdd<3, DIR, V>::run(t);
#pragma unroll
      for (int k1 = 0; k1 < 3; ++k1) {
        const int m = n2 * k1;  // 0, 1, 2 or 4
        const int i = m == 4 ? 2 : (m == 2 ? 1 : 0);
        y[n2 * 3 + k1] = m ? rotc<DIR>(t[k1], kCos9[i], kSin9[i]) : t[k1];
      }
    }
#pragma unroll
    for (int k1 = 0; k1 < 3; ++k1) {
      V t[3] = {y[k1], y[3 + k1], y[6 + k1]};
      DftOdd<3, DIR, V>::run(t);
#pragma unroll
      for (int k2 = 0; k2 < 3; ++k2) v[k1 + 3 * k2] = t[k2];
    }
  }
};
template <int DIR, typename V> struct Dft<5, DIR, V> : DftOdd<5, DIR, V> {};
template <int DIR, typename V> struct Dft<7, DIR, V> : DftOdd<7, DIR, V> {};

// cos and sin of 2*pi*m/R for integers m, R (R > 0), evaluated at compile
// time: exact reduction of 4m mod 4R to a quadrant and a remainder r < R,
// then the Taylor series of sin and cos on theta = (pi/2) * r / R <= pi/2
// (terms to degree 29: truncation far below double rounding).
struct UnitRoot { double c, s; };
__host__ __device__ constexpr UnitRoot unit_root(int m, int R) {
  const long long m4 = ((4LL * m) % (4LL * R) + 4LL * R) % (4LL * R);
  const long long q = m4 / R, r = m4 - q * R;
  const double th = 1.57079632679489661923132169163975144 * ((double)r / (double)R);
  double sn = th, cs = 1.0, ts = th, tc = 1.0;
  for (int k = 1; k <= 14; ++k) {
    ts *= -th * th / (double)((2 * k) * (2 * k + 1));
    tc *= -th * th / (double)((2 * k - 1) * (2 * k));
    sn += ts;
    cs += tc;
  }
  if (r == 0) sn = 0.0, cs = 1.0;
  switch (q) {
    case 0: return UnitRoot{cs, sn};
    case 1: return UnitRoot{-sn, cs};
    case 2: return UnitRoot{-cs, -sn};
    default: return UnitRoot{sn, -cs};
  }
}

// a * w_R^m (w_R = exp(DIR * 2*pi*i / R)) for an m known after unrolling.
template <int R, int DIR, typename V> __device__ __forceinline__ V wconst(V a, int m) {
  m %= R;
  if (m == 0) return a;
  if (2 * m == R) return V{-a.x, -a.y};
  if (4 * m == R) return rot90<DIR>(a);
  if (4 * m == 3 * R) return rot90<-DIR>(a);
  const UnitRoot u = unit_root(m, R);
  return rotc<DIR>(a, u.c, u.s);
}

// Composite radix R = A * B in registers (two levels, like Dft<16> and
// Dft<9>): x[B*n1 + n2]; DFT_A over n1, twiddle w_R^(n2*k1), DFT_B over n2,
// output X[k1 + A*k2].  One pass of such a radix replaces two LDS passes.
template <int A, int B, int DIR, typename V> struct DftAB {
  static __device__ __forceinline__ void run(V* v) {
    constexpr int R = A * B;
    V y[R];
#pragma unroll
    for (int n2 = 0; n2 < B; ++n2) {
      V t[A];
#pragma unroll
      for (int n1 = 0; n1 < A; ++n1) t[n1] = v[B * n1 + n2];
      Dft<A, DIR, V>::run(t);
#pragma unroll
      for (int k1 = 0; k1 < A; ++k1) y[n2 * A + k1] = wconst<R, DIR>(t[k1], n2 * k1);
    }
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1) {
      V t[B];
#pragma unroll
      for (int n2 = 0; n2 < B; ++n2) t[n2] = y[n2 * A + k1];
      Dft<B, DIR, V>::run(t);
#pragma unroll
      for (int k2 = 0; k2 < B; ++k2) v[k1 + A * k2] = t[k2];
    }
  }
};
// The composite radices of the runtime passes and their two factors.
template <int R> struct Composite { static constexpr int A = 0, B = 0; };
template <> struct Composite<6> { static constexpr int A = 2, B = 3; };
template <> struct Composite<10> { static constexpr int A = 2, B = 5; };
template <> struct Composite<12> { static constexpr int A = 4, B = 3; };
template <> struct Composite<14> { static constexpr int A = 2, B = 7; };
template <> struct Composite<15> { static constexpr int A = 3, B = 5; };
template <> struct Composite<20> { static constexpr int A = 4, B = 5; };
template <> struct Composite<21> { static constexpr int A = 3, B = 7; };
template <> struct Composite<25> { static constexpr int A = 5, B = 5; };
template <int DIR, typename V> struct Dft<6, DIR, V> : DftAB<2, 3, DIR, V> {};
template <int DIR, typename V> struct Dft<10, DIR, V> : DftAB<2, 5, DIR, V> {};
template <int DIR, typename V> struct Dft<12, DIR, V> : DftAB<4, 3, DIR, V> {};
template <int DIR, typename V> struct Dft<14, DIR, V> : DftAB<2, 7, DIR, V> {};
template <int DIR, typename V> struct Dft<15, DIR, V> : DftAB<3, 5, DIR, V> {};
template <int DIR, typename V> struct Dft<20, DIR, V> : DftAB<4, 5, DIR, V> {};
template <int DIR, typename V> struct Dft<21, DIR, V> : DftAB<3, 7, DIR, V> {};
template <int DIR, typename V> struct Dft<25, DIR, V> : DftAB<5, 5, DIR, V> {};

// ---------------------------------------------------------------- twiddles
// Multiply inputs r = 1..R-1 of one butterfly by w^(k*r), where t points at
// the pass table slice [w^k, w^2k, ..., w^((R-1)k)].  For R = 7, 8, 9 and 16
// a subset of the powers is loaded and each other power is ONE product of
// two loaded ones (R = 16: w^a * w^b with a in {1,2,3}, b in {4,8,12};
// R = 9: a in {1,2}, b in {3,6}; R = 8: w^1 times {2,4,6}; R = 7: w^1 or
// w^2 times w^4), so every factor carries at most three roundings; chains
// of two or three products measurably raised the transform's error
// (N = 4096 forward: 1.36e-7 rel-L2 with w^15 = ((w^1 w^2) w^4) w^8, 1.195e-7
// with the depth-one products and the split rotation constants of rotc).
// R = 2..5 load every factor.  Loading all fifteen factors of a radix-16
// butterfly instead was measured 3-4 % slower at N = 4096 (1.16e-7), even
// for pass 1, whose 240-entry table slice stays in the L1 cache, and so was
// loading the factors of passes 1 and 2 before pass 0.
template <int R, int DIR, typename V>
__device__ __forceinline__ void apply_twiddles(V* v, const V* __restrict__ t) {
  if constexpr (Composite<R>::A > 0) {
    // R = A * B: w^(a + A*b) = w^a * w^(A*b), both loaded (depth-one products)
    constexpr int A = Composite<R>::A, B = Composite<R>::B;
    V lo[A], hi[B];
    lo[0] = hi[0] = V{1, 0};
#pragma unroll
    for (int a = 1; a < A; ++a) lo[a] = t[a - 1];
#pragma unroll
    for (int b = 1; b < B; ++b) hi[b] = t[A * b - 1];
#pragma unroll
    for (int b = 0; b < B; ++b)
#pragma unroll
      for (int a = 0; a < A; ++a) {
        const int r = a + A * b;
        if (r == 0) continue;
        v[r] = twmul<DIR>(v[r], a == 0 ? hi[b] : (b == 0 ? lo[a] : cmul(lo[a], hi[b])));
      }
  } else if constexpr (R <= 5) {
#pragma unroll
    for (int r = 1; r < R; ++r) v[r] = twmul<DIR>(v[r], t[r - 1]);
  } else if constexpr (R == 16) {
    const V w1 = t[0], w2 = t[1], w3 = t[2], w4 = t[3], w8 = t[7], w12 = t[11];
    const V w[16] = {V{1, 0}, w1, w2, w3, w4, cmul(w1, w4), cmul(w2, w4), cmul(w3, w4), w8,
                     cmul(w1, w8), cmul(w2, w8), cmul(w3, w8), w12, cmul(w1, w12), cmul(w2, w12),
                     cmul(w3, w12)};
#pragma unroll
    for (int r = 1; r < 16; ++r) v[r] = twmul<DIR>(v[r], w[r]);
  } else if constexpr (R == 9) {
    const V w1 = t[0], w2 = t[1], w3 = t[2], w6 = t[5];
    const V w[9] = {V{1, 0}, w1, w2, w3, cmul(w1, w3), cmul(w2, w3), w6, cmul(w1, w6), cmul(w2, w6)};
#pragma unroll
    for (int r = 1; r < 9; ++r) v[r] = twmul<DIR>(v[r], w[r]);
  } else if constexpr (R == 8) {
    const V w1 = t[0], w2 = t[1], w4 = t[3], w6 = t[5];
    const V w[8] = {V{1, 0}, w1, w2, cmul(w1, w2), w4, cmul(w1, w4), w6, cmul(w1, w6)};
#pragma unroll
    for (int r = 1; r < 8; ++r) v[r] = twmul<DIR>(v[r], w[r]);
  } else {  // R == 7
    const V w1 = t[0], w2 = t[1], w4 = t[3];
    const V w[7] = {V{1, 0}, w1, w2, cmul(w1, w2), w4, cmul(w1, w4), cmul(w2, w4)};
#pragma unroll
    for (int r = 1; r < 7; ++r) v[r] = twmul<DIR>(v[r], w[r]);
  }
}
// ---------------------------------------------------------------- uniform values
// Copy of a trivially copyable descriptor with every 32-bit word taken from
// the first active lane: the compiler then keeps it in scalar registers
// (a descriptor reached through a pointer argument is otherwise per-lane).
template <typename T> __device__ __forceinline__ T make_uniform(const T& v) {
  static_assert(sizeof(T) % 4 == 0, "4-byte multiple expected");
  struct W { int w[sizeof(T) / 4]; };
  const W src = __builtin_bit_cast(W, v);
  W dst;
#pragma unroll
  for (int i = 0; i < (int)(sizeof(T) / 4); ++i) dst.w[i] = __builtin_amdgcn_readfirstlane(src.w[i]);
  return __builtin_bit_cast(T, dst);
}

// ---------------------------------------------------------------- XCD mapping
// The dispatcher hands workgroup b of a launch to XCD (b mod 8).  For passes
// where several consecutive logical workgroups work on one transform (the
// four-step column/row groups, the Rader gather and scatter), remapping gives
// each XCD one contiguous range of logical workgroups, so a transform's
// cache lines are fetched into one L2 instead of eight.  It is a bijection
// on [0, total) for any total.
constexpr int kXcds = 8;
// Each XCD walks its range from a different eighth of it, so the workgroups
// that run at the same moment on different XCDs are not a power-of-two
// number of transforms apart (measured 2-5 % faster for the four-step,
// column and convolution passes: N = 65536 x 2048 0.858 -> 0.843 ms,
// Bluestein N = 1000003 x 16 0.534 -> 0.506 ms).
__device__ __forceinline__ int xcd_remap(int b, int total) {
  const int per = total / kXcds, rem = total % kXcds;
  const int x = b % kXcds, i = b / kXcds;
  const int len = x < rem ? per + 1 : per;
  const int start = x < rem ? x * (per + 1) : rem * (per + 1) + (x - rem) * per;
  int j = i + (int)(((long long)len * x) / kXcds);
  if (j >= len) j -= len;
  return start + j;
}

// ---------------------------------------------------------------- address spaces
// Pointers that cross a call boundary lose their address space; passing LDS
// as address_space(3) and tables as address_space(1) keeps ds_* / global_*
// instructions instead of flat ones in out-of-line code.
template <typename V> using lds_ptr = __attribute__((address_space(3))) V*;
template <typename V> using lds_cptr = const __attribute__((address_space(3))) V*;
template <typename V> using gbl_cptr = const __attribute__((address_space(1))) V*;
// The Stockham twiddle table of a runtime pass loop: global memory, or (TL)
// a copy in LDS that the workgroup staged once for all its transforms.
template <bool TL, typename V> struct TwTable { using type = gbl_cptr<V>; };
template <typename V> struct TwTable<true, V> { using type = lds_cptr<V>; };
template <bool TL, typename V> using tw_ptr = typename TwTable<TL, V>::type;
template <typename V> __device__ __forceinline__ V* as_lds(V* p) { return (V*)(lds_ptr<V>)p; }

// ---------------------------------------------------------------- LDS layout
// One pad element every 16 complex values breaks the power-of-two strides
// of the pass read/write patterns across LDS banks.
constexpr int kPadShift = 4;
__host__ __device__ __forceinline__ constexpr int lds_len(int n, int ps) { return (n - 1) + ((n - 1) >> ps) + 1; }

constexpr int kMaxPasses = 16;
struct PassDesc {
  int npass;
  int radix[kMaxPasses];
};

// The pass list packed 5 bits per radix (radices are at most 25, stored as
// R - 1; at most kMaxPackedPasses passes) so that the runtime pass loops
// read the radix of pass p with a shift instead of indexing the array, which
// would keep the descriptor in scratch memory.
constexpr int kMaxPackedPasses = 12;
__device__ __forceinline__ unsigned long long pack_radices(const PassDesc& pd) {
  unsigned long long code = 0;
#pragma unroll
  for (int i = 0; i < kMaxPackedPasses; ++i) code |= (unsigned long long)((pd.radix[i] - 1) & 31) << (5 * i);
  return code;
}
__device__ __forceinline__ int radix_at(unsigned long long code, int p) { return (int)((code >> (5 * p)) & 31) + 1; }

__host__ __device__ constexpr int cdiv(int a, int b) { return (a + b - 1) / b; }
// FP32 runtime-pass kernels bounded at 8 points per thread serve pass lists
// without radix 16 only (the planner's small-radix geometry).
template <typename V, int PPT> constexpr bool kNo16 = (sizeof(V) == 8 && PPT <= 8) || (sizeof(V) == 16 && PPT <= 4);

// ---------------------------------------------------------------- pass
// One Stockham pass of radix R over the transform of length N held by T
// threads (this thread is `tid`).  SRC_G / DST_G select global (HBM) or LDS
// addressing (LDS index i is stored at i + (i >> ps)).  AM, the LDS address
// form: 0 the padded index per point; 1 the transform is stored unpadded
// (ps >= 30) and the index is used as is; 2 padded, where the pass's strides
// keep a butterfly's points at one padded base plus uniform steps (N / R a
// multiple of 2^ps for the reads; Ls a multiple of 2^ps, or R * Ls dividing
// 2^ps, for the writes: fast_pad_pass).  Forms 1 and 2 address each point
// with one add.  Each thread owns at most MAXB butterflies.  When source and destination are both LDS the pass is in
// place: every read completes (barrier) before any write.  A barrier
// follows LDS writes so the next pass sees them.  `scale` is applied only
// on global stores.
template <typename V, int R, int MAXB, int DIR, bool SRC_G, bool DST_G, int AM = 0>
__device__ __forceinline__ void stockham_pass(const V* src, V* dst, int N, int Ls, int T, int tid,
                                              const V* __restrict__ tw, bool active, S<V> scale,
                                              int ps) {
  const int nb = N / R;
  V v[MAXB][R];
  // k = j mod Ls and q = j / Ls for j = tid + b*T, stepped without a
  // division per butterfly: two divisions per pass.
  int q0 = 0, k0 = 0, tq = 0, tk = 0;
  if (Ls > 1) {
    q0 = tid / Ls;
    k0 = tid - q0 * Ls;
    tq = T / Ls;
    tk = T - tq * Ls;
  }
  {
    int k = k0;
#pragma unroll
    for (int b = 0; b < MAXB; ++b) {
      const int j = tid + b * T;
      if (active && j < nb) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int idx = j + r * nb;
          v[b][r] = SRC_G ? gload(src + idx)
                    : src[AM == 1 ? idx : AM == 2 ? j + (j >> ps) + r * (nb + (nb >> ps)) : idx + (idx >> ps)];
        }
        if (Ls > 1) apply_twiddles<R, DIR>(v[b], tw + (Ls - 1) + k * (R - 1));
        Dft<R, DIR, V>::run(v[b]);
      }
      if (Ls > 1) {
        k += tk;
        if (k >= Ls) k -= Ls;
      }
    }
  }
  if (!SRC_G && !DST_G) __syncthreads();
  {
    int q = q0, k = k0;
#pragma unroll
    for (int b = 0; b < MAXB; ++b) {
      const int j = tid + b * T;
      if (active && j < nb) {
        const int base = (Ls > 1) ? q * Ls * R + k : j * R;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int idx = base + r * Ls;
          if (DST_G) gstore(dst + idx, cscale(v[b][r], scale));
          else if (AM == 2) dst[base + (base >> ps) + r * (Ls + (((Ls >> ps) << ps) == Ls ? Ls >> ps : 0))] = v[b][r];
          else dst[AM == 1 ? idx : idx + (idx >> ps)] = v[b][r];
        }
      }
      if (Ls > 1) {
        q += tq;
        k += tk;
        if (k >= Ls) { k -= Ls; ++q; }
      }
    }
  }
  if (!DST_G) __syncthreads();
}

template <int R> struct RadixC { static constexpr int value = R; };
// NO16: the pass list has no radix 16 (the caller guarantees it), so that
// case is not instantiated and the registers are sized for radix 9 at most.
// CMP: the composite radices 6, 10, 12, 14, 15, 20, 21, 25 are cases too
// (FP32 four-step sub-transforms at 16 points per thread; the planner emits
// them only for those).
template <bool NO16 = false, bool CMP = false, typename F> __device__ __forceinline__ void radix_switch(int R, F f) {
  if constexpr (CMP) {
    switch (R) {
      case 6: f(RadixC<6>{}); return;
      case 10: f(RadixC<10>{}); return;
      case 12: f(RadixC<12>{}); return;
      case 14: f(RadixC<14>{}); return;
      case 15: f(RadixC<15>{}); return;
      case 20: f(RadixC<20>{}); return;
      case 21: f(RadixC<21>{}); return;
      case 25: f(RadixC<25>{}); return;
      default: break;
    }
  }
  switch (R) {
    case 2: f(RadixC<2>{}); break;
    case 3: f(RadixC<3>{}); break;
    case 4: f(RadixC<4>{}); break;
    case 5: f(RadixC<5>{}); break;
    case 7: f(RadixC<7>{}); break;
    case 8: f(RadixC<8>{}); break;
    case 9: f(RadixC<9>{}); break;
    default:
      if constexpr (NO16) f(RadixC<8>{});
      else f(RadixC<16>{});
      break;
  }
}

// One LDS -> LDS pass kept out of line: inlining every radix case into the
// pass loop of lds_fft makes the register allocator keep all cases' arrays
// live at once and spill; as separate functions each gets its own budget.
template <typename V, int R, int MAXB, int DIR, bool TL, int AM>
__device__ __noinline__ void lds_pass_ool(lds_ptr<V> lds_s, int N, int Ls, int T, int tid,
                                          tw_ptr<TL, V> tw_g, bool active, int ps) {
  V* lds = (V*)lds_s;
  stockham_pass<V, R, MAXB, DIR, false, false, AM>(lds, lds, N, Ls, T, tid, (const V*)tw_g, active, 1, ps);
}

// Whether a padded radix-R pass at stride Ls over N points keeps each
// butterfly's padded addresses linear (stockham_pass address form 2).
__device__ __forceinline__ bool fast_pad_pass(int N, int Ls, int R, int ps) {
  const int m = (1 << ps) - 1;
  return ((N / R) & m) == 0 && ((Ls & m) == 0 || ((m + 1) % (Ls * R)) == 0);
}

// Run every pass of `pd` on an LDS-resident transform (LDS -> LDS).
// PPT bounds the points per thread (N <= PPT * T).  TL: tw_gen points at an
// LDS copy of the table.
template <typename V, int PPT, int DIR, bool TL = false, bool CMP = false>
__device__ __forceinline__ void lds_fft(V* lds_gen, int N, int T, int tid, const PassDesc& pd,
                                        const V* __restrict__ tw_gen, bool active, int ps) {
  const lds_ptr<V> lds = (lds_ptr<V>)lds_gen;
  const tw_ptr<TL, V> tw = (tw_ptr<TL, V>)tw_gen;
  const unsigned long long code = pack_radices(pd);
  int Ls = 1;
  for (int p = 0; p < pd.npass; ++p) {
    const int R = radix_at(code, p);
    radix_switch<kNo16<V, PPT>, CMP>(R, [&](auto rc) {
      constexpr int RR = decltype(rc)::value;
      if (ps >= 30) lds_pass_ool<V, RR, cdiv(PPT, RR), DIR, TL, 1>(lds, N, Ls, T, tid, tw, active, ps);
      else if (fast_pad_pass(N, Ls, RR, ps)) lds_pass_ool<V, RR, cdiv(PPT, RR), DIR, TL, 2>(lds, N, Ls, T, tid, tw, active, ps);
      else lds_pass_ool<V, RR, cdiv(PPT, RR), DIR, TL, 0>(lds, N, Ls, T, tid, tw, active, ps);
    });
    Ls *= R;
  }
}

// ---------------------------------------------------------------- I/O-fused passes
// The first pass of an LDS-resident transform can read its inputs through a
// source functor (HBM, a gather, a chirp pre-multiply ...) straight into
// registers, and the last pass can hand its outputs to a sink functor (HBM,
// a scatter, a kernel multiply back into LDS): that saves the LDS round trip
// and the barrier at either end.  kFromLds / kToLds mark functors that read /
// write the transform's LDS buffer: a pass whose source and sink both touch
// it is in place and completes every read (barrier) before any write, and a
// sink that writes LDS is followed by a barrier.
template <typename V> struct LdsSrc {
  static constexpr bool kFromLds = true;
  lds_ptr<V> p;
  int ps;
  __device__ V operator()(int i) const { return ((const V*)p)[i + (i >> ps)]; }
};
template <typename V> struct LdsSink {
  static constexpr bool kToLds = true;
  lds_ptr<V> p;
  int ps;
  __device__ void operator()(int i, V v) const { ((V*)p)[i + (i >> ps)] = v; }
};

// A source with kSplit = true takes the index of a first-pass point as
// (j, u): j the butterfly (per thread) and u = r * (N / R), uniform across
// the workgroup and a compile-time constant in the fixed pass lists, and
// works in two phases: load(j, u) for every point of the butterfly (its
// loads all in flight), then finish(raw, j, u).
template <typename Src, typename = void> struct SplitSrc { static constexpr bool value = false; };
template <typename Src> struct SplitSrc<Src, decltype((void)Src::kSplit)> {
  static constexpr bool value = Src::kSplit;
};

template <typename V, int R, int MAXB, int DIR, typename Src, typename Dst>
__device__ __forceinline__ void io_pass(Src src, Dst dst, int N, int Ls, int T, int tid,
                                        const V* __restrict__ tw, bool active) {
  const int nb = N / R;
  V v[MAXB][R];
  int q0 = 0, k0 = 0, tq = 0, tk = 0;
  if (Ls > 1) {
    q0 = tid / Ls;
    k0 = tid - q0 * Ls;
    tq = T / Ls;
    tk = T - tq * Ls;
  }
  {
    int k = k0;
#pragma unroll
    for (int b = 0; b < MAXB; ++b) {
      const int j = tid + b * T;
      if (active && j < nb) {
        if constexpr (SplitSrc<Src>::value) {
          typename Src::Raw raw[R];
#pragma unroll
          for (int r = 0; r < R; ++r) raw[r] = src.load(j, r * nb);
#pragma unroll
          for (int r = 0; r < R; ++r) v[b][r] = src.finish(raw[r], j, r * nb);
        } else {
#pragma unroll
          for (int r = 0; r < R; ++r) v[b][r] = src(j + r * nb);
        }
        if (Ls > 1) apply_twiddles<R, DIR>(v[b], tw + (Ls - 1) + k * (R - 1));
        Dft<R, DIR, V>::run(v[b]);
      }
      if (Ls > 1) {
        k += tk;
        if (k >= Ls) k -= Ls;
      }
    }
  }
  if constexpr (Src::kFromLds && Dst::kToLds) __syncthreads();
  {
    int q = q0, k = k0;
#pragma unroll
    for (int b = 0; b < MAXB; ++b) {
      const int j = tid + b * T;
      if (active && j < nb) {
        const int base = (Ls > 1) ? q * Ls * R + k : j * R;
#pragma unroll
        for (int r = 0; r < R; ++r) dst(base + r * Ls, v[b][r]);
      }
      if (Ls > 1) {
        q += tq;
        k += tk;
        if (k >= Ls) { k -= Ls; ++q; }
      }
    }
  }
  if constexpr (Dst::kToLds) __syncthreads();
}

// Out-of-line form for the runtime pass loops.
template <typename V, int R, int MAXB, int DIR, bool TL, typename Src, typename Dst>
__device__ __noinline__ void io_pass_ool(Src src, Dst dst, int N, int Ls, int T, int tid,
                                         tw_ptr<TL, V> tw_g, bool active) {
  io_pass<V, R, MAXB, DIR>(src, dst, N, Ls, T, tid, (const V*)tw_g, active);
}

// Compile-time pass list (N, T and the radices fixed, Ls starting at 1) with
// the first pass reading `src` and the last writing `dst`; `lds` holds the
// transform between passes (pad shift ps).
template <typename V, int DIR, int N, int T, int Ls, typename Src, typename Dst, int R, int... Rest>
__device__ __forceinline__ void fixed_io_passes(Src src, Dst dst, V* lds, int ps, int tid,
                                                const V* __restrict__ tw, bool active) {
  constexpr bool first = Ls == 1, last = sizeof...(Rest) == 0;
  constexpr int MAXB = cdiv(N / R, T);
  const LdsSrc<V> ls{(lds_ptr<V>)lds, ps};
  const LdsSink<V> lk{(lds_ptr<V>)lds, ps};
  if constexpr (first && last) io_pass<V, R, MAXB, DIR>(src, dst, N, Ls, T, tid, tw, active);
  else if constexpr (first) io_pass<V, R, MAXB, DIR>(src, lk, N, Ls, T, tid, tw, active);
  else if constexpr (last) io_pass<V, R, MAXB, DIR>(ls, dst, N, Ls, T, tid, tw, active);
  else io_pass<V, R, MAXB, DIR>(ls, lk, N, Ls, T, tid, tw, active);
  if constexpr (!last) fixed_io_passes<V, DIR, N, T, Ls * R, Src, Dst, Rest...>(src, dst, lds, ps, tid, tw, active);
}

// Every pass of `pd` with the first reading `src` and the last writing
// `dst`; `lds` (padded by ps) holds the transform between passes.  Every
// thread of the workgroup must call it (barriers).
template <typename V, int PPT, int DIR, bool TL = false, bool CMP = false, typename Src, typename Dst>
__device__ __forceinline__ void lds_fft_io(Src src, Dst dst, V* lds_gen, int N, int T, int tid,
                                           const PassDesc& pd, const V* __restrict__ tw_gen,
                                           bool active, int ps) {
  const lds_ptr<V> lds = (lds_ptr<V>)lds_gen;
  const tw_ptr<TL, V> tw = (tw_ptr<TL, V>)tw_gen;
  const LdsSrc<V> ls{lds, ps};
  const LdsSink<V> lk{lds, ps};
  const int P = pd.npass;
  if (P <= 1) {
    if (P == 1)
      radix_switch<kNo16<V, PPT>, CMP>(radix_at(pack_radices(pd), 0), [&](auto rc) {
        constexpr int R = decltype(rc)::value;
        io_pass_ool<V, R, cdiv(PPT, R), DIR, TL>(src, lk, N, 1, T, tid, tw, active);
      });
    else if (active)
      for (int i = tid; i < N; i += T) lk(i, src(i));
    if (P == 0) __syncthreads();
    if (active)
      for (int i = tid; i < N; i += T) dst(i, ls(i));
    if constexpr (Dst::kToLds) __syncthreads();
    return;
  }
  const unsigned long long code = pack_radices(pd);
  int Ls = 1;
  for (int p = 0; p < P; ++p) {
    radix_switch<kNo16<V, PPT>, CMP>(radix_at(code, p), [&](auto rc) {
      constexpr int R = decltype(rc)::value;
      if (p == 0) io_pass_ool<V, R, cdiv(PPT, R), DIR, TL>(src, lk, N, Ls, T, tid, tw, active);
      else if (p == P - 1) io_pass_ool<V, R, cdiv(PPT, R), DIR, TL>(ls, dst, N, Ls, T, tid, tw, active);
      else io_pass_ool<V, R, cdiv(PPT, R), DIR, TL>(ls, lk, N, Ls, T, tid, tw, active);
    });
    Ls *= radix_at(code, p);
  }
}

// w_n^m = exp(-2*pi*i*m/n) for 0 <= m < n <= 2^24 in FP32 (conjugated when
// DIR > 0), computed instead of looked up: an exact integer reduction of 8m
// to an octant q and remainder r (odd octants reflected, r -> n - r, so the
// angle left is x * pi/4 with x = r/n in [0, 1]), x formed with one
// residual correction (about half an ulp), then the Taylor polynomials of
// sin(pi*x/4) and cos(pi*x/4) to degree 9 and 10 (truncation below 2e-9)
// and the octant's symmetry.  The four-step twiddle w_N^(n2*k1) is needed
// once per element at scattered exponents; two table lookups per element
// (hi * lo) made the passes that apply it L2-request bound.
template <int DIR>
__device__ __forceinline__ float2 root_f32(int m, int n, float inv_n) {
  const int t = m << 3;                                  // 8m < 2^27
  int q = (int)((float)t * inv_n);                       // octant estimate, off by at most 1
  int r = t - q * n;
  if (r < 0) { r += n; --q; }
  if (r >= n) { r -= n; ++q; }
  const bool odd = q & 1;
  const int rr = odd ? n - r : r;                        // reflect odd octants
  const float rf = (float)rr, nf = (float)n;
  float x = rf * inv_n;
  x = __builtin_fmaf(__builtin_fmaf(-x, nf, rf), inv_n, x);
  const float x2 = x * x;
  // sin(pi x / 4), cos(pi x / 4): Taylor coefficients (pi/4)^k / k! with alternating signs
  const float sn = x * __builtin_fmaf(x2, __builtin_fmaf(x2, __builtin_fmaf(x2, __builtin_fmaf(x2,
                   (float)3.13361689037812e-07, (float)-3.657620418217724e-05), (float)0.00249039457019272),
                   (float)-0.08074551218828077), (float)0.7853981633974483);
  const float cs = __builtin_fmaf(x2, __builtin_fmaf(x2, __builtin_fmaf(x2, __builtin_fmaf(x2, __builtin_fmaf(x2,
                   (float)-2.461136950494199e-08, (float)3.590860448591509e-06), (float)-0.00032599188692738996),
                   (float)0.0158543442438155), (float)-0.30842513753404244), 1.0f);
  const float c = odd ? sn : cs, s = odd ? cs : sn;      // cos / sin of the angle within its quadrant
  const int qd = (q >> 1) & 3;
  float re, im;
  if (qd == 0) { re = c; im = s; }
  else if (qd == 1) { re = -s; im = c; }
  else if (qd == 2) { re = -c; im = -s; }
  else { re = s; im = -c; }
  return DIR < 0 ? float2{re, -im} : float2{re, im};
}

// ---------------------------------------------------------------- roots
// exp(sign * 2*pi*i * m / D) in FP64 with an exact integer reduction of m to
// a quadrant, so the only rounding is in r/D and in sincospi (|angle| <=
// pi/2 after reduction).
__device__ __forceinline__ double2 root_fp64(long long m, long long D, int sign) {
  m %= D;
  if (m < 0) m += D;
  const long long m4 = 4 * m;
  const long long q = m4 / D;          // quadrant 0..3
  const long long r = m4 - q * D;      // 0 <= r < D
  double s, c;
  sincospi(0.5 * ((double)r / (double)D), &s, &c);
  double re, im;  // exp(+i*(q*pi/2 + theta))
  switch (q) {
    case 0: re = c; im = s; break;
    case 1: re = -s; im = c; break;
    case 2: re = -c; im = -s; break;
    default: re = s; im = -c; break;
  }
  return double2{re, sign < 0 ? -im : im};
}

}  // namespace dev
}  // namespace fyfft
