// Mode mixing of the spectral convolution: one complex GEMM per kept mode,
// on the matrix cores: v_mfma_f32_16x16x4f32 for complex float,
// v_mfma_f64_16x16x4f64 for complex double (same 16x16x4 block shape and
// operand layout; the result rows are interleaved differently, below).
//
//   C_k = op(A_k) * op(B_k),  k < K,   op = identity or conjugate, any strides
//
// A complex product is four real MFMAs per 16x16x4 block:
//   Cr += Ar*Br - Ai*Bi,  Ci += Ar*Bi + Ai*Br
// (the sign of Ai is flipped in registers for the second one).  A workgroup
// (4 waves) computes a 64x64 tile of one mode's C, each wave a 32x32 quarter
// as 2x2 blocks of 16x16; the L dimension goes through LDS in chunks of 16,
// staged with coalesced loads whichever of the operand's strides is 1 and
// the next chunk's loads in flight while the current one is multiplied.
// Lane layout of the 16x16x4 instructions (probed on the device): lane l
// holds A[l%16][l/16] and B[l/16][l%16]; result element i of lane l is
// C[4*(l/16) + i][l%16] for the FP32 instruction and C[l/16 + 4*i][l%16]
// for the FP64 one.
#include "bodies.hpp"

namespace fyfft {
using namespace dev;

namespace {

constexpr int kTile = 64, kChunk = 16;
constexpr int kAStride = kChunk + 1;  // A tile rows padded (16 lanes read 16 rows)
constexpr int kBStride = kTile + 16;  // B tile rows padded by 32 banks

typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v4d __attribute__((ext_vector_type(4)));

template <typename T> struct Acc;
template <> struct Acc<float> { using type = v4f; };
template <> struct Acc<double> { using type = v4d; };

__device__ __forceinline__ v4f mfma(float a, float b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ v4d mfma(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

template <typename V> struct Op {
  const V* p;
  long long sk, s0, s1;  // element (i0, i1) of batch k at p[k*sk + i0*s0 + i1*s1]
  int n0, n1;            // extents
  bool conj;
};

// Tile element e (0..1023) of a rows x cols = R x Cc tile whose storage is
// fast along dimension `fast1` (true: the second index): the e -> (i0, i1)
// mapping puts consecutive threads on consecutive addresses.
template <int R, int Cc>
__device__ __forceinline__ void tile_coord(int e, bool fast1, int* i0, int* i1) {
  if (fast1) {
    *i0 = e / Cc;
    *i1 = e % Cc;
  } else {
    *i0 = e % R;
    *i1 = e / R;
  }
}

}  // namespace

// A: (m, l), B: (l, n), C: (m, n).  grid: x = tiles of M x N, y = K.
template <typename V>
__global__ __launch_bounds__(256) void k_cgemm_modes(Op<V> A, Op<V> B, V* C, long long csk, long long csm,
                                                     long long csn, int M, int N, int L) {
  using T = S<V>;
  using VA = typename Acc<T>::type;
  __shared__ V As[kTile * kAStride];   // [m][l]
  __shared__ V Bs[kChunk * kBStride];  // [l][n]
  const int tiles_n = (N + kTile - 1) / kTile;
  const int m0 = (int)(blockIdx.x / tiles_n) * kTile, n0 = (int)(blockIdx.x % tiles_n) * kTile;
  const long long k = blockIdx.y;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  const bool a_fast_l = A.s1 == 1, b_fast_n = B.s1 == 1;
  const V* Ap = A.p + k * A.sk;
  const V* Bp = B.p + k * B.sk;

  // one chunk of both tiles: 4 + 4 elements per thread (zero outside)
  V ra[4], rb[4];
  auto load = [&](int l0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      int i, j;
      tile_coord<kTile, kChunk>(tid + q * 256, a_fast_l, &i, &j);  // (m, l)
      const int m = m0 + i, l = l0 + j;
      ra[q] = (m < M && l < L) ? Ap[m * A.s0 + l * A.s1] : V{0, 0};
      tile_coord<kChunk, kTile>(tid + q * 256, b_fast_n, &i, &j);  // (l, n)
      const int lb = l0 + i, n = n0 + j;
      rb[q] = (lb < L && n < N) ? Bp[lb * B.s0 + n * B.s1] : V{0, 0};
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      int i, j;
      tile_coord<kTile, kChunk>(tid + q * 256, a_fast_l, &i, &j);
      As[i * kAStride + j] = ra[q];
      tile_coord<kChunk, kTile>(tid + q * 256, b_fast_n, &i, &j);
      Bs[i * kBStride + j] = rb[q];
    }
  };

  VA cr[2][2], ci[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) cr[i][j] = ci[i][j] = VA{0, 0, 0, 0};
  const T sa = A.conj ? -1 : 1, sb = B.conj ? -1 : 1;

  load(0);
  for (int l0 = 0; l0 < L; l0 += kChunk) {
    __syncthreads();  // every read of the previous chunk is done
    store();
    __syncthreads();
    if (l0 + kChunk < L) load(l0 + kChunk);  // in flight during the products below
#pragma unroll
    for (int s = 0; s < kChunk / 4; ++s) {
      const int kk = s * 4 + (lane >> 4), r = lane & 15;
      V a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        a[i] = As[(wm + i * 16 + r) * kAStride + kk];
        a[i].y *= sa;
      }
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        b[j] = Bs[kk * kBStride + wn + j * 16 + r];
        b[j].y *= sb;
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          cr[i][j] = mfma(a[i].x, b[j].x, cr[i][j]);
          cr[i][j] = mfma(-a[i].y, b[j].y, cr[i][j]);
          ci[i][j] = mfma(a[i].x, b[j].y, ci[i][j]);
          ci[i][j] = mfma(a[i].y, b[j].x, ci[i][j]);
        }
    }
  }
  V* Cp = C + k * csk;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int row = sizeof(T) == 8 ? (lane >> 4) + 4 * q : 4 * (lane >> 4) + q;
        const int m = m0 + wm + i * 16 + row, n = n0 + wn + j * 16 + (lane & 15);
        if (m < M && n < N) Cp[m * csm + n * csn] = V{cr[i][j][q], ci[i][j][q]};
      }
}

// Mode-major operands: modes [K][rows][C], weights [K][cin][cout].
template <typename V>
static hipError_t go_mix(int kind, const void* p, const void* q, void* out, long long K, long long rows,
                         long long cin, long long cout, hipStream_t s) {
  Op<V> A, B;
  long long csk, csm, csn;
  int M, N, L;
  const V* P = static_cast<const V*>(p);
  const V* Q = static_cast<const V*>(q);
  if (kind == kMixForward) {  // out[k][r][o] = sum_i p[k][r][i] q[k][i][o]
    A = Op<V>{P, rows * cin, cin, 1, (int)rows, (int)cin, false};
    B = Op<V>{Q, cin * cout, cout, 1, (int)cin, (int)cout, false};
    M = (int)rows, N = (int)cout, L = (int)cin;
    csk = rows * cout, csm = cout, csn = 1;
  } else if (kind == kMixAdjointX) {  // out[k][r][i] = sum_o p[k][r][o] conj(q[k][i][o])
    A = Op<V>{P, rows * cout, cout, 1, (int)rows, (int)cout, false};
    B = Op<V>{Q, cin * cout, 1, cout, (int)cout, (int)cin, true};
    M = (int)rows, N = (int)cin, L = (int)cout;
    csk = rows * cin, csm = cin, csn = 1;
  } else if (kind == kMixGradW) {  // out[k][i][o] = sum_r conj(p[k][r][i]) q[k][r][o]
    A = Op<V>{P, rows * cin, 1, cin, (int)cin, (int)rows, true};
    B = Op<V>{Q, rows * cout, cout, 1, (int)rows, (int)cout, false};
    M = (int)cin, N = (int)cout, L = (int)rows;
    csk = cin * cout, csm = cout, csn = 1;
  } else {
    return hipErrorInvalidValue;
  }
  if (K == 0 || M == 0 || N == 0) return hipSuccess;
  const long long tiles = (long long)((M + kTile - 1) / kTile) * ((N + kTile - 1) / kTile);
  if (tiles > 0x7fffffffLL || K > 65535) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_cgemm_modes<V>, dim3((unsigned)tiles, (unsigned)K), dim3(256), 0, s, A, B, (V*)out, csk,
                     csm, csn, M, N, L);
  return hipGetLastError();
}

hipError_t launch_mix_modes(int prec, int kind, const void* p, const void* q, void* out, long long K,
                            long long rows, long long cin, long long cout, hipStream_t s) {
  if (prec == kF64) return go_mix<double2>(kind, p, q, out, K, rows, cin, cout, s);
  if (prec == kF32) return go_mix<float2>(kind, p, q, out, K, rows, cin, cout, s);
  return hipErrorInvalidValue;
}

}  // namespace fyfft
