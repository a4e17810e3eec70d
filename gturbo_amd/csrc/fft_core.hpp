// fft_core.hpp -- FP64 complex FFT building blocks for gfx950.
//
// A line of N = 2^LOGN complex doubles lives in LDS (padded: element e sits
// at e + e/16, which spreads the stride-16 writes of the first pass over the
// banks).  TPL = N/16 threads own one line.  Each Stockham pass reads 16
// elements per thread at stride N/R, multiplies by twiddles, runs a radix-R
// DFT in registers and writes them back at stride Ns, so the result comes
// out in natural order with no bit reversal.  N = 16^a * r uses a radix-16
// passes and one radix-r pass (r in {2,4,8}).
//
// Twiddles come from a table W[m] = exp(-2 pi i m / N), m < N, built on the
// host in extended precision and rounded once, so every twiddle is the
// correctly rounded value.
#pragma once
#include <hip/hip_runtime.h>

namespace gt {

__device__ __forceinline__ double2 cadd(double2 a, double2 b) {
  return make_double2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ double2 csub(double2 a, double2 b) {
  return make_double2(a.x - b.x, a.y - b.y);
}
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {
  return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
// a * (i*S), S = +-1
template <int S>
__device__ __forceinline__ double2 mul_is(double2 a) {
  return S > 0 ? make_double2(-a.y, a.x) : make_double2(a.y, -a.x);
}

constexpr double kC8 = 0.92387953251128675613;  // cos(pi/8)
constexpr double kS8 = 0.38268343236508977173;  // sin(pi/8)
constexpr double kR2 = 0.70710678118654752440;  // cos(pi/4)

// a * exp(S*2*pi*i*M/16)
template <int S, int M>
__device__ __forceinline__ double2 w16(double2 a) {
  constexpr int m = M & 15;
  if constexpr (m == 0) return a;
  else if constexpr (m == 4) return mul_is<S>(a);
  else if constexpr (m == 8) return make_double2(-a.x, -a.y);
  else if constexpr (m == 12) return mul_is<-S>(a);
  else {
    constexpr double c = (m == 1 || m == 15) ? kC8 : (m == 2 || m == 14) ? kR2
                       : (m == 3 || m == 13) ? kS8 : (m == 5 || m == 11) ? -kS8
                       : (m == 6 || m == 10) ? -kR2 : -kC8;
    constexpr double s0 = (m == 1 || m == 7) ? kS8 : (m == 2 || m == 6) ? kR2
                        : (m == 3 || m == 5) ? kC8 : (m == 9 || m == 15) ? -kS8
                        : (m == 10 || m == 14) ? -kR2 : -kC8;
    constexpr double s = S > 0 ? s0 : -s0;
    return make_double2(a.x * c - a.y * s, a.x * s + a.y * c);
  }
}

template <int S>
__device__ __forceinline__ void dft2(double2& a0, double2& a1) {
  double2 t = csub(a0, a1);
  a0 = cadd(a0, a1);
  a1 = t;
}

template <int S>
__device__ __forceinline__ void dft4(double2& a0, double2& a1, double2& a2,
                                     double2& a3) {
  double2 b0 = cadd(a0, a2), b1 = csub(a0, a2);
  double2 b2 = cadd(a1, a3), b3 = mul_is<S>(csub(a1, a3));
  a0 = cadd(b0, b2);
  a1 = cadd(b1, b3);
  a2 = csub(b0, b2);
  a3 = csub(b1, b3);
}

// Radix-8 as 4 x 2 (n = 2 n1 + n2, k = k1 + 4 k2), natural-order output.
template <int S>
__device__ __forceinline__ void dft8(double2* a) {
  dft4<S>(a[0], a[2], a[4], a[6]);
  dft4<S>(a[1], a[3], a[5], a[7]);
  a[3] = w16<S, 2>(a[3]);
  a[5] = w16<S, 4>(a[5]);
  a[7] = w16<S, 6>(a[7]);
  double2 y[8];
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    y[k1] = cadd(a[2 * k1], a[2 * k1 + 1]);
    y[k1 + 4] = csub(a[2 * k1], a[2 * k1 + 1]);
  }
#pragma unroll
  for (int i = 0; i < 8; ++i) a[i] = y[i];
}

// Radix-16 as 4 x 4 (n = 4 n1 + n2, k = k1 + 4 k2), natural-order output.
template <int S>
__device__ __forceinline__ void dft16(double2* a) {
  dft4<S>(a[0], a[4], a[8], a[12]);
  dft4<S>(a[1], a[5], a[9], a[13]);
  dft4<S>(a[2], a[6], a[10], a[14]);
  dft4<S>(a[3], a[7], a[11], a[15]);
  // a[4*k1 + n2] *= W16^(n2*k1)
  a[5] = w16<S, 1>(a[5]);
  a[9] = w16<S, 2>(a[9]);
  a[13] = w16<S, 3>(a[13]);
  a[6] = w16<S, 2>(a[6]);
  a[10] = w16<S, 4>(a[10]);
  a[14] = w16<S, 6>(a[14]);
  a[7] = w16<S, 3>(a[7]);
  a[11] = w16<S, 6>(a[11]);
  a[15] = w16<S, 9>(a[15]);
  dft4<S>(a[0], a[1], a[2], a[3]);
  dft4<S>(a[4], a[5], a[6], a[7]);
  dft4<S>(a[8], a[9], a[10], a[11]);
  dft4<S>(a[12], a[13], a[14], a[15]);
  // y[k1 + 4 k2] sits at a[4 k1 + k2]: transpose the 4x4 block.
  double2 y[16];
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1)
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) y[k1 + 4 * k2] = a[4 * k1 + k2];
#pragma unroll
  for (int i = 0; i < 16; ++i) a[i] = y[i];
}

template <int R, int S>
__device__ __forceinline__ void dft(double2* a) {
  if constexpr (R == 2) dft2<S>(a[0], a[1]);
  else if constexpr (R == 4) dft4<S>(a[0], a[1], a[2], a[3]);
  else if constexpr (R == 8) dft8<S>(a);
  else dft16<S>(a);
}

__device__ __forceinline__ int lpad(int e) { return e + (e >> 4); }

// Elements of LDS needed by one padded line.
template <int LOGN>
struct Line {
  static constexpr int N = 1 << LOGN;
  static constexpr int TPL = N / 16;          // threads per line
  static constexpr int PADDED = N + N / 16;   // double2 per line in LDS
  static constexpr int NPASS16 = LOGN / 4;    // radix-16 passes
  static constexpr int RLAST = 1 << (LOGN % 4);  // trailing radix (1 = none)
};

// One Stockham pass of radix R with stride Ns over a padded LDS line.
template <int LOGN, int S, int R, int NS>
__device__ __forceinline__ void stockham_pass(double2* buf, int j,
                                              const double2* __restrict__ tw) {
  constexpr int N = 1 << LOGN;
  constexpr int TPL = N / 16;
  constexpr int NB = 16 / R;  // butterflies per thread
  double2 a[16];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int jj = j + TPL * b;
#pragma unroll
    for (int r = 0; r < R; ++r) a[b * R + r] = buf[lpad(jj + r * (N / R))];
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int jj = j + TPL * b;
    const int k = jj % NS;
    if constexpr (NS > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) {
        double2 w = tw[r * k * (N / (NS * R))];
        a[b * R + r] = S < 0 ? cmul(a[b * R + r], w) : cmulc(a[b * R + r], w);
      }
    }
    dft<R, S>(a + b * R);
    const int idx = (jj / NS) * NS * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) buf[lpad(idx + r * NS)] = a[b * R + r];
  }
  __syncthreads();
}

// Full in-LDS transform of one line; S = -1 forward, +1 inverse
// (unnormalised).  Every thread of the workgroup must call it (it holds
// barriers); j is the thread's index within its line.
template <int LOGN, int S>
__device__ __forceinline__ void fft_line(double2* buf, int j,
                                         const double2* __restrict__ tw) {
  using L = Line<LOGN>;
  if constexpr (L::NPASS16 >= 1) stockham_pass<LOGN, S, 16, 1>(buf, j, tw);
  if constexpr (L::NPASS16 >= 2) stockham_pass<LOGN, S, 16, 16>(buf, j, tw);
  if constexpr (L::NPASS16 >= 3) stockham_pass<LOGN, S, 16, 256>(buf, j, tw);
  if constexpr (L::NPASS16 >= 4) stockham_pass<LOGN, S, 16, 4096>(buf, j, tw);
  if constexpr (L::RLAST > 1) {
    constexpr int NS = 1 << (4 * L::NPASS16);
    stockham_pass<LOGN, S, L::RLAST, NS>(buf, j, tw);
  }
}

}  // namespace gt
