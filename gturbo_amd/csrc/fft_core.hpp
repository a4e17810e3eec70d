// fft_core.hpp -- FP64 complex FFT building blocks for gfx950.
//
// A line of N = 2^LOGN complex doubles is spread over TPL = N/8 threads:
// thread j holds a[r] = x[j + TPL*r], r = 0..7, in registers, before and
// after the transform.  The transform is a Stockham sequence of radix-8
// passes (plus one radix-2 or radix-4 pass when log2 N is not a multiple of
// 3).  Every pass reads exactly that per-thread set; the last pass
// (Ns*R = N) writes it back in natural order, so only the transposes between
// passes go through LDS (one padded line per transform: element e sits at
// e + e/16, which spreads the strided writes over the banks).
//
// Eight elements per thread keep a transform at 32 VGPRs of data per lane,
// which lets the passes that carry a second field run at 4 waves per SIMD.
//
// Twiddles come from a table W[m] = exp(-2 pi i m / N), m < N, built on the
// host in extended precision and rounded once.  A butterfly loads only W^k
// and forms W^2k ... W^7k by complex products (at most four extra
// roundings): on gfx950 these passes are bound by memory-instruction
// latency, not by FP64 arithmetic, and one load per butterfly measured
// faster than three (and three faster than seven).
#pragma once
#include <hip/hip_runtime.h>

namespace gt {

constexpr int kE = 8;  // complex elements per thread and line

__device__ __forceinline__ double2 cadd(double2 a, double2 b) {
  return make_double2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ double2 csub(double2 a, double2 b) {
  return make_double2(a.x - b.x, a.y - b.y);
}
// Products with explicit fused multiply-adds.  The kernels are compiled
// without automatic contraction (-ffp-contract=off), so each product rounds
// the same way in every kernel it is inlined into: a member of a batched
// run and the same run alone give bit-identical fields.
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}
// a * conj(b)
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {
  return make_double2(fma(a.x, b.x, a.y * b.y), fma(a.y, b.x, -a.x * b.y));
}
// a * (i*S), S = +-1
template <int S>
__device__ __forceinline__ double2 mul_is(double2 a) {
  return S > 0 ? make_double2(-a.y, a.x) : make_double2(a.y, -a.x);
}

constexpr double kR2 = 0.70710678118654752440;  // cos(pi/4)

// a * exp(S*2*pi*i*M/8)
template <int S, int M>
__device__ __forceinline__ double2 w8(double2 a) {
  constexpr int m = M & 7;
  if constexpr (m == 0) return a;
  else if constexpr (m == 2) return mul_is<S>(a);
  else if constexpr (m == 4) return make_double2(-a.x, -a.y);
  else if constexpr (m == 6) return mul_is<-S>(a);
  else {
    constexpr double c = (m == 1 || m == 7) ? kR2 : -kR2;
    constexpr double s0 = (m == 1 || m == 3) ? kR2 : -kR2;
    constexpr double s = S > 0 ? s0 : -s0;
    return make_double2(fma(a.x, c, -a.y * s), fma(a.x, s, a.y * c));
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
  a[3] = w8<S, 1>(a[3]);
  a[5] = w8<S, 2>(a[5]);
  a[7] = w8<S, 3>(a[7]);
  double2 y[8];
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    y[k1] = cadd(a[2 * k1], a[2 * k1 + 1]);
    y[k1 + 4] = csub(a[2 * k1], a[2 * k1 + 1]);
  }
#pragma unroll
  for (int i = 0; i < 8; ++i) a[i] = y[i];
}

template <int R, int S>
__device__ __forceinline__ void dft(double2* a) {
  if constexpr (R == 2) dft2<S>(a[0], a[1]);
  else if constexpr (R == 4) dft4<S>(a[0], a[1], a[2], a[3]);
  else dft8<S>(a);
}

__device__ __forceinline__ int lpad(int e) { return e + (e >> 4); }
// lpad(e + c) - lpad(e) for a power-of-two-aligned run: valid whenever c is a
// multiple of a power of two P and e mod P + c lies in the same 16-element
// block as c (the transposes below only form such sums), so one base address
// per thread plus compile-time offsets replaces a padded index per access.
__host__ __device__ constexpr int lpad_off(int c) { return c + c / 16; }

template <int LOGN>
struct Line {
  static constexpr int N = 1 << LOGN;
  static constexpr int TPL = N / kE;          // threads per line
  static constexpr int PADDED = N + N / 16;   // double2 per line in LDS
  static constexpr int NPASS8 = LOGN / 3;     // radix-8 passes
  static constexpr int RLAST = 1 << (LOGN % 3);  // trailing radix (1 = none)
};

// Entries of the twiddle table the passes of a line read: W^(k Ns') with
// k Ns' < N/8 for the radix-8 passes and < N/R for a trailing radix-R pass.
template <int LOGN>
constexpr int tw_span() {
  return (1 << LOGN) / (Line<LOGN>::RLAST > 1 ? Line<LOGN>::RLAST : 8);
}
// Up to N = 4096 the kernels keep that part of the table in LDS (at most
// 512 entries, beside their line buffers); larger lines read it from memory.
template <int LOGN>
constexpr bool tw_in_lds() { return LOGN <= 12; }
template <int LOGN>
constexpr int tw_lds_size() { return tw_in_lds<LOGN>() ? tw_span<LOGN>() : 1; }

// Copies the table's first tw_span entries to LDS (every thread of the
// workgroup calls it; ends with a barrier) and returns the table the
// transforms should read.
template <int LOGN>
__device__ __forceinline__ const double2* stage_tw(double2* twl, const double2* __restrict__ tw,
                                                   int tid, int nthr) {
  if constexpr (tw_in_lds<LOGN>()) {
    for (int i = tid; i < tw_span<LOGN>(); i += nthr) twl[i] = tw[i];
    __syncthreads();
    return twl;
  } else {
    return tw;
  }
}

// W^k of a radix-8 pass of stride NS for the butterfly of thread j.
template <int LOGN, int NS>
__device__ __forceinline__ double2 tw_of(const double2* tw, int j) {
  return tw[(j % NS) * ((1 << LOGN) / (NS * 8))];
}

// One Stockham pass: radix R, stride Ns.  Butterfly jj (< N/R) reads
// x[jj + (N/R) r'] and writes y[(jj/Ns) Ns R + jj%Ns + Ns r'].  With PRE the
// pass takes its W^k from wc (loaded by the previous pass); with NSN > 0 it
// loads the next radix-8 pass's W^k into wc before its transpose, so that
// the load is in flight across the barriers.
template <int LOGN, int S, int R, int NS, bool LAST, bool PRE = false, int NSN = 0>
__device__ __forceinline__ void reg_pass(double2* a, double2* buf, int j,
                                         const double2* tw, double2& wc) {
  constexpr int N = 1 << LOGN;
  constexpr int TPL = N / kE;
  constexpr int NB = kE / R;  // butterflies per thread
  static_assert(!PRE || (R == kE && NS > 1), "a carried twiddle needs one butterfly");
  // Opaque copy of the thread index: the addresses below are then computed
  // inside each pass instead of being shared (and kept live) across every
  // transform of the kernel.
  asm volatile("" : "+v"(j));
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int jj = j + TPL * b;
    const int k = jj % NS;
    double2 t[R];
#pragma unroll
    for (int r = 0; r < R; ++r) t[r] = a[b + NB * r];
    if constexpr (NS > 1) {
      // W^(r k), r < R, from the table entries for r = 1, 2, 4
      constexpr int ST = N / (NS * R);
      double2 w[R];
      if constexpr (PRE) w[1] = wc;
      else w[1] = tw[k * ST];
      if constexpr (R >= 4) {
        w[2] = cmul(w[1], w[1]);
        w[3] = cmul(w[1], w[2]);
      }
      if constexpr (R == 8) {
        w[4] = cmul(w[2], w[2]);
        w[5] = cmul(w[1], w[4]);
        w[6] = cmul(w[2], w[4]);
        w[7] = cmul(w[3], w[4]);
      }
#pragma unroll
      for (int r = 1; r < R; ++r) t[r] = S < 0 ? cmul(t[r], w[r]) : cmulc(t[r], w[r]);
    }
    dft<R, S>(t);
    if constexpr (NSN > 0) {
      if (b == NB - 1) wc = tw_of<LOGN, NSN>(tw, j);
    }
    if constexpr (LAST) {
#pragma unroll
      for (int r = 0; r < R; ++r) a[b + NB * r] = t[r];
    } else {
      // idx = (jj/NS) NS R + k with k < NS: idx + r NS never crosses into
      // another 16-element block than r NS does, so lpad splits.
      double2* p = buf + lpad((jj / NS) * NS * R + k);
#pragma unroll
      for (int r = 0; r < R; ++r) p[lpad_off(r * NS)] = t[r];
    }
  }
  if constexpr (!LAST) {
    __syncthreads();
    const double2* p = buf + lpad(j);  // j < TPL, TPL a power of two
#pragma unroll
    for (int r = 0; r < kE; ++r) a[r] = p[lpad_off(TPL * r)];
    __syncthreads();
  }
}

// Transform of one line held in registers; S = -1 forward, +1 inverse
// (unnormalised).  Every thread of the workgroup must call it (the
// transposes hold barriers); j is the thread's index within its line.
// TWPF: each radix-8 pass loads the next one's twiddle before its transpose
// (for a table in memory, whose load latency the barriers then hide; a
// table in LDS gains nothing from it and pays the extra registers).
// SCHED picks the radix order when log2 N is not a multiple of 3: 0 runs
// the radix-8 passes first and the radix-2/4 pass last, 1 the other way
// round (same result up to rounding; the faster one depends on N).
template <int LOGN, int S, int SCHED = 0, bool TWPF = false>
__device__ __forceinline__ void fft_reg(double2* a, double2* buf, int j,
                                        const double2* tw) {
  using L = Line<LOGN>;
  constexpr int NP = L::NPASS8 + (L::RLAST > 1 ? 1 : 0);
  constexpr int P8 = L::NPASS8;
  constexpr bool P = TWPF;
  // next radix-8 stride when this pass hands a twiddle on, else 0
  constexpr auto nx = [](bool has_next, int ns) { return TWPF && has_next ? ns : 0; };
  double2 wc;  // W^k of the next radix-8 pass, carried across a transpose
  if constexpr (SCHED == 0 || L::RLAST == 1) {
    if constexpr (P8 >= 1) reg_pass<LOGN, S, 8, 1, NP == 1, false, nx(P8 >= 2, 8)>(a, buf, j, tw, wc);
    if constexpr (P8 >= 2) reg_pass<LOGN, S, 8, 8, NP == 2, P, nx(P8 >= 3, 64)>(a, buf, j, tw, wc);
    if constexpr (P8 >= 3) reg_pass<LOGN, S, 8, 64, NP == 3, P, nx(P8 >= 4, 512)>(a, buf, j, tw, wc);
    if constexpr (P8 >= 4) reg_pass<LOGN, S, 8, 512, NP == 4, P, 0>(a, buf, j, tw, wc);
    if constexpr (L::RLAST > 1) {
      constexpr int NS = 1 << (3 * P8);
      reg_pass<LOGN, S, L::RLAST, NS, true>(a, buf, j, tw, wc);
    }
  } else {
    constexpr int R0 = L::RLAST;
    reg_pass<LOGN, S, R0, 1, NP == 1>(a, buf, j, tw, wc);
    if constexpr (P8 >= 1) reg_pass<LOGN, S, 8, R0, NP == 2, false, nx(P8 >= 2, R0 * 8)>(a, buf, j, tw, wc);
    if constexpr (P8 >= 2) reg_pass<LOGN, S, 8, R0 * 8, NP == 3, P, nx(P8 >= 3, R0 * 64)>(a, buf, j, tw, wc);
    if constexpr (P8 >= 3) reg_pass<LOGN, S, 8, R0 * 64, NP == 4, P, nx(P8 >= 4, R0 * 512)>(a, buf, j, tw, wc);
    if constexpr (P8 >= 4) reg_pass<LOGN, S, 8, R0 * 512, NP == 5, P, 0>(a, buf, j, tw, wc);
  }
}

// ---------------------------------------------------------------------------
// In-place radix-8 transforms of a 4096-point line over 512 threads (eight
// whole waves, thread j = lane + 64 wave).  The first pass combines the
// elements j + 512 r that thread j holds; its outputs go to LDS, and each
// later pass reads, transforms and writes back one run of the line that
// belongs to a single wave (stride 64: wave j/64's 512 elements; stride 8
// and 1: runs inside them).  Only the first exchange crosses waves, so a
// transform needs one workgroup barrier plus one before its first write
// (when LDS may still be read by another wave), instead of two per pass.
//   fft_dif: natural order in, out of order: thread j, register r holds
//            X[D(j) + 512 r], D(j) = 64 (j mod 8) + 8 (j/8 mod 8) + j/64
//            (the octal digits of j reversed).
//   fft_dit: the transposed algorithm (same twiddles, applied before each
//            radix-8 DFT, passes in reverse): input in that order, natural
//            order out.
// LDS element e sits at e + e/8: every pass then addresses one base per
// thread plus constants, and a wave's accesses at stride 8 or 9 spread over
// the banks.
// ---------------------------------------------------------------------------
constexpr int kIpLine = 4096 + 4096 / 8;  // double2 per line in LDS

__device__ __forceinline__ int ip_pad(int e) { return e + (e >> 3); }

// Orders a wave's LDS writes before its later reads (LDS executes one wave's
// operations in order; this keeps the compiler from moving them).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// t[r] *= W^(k r) (conjugated for S > 0), from w1 = W^k; each power is
// applied as soon as it is formed (at most four live at a time).
template <int S>
__device__ __forceinline__ void apply_tw8(double2* t, double2 w1) {
  auto ap = [](double2& x, double2 w) { x = S < 0 ? cmul(x, w) : cmulc(x, w); };
  ap(t[1], w1);
  const double2 w2 = cmul(w1, w1);
  ap(t[2], w2);
  const double2 w3 = cmul(w1, w2);
  ap(t[3], w3);
  const double2 w4 = cmul(w2, w2);
  ap(t[4], w4);
  ap(t[5], cmul(w1, w4));
  ap(t[6], cmul(w2, w4));
  ap(t[7], cmul(w3, w4));
}

// Twiddles of the three twiddled passes of a 4096-point in-place transform
// in thread j: W^j (stride 512), W^(8 (j % 64)) (stride 64), W^(64 (j % 8))
// (stride 8), from the table read at stride TS.  The same for every
// transform a thread runs, so a kernel running several may load them once.
struct IpTw {
  double2 w1, w2, w3;
};
template <int TS = 1>
__device__ __forceinline__ IpTw ip_tw(const double2* __restrict__ tw, int j) {
  return {tw[TS * j], tw[TS * 8 * (j & 63)], tw[TS * 64 * (j & 7)]};
}
// Twiddle sources for fft_dif / fft_dit: operator()(P) gives the twiddle of
// pass P (1-3), called where the transform first needs it in flight.
template <int TS>
struct TwFromTable {
  const double2* __restrict__ tw;
  int j;
  template <int P>
  __device__ __forceinline__ double2 at() const {
    return P == 1 ? tw[TS * j] : P == 2 ? tw[TS * 8 * (j & 63)] : tw[TS * 64 * (j & 7)];
  }
};
struct TwHeld {
  const IpTw& w;
  template <int P>
  __device__ __forceinline__ double2 at() const {
    return P == 1 ? w.w1 : P == 2 ? w.w2 : w.w3;
  }
};

// bar: barrier before the first LDS write (another wave may still read
// LDS); a uniform run-time flag, so that a loop holds one copy of the code.
template <int S, class TW>
__device__ __forceinline__ void fft_dif_t(double2* a, double2* buf, int j, const TW& T,
                                          bool bar) {
  double2 wn = T.template at<1>();  // pass 1: W^(j r)
  dft<8, S>(a);
  apply_tw8<S>(a, wn);
  wn = T.template at<2>();  // pass 2, in flight across the barrier
  if (bar) __syncthreads();
  {
    double2* p = buf + ip_pad(j);
#pragma unroll
    for (int r = 0; r < 8; ++r) p[576 * r] = a[r];  // ip_pad(j + 512 r)
  }
  __syncthreads();
  {  // pass 2, stride 64: wave j/64's elements
    double2* p = buf + ip_pad((j >> 6) * 512 + (j & 63));
#pragma unroll
    for (int r = 0; r < 8; ++r) a[r] = p[72 * r];
    dft<8, S>(a);
    apply_tw8<S>(a, wn);
    wn = T.template at<3>();
#pragma unroll
    for (int r = 0; r < 8; ++r) p[72 * r] = a[r];
  }
  wave_sync();
  {  // pass 3, stride 8
    double2* p = buf + ip_pad((j >> 3) * 64 + (j & 7));
#pragma unroll
    for (int r = 0; r < 8; ++r) a[r] = p[9 * r];
    dft<8, S>(a);
    apply_tw8<S>(a, wn);
#pragma unroll
    for (int r = 0; r < 8; ++r) p[9 * r] = a[r];
  }
  wave_sync();
  {  // pass 4, stride 1
    const double2* p = buf + 9 * j;  // ip_pad(8 j)
#pragma unroll
    for (int r = 0; r < 8; ++r) a[r] = p[r];
    dft<8, S>(a);
  }
}
template <int S, int TS = 1>
__device__ __forceinline__ void fft_dif(double2* a, double2* buf, int j,
                                        const double2* __restrict__ tw, bool bar) {
  asm volatile("" : "+v"(j));
  fft_dif_t<S>(a, buf, j, TwFromTable<TS>{tw, j}, bar);
}
// Held twiddles enter each transform opaque, so that the compiler keeps
// forming the powers inside it rather than holding them across transforms.
__device__ __forceinline__ IpTw opaque(IpTw w) {
  asm volatile("" : "+v"(w.w1.x), "+v"(w.w1.y), "+v"(w.w2.x), "+v"(w.w2.y), "+v"(w.w3.x),
               "+v"(w.w3.y));
  return w;
}
template <int S>
__device__ __forceinline__ void fft_dif(double2* a, double2* buf, int j, const IpTw& w,
                                        bool bar) {
  asm volatile("" : "+v"(j));
  const IpTw lw = opaque(w);
  fft_dif_t<S>(a, buf, j, TwHeld{lw}, bar);
}

// Input as fft_dif leaves it; the thread's own LDS run is free to write
// (its last reader was this wave).  Ends after reads of every wave's
// elements: a barrier must precede the next LDS write.
// BAR: start with a workgroup barrier (another wave may still read this
// wave's run of LDS, e.g. after a preceding fft_dit).
template <int S, bool BAR, class TW>
__device__ __forceinline__ void fft_dit_t(double2* a, double2* buf, int j, const TW& T) {
  double2 wn = T.template at<3>();
  dft<8, S>(a);  // pass 4, stride 1
  if constexpr (BAR) __syncthreads();
  {
    double2* p = buf + 9 * j;
#pragma unroll
    for (int r = 0; r < 8; ++r) p[r] = a[r];
  }
  wave_sync();
  {  // pass 3, stride 8
    double2* p = buf + ip_pad((j >> 3) * 64 + (j & 7));
#pragma unroll
    for (int r = 0; r < 8; ++r) a[r] = p[9 * r];
    apply_tw8<S>(a, wn);
    wn = T.template at<2>();
    dft<8, S>(a);
#pragma unroll
    for (int r = 0; r < 8; ++r) p[9 * r] = a[r];
  }
  wave_sync();
  {  // pass 2, stride 64
    double2* p = buf + ip_pad((j >> 6) * 512 + (j & 63));
#pragma unroll
    for (int r = 0; r < 8; ++r) a[r] = p[72 * r];
    apply_tw8<S>(a, wn);
    wn = T.template at<1>();
    dft<8, S>(a);
#pragma unroll
    for (int r = 0; r < 8; ++r) p[72 * r] = a[r];
  }
  __syncthreads();
  {  // pass 1, stride 512: natural order out
    const double2* p = buf + ip_pad(j);
#pragma unroll
    for (int r = 0; r < 8; ++r) a[r] = p[576 * r];
    apply_tw8<S>(a, wn);
    dft<8, S>(a);
  }
}
template <int S, bool BAR = false, int TS = 1>
__device__ __forceinline__ void fft_dit(double2* a, double2* buf, int j,
                                        const double2* __restrict__ tw) {
  asm volatile("" : "+v"(j));
  fft_dit_t<S, BAR>(a, buf, j, TwFromTable<TS>{tw, j});
}
template <int S, bool BAR = false>
__device__ __forceinline__ void fft_dit(double2* a, double2* buf, int j, const IpTw& w) {
  asm volatile("" : "+v"(j));
  const IpTw lw = opaque(w);
  fft_dit_t<S, BAR>(a, buf, j, TwHeld{lw});
}

// 8192-point lines over 1024 threads (thread t holds x[t + 1024 r]): a
// radix-2 pass (stride 4096) on the thread's own elements splits the line
// into two 4096-point problems, one exchange hands problem h to threads
// 512h..512h+511 in natural order, and each half runs the in-place
// transform above with the 8192-point table read at stride 2.  fft_dif2
// leaves thread (h, j), register r with X[2 (D(j) + 512 r) + h]; fft_dit2
// takes that order and returns natural order.  buf: two kIpLine regions.
template <int S>
__device__ __forceinline__ void fft_dif2(double2* a, double2* buf, int t,
                                         const double2* __restrict__ tw, bool bar) {
  asm volatile("" : "+v"(t));
#pragma unroll
  for (int r = 0; r < 4; ++r) {  // Y0[n] = x[n] + x[n+4096], Y1[n] = (x[n] - x[n+4096]) W^n
    const double2 d = csub(a[r], a[r + 4]);
    a[r] = cadd(a[r], a[r + 4]);
    const double2 w = tw[t + 1024 * r];
    a[r + 4] = S < 0 ? cmul(d, w) : cmulc(d, w);
  }
  if (bar) __syncthreads();
#pragma unroll
  for (int r = 0; r < 8; ++r)  // Y_(r/4)[t + 1024 (r%4)] into region r/4
    buf[(r >> 2) * kIpLine + ip_pad(t + 1024 * (r & 3))] = a[r];
  __syncthreads();
  const int j = t & 511;
  double2* hb = buf + (t >> 9) * kIpLine;
  {
    const double2* p = hb + ip_pad(j);
#pragma unroll
    for (int r = 0; r < 8; ++r) a[r] = p[576 * r];  // ip_pad(j + 512 r)
  }
  // the first pass of fft_dif writes back exactly the positions just read
  fft_dif<S, 2>(a, hb, j, tw, false);
}

// Ends after reads of every wave's elements: a barrier must precede the
// next LDS write.  BAR: start with a barrier (another wave may still read
// this wave's run, e.g. after a preceding fft_dit2).
template <int S, bool BAR = false>
__device__ __forceinline__ void fft_dit2(double2* a, double2* buf, int t,
                                         const double2* __restrict__ tw) {
  asm volatile("" : "+v"(t));
  const int j = t & 511;
  double2* hb = buf + (t >> 9) * kIpLine;
  fft_dit<S, BAR, 2>(a, hb, j, tw);  // natural order within the half
  {
    double2* p = hb + ip_pad(j);  // the positions fft_dit's last pass read
#pragma unroll
    for (int r = 0; r < 8; ++r) p[576 * r] = a[r];
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 4; ++r) {  // x[n] = Y0[n] + W^n Y1[n], x[n+4096] = Y0[n] - W^n Y1[n]
    const int e = ip_pad(t + 1024 * r);
    const double2 y0 = buf[e];
    const double2 w = tw[t + 1024 * r];
    const double2 y1 = S < 0 ? cmul(buf[kIpLine + e], w) : cmulc(buf[kIpLine + e], w);
    a[r] = cadd(y0, y1);
    a[r + 4] = csub(y0, y1);
  }
}

}  // namespace gt
