// gturbo_kernels.hip -- gfx950 kernels of the pseudospectral step.
//
// Nonlinear term N(q) = -ik.(q u)^ (src/nlt.f90) on the compact layout, as two
// kinds of pass per evaluation:
//
//  * y pass (c_pass): one workgroup per tile of BT kx-columns.  Forward part:
//    forward y transform of the two product fields, -ik. combine, 2/3
//    truncation and the RK stage update (src/rkstep.f90 stage AXPYs).  Inverse
//    part: from the next stage input, the spectral operator |k|^-alpha
//    (src/nlt.f90:22-36) for q, u, v and their inverse y transforms.
//  * x pass (r_pass): one workgroup per RP y rows.  Complex-to-real x
//    transforms of q+iv and q+iu (two real sequences per complex transform),
//    the products q u, q v (src/produ.f90), and one real-to-complex transform
//    of q u + i q v.  At N = 8192 a workgroup takes one row, as the 4096-point
//    complex transform of its even and odd samples (r_pass_row).
//
// Between the passes the three (resp. two) fields live in HBM in blocks of
// kYR rows x BT columns, [field][y / kYR][tile][y % kYR][kx % BT]: the y
// pass writes whole blocks (a 128-byte line at N = 4096), the x pass reads a
// row's pieces of consecutive tiles from one DRAM page, and its workgroup
// order keeps neighbouring rows on one XCD so that they share L2 lines.
#include <mutex>

#include "fft_core.hpp"
#include "gt_internal.hpp"

namespace gt {

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------

__device__ __forceinline__ bool cut_mode(const Phys& ph, double fkxs,
                                         double fkys) {
  // src/trunc.f90:13-15
  double k1 = fkys / ph.fkysmax;
  double k2 = k1 + fkxs / ph.fkxsmax;
  return k2 > ph.rtrunc2;
}

// ka = k2**(alpv/2) (src/nlt.f90:27): k2 itself for alpv = 2, correctly
// rounded square roots for alpv = 1 and 1/2, the power otherwise.
__device__ __forceinline__ double kalpha(const Phys& ph, double k2) {
  switch (ph.akind) {
    case 1: return k2;
    case 2: return sqrt(k2);
    case 3: return sqrt(sqrt(k2));
    default: return pow(k2, ph.ahalf);
  }
}

// ka in a y pass specialised on the alpha kind: AK = 1 when alpv = 2 (ka =
// k2, nothing else in the kernel), AK = 0 a per-thread table of ka built
// once per column, AK = 2 ka where it is used (the batched passes, whose
// members differ in alpha): k2 for alpv = 2, else the member's table of ka
// in the compact layout, built on the host with the same expressions as
// kalpha (no power or root code in those kernels).
template <int AK>
__device__ __forceinline__ double kalpha_t(const PassArgs& a, double k2, unsigned q) {
  if constexpr (AK == 1) return k2;
  else if constexpr (AK == 0) return kalpha(a.ph, k2);
  else return a.ph.akind == 1 ? k2 : a.ka[q];
}

// With the 2/3 rule the kept rectangle stays below |k| = 3N/8, so of a
// thread's positions p = j + r*N/8 the ones with r = 3, 4 are never kept:
// those inputs are zero and those outputs unused.  TR = 1 makes that a
// compile-time fact.
template <int TR>
__device__ __forceinline__ constexpr bool live_r(int r) {
  return TR == 0 || (r != 3 && r != 4);
}

// Cache use of the y passes at N = 4096 and 8192, where a pass moves more
// than the 256 MB Infinity Cache holds.  The state z, the RK4 accumulator
// znew and the integrating factor emk are read and written once per y pass
// and next used a whole x pass later, by when they have left the caches;
// the products Y2 are read once.  The y passes access them without
// allocating cache lines (non-temporal), which leaves the caches to the
// fields Y1 they write and the next x pass reads at once.  Measured against
// the same kernels with ordinary accesses, one process on one box, results
// bit-identical: at 4096^2 the state accesses alone y stages -8%, RK4 step
// -2.4%; the Y2 reads on top of that x pass -4%, step -1.3%; both together
// RK4 -3.5%, RK2 -2%.  At 8192^2 the state accesses RK4 step -1.7%, but
// the Y2 reads on top +2.8% (there a one-column y workgroup reads half of
// each Y2 line and the next column's workgroup the other half), so there
// only the state bypasses the caches.  Non-temporal loads in the x pass
// slow it by 9-22%.
typedef double nt_d2 __attribute__((ext_vector_type(2)));
template <bool NT>
__device__ __forceinline__ double2 ld_state(const double2* p) {
  if constexpr (NT) {
    const nt_d2 v = __builtin_nontemporal_load(reinterpret_cast<const nt_d2*>(p));
    return make_double2(v.x, v.y);
  } else {
    return *p;
  }
}
template <bool NT>
__device__ __forceinline__ double ld_state(const double* p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}
template <bool NT>
__device__ __forceinline__ void st_state(double2* p, double2 v) {
  if constexpr (NT) {
    nt_d2 w;
    w.x = v.x;
    w.y = v.y;
    __builtin_nontemporal_store(w, reinterpret_cast<nt_d2*>(p));
  } else {
    *p = v;
  }
}

template <int LOGN, int BT>
__device__ __forceinline__ void load_tile(double2* lds,
                                          const double2* __restrict__ src,
                                          int tid) {
  constexpr int N = 1 << LOGN, NTHR = BT * (N / 16), LINE = Line<LOGN>::PADDED;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    int e = tid + NTHR * i;
    lds[(e % BT) * LINE + lpad(e / BT)] = src[e];
  }
}

template <int LOGN, int BT>
__device__ __forceinline__ void store_tile(double2* __restrict__ dst,
                                           const double2* lds, int tid) {
  constexpr int N = 1 << LOGN, NTHR = BT * (N / 16), LINE = Line<LOGN>::PADDED;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    int e = tid + NTHR * i;
    dst[e] = lds[(e % BT) * LINE + lpad(e / BT)];
  }
}

// Workgroup -> tile remap so that consecutive tiles run on one XCD
// (workgroups are dealt round-robin over the 8 XCDs, each with its own L2):
// XCD x takes a contiguous run of tiles, the first nblk % 8 XCDs one more.
__device__ __forceinline__ int xcd_remap(int bid, int nblk) {
  const int q = nblk / 8, rem = nblk % 8, x = bid % 8, i = bid / 8;
  return (x < rem ? x * (q + 1) : rem * (q + 1) + (x - rem) * q) + i;
}

// Thread -> (line, index in line) for NL lines per workgroup.  Lines
// alternate every 16 lanes: one 16-lane LDS access group then touches a
// single line (contiguous indices: no bank conflicts between lines), while a
// wave's global access still covers all NL lines of a tile row or run.
template <int TPL, int NL>
__device__ __forceinline__ void line_map(int tid, int& line, int& j) {
  constexpr int G = TPL < 16 ? TPL : 16;
  const int c = tid / G;
  line = c % NL;
  j = (c / NL) * G + tid % G;
}

// ---------------------------------------------------------------------------
// y pass
// ---------------------------------------------------------------------------

// The forcing of column c (fused step boundary) on the thread's values
// s[r] of rows p = j + (N/8) r: the same operations as forcing_k's
// force_one on each forced element, in the order of the forced-mode table.
template <int LOGN, int TR>
__device__ __forceinline__ void force_column(const PassArgs& a, int c, bool colok, int j,
                                             double2* s) {
#pragma clang fp contract(off)
  constexpr int N = 1 << LOGN, TPL = N / kE;
  if (a.fcs == nullptr || !colok) return;
  const double2* ph = a.fph;
  double amp = a.famp;
  if (a.fens != nullptr) {  // a batched ensemble member: this step's phases
    ph = a.fens->phases + (size_t)(*a.fstep) * a.fens->nf;
    amp = a.fens->amp;
  }
  const int e1 = a.fcs[c + 1];
  for (int e = a.fcs[c]; e < e1; ++e) {
    const FEntry f = a.fce[e];
    const double2 cs = ph[f.k];
    const double a1 = amp * f.ff1 * cs.x, a2 = amp * f.ff2 * cs.y;
#pragma unroll
    for (int r = 0; r < kE; ++r) {
      if (live_r<TR>(r) && row_of(j + TPL * r, a.rc, N) == f.row) {
        s[r].x = s[r].x + a1;
        s[r].y = f.conj ? s[r].y - a2 : s[r].y + a2;
      }
    }
  }
}

// One tile of BT kx-columns; lds holds BT padded lines.  SCHED is the radix
// order of the transforms (fft_reg).
// IP (4096-point lines only): the y buffers hold the rows of a column in
// the order the in-place transforms use (slot j + 512 r holds physical row
// D(j) + 512 r, D the octal digit reversal of j); the forward transforms
// (fft_dit) take the loaded slots and return natural ky order, the inverse
// ones (fft_dif) return slot order for the stores.  Each line is eight
// whole waves: line b = wave mod 2, j = 64 (wave / 2) + lane.  The x pass
// transforms rows, whichever rows a slot pair holds.
template <int LOGN, int MODE, int AK, int TR, int BT, int SCHED, bool IP = false>
__device__ __forceinline__ void c_body(const PassArgs& a, int tile, int tid,
                                       double2* lds, const double2* tw) {
  constexpr int N = 1 << LOGN, TPL = N / kE;
  constexpr int LINE = IP ? (LOGN == 13 ? 2 : 1) * kIpLine : Line<LOGN>::PADDED;
  static_assert(!IP || (LOGN == 12 && BT == 2) || (LOGN == 13 && BT == 1),
                "in-place y pass: two 4096-point lines or one 8192-point line");
  constexpr bool FWD = MODE != C_START;
  constexpr bool INV = MODE == C_START || MODE == C_RK4_1 || MODE == C_RK4_2 ||
                       MODE == C_RK4_3 || MODE == C_RK2_1 || MODE == C_RK4_4S ||
                       MODE == C_RK2_2S;
  constexpr bool NT = LOGN >= 12;    // state accesses that bypass the caches
  constexpr bool NT2 = LOGN == 12;   // Y2 reads likewise, at 4096 only
  int b, j;
  if constexpr (IP && BT == 2) {
    b = (tid >> 6) & 1;
    j = ((tid >> 7) << 6) | (tid & 63);
  } else if constexpr (IP) {
    b = 0;
    j = tid;
  } else {
    line_map<TPL, BT>(tid, b, j);
  }
  const int c = tile * BT + b;
  const Rect rc = a.rc;
  const bool colok = c <= rc.kx;
  double2* buf = lds + b * LINE;
  const Phys& ph = a.ph;
  const double fkx = (double)c * ph.dkx;
  const double fkxs = fkx * fkx;
  const unsigned cbase = (unsigned)c * (unsigned)a.kyp;
  double2 s[kE];
  double2 z[kE];
  // 4096-point in-place transforms: their three twiddles, loaded once.
  IpTw ytw{};
  if constexpr (IP && LOGN == 12) ytw = ip_tw(tw, j);

  if constexpr (FWD) {
    double2 A[kE];
    const unsigned colo = ycol_off(c, BT);
    const double2* srcA = a.y2;
#pragma unroll
    for (int r = 0; r < kE; ++r) z[r] = ld_state<NT2>(&srcA[yrow_off(j + TPL * r, a.ntiles, BT) + colo]);
    if constexpr (IP && LOGN == 12) fft_dit<-1, false>(z, buf, j, ytw);
    else if constexpr (IP) fft_dit2<-1, false>(z, buf, j, tw);
    else fft_reg<LOGN, -1, SCHED>(z, buf, j, tw);
#pragma unroll
    for (int r = 0; r < kE; ++r) A[r] = live_r<TR>(r) ? z[r] : make_double2(0.0, 0.0);
    const double2* srcB = a.y2 + (size_t)a.ntiles * N * BT;
#pragma unroll
    for (int r = 0; r < kE; ++r) z[r] = ld_state<NT2>(&srcB[yrow_off(j + TPL * r, a.ntiles, BT) + colo]);
    if constexpr (IP && LOGN == 12) fft_dit<-1, true>(z, buf, j, ytw);
    else if constexpr (IP) fft_dit2<-1, true>(z, buf, j, tw);
    else fft_reg<LOGN, -1, SCHED>(z, buf, j, tw);
    int je = j;
    // The epilogue's indices (and so its loads) depend on the transform's
    // output: nothing of the epilogue is scheduled inside the last pass.
    asm volatile("" : "+v"(je) : "v"(z[0].x), "v"(z[kE - 1].y));
#pragma unroll
    for (int r = 0; r < kE; ++r) {
      const int p = je + TPL * r;
      const int row = row_of(p, rc, N);
      s[r] = make_double2(0.0, 0.0);
      if (live_r<TR>(r) && colok && row >= 0) {
        const double fky = (double)kyi_of(p, N) * ph.dky;
        const double fkys = fky * fky;
        // fft_dir normalisation /NY/NX (src/fft_dir.f90:19), exact for N = 2^k
        const double2 Av = make_double2(A[r].x * ph.inv_n2, A[r].y * ph.inv_n2);
        const double2 Bv = make_double2(z[r].x * ph.inv_n2, z[r].y * ph.inv_n2);
        // src/nlt.f90:57-58
        double2 k = make_double2(fkx * Av.y + fky * Bv.y, -fkx * Av.x - fky * Bv.x);
        if (ph.truncate == 1 && cut_mode(ph, fkxs, fkys)) k = make_double2(0.0, 0.0);
        const unsigned q = cbase + (unsigned)row;
        const bool mean = (c == 0 && p == 0);
        if constexpr (MODE == C_NLT) {
          if (mean) k = make_double2(0.0, 0.0);  // trunc.f90:24
          a.out[q] = k;
        } else if constexpr (MODE == C_RK4_1) {
          // rkstep.f90:96,101
          const double E = ld_state<NT>(&a.emk[q]);
          const double2 zz = ld_state<NT>(&a.z[q]);
          const double EE = E * E;
          st_state<NT>(&a.znew[q], make_double2(EE * (zz.x + ph.dt6 * k.x), EE * (zz.y + ph.dt6 * k.y)));
          s[r] = make_double2(E * (zz.x + ph.dt2 * k.x), E * (zz.y + ph.dt2 * k.y));
        } else if constexpr (MODE == C_RK4_2) {
          // rkstep.f90:105,110
          const double E = ld_state<NT>(&a.emk[q]);
          const double2 zz = ld_state<NT>(&a.z[q]);
          const double2 zn = ld_state<NT>(&a.znew[q]);
          const double d3E = ph.dt3 * E;
          st_state<NT>(&a.znew[q], make_double2(zn.x + d3E * k.x, zn.y + d3E * k.y));
          s[r] = make_double2(E * zz.x + ph.dt2 * k.x, E * zz.y + ph.dt2 * k.y);
        } else if constexpr (MODE == C_RK4_3) {
          // rkstep.f90:114,119
          const double E = ld_state<NT>(&a.emk[q]);
          const double2 zz = ld_state<NT>(&a.z[q]);
          const double2 zn = ld_state<NT>(&a.znew[q]);
          const double d3E = ph.dt3 * E;
          st_state<NT>(&a.znew[q], make_double2(zn.x + d3E * k.x, zn.y + d3E * k.y));
          s[r] = make_double2(E * (E * zz.x + ph.dt * k.x), E * (E * zz.y + ph.dt * k.y));
        } else if constexpr (MODE == C_RK4_4) {
          // rkstep.f90:123,127
          const double2 zn = ld_state<NT>(&a.znew[q]);
          double2 zz = make_double2(zn.x + ph.dt6 * k.x, zn.y + ph.dt6 * k.y);
          if (mean) zz = make_double2(0.0, 0.0);
          st_state<NT>(&a.z[q], zz);
        } else if constexpr (MODE == C_RK4_4S) {
          // rkstep.f90:123,127; forced and stored below
          const double2 zn = ld_state<NT>(&a.znew[q]);
          double2 zz = make_double2(zn.x + ph.dt6 * k.x, zn.y + ph.dt6 * k.y);
          if (mean) zz = make_double2(0.0, 0.0);
          s[r] = zz;
        } else if constexpr (MODE == C_RK2_2S) {
          // rkstep.f90:54/64 and :72; forced and stored below
          const double E = ld_state<NT>(&a.emk[q]);
          const double2 zz = ld_state<NT>(&a.z[q]);
          double2 zo = make_double2(E * (E * zz.x + ph.dt * k.x), E * (E * zz.y + ph.dt * k.y));
          if (mean) zo = make_double2(0.0, 0.0);
          s[r] = zo;
        } else if constexpr (MODE == C_RK2_1) {
          // rkstep.f90:24/34 and :42
          const double E = ld_state<NT>(&a.emk[q]);
          const double2 zz = ld_state<NT>(&a.z[q]);
          s[r] = make_double2(E * (zz.x + ph.dt2 * k.x), E * (zz.y + ph.dt2 * k.y));
          if (mean) s[r] = make_double2(0.0, 0.0);
        } else if constexpr (MODE == C_RK2_2) {
          // rkstep.f90:54/64 and :72
          const double E = ld_state<NT>(&a.emk[q]);
          const double2 zz = ld_state<NT>(&a.z[q]);
          double2 zo = make_double2(E * (E * zz.x + ph.dt * k.x), E * (E * zz.y + ph.dt * k.y));
          if (mean) zo = make_double2(0.0, 0.0);
          st_state<NT>(&a.z[q], zo);
        }
      }
    }
    if constexpr (MODE == C_RK4_4S || MODE == C_RK2_2S) {
      // the step's forcing on this column (forcing.f90:16-33), then the
      // forced state stored: it is also the next step's stage input
      force_column<LOGN, TR>(a, c, colok, je, s);
#pragma unroll
      for (int r = 0; r < kE; ++r) {
        const int row = row_of(je + TPL * r, rc, N);
        if (live_r<TR>(r) && colok && row >= 0) st_state<NT>(&a.z[cbase + (unsigned)row], s[r]);
      }
    }
  } else {
#pragma unroll
    for (int r = 0; r < kE; ++r) {
      const int p = j + TPL * r;
      const int row = row_of(p, rc, N);
      s[r] = (live_r<TR>(r) && colok && row >= 0) ? ld_state<NT>(&a.z[cbase + (unsigned)row])
                                                   : make_double2(0.0, 0.0);
    }
  }

  if constexpr (INV) {
    if (a.sst != nullptr) {  // stage input kept for the adjoint's recompute
#pragma unroll
      for (int r = 0; r < kE; ++r) {
        const int row = row_of(j + TPL * r, rc, N);
        if (live_r<TR>(r) && colok && row >= 0) a.sst[cbase + (unsigned)row] = s[r];
      }
    }
    // ka is tabled only for AK = 0; for AK = 1 it is k2 and for AK = 2 it is
    // evaluated where needed.
    double ka[AK == 0 ? kE : 1];
    if constexpr (AK == 0) {
#pragma unroll
      for (int r = 0; r < kE; ++r) {
        const int p = j + TPL * r;
        const double fky = (double)kyi_of(p, N) * ph.dky;
        const bool use = live_r<TR>(r) && colok && row_of(p, rc, N) >= 0 && !(c == 0 && p == 0);
        ka[r] = use ? kalpha(ph, fkxs + fky * fky) : 1.0;
      }
    }
#pragma unroll 1
    for (int f = 0; f < 3; ++f) {
      int jo = j;
      asm volatile("" : "+v"(jo));  // keep the per-row factors inside the loop
#pragma unroll
      for (int r = 0; r < kE; ++r) {
        const int p = jo + TPL * r;
        double2 v = make_double2(0.0, 0.0);
        if (live_r<TR>(r) && colok && row_of(p, rc, N) >= 0) {
          if (f == 0) {
            v = s[r];
          } else if (!(c == 0 && p == 0)) {  // nlt.f90:25
            const double fky = (double)kyi_of(p, N) * ph.dky;
            const double kar =
                AK == 0 ? ka[AK == 0 ? r : 0]
                        : kalpha_t<AK>(a, fkxs + fky * fky, cbase + (unsigned)row_of(p, rc, N));
            if (f == 1) {
              const double kyn = fky / kar;
              v = make_double2(-kyn * s[r].y, kyn * s[r].x);  // nlt.f90:31-32
            } else {
              const double kxn = fkx / kar;
              v = make_double2(kxn * s[r].y, -kxn * s[r].x);  // nlt.f90:33-34
            }
          }
        }
        z[r] = v;
      }
      if constexpr (IP) {
        // a barrier first unless nothing has used LDS since the table staging
        if constexpr (LOGN == 12) fft_dif<+1>(z, buf, j, ytw, FWD || f > 0);
        else fft_dif2<+1>(z, buf, j, tw, FWD || f > 0);
      } else {
        fft_reg<LOGN, +1, SCHED>(z, buf, j, tw);
      }
      if constexpr (LOGN == 12) {  // row-pair blocks (y1_off)
        const int cols = y1_cols(rc.kc());
        double2* dst = a.y1 + (size_t)f * N * cols;
#pragma unroll
        for (int r = 0; r < kE; ++r) dst[y1_off(j + TPL * r, c, cols)] = z[r];
      } else {
        double2* dst = a.y1 + (size_t)f * a.ntiles * N * BT;
        const unsigned colo = ycol_off(c, BT);
#pragma unroll
        for (int r = 0; r < kE; ++r) dst[yrow_off(j + TPL * r, a.ntiles, BT) + colo] = z[r];
      }
    }
  }
}

template <int LOGN, int MODE, int AK, int TR>
__global__ void __launch_bounds__(bt_of(LOGN) * (1 << LOGN) / kE, AK == 1 ? 4 : 1)
    c_pass(PassArgs a) {
  constexpr int BT = bt_of(LOGN);
  // Row slots in transform order (in-place transforms) for alpha = 2 on the
  // 2/3 rectangle; every y pass of a run has the same AK and TR, so the
  // buffers' row order is the same in all of them.
  constexpr bool IP = (LOGN == 12 || LOGN == 13) && AK == 1 && TR == 1;
  // 4096-point in-place transforms hold their three twiddles in registers
  // (read once from memory): no table in LDS.
  constexpr bool HELD = IP && LOGN == 12;
  __shared__ double2 lds[IP ? 2 * kIpLine : BT * Line<LOGN>::PADDED];
  __shared__ double2 twl[HELD ? 1 : tw_lds_size<LOGN>()];
  const double2* tw = HELD ? a.tw : stage_tw<LOGN>(twl, a.tw, threadIdx.x, blockDim.x);
  // 8192: a one-column workgroup shares each line of Y1 and Y2 with the next
  // column's, so neighbouring columns run on one XCD (y stages -5.6%).
  const int tile = LOGN == 13 ? xcd_remap(blockIdx.x, gridDim.x) : (int)blockIdx.x;
  c_body<LOGN, MODE, AK, TR, BT, 0, IP>(a, tile, threadIdx.x, lds, tw);
}

// ---------------------------------------------------------------------------
// x pass
// ---------------------------------------------------------------------------

// Offset of column k of a y-row in the blocked layout.
template <int LOGN, int BT>
__device__ __forceinline__ unsigned koff(int k) {
  return ycol_off(k, BT);
}

// Z = F + i G over the full x-spectrum at the thread's positions
// p = j + TPL*r, from the half-spectra F, G of two real sequences (kx <=
// kx_max; the imaginary parts at kx = 0 and N/2 are not used, as in a C2R
// transform).  The inverse transform of Z is f + i g.  Built in two
// halves: the loads (zeros outside the kept half-spectrum), all issued
// before any is used so that they are in flight together, and the packing.
// Y1: the input is in row-pair blocks (y1_off) rather than tiles.
template <int LOGN, int TR, int BT, bool Y1 = false>
__device__ __forceinline__ void load_two(double2* x0, double2* x1, int j,
                                         const double2* __restrict__ y,
                                         unsigned off_f, unsigned off_g, int kxm) {
  constexpr int N = 1 << LOGN, TPL = N / kE;
#pragma unroll
  for (int r = 0; r < kE; ++r) {
    const int p = j + TPL * r;
    const int k = p <= N / 2 ? p : N - p;
    const unsigned ko = Y1 ? y1_off(0, k, 0) : koff<LOGN, BT>(k);
    x0[r] = x1[r] = make_double2(0.0, 0.0);
    if (live_r<TR>(r) && k <= kxm) {
      x0[r] = y[off_f + ko];
      x1[r] = y[off_g + ko];
    }
  }
}
template <int LOGN>
__device__ __forceinline__ void pack_two(double2* z, int j, const double2* x0,
                                         const double2* x1) {
  constexpr int N = 1 << LOGN, TPL = N / kE;
#pragma unroll
  for (int r = 0; r < kE; ++r) {
    const int p = j + TPL * r;
    if (p == 0 || p == N / 2) z[r] = make_double2(x0[r].x, x1[r].x);
    else if (p < N / 2) z[r] = make_double2(x0[r].x - x1[r].y, x0[r].y + x1[r].x);
    else z[r] = make_double2(x0[r].x + x1[r].y, x1[r].x - x0[r].y);
  }
}
template <int LOGN, int TR, int BT, bool Y1 = false>
__device__ __forceinline__ void build_two(double2* z, int j,
                                          const double2* __restrict__ y,
                                          unsigned off_f, unsigned off_g, int kxm) {
  double2 x0[kE], x1[kE];
  load_two<LOGN, TR, BT, Y1>(x0, x1, j, y, off_f, off_g, kxm);
  pack_two<LOGN>(z, j, x0, x1);
}
// Element (N - k) mod N, k = j + TPL r (r a constant), of a line in LDS;
// pm = buf + lpad(TPL - j).  For TPL >= 16 the index TPL (7 - r) + TPL - j
// splits into that base and a constant offset; k = 0 maps to element 0.
template <int LOGN>
__device__ __forceinline__ double2 mirror_at(const double2* buf, const double2* pm, int j,
                                             int r) {
  constexpr int N = 1 << LOGN, TPL = N / kE;
  if constexpr (TPL % 16 != 0) {
    return buf[lpad((N - j - TPL * r) & (N - 1))];
  } else {
    if (r == 0) return (j == 0 ? buf : pm + lpad_off(7 * TPL))[0];
    return pm[lpad_off(TPL * (7 - r))];
  }
}

// Split W = FFT(a + i b) (held at k = j + TPL*r) into the half-spectra of
// the real sequences a and b and store kx <= kx_max.  W[N-k] comes through
// LDS.
template <int LOGN, int TR, int BT>
__device__ __forceinline__ void split_store(const double2* w, double2* buf, int j,
                                            double2* __restrict__ y,
                                            unsigned off_a, unsigned off_b, int kxm) {
  constexpr int N = 1 << LOGN, TPL = N / kE;
  double2* pj = buf + lpad(j);
#pragma unroll
  for (int r = 0; r < kE; ++r)
    if (live_r<TR>(r)) pj[lpad_off(TPL * r)] = w[r];
  __syncthreads();
  const double2* pm = buf + lpad(TPL - j);
#pragma unroll
  for (int r = 0; r <= kE / 2; ++r) {  // k <= N/2 (r = kE/2 only for k = N/2)
    const int k = j + TPL * r;
    if (live_r<TR>(r) && k <= kxm && k <= N / 2) {
      const double2 W = w[r];
      const double2 Wc = mirror_at<LOGN>(buf, pm, j, r);
      y[off_a + koff<LOGN, BT>(k)] = make_double2(0.5 * (W.x + Wc.x), 0.5 * (W.y - Wc.y));
      y[off_b + koff<LOGN, BT>(k)] = make_double2(0.5 * (W.y + Wc.y), -0.5 * (W.x - Wc.x));
    }
  }
  __syncthreads();
}

// One y row per line, NL lines per workgroup (group g holds rows g*NL ..):
// q+iv and q+iu come out of one complex transform each, so each product is
// formed the moment its factors exist and only q*v is carried (as 8
// doubles) into the R2C transform of W = q*u + i q*v.  BT is the tile width
// of the y-pass buffers.  Lines alternate every 16 lanes: with two rows over
// two-column tiles a wave's access covers whole 64-byte runs of a tile.
template <int LOGN, int TR, int BT, int NL, int SCHED>
__device__ __forceinline__ void r_body(const PassArgs& a, int g, int tid,
                                       double2* lds, const double2* tw) {
  constexpr int N = 1 << LOGN, TPL = N / kE;
  constexpr int LINE = Line<LOGN>::PADDED;
  int pl, j;
  line_map<TPL, NL>(tid, pl, j);
  const int yrow = g * NL + pl;
  double2* buf = lds + pl * LINE;
  const int kxm = a.rc.kx;
  const unsigned fstride = (unsigned)a.ntiles * N * BT;  // one field
  const unsigned offQ = yrow_off(yrow, a.ntiles, BT), offU = offQ + fstride,
                 offV = offU + fstride;

  double2 z[kE];
  double qv[kE];
  build_two<LOGN, TR, BT>(z, j, a.y1, offQ, offV, kxm);
  fft_reg<LOGN, +1, SCHED>(z, buf, j, tw);
#pragma unroll
  for (int r = 0; r < kE; ++r) qv[r] = z[r].x * z[r].y;  // produ.f90:16
  build_two<LOGN, TR, BT>(z, j, a.y1, offQ, offU, kxm);
  fft_reg<LOGN, +1, SCHED>(z, buf, j, tw);
#pragma unroll
  for (int r = 0; r < kE; ++r) z[r] = make_double2(z[r].x * z[r].y, qv[r]);  // produ.f90:15
  fft_reg<LOGN, -1, SCHED>(z, buf, j, tw);
  split_store<LOGN, TR, BT>(z, buf, j, a.y2, offQ, offQ + fstride, kxm);
}

// Row-pair variant: each line of N/8 threads handles rows y0 and y0+1
// (group g holds pairs g*NL ..).  The complex transform of Q0 + iQ1 gives q
// of both rows; V0 + iV1 and U0 + iU1 likewise, and the products of the two
// rows are again packed two-for-one for the forward transforms: five
// transforms per row pair, only q carried across them.
template <int LOGN, int TR, int BT, int NL, int SCHED, bool TWPF = false>
__device__ __forceinline__ void pair_body(const PassArgs& a, int g, int tid,
                                          double2* lds, const double2* tw) {
  constexpr int N = 1 << LOGN, TPL = N / kE;
  constexpr int LINE = Line<LOGN>::PADDED;
  int pl, j;
  line_map<TPL, NL>(tid, pl, j);
  double2* buf = lds + pl * LINE;
  const int y0 = 2 * (g * NL + pl);
  const int kxm = a.rc.kx;
  const unsigned fstride = (unsigned)a.ntiles * N * BT;  // one field of y2
  const unsigned r0 = yrow_off(y0, a.ntiles, BT), r1 = r0 + BT;  // rows y0, y0+1 of y2
  // y1: row-pair blocks at N = 4096 (the y pass's layout there), else tiles
  constexpr bool Y1 = LOGN == 12;
  const unsigned cols = Y1 ? (unsigned)y1_cols(a.rc.kc()) : 0u;
  const unsigned fs1 = Y1 ? N * cols : fstride;
  const unsigned i0 = Y1 ? y1_off(y0, 0, cols) : r0, i1 = Y1 ? y1_off(y0 + 1, 0, cols) : r1;
  const unsigned fq = 0, fu = fs1, fv = 2 * fs1;

  // 4096-point lines of a one-pair workgroup use the in-place transforms:
  // the inverse ones leave x out of order, which the pointwise products do
  // not mind, and the forward ones restore natural order for split_store.
  constexpr bool IP = LOGN == 12 && NL == 1 && TWPF;
  constexpr bool IP2 = LOGN == 13 && NL == 1 && TWPF;  // 8192: radix 2, then two 4096s
  // 4096: the three twiddles every transform of the thread uses, loaded once.
  IpTw w{};
  if constexpr (IP) w = ip_tw(tw, j);
  // bar: another wave may still read LDS (ignored by the Stockham transforms,
  // which synchronise on every transpose)
  auto inv = [&](double2* v, bool bar) {
    if constexpr (IP) fft_dif<+1>(v, buf, j, w, bar);
    else if constexpr (IP2) fft_dif2<+1>(v, buf, j, tw, bar);
    else fft_reg<LOGN, +1, SCHED, TWPF>(v, buf, j, tw);
  };
  auto fwd = [&](double2* v) {
    if constexpr (IP || IP2) {
      if constexpr (IP) fft_dit<-1>(v, buf, j, w);
      else fft_dit2<-1>(v, buf, j, tw);
      __syncthreads();  // every wave's reads done before split_store writes
    } else {
      fft_reg<LOGN, -1, SCHED, TWPF>(v, buf, j, tw);
    }
  };
  double2 z[kE], qq[kE];
  build_two<LOGN, TR, BT, Y1>(z, j, a.y1, fq + i0, fq + i1, kxm);  // Q0 + i Q1
  inv(z, false);  // no LDS reader yet
#pragma unroll
  for (int r = 0; r < kE; ++r) qq[r] = z[r];                   // q0 + i q1
  build_two<LOGN, TR, BT, Y1>(z, j, a.y1, fv + i0, fv + i1, kxm);  // V0 + i V1
  inv(z, true);
#pragma unroll
  for (int r = 0; r < kE; ++r) z[r] = make_double2(qq[r].x * z[r].x, qq[r].y * z[r].y);
  fwd(z);                                                      // (q v)0 + i (q v)1
  split_store<LOGN, TR, BT>(z, buf, j, a.y2, fstride + r0, fstride + r1, kxm);
  build_two<LOGN, TR, BT, Y1>(z, j, a.y1, fu + i0, fu + i1, kxm);  // U0 + i U1
  inv(z, false);  // split_store ends with a barrier
#pragma unroll
  for (int r = 0; r < kE; ++r) z[r] = make_double2(qq[r].x * z[r].x, qq[r].y * z[r].y);
  fwd(z);                                                      // (q u)0 + i (q u)1
  split_store<LOGN, TR, BT>(z, buf, j, a.y2, r0, r1, kxm);
}

// Not launched for N >= 2048 (r_pass_pair there; at 4096 the y passes
// write y1 in row-pair blocks, which only pair_body reads).
template <int LOGN, int TR>
__global__ void __launch_bounds__(rows_of(LOGN) * (1 << LOGN) / kE)
    r_pass(PassArgs a) {
  constexpr int RP = rows_of(LOGN);
  __shared__ double2 lds[RP * Line<LOGN>::PADDED];
  __shared__ double2 twl[tw_lds_size<LOGN>()];
  const double2* tw = stage_tw<LOGN>(twl, a.tw, threadIdx.x, blockDim.x);
  r_body<LOGN, TR, bt_of(LOGN), RP, 0>(a, xcd_remap(blockIdx.x, gridDim.x), threadIdx.x, lds,
                                       tw);
}

// N = 8192, one row per workgroup.  A real row x of N points is the
// M = N/2-point complex sequence z[m] = x[2m] + i x[2m+1]; with
// W = exp(-2 pi i / N) its transform is
//   Z[k] = E[k] + i O[k],  E[k] = X[k] + conj(X[M-k]),
//                          O[k] = (X[k] - conj(X[M-k])) W^-k,
// so the C2R transform of a row is one 4096-point complex transform, and
// the R2C transform of a product p the same backwards: from
// ZP = FFT(p_even + i p_odd),
//   P[k] = PE[k] + W^k PO[k],  PE[k] = (ZP[k] + conj(ZP[M-k])) / 2,
//                              PO[k] = (ZP[k] - conj(ZP[M-k])) / 2i.
// Products stay pointwise in z (even samples in .x, odd in .y), as for row
// pairs.  A row needs one 4096-point line of LDS and 512 threads: two
// workgroups share a CU, where a row pair of 8192-point lines fills one.
// The rectangle lies below 3N/8 (X[k] = 0 from k = 3M/4 on).

// W^(j + 512 r) from wj = W^j (8192-point): W^(512 r) = exp(-i pi r / 8).
template <int R>
__device__ __forceinline__ double2 w_step(double2 wj) {
  constexpr double C1 = 0.92387953251128675613, S1 = 0.38268343236508977173;
  constexpr double H = 0.70710678118654752440;
  if constexpr (R == 0) return wj;
  if constexpr (R == 4) return make_double2(wj.y, -wj.x);
  constexpr double c = R == 1 ? C1 : R == 2 ? H : R == 3 ? S1 : R == 5 ? -S1 : R == 6 ? -H : -C1;
  constexpr double s = R == 1 ? -S1 : R == 2 ? -H : R == 3 ? -C1 : R == 5 ? -C1 : R == 6 ? -H : -S1;
  return cmul(wj, make_double2(c, s));
}
// z[R..7] = E + i O from A = X[k], B = X[M-k], k = j + 512 R.
template <int R>
__device__ __forceinline__ void row_pack(double2* z, const double2* A, const double2* B,
                                         double2 wj) {
  if constexpr (R < kE) {
    const double2 a = A[R], b = B[R];
    const double2 e = make_double2(a.x + b.x, a.y - b.y);
    const double2 o = cmulc(make_double2(a.x - b.x, a.y + b.y), w_step<R>(wj));
    z[R] = make_double2(e.x - o.y, e.y + o.x);
    row_pack<R + 1>(z, A, B, wj);
  }
}
// Z of one row of a field (element k at off + kYR k: tiles of one column).
__device__ __forceinline__ void row_build(double2* z, int j, const double2* __restrict__ y,
                                          unsigned off, int kxm,
                                          const double2* __restrict__ tw) {
  constexpr int M = 4096;
  asm volatile("" : "+v"(j));  // offsets formed here, not held from another field's
  const double2 wj = tw[j];
  double2 A[kE], B[kE];
#pragma unroll
  for (int r = 0; r < kE; ++r) {  // all loads in flight together
    const int k = j + 512 * r, mk = M - k;
    A[r] = B[r] = make_double2(0.0, 0.0);
    if (r < 6 && k <= kxm) A[r] = y[off + ycol_off(k, 1)];    // k < 3M/4
    if (r >= 2 && mk <= kxm) B[r] = y[off + ycol_off(mk, 1)];  // M - k < 3M/4
  }
  if (j == 0) A[0].y = 0.0;  // Im X[0] is not used (C2R)
  row_pack<0>(z, A, B, wj);
}
// Store P[k], k <= kx_max, of the product whose ZP the thread holds in
// natural order (k = j + 512 r).  ZP[M-k] comes through LDS.
__device__ __forceinline__ void row_split_store(const double2* w, double2* buf, int j,
                                                double2* __restrict__ y, unsigned off,
                                                int kxm, const double2* __restrict__ tw) {
  constexpr int TPL = 512;
  asm volatile("" : "+v"(j));
  const double2 wj = tw[j];
  double2* pj = buf + lpad(j);
#pragma unroll
  for (int r = 0; r < kE; ++r) pj[lpad_off(TPL * r)] = w[r];
  __syncthreads();
  const double2* pm = buf + lpad(TPL - j);
  auto one = [&](auto RC) {
    constexpr int r = decltype(RC)::value;
    const int k = j + TPL * r;
    if (k <= kxm) {
      const double2 W = w[r];
      const double2 Wc = mirror_at<12>(buf, pm, j, r);
      const double2 pe = make_double2(0.5 * (W.x + Wc.x), 0.5 * (W.y - Wc.y));
      const double2 po = make_double2(0.5 * (W.y + Wc.y), -0.5 * (W.x - Wc.x));
      y[off + ycol_off(k, 1)] = cadd(pe, cmul(w_step<r>(wj), po));
    }
  };
  one(std::integral_constant<int, 0>{});
  one(std::integral_constant<int, 1>{});
  one(std::integral_constant<int, 2>{});
  one(std::integral_constant<int, 3>{});
  one(std::integral_constant<int, 4>{});
  one(std::integral_constant<int, 5>{});
  __syncthreads();
}
__device__ __forceinline__ void row_body13(const PassArgs& a, int y, int j, double2* buf,
                                           const double2* __restrict__ tw) {
  constexpr int N = 8192;
  const int kxm = a.rc.kx;
  const unsigned fstride = (unsigned)a.ntiles * N;  // one field (tiles of one column)
  const unsigned ro = yrow_off(y, a.ntiles, 1);
  double2 z[kE], qq[kE];
  row_build(z, j, a.y1, ro, kxm, tw);  // Q
  fft_dif<+1, 2>(z, buf, j, tw, false);
#pragma unroll
  for (int r = 0; r < kE; ++r) qq[r] = z[r];  // q_even + i q_odd
  row_build(z, j, a.y1, 2 * fstride + ro, kxm, tw);  // V
  fft_dif<+1, 2>(z, buf, j, tw, true);
#pragma unroll
  for (int r = 0; r < kE; ++r) z[r] = make_double2(qq[r].x * z[r].x, qq[r].y * z[r].y);
  fft_dit<-1, false, 2>(z, buf, j, tw);
  __syncthreads();  // every wave's reads done before the split's writes
  row_split_store(z, buf, j, a.y2, fstride + ro, kxm, tw);  // q v
  row_build(z, j, a.y1, fstride + ro, kxm, tw);  // U
  fft_dif<+1, 2>(z, buf, j, tw, false);  // the split ends with a barrier
#pragma unroll
  for (int r = 0; r < kE; ++r) z[r] = make_double2(qq[r].x * z[r].x, qq[r].y * z[r].y);
  fft_dit<-1, false, 2>(z, buf, j, tw);
  __syncthreads();
  row_split_store(z, buf, j, a.y2, ro, kxm, tw);  // q u
}
template <int LOGN>
__global__ void __launch_bounds__(512, 4) r_pass_row(PassArgs a) {
  static_assert(LOGN == 13, "one-row x pass: 8192-point rows only");
  __shared__ double2 lds[kIpLine];
  row_body13(a, xcd_remap(blockIdx.x, gridDim.x), threadIdx.x, lds, a.tw);
}

// One row pair per workgroup: two workgroups share a CU up to N = 4096 (at
// 8192 one fills it: the rectangle wider than 3N/8 only).  Twiddles come
// from memory, each transform pass loading the next pass's entry ahead.
template <int LOGN, int TR>
__global__ void __launch_bounds__((1 << LOGN) / kE, 4)
    r_pass_pair(PassArgs a) {
  // Twiddles from memory: these workgroups are short-lived (one row pair),
  // and staging the table in LDS measured slower here.
  __shared__ double2 lds[LOGN == 12 ? kIpLine : LOGN == 13 ? 2 * kIpLine : Line<LOGN>::PADDED];
  pair_body<LOGN, TR, bt_of(LOGN), 1, 0, true>(a, xcd_remap(blockIdx.x, gridDim.x),
                                               threadIdx.x, lds, a.tw);
}

// ---------------------------------------------------------------------------
// Reference-layout 2D transforms (src/fft_dir.f90, src/fft_inv.f90):
// real rows of N+2 doubles, half-spectrum rows of N/2+1 complex.
// ---------------------------------------------------------------------------

template <int LOGN>
__global__ void __launch_bounds__(rp_of(LOGN) * (1 << LOGN) / kE)
    ref_rows_fwd(double* z, const double2* __restrict__ tw) {
  constexpr int N = 1 << LOGN, TPL = N / kE, RP = rp_of(LOGN);
  constexpr int LINE = Line<LOGN>::PADDED;
  __shared__ double2 lds[RP * LINE];
  const int pl = threadIdx.x / TPL, j = threadIdx.x % TPL;
  const int y0 = 2 * (blockIdx.x * RP + pl);
  double2* buf = lds + pl * LINE;
  const size_t ld = (size_t)N + 2;
  const double* r0 = z + ld * y0;
  const double* r1 = r0 + ld;
  double2 w[kE];
#pragma unroll
  for (int r = 0; r < kE; ++r) {
    const int x = j + TPL * r;
    w[r] = make_double2(r0[x], r1[x]);
  }
  fft_reg<LOGN, -1>(w, buf, j, tw);
  double2* pj = buf + lpad(j);
#pragma unroll
  for (int r = 0; r < kE; ++r) pj[lpad_off(TPL * r)] = w[r];
  __syncthreads();  // also orders every read of the real rows before the writes
  const double2* pm = buf + lpad(TPL - j);
  double2* c0 = reinterpret_cast<double2*>(z + ld * y0);
  double2* c1 = c0 + (N / 2 + 1);
#pragma unroll
  for (int r = 0; r < kE; ++r) {
    const int k = j + TPL * r;
    if (k <= N / 2) {
      const double2 W = w[r];
      const double2 Wc = mirror_at<LOGN>(buf, pm, j, r);
      c0[k] = make_double2(0.5 * (W.x + Wc.x), 0.5 * (W.y - Wc.y));
      c1[k] = make_double2(0.5 * (W.y + Wc.y), -0.5 * (W.x - Wc.x));
    }
  }
}

template <int LOGN>
__global__ void __launch_bounds__(rp_of(LOGN) * (1 << LOGN) / kE)
    ref_rows_inv(double* z, const double2* __restrict__ tw) {
  constexpr int N = 1 << LOGN, TPL = N / kE, RP = rp_of(LOGN);
  constexpr int LINE = Line<LOGN>::PADDED;
  __shared__ double2 lds[RP * LINE];
  const int pl = threadIdx.x / TPL, j = threadIdx.x % TPL;
  const int y0 = 2 * (blockIdx.x * RP + pl);
  double2* buf = lds + pl * LINE;
  const size_t ld = (size_t)N + 2;
  const double2* c0 = reinterpret_cast<const double2*>(z + ld * y0);
  const double2* c1 = c0 + (N / 2 + 1);
  double2 w[kE];
#pragma unroll
  for (int r = 0; r < kE; ++r) {
    const int p = j + TPL * r;
    double2 v;
    if (p == 0 || p == N / 2) {
      v = make_double2(c0[p].x, c1[p].x);
    } else if (p < N / 2) {
      const double2 x0 = c0[p], x1 = c1[p];
      v = make_double2(x0.x - x1.y, x0.y + x1.x);
    } else {
      const double2 x0 = c0[N - p], x1 = c1[N - p];
      v = make_double2(x0.x + x1.y, x1.x - x0.y);
    }
    w[r] = v;
  }
  __syncthreads();  // every read of the complex rows before any write
  fft_reg<LOGN, +1>(w, buf, j, tw);
  double* r0 = z + ld * y0;
  double* r1 = r0 + ld;
#pragma unroll
  for (int r = 0; r < kE; ++r) {
    const int x = j + TPL * r;
    r0[x] = w[r].x;
    r1[x] = w[r].y;
  }
  if (j < 2) {
    r0[N + j] = 0.0;
    r1[N + j] = 0.0;
  }
}

// y transform of every kx column in place; SCALE applies /NY/NX.
template <int LOGN, int S, int SCALE>
__global__ void __launch_bounds__(bt_of(LOGN) * (1 << LOGN) / kE)
    ref_cols(double* z, const double2* __restrict__ tw) {
  constexpr int N = 1 << LOGN, TPL = N / kE, BT = bt_of(LOGN);
  constexpr int LINE = Line<LOGN>::PADDED, NH = N / 2 + 1;
  __shared__ double2 lds[BT * LINE];
  const int tid = threadIdx.x;
  int b, j;
  line_map<TPL, BT>(tid, b, j);
  const int kx = blockIdx.x * BT + b;
  const bool ok = kx < NH;
  double2* zc = reinterpret_cast<double2*>(z);
  double2 w[kE];
#pragma unroll
  for (int r = 0; r < kE; ++r)
    w[r] = ok ? zc[kx + (size_t)NH * (j + TPL * r)] : make_double2(0.0, 0.0);
  fft_reg<LOGN, S>(w, lds + b * LINE, j, tw);
  if (ok) {
#pragma unroll
    for (int r = 0; r < kE; ++r) {
      double2 v = w[r];
      if (SCALE) {
        v.x = v.x / (double)N / (double)N;
        v.y = v.y / (double)N / (double)N;
      }
      zc[kx + (size_t)NH * (j + TPL * r)] = v;
    }
  }
}

// ---------------------------------------------------------------------------
// Adjoint passes.  With q, u, v the physical fields of a stage state s and
// lambda_x,y = G(i k_x,y a) those of the adjoint input a (G the unscaled
// inverse transform, F = G^-1), the transpose of the linearised nonlinear
// term in the Parseval inner product is
//   J(s)^T a = P [ F(u lambda_x + v lambda_y)
//                  + (-i k_y F(q lambda_x) + i k_x F(q lambda_y)) / |k|^alpha ]
// (P the 2/3 mask, the mean mode zero).  The adjoint y pass holds the
// forward y transforms of the three products and their combination, the
// reversed RK stage (JMode) and the inverse y transforms of lambda_x,
// lambda_y, q, u, v for the next transposed term; the adjoint x pass forms
// the products.
// ---------------------------------------------------------------------------

template <int LOGN, int MODE, int AK, int TR>
__global__ void __launch_bounds__(bt_of(LOGN) * (1 << LOGN) / kE, AK == 1 ? 4 : 1)
    j_pass(AdjArgs a) {
  constexpr int N = 1 << LOGN, TPL = N / kE, BT = bt_of(LOGN);
  constexpr int LINE = Line<LOGN>::PADDED;
  constexpr bool FWD = MODE != J_START4 && MODE != J_START2;
  constexpr bool INV = MODE != J_END;
  __shared__ double2 lds[BT * LINE];
  __shared__ double2 twl[tw_lds_size<LOGN>()];
  const double2* tw = stage_tw<LOGN>(twl, a.tw, threadIdx.x, blockDim.x);
  int b, j;
  line_map<TPL, BT>(threadIdx.x, b, j);
  const int tile = blockIdx.x;
  const int c = tile * BT + b;
  const Rect rc = a.rc;
  const bool colok = c <= rc.kx;
  double2* buf = lds + b * LINE;
  const Phys& ph = a.ph;
  const double fkx = (double)c * ph.dkx;
  const double fkxs = fkx * fkx;
  const unsigned cbase = (unsigned)c * (unsigned)a.kyp;
  double2 z[kE];
  double2 av[kE];  // adjoint input of the next transposed term

  if constexpr (FWD) {
    double2 C[kE];
    const double2* y2 = a.y2;
    const size_t fs = (size_t)a.ntiles * N * BT;
    const unsigned colo = ycol_off(c, BT);
#pragma unroll
    for (int r = 0; r < kE; ++r) z[r] = y2[fs + yrow_off(j + TPL * r, a.ntiles, BT) + colo];
    fft_reg<LOGN, -1>(z, buf, j, tw);  // F(q lambda_x)
#pragma unroll
    for (int r = 0; r < kE; ++r) C[r] = live_r<TR>(r) ? z[r] : make_double2(0.0, 0.0);
#pragma unroll
    for (int r = 0; r < kE; ++r) z[r] = y2[2 * fs + yrow_off(j + TPL * r, a.ntiles, BT) + colo];
    fft_reg<LOGN, -1>(z, buf, j, tw);  // F(q lambda_y)
    int jc = j;
    asm volatile("" : "+v"(jc) : "v"(z[0].x), "v"(z[kE - 1].y));
#pragma unroll
    for (int r = 0; r < kE; ++r) {
      const int p = jc + TPL * r;
      double2 v = make_double2(0.0, 0.0);
      if (live_r<TR>(r) && colok && row_of(p, rc, N) >= 0 && !(c == 0 && p == 0)) {
        const double fky = (double)kyi_of(p, N) * ph.dky;
        const double k2 = fkxs + fky * fky;
        const double kar = AK == 1 ? k2 : kalpha(ph, k2);
        v = make_double2((fky * C[r].y - fkx * z[r].y) / kar,
                         (fkx * z[r].x - fky * C[r].x) / kar);
      }
      C[r] = v;
    }
#pragma unroll
    for (int r = 0; r < kE; ++r) z[r] = y2[yrow_off(j + TPL * r, a.ntiles, BT) + colo];
    fft_reg<LOGN, -1>(z, buf, j, tw);  // F(u lambda_x + v lambda_y)
    int je = j;
    asm volatile("" : "+v"(je) : "v"(z[0].x), "v"(z[kE - 1].y));
#pragma unroll
    for (int r = 0; r < kE; ++r) {
      const int p = je + TPL * r;
      const int row = row_of(p, rc, N);
      av[r] = make_double2(0.0, 0.0);
      if (live_r<TR>(r) && colok && row >= 0) {
        const double fky = (double)kyi_of(p, N) * ph.dky;
        const bool mean = (c == 0 && p == 0);
        // the transposed term (fft_dir's 1/N^2 on the forward transforms)
        double2 t = make_double2((z[r].x + C[r].x) * ph.inv_n2, (z[r].y + C[r].y) * ph.inv_n2);
        if ((ph.truncate == 1 && cut_mode(ph, fkxs, fky * fky)) || mean)
          t = make_double2(0.0, 0.0);
        const unsigned q = cbase + (unsigned)row;
        const double E = a.emk[q];
        if constexpr (MODE == J_END) {
          const double2 l = a.lacc[q];
          a.lam[q] = make_double2(l.x + t.x, l.y + t.y);
        } else {
          const double2 mu = mean ? make_double2(0.0, 0.0) : a.lam[q];
          double2 l = a.lacc[q];
          if constexpr (MODE == J_RK4_3) {
            const double EE = E * E;
            l = make_double2(l.x + EE * t.x, l.y + EE * t.y);
            av[r] = make_double2(ph.dt3 * E * mu.x + ph.dt * E * t.x,
                                 ph.dt3 * E * mu.y + ph.dt * E * t.y);
          } else if constexpr (MODE == J_RK4_2) {
            l = make_double2(l.x + E * t.x, l.y + E * t.y);
            av[r] = make_double2(ph.dt3 * E * mu.x + ph.dt2 * t.x,
                                 ph.dt3 * E * mu.y + ph.dt2 * t.y);
          } else if constexpr (MODE == J_RK4_1) {
            const double EE = E * E;
            l = make_double2(l.x + E * t.x, l.y + E * t.y);
            av[r] = make_double2(ph.dt6 * EE * mu.x + ph.dt2 * E * t.x,
                                 ph.dt6 * EE * mu.y + ph.dt2 * E * t.y);
          } else if constexpr (MODE == J_RK2_1) {
            l = make_double2(l.x + E * t.x, l.y + E * t.y);
            av[r] = make_double2(ph.dt2 * E * t.x, ph.dt2 * E * t.y);
          }
          a.lacc[q] = l;
        }
      }
    }
  } else {
#pragma unroll
    for (int r = 0; r < kE; ++r) {
      const int p = j + TPL * r;
      const int row = row_of(p, rc, N);
      av[r] = make_double2(0.0, 0.0);
      if (live_r<TR>(r) && colok && row >= 0) {
        const unsigned q = cbase + (unsigned)row;
        const double2 mu = (c == 0 && p == 0) ? make_double2(0.0, 0.0) : a.lam[q];
        const double E = a.emk[q];
        const double EE = E * E;
        a.lacc[q] = make_double2(EE * mu.x, EE * mu.y);
        if constexpr (MODE == J_START4) av[r] = make_double2(ph.dt6 * mu.x, ph.dt6 * mu.y);
        else av[r] = make_double2(ph.dt * E * mu.x, ph.dt * E * mu.y);
      }
    }
  }

  if constexpr (INV) {
#pragma unroll 1
    for (int f = 0; f < 5; ++f) {
      int jo = j;
      asm volatile("" : "+v"(jo));
#pragma unroll
      for (int r = 0; r < kE; ++r) {
        const int p = jo + TPL * r;
        const int row = row_of(p, rc, N);
        double2 v = make_double2(0.0, 0.0);
        if (live_r<TR>(r) && colok && row >= 0) {
          const double fky = (double)kyi_of(p, N) * ph.dky;
          if (f == 0) {
            v = make_double2(-fkx * av[r].y, fkx * av[r].x);  // i kx a
          } else if (f == 1) {
            v = make_double2(-fky * av[r].y, fky * av[r].x);  // i ky a
          } else {
            const double2 sv = a.snext[cbase + (unsigned)row];
            if (f == 2) {
              v = sv;
            } else if (!(c == 0 && p == 0)) {  // u, v as src/nlt.f90:25-34
              const double k2 = fkxs + fky * fky;
              const double kar = AK == 1 ? k2 : kalpha(ph, k2);
              if (f == 3) {
                const double kyn = fky / kar;
                v = make_double2(-kyn * sv.y, kyn * sv.x);
              } else {
                const double kxn = fkx / kar;
                v = make_double2(kxn * sv.y, -kxn * sv.x);
              }
            }
          }
        }
        z[r] = v;
      }
      fft_reg<LOGN, +1>(z, buf, j, tw);
      double2* dst = a.y1 + (size_t)f * a.ntiles * N * BT;
      const unsigned colo = ycol_off(c, BT);
#pragma unroll
      for (int r = 0; r < kE; ++r) dst[yrow_off(j + TPL * r, a.ntiles, BT) + colo] = z[r];
    }
  }
}

// Half-spectrum (kx <= kx_max) of one real row -> its full x-spectrum at
// the thread's positions (as build_two with g = 0).
template <int LOGN, int TR, int BT>
__device__ __forceinline__ void build_one(double2* z, int j, const double2* __restrict__ y,
                                          unsigned off, int kxm) {
  constexpr int N = 1 << LOGN, TPL = N / kE;
#pragma unroll
  for (int r = 0; r < kE; ++r) {
    const int p = j + TPL * r;
    double2 v = make_double2(0.0, 0.0);
    const int k = p <= N / 2 ? p : N - p;
    if (live_r<TR>(r) && k <= kxm) {
      const double2 x0 = y[off + koff<LOGN, BT>(k)];
      if (p == 0 || p == N / 2) v = make_double2(x0.x, 0.0);
      else if (p < N / 2) v = x0;
      else v = make_double2(x0.x, -x0.y);
    }
    z[r] = v;
  }
}

// Adjoint x pass, one row per line: C2R of lambda_x + i lambda_y, of q + i u
// and of v; R2C of q lambda_x + i q lambda_y and of u lambda_x + v lambda_y.
template <int LOGN, int TR>
__global__ void __launch_bounds__(rows_of(LOGN) * (1 << LOGN) / kE)
    jr_pass(AdjArgs a) {
  constexpr int N = 1 << LOGN, TPL = N / kE, RP = rows_of(LOGN), BT = bt_of(LOGN);
  constexpr int LINE = Line<LOGN>::PADDED;
  __shared__ double2 lds[RP * LINE];
  __shared__ double2 twl[tw_lds_size<LOGN>()];
  const double2* tw = stage_tw<LOGN>(twl, a.tw, threadIdx.x, blockDim.x);
  int pl, j;
  line_map<TPL, RP>(threadIdx.x, pl, j);
  const int yrow = xcd_remap(blockIdx.x, gridDim.x) * RP + pl;
  double2* buf = lds + pl * LINE;
  const int kxm = a.rc.kx;
  const unsigned fs = (unsigned)a.ntiles * N * BT;
  const unsigned off = yrow_off(yrow, a.ntiles, BT);

  double2 z[kE], l[kE];
  double ap[kE];
  build_two<LOGN, TR, BT>(z, j, a.y1, off, off + fs, kxm);  // LX + i LY
  fft_reg<LOGN, +1>(z, buf, j, tw);
#pragma unroll
  for (int r = 0; r < kE; ++r) l[r] = z[r];                   // lambda_x + i lambda_y
  build_two<LOGN, TR, BT>(z, j, a.y1, off + 2 * fs, off + 3 * fs, kxm);  // Q + i U
  fft_reg<LOGN, +1>(z, buf, j, tw);
#pragma unroll
  for (int r = 0; r < kE; ++r) {
    ap[r] = z[r].y * l[r].x;                                  // u lambda_x
    z[r] = make_double2(z[r].x * l[r].x, z[r].x * l[r].y);    // q lambda_x + i q lambda_y
  }
  fft_reg<LOGN, -1>(z, buf, j, tw);
  split_store<LOGN, TR, BT>(z, buf, j, a.y2, off + fs, off + 2 * fs, kxm);
  build_one<LOGN, TR, BT>(z, j, a.y1, off + 4 * fs, kxm);   // V
  fft_reg<LOGN, +1>(z, buf, j, tw);
#pragma unroll
  for (int r = 0; r < kE; ++r) z[r] = make_double2(ap[r] + z[r].x * l[r].y, 0.0);
  fft_reg<LOGN, -1>(z, buf, j, tw);                          // F(u lambda_x + v lambda_y)
#pragma unroll
  for (int r = 0; r <= kE / 2; ++r) {
    const int k = j + TPL * r;
    if (live_r<TR>(r) && k <= kxm && k <= N / 2) a.y2[off + koff<LOGN, BT>(k)] = z[r];
  }
}

// ---------------------------------------------------------------------------
// launchers.  Each grid size is instantiated in its own translation unit
// (compiled with -DGT_INST_LOGN=L); the dispatchers and the size-independent
// kernels are in the one compiled with -DGT_INST_COMMON.
// ---------------------------------------------------------------------------

template <int L>
hipError_t c_pass_launch(int mode, const PassArgs& a, hipStream_t s);
template <int L>
hipError_t r_pass_launch(const PassArgs& a, hipStream_t s);
template <int L>
hipError_t ref_fft_launch(int dir, double* z, const double2* tw, hipStream_t s);
template <int L>
hipError_t j_pass_launch(int mode, const AdjArgs& a, hipStream_t s);
template <int L>
hipError_t jr_pass_launch(const AdjArgs& a, hipStream_t s);
template <int T, int SCHED>
hipError_t ens_c_launch(int mode, const PassArgs* d, const uint32_t* it, int n, int grid,
                        hipStream_t s);
template <int TX, int SCHED>
hipError_t ens_r_launch(int ty, const PassArgs* d, const uint32_t* it, int n, int grid,
                        hipStream_t s);

#if defined(GT_INST_LOGN)
template <int L, int AK, int TR>
static hipError_t c_pass_launch_ak(int mode, const PassArgs& a, hipStream_t s) {
  constexpr int T = bt_of(L) * (1 << L) / kE;
  dim3 g(a.ntiles), b(T);
  switch (mode) {
    case C_START: c_pass<L, C_START, AK, TR><<<g, b, 0, s>>>(a); break;
    case C_NLT: c_pass<L, C_NLT, AK, TR><<<g, b, 0, s>>>(a); break;
    case C_RK4_1: c_pass<L, C_RK4_1, AK, TR><<<g, b, 0, s>>>(a); break;
    case C_RK4_2: c_pass<L, C_RK4_2, AK, TR><<<g, b, 0, s>>>(a); break;
    case C_RK4_3: c_pass<L, C_RK4_3, AK, TR><<<g, b, 0, s>>>(a); break;
    case C_RK4_4: c_pass<L, C_RK4_4, AK, TR><<<g, b, 0, s>>>(a); break;
    case C_RK2_1: c_pass<L, C_RK2_1, AK, TR><<<g, b, 0, s>>>(a); break;
    case C_RK2_2: c_pass<L, C_RK2_2, AK, TR><<<g, b, 0, s>>>(a); break;
    case C_RK4_4S: c_pass<L, C_RK4_4S, AK, TR><<<g, b, 0, s>>>(a); break;
    case C_RK2_2S: c_pass<L, C_RK2_2S, AK, TR><<<g, b, 0, s>>>(a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// TR = 1 when the rectangle lies below 3N/8 on both axes (always so with
// the 2/3 rule); the position set of a thread is j + r*N/8.
__host__ __device__ inline int rect_tr(const Rect& rc, int n) {
  const int lim = 3 * (n / kE);
  return (rc.kx < lim && rc.p < lim && rc.q < lim && n >= 16) ? 1 : 0;
}

template <int L>
hipError_t c_pass_launch(int mode, const PassArgs& a, hipStream_t s) {
  const int tr = rect_tr(a.rc, 1 << L);
  if (a.ph.akind == 1)
    return tr ? c_pass_launch_ak<L, 1, 1>(mode, a, s) : c_pass_launch_ak<L, 1, 0>(mode, a, s);
  return tr ? c_pass_launch_ak<L, 0, 1>(mode, a, s) : c_pass_launch_ak<L, 0, 0>(mode, a, s);
}

template <int L>
hipError_t r_pass_launch(const PassArgs& a, hipStream_t s) {
  if constexpr (L >= 11) {  // one row pair per workgroup (8192: one row)
    constexpr int T = (1 << L) / kE;
    if (rect_tr(a.rc, 1 << L)) {
      if constexpr (L == 13)
        r_pass_row<L><<<dim3(1 << L), dim3(T / 2), 0, s>>>(a);
      else
        r_pass_pair<L, 1><<<dim3((1 << L) / 2), dim3(T), 0, s>>>(a);
    } else {
      r_pass_pair<L, 0><<<dim3((1 << L) / 2), dim3(T), 0, s>>>(a);
    }
    return hipGetLastError();
  }
  constexpr int RP = rows_of(L), T = RP * (1 << L) / kE;
  if (rect_tr(a.rc, 1 << L))
    r_pass<L, 1><<<dim3((1 << L) / RP), dim3(T), 0, s>>>(a);
  else
    r_pass<L, 0><<<dim3((1 << L) / RP), dim3(T), 0, s>>>(a);
  return hipGetLastError();
}

template <int L>
hipError_t ref_fft_launch(int dir, double* z, const double2* tw, hipStream_t s) {
  constexpr int N = 1 << L, RP = rp_of(L), BT = bt_of(L);
  const dim3 gr(N / 2 / RP), br(RP * N / kE);
  const dim3 gc((N / 2 + 1 + BT - 1) / BT), bc(BT * N / kE);
  if (dir < 0) {  // fft_dir: rows R2C, columns forward, /NY/NX
    ref_rows_fwd<L><<<gr, br, 0, s>>>(z, tw);
    ref_cols<L, -1, 1><<<gc, bc, 0, s>>>(z, tw);
  } else {  // fft_inv: columns inverse, rows C2R
    ref_cols<L, +1, 0><<<gc, bc, 0, s>>>(z, tw);
    ref_rows_inv<L><<<gr, br, 0, s>>>(z, tw);
  }
  return hipGetLastError();
}

template <int L, int AK>
static hipError_t j_pass_launch_ak(int mode, const AdjArgs& a, hipStream_t s) {
  constexpr int T = bt_of(L) * (1 << L) / kE;
  const dim3 g(a.ntiles), b(T);
  switch (mode) {
    case J_START4: j_pass<L, J_START4, AK, 1><<<g, b, 0, s>>>(a); break;
    case J_RK4_3: j_pass<L, J_RK4_3, AK, 1><<<g, b, 0, s>>>(a); break;
    case J_RK4_2: j_pass<L, J_RK4_2, AK, 1><<<g, b, 0, s>>>(a); break;
    case J_RK4_1: j_pass<L, J_RK4_1, AK, 1><<<g, b, 0, s>>>(a); break;
    case J_END: j_pass<L, J_END, AK, 1><<<g, b, 0, s>>>(a); break;
    case J_START2: j_pass<L, J_START2, AK, 1><<<g, b, 0, s>>>(a); break;
    case J_RK2_1: j_pass<L, J_RK2_1, AK, 1><<<g, b, 0, s>>>(a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// The adjoint needs the 2/3 rectangle (TR = 1).
template <int L>
hipError_t j_pass_launch(int mode, const AdjArgs& a, hipStream_t s) {
  if (!rect_tr(a.rc, 1 << L)) return hipErrorInvalidValue;
  return a.ph.akind == 1 ? j_pass_launch_ak<L, 1>(mode, a, s)
                         : j_pass_launch_ak<L, 0>(mode, a, s);
}

template <int L>
hipError_t jr_pass_launch(const AdjArgs& a, hipStream_t s) {
  if (!rect_tr(a.rc, 1 << L)) return hipErrorInvalidValue;
  constexpr int RP = rows_of(L), T = RP * (1 << L) / kE;
  jr_pass<L, 1><<<dim3((1 << L) / RP), dim3(T), 0, s>>>(a);
  return hipGetLastError();
}

template hipError_t j_pass_launch<GT_INST_LOGN>(int, const AdjArgs&, hipStream_t);
template hipError_t jr_pass_launch<GT_INST_LOGN>(const AdjArgs&, hipStream_t);
template hipError_t c_pass_launch<GT_INST_LOGN>(int, const PassArgs&, hipStream_t);
template hipError_t r_pass_launch<GT_INST_LOGN>(const PassArgs&, hipStream_t);
template hipError_t ref_fft_launch<GT_INST_LOGN>(int, double*, const double2*, hipStream_t);
#endif  // GT_INST_LOGN

#if defined(GT_INST_ENS)
// ---------------------------------------------------------------------------
// Batched ensemble passes.  One launch covers the pass of every member,
// whatever its N and alpha: the workgroups loop over a list of work items
// (member, tile) and run the pass body of the member's size.  For T threads
// a workgroup holds T*8/N lines of N points, i.e. T*8*17/16 double2 of LDS
// for every N, so the sizes share one LDS allocation.  Alpha enters through
// kalpha_t<2>, the x-pass packing per member through PassArgs::xvar, the
// radix order (SCHED) per launch.  Every member has the 2/3 rectangle
// (TR = 1).  Compiled once per (T, SCHED): -DGT_INST_ENS=T
// -DGT_ENS_SCHED=SCHED.
// ---------------------------------------------------------------------------

// Twiddles of the item's size in LDS: staged again only when the member's
// table changes (barriers on both sides: the previous item's last pass may
// still read the old table).
__device__ __forceinline__ const double2* ens_stage_tw(double2* twl, const PassArgs& a,
                                                       const double2*& cur) {
  if (a.tw != cur) {
    __syncthreads();
    const int n = a.ph.n;
    const int span = n == 128 ? tw_span<7>() : n == 256 ? tw_span<8>()
                   : n == 512 ? tw_span<9>() : n == 1024 ? tw_span<10>() : tw_span<11>();
    for (int i = threadIdx.x; i < span; i += blockDim.x) twl[i] = a.tw[i];
    __syncthreads();
    cur = a.tw;
  }
  return twl;
}

template <int T, int MODE, int SCHED, int L>
__device__ __forceinline__ void ens_c_one(const PassArgs& a, int tile, int tid,
                                          double2* lds, const double2* tw) {
  c_body<L, MODE, 2, 1, ens_lines(T, L), SCHED>(a, tile, tid, lds, tw);
}

template <int T, int MODE, int SCHED>
__global__ void __launch_bounds__(T, 4)
    ens_c_pass(const PassArgs* __restrict__ descs, const uint32_t* __restrict__ items,
               int nitems) {
  __shared__ double2 lds[T * 17 / 2];
  __shared__ double2 twl[tw_span<kEnsMaxLogn>()];
  const double2* cur = nullptr;
  for (int w = blockIdx.x; w < nitems; w += gridDim.x) {
    const uint32_t it = items[w];
    const PassArgs& a = descs[it >> 16];
    const int tile = (int)(it & 0xFFFFu);
    const double2* tw = ens_stage_tw(twl, a, cur);
    // Opaque per item: nothing derived from the thread index is hoisted out
    // of the loop (and kept live over every size's code).
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    switch (a.ph.n) {
      case 128: ens_c_one<T, MODE, SCHED, 7>(a, tile, tid, lds, tw); break;
      case 256: ens_c_one<T, MODE, SCHED, 8>(a, tile, tid, lds, tw); break;
      case 512: ens_c_one<T, MODE, SCHED, 9>(a, tile, tid, lds, tw); break;
      case 1024: ens_c_one<T, MODE, SCHED, 10>(a, tile, tid, lds, tw); break;
      case 2048: ens_c_one<T, MODE, SCHED, 11>(a, tile, tid, lds, tw); break;
      default: break;
    }
  }
}

template <int TX, int TY, int SCHED, int L>
__device__ __forceinline__ void ens_r_one(const PassArgs& a, int g, int tid, double2* lds,
                                          const double2* tw) {
  constexpr int NL = ens_lines(TX, L), BT = ens_lines(TY, L);
  if (a.xvar) pair_body<L, 1, BT, NL, SCHED>(a, g, tid, lds, tw);
  else r_body<L, 1, BT, NL, SCHED>(a, g, tid, lds, tw);
}

template <int TX, int TY, int SCHED>
__global__ void __launch_bounds__(TX, 4)
    ens_r_pass(const PassArgs* __restrict__ descs, const uint32_t* __restrict__ items,
               int nitems) {
  __shared__ double2 lds[TX * 17 / 2];
  __shared__ double2 twl[tw_span<kEnsMaxLogn>()];
  const double2* cur = nullptr;
  for (int w = blockIdx.x; w < nitems; w += gridDim.x) {
    const uint32_t it = items[w];
    const PassArgs& a = descs[it >> 16];
    const int g = (int)(it & 0xFFFFu);
    const double2* tw = ens_stage_tw(twl, a, cur);
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    switch (a.ph.n) {
      case 128: ens_r_one<TX, TY, SCHED, 7>(a, g, tid, lds, tw); break;
      case 256: ens_r_one<TX, TY, SCHED, 8>(a, g, tid, lds, tw); break;
      case 512: ens_r_one<TX, TY, SCHED, 9>(a, g, tid, lds, tw); break;
      case 1024: ens_r_one<TX, TY, SCHED, 10>(a, g, tid, lds, tw); break;
      case 2048: ens_r_one<TX, TY, SCHED, 11>(a, g, tid, lds, tw); break;
      default: break;
    }
  }
}

template <int T, int SCHED>
hipError_t ens_c_launch(int mode, const PassArgs* d, const uint32_t* it, int n, int grid,
                        hipStream_t s) {
  const dim3 g(grid), b(T);
  switch (mode) {
    case C_START: ens_c_pass<T, C_START, SCHED><<<g, b, 0, s>>>(d, it, n); break;
    case C_RK4_1: ens_c_pass<T, C_RK4_1, SCHED><<<g, b, 0, s>>>(d, it, n); break;
    case C_RK4_2: ens_c_pass<T, C_RK4_2, SCHED><<<g, b, 0, s>>>(d, it, n); break;
    case C_RK4_3: ens_c_pass<T, C_RK4_3, SCHED><<<g, b, 0, s>>>(d, it, n); break;
    case C_RK4_4: ens_c_pass<T, C_RK4_4, SCHED><<<g, b, 0, s>>>(d, it, n); break;
    case C_RK2_1: ens_c_pass<T, C_RK2_1, SCHED><<<g, b, 0, s>>>(d, it, n); break;
    case C_RK2_2: ens_c_pass<T, C_RK2_2, SCHED><<<g, b, 0, s>>>(d, it, n); break;
    case C_RK4_4S: ens_c_pass<T, C_RK4_4S, SCHED><<<g, b, 0, s>>>(d, it, n); break;
    case C_RK2_2S: ens_c_pass<T, C_RK2_2S, SCHED><<<g, b, 0, s>>>(d, it, n); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int TX, int SCHED>
hipError_t ens_r_launch(int ty, const PassArgs* d, const uint32_t* it, int n, int grid,
                        hipStream_t s) {
  const dim3 g(grid), b(TX);
  if (ty == 256) ens_r_pass<TX, 256, SCHED><<<g, b, 0, s>>>(d, it, n);
  else if (ty == 512) ens_r_pass<TX, 512, SCHED><<<g, b, 0, s>>>(d, it, n);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

template hipError_t ens_c_launch<GT_INST_ENS, GT_ENS_SCHED>(int, const PassArgs*,
                                                            const uint32_t*, int, int,
                                                            hipStream_t);
template hipError_t ens_r_launch<GT_INST_ENS, GT_ENS_SCHED>(int, const PassArgs*,
                                                            const uint32_t*, int, int,
                                                            hipStream_t);
#endif  // GT_INST_ENS

#if defined(GT_INST_COMMON)
// ---------------------------------------------------------------------------
// Elementwise operators
// ---------------------------------------------------------------------------

// src/produ.f90:11-23
__global__ void produ_k(int n, const double* __restrict__ z, double* u,
                        double* v) {
  const size_t ld = (size_t)n + 2, tot = ld * n;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < tot;
       e += (size_t)gridDim.x * blockDim.x) {
    const size_t i = e % ld;
    if (i < (size_t)n) {
      u[e] = z[e] * u[e];
      v[e] = z[e] * v[e];
    } else {
      u[e] = 0.0;
      v[e] = 0.0;
    }
  }
}

// src/trunc.f90:10-25 on the reference layout.
__global__ void trunc_ref_k(Phys ph, double2* z) {
  const int n = ph.n, nh = n / 2 + 1;
  const size_t tot = (size_t)nh * n;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < tot;
       e += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(e % nh), jj = (int)(e / nh);
    const double fkx = (double)i * ph.dkx, fky = (double)kyi_of(jj, n) * ph.dky;
    if (cut_mode(ph, fkx * fkx, fky * fky) || e == 0) z[e] = make_double2(0.0, 0.0);
  }
}

__global__ void trunc_cmp_k(Phys ph, Rect rc, int kyp, double2* z) {
  const int n = ph.n, ky = rc.ky();
  const size_t tot = (size_t)rc.kc() * ky;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < tot;
       e += (size_t)gridDim.x * blockDim.x) {
    const int c = (int)(e / ky), r = (int)(e % ky);
    const int p = fft_of(r, rc, n);
    const double fkx = (double)c * ph.dkx, fky = (double)kyi_of(p, n) * ph.dky;
    if (cut_mode(ph, fkx * fkx, fky * fky) || (c == 0 && r == 0))
      z[(size_t)c * kyp + r] = make_double2(0.0, 0.0);
  }
}

// Reference half-spectrum -> compact rectangle (optionally truncated).
__global__ void gather_k(Phys ph, Rect rc, int kyp, const double2* __restrict__ ref,
                         double2* cmp, int apply_trunc) {
  const int n = ph.n, nh = n / 2 + 1, ky = rc.ky(), kc = rc.kc();
  const size_t tot = (size_t)kc * ky;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < tot;
       e += (size_t)gridDim.x * blockDim.x) {
    const int c = (int)(e % kc), r = (int)(e / kc);
    const int p = fft_of(r, rc, n);
    double2 v = ref[c + (size_t)nh * p];
    if (apply_trunc) {
      const double fkx = (double)c * ph.dkx, fky = (double)kyi_of(p, n) * ph.dky;
      if (cut_mode(ph, fkx * fkx, fky * fky) || (c == 0 && p == 0))
        v = make_double2(0.0, 0.0);
    }
    cmp[(size_t)c * kyp + r] = v;
  }
}

// ka = k2**(alpv/2) over the compact rectangle, with the expressions of the
// y pass (c_body, AK = 0), so that a batched member reading the table and a
// single run computing ka in place get the same values.
__global__ void ka_table_k(Phys ph, Rect rc, int kyp, double* ka) {
  const int n = ph.n, ky = rc.ky();
  const size_t tot = (size_t)rc.kc() * ky;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < tot;
       e += (size_t)gridDim.x * blockDim.x) {
    const int c = (int)(e / ky), r = (int)(e % ky);
    const double fkx = (double)c * ph.dkx;
    const double fkxs = fkx * fkx;
    const double fky = (double)kyi_of(fft_of(r, rc, n), n) * ph.dky;
    ka[(size_t)c * kyp + r] = (c == 0 && r == 0) ? 1.0 : kalpha(ph, fkxs + fky * fky);
  }
}

// Flags a nonzero reference mode outside the rectangle.
__global__ void outside_k(Phys ph, Rect rc, const double2* __restrict__ ref,
                          int* flag) {
  const int n = ph.n, nh = n / 2 + 1;
  const size_t tot = (size_t)nh * n;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < tot;
       e += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(e % nh), p = (int)(e / nh);
    if (i > rc.kx || row_of(p, rc, n) < 0) {
      const double2 v = ref[e];
      if (v.x != 0.0 || v.y != 0.0) atomicOr(flag, 1);
    }
  }
}

// Compact rectangle -> reference half-spectrum (zeros outside).
__global__ void scatter_k(Phys ph, Rect rc, int kyp, const double2* __restrict__ cmp,
                          double2* ref) {
  const int n = ph.n, nh = n / 2 + 1;
  const size_t tot = (size_t)nh * n;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < tot;
       e += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(e % nh), p = (int)(e / nh);
    const int r = row_of(p, rc, n);
    ref[e] = (i <= rc.kx && r >= 0) ? cmp[(size_t)i * kyp + r]
                                    : make_double2(0.0, 0.0);
  }
}

// src/forcing.f90:16-33.  One thread per forced mode: no two modes share an
// element, and a mode's own updates (itself, then its Hermitian partner) run
// in the reference's order.  amp = sdt*famp.
// z(i,j) += sdt*famp*(ff1 cos te, ff2 sin te) and the conjugate update of
// the Hermitian partner (src/forcing.f90:25-31).
__device__ __forceinline__ void force_one(const ForcedMode& f, double2 cs, double amp,
                                          double2* z) {
#pragma clang fp contract(off)
  const double a1 = amp * f.ff1 * cs.x, a2 = amp * f.ff2 * cs.y;
  double2 v = z[f.idx];
  v.x = v.x + a1;
  v.y = v.y + a2;
  z[f.idx] = v;
  if (f.pidx >= 0) {
    double2 w = z[f.pidx];
    w.x = w.x + a1;
    w.y = w.y - a2;
    z[f.pidx] = w;
  }
}

__global__ void forcing_k(int nf, const ForcedMode* __restrict__ fm,
                          const double2* __restrict__ ph, double amp,
                          double2* z) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nf) return;
  force_one(fm[k], ph[k], amp, z);
}

// Forcing of every ensemble member in one launch, one thread per forced
// mode; the phases of chunk step *step are used.
__global__ void ens_forcing_k(const EnsForce* __restrict__ f, int nmem, int total,
                              const int* __restrict__ step) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  int lo = 0, hi = nmem - 1;  // last member whose first mode is <= t
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (f[mid].first <= t) lo = mid;
    else hi = mid - 1;
  }
  const EnsForce& e = f[lo];
  const int k = t - e.first;
  if (k >= e.nf) return;
  force_one(e.modes[k], e.phases[(size_t)(*step) * e.nf + k], e.amp, e.z);
}

__global__ void ens_tick_k(int* step) { step[0] = step[0] + 1; }

// ---------------------------------------------------------------------------
// Diagnostics (src/diagnostics.f90), deterministic, in two launches each.
//
// Spectrum and fluxes: one workgroup (4 waves) per kx column bins the
// column's modes into LDS bins, one per shell.  In each half of a column
// (ky >= 0: rows 0..p; ky < 0: rows p+1..) the shell index is monotone in
// the row, so a chunk of 64 consecutive rows of one wave holds runs of equal
// shells: a segmented wavefront scan (lane shuffles, a fixed tree) sums each
// run, the run's last lane carries the sum, and runs continued across the
// four waves of a 256-row chunk are joined in wave order by the run's first
// lane, which alone adds it to the shell's LDS bin.  The column's bins then
// go to part[c][b].  A second launch sums part[.][b] over the columns, one
// wavefront per shell: lanes take columns c = lane, lane + 64, ... in
// order, then a fixed shuffle tree.  The order of every sum is fixed, so
// the result does not change from run to run; it replaces the reference's
// unprotected accumulations on the GPU (diagnostics.f90:25-39,84-101).
// ---------------------------------------------------------------------------

__host__ __device__ __forceinline__ int shell_of(const Phys& ph, double fkxs, int p) {
  const double fky = (double)kyi_of(p, ph.n) * ph.dky;
  const double k = sqrt(fkxs + fky * fky);
  return (int)(k / ph.dkx) + 1;  // diagnostics.f90:32-34 (dk = dkx)
}

// Shells a column's kept rows fall in: [lo, hi], hi below NBIN
// (diagnostics.f90:35 bins only ib < NBIN).
__host__ __device__ __forceinline__ void column_shells(const Phys& ph, const Rect& rc, int c,
                                                       int& lo, int& hi) {
  const int n = ph.n;
  const double fkx = (double)c * ph.dkx, fkxs = fkx * fkx;
  lo = shell_of(ph, fkxs, 0);
  hi = shell_of(ph, fkxs, fft_of(rc.p, rc, n));
  if (rc.ky() > rc.p + 1) {
    const int h2 = shell_of(ph, fkxs, fft_of(rc.p + 1, rc, n));
    hi = h2 > hi ? h2 : hi;
  }
  hi = hi < n / 2 - 1 ? hi : n / 2 - 1;
}

constexpr int kDiagT = 256;  // threads of a diagnostics workgroup (4 waves)

// Inclusive sums of v1, v2 over each run of equal keys in the wave (runs
// are contiguous); returns true on a run's last lane.
__device__ __forceinline__ bool wave_runs(int key, double& v1, double& v2) {
  const int lane = (int)(threadIdx.x & 63);
  const int kprev = __shfl_up(key, 1);
  const bool head = lane == 0 || kprev != key;
  const unsigned long long heads = __ballot(head);
  // run number: heads at or below this lane
  const int run = __popcll(heads & (~0ull >> (63 - lane)));
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const double t1 = __shfl_up(v1, d), t2 = __shfl_up(v2, d);
    const int rr = __shfl_up(run, d);
    if (lane >= d && rr == run) {
      v1 = v1 + t1;
      v2 = v2 + t2;
    }
  }
  const bool next_head = lane == 63 ? true : ((heads >> (lane + 1)) & 1ull) != 0;
  return next_head;
}

// which = 0: spectrum (sp, sz);  which = 1: fluxes (flzm, flem)
// nbl: the most shells of any column (LDS bins per sum).
__global__ void __launch_bounds__(kDiagT) shell_bin_k(DiagArgs d, int which, int nbl) {
  extern __shared__ double dsm[];
  const Phys& ph = d.ph;
  const Rect rc = d.rc;
  const int n = ph.n, nbin = n / 2, c = blockIdx.x, tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  double* bin1 = dsm;                  // shell b of the column in bin b - lo
  double* bin2 = bin1 + nbl;
  double* sv1 = bin2 + nbl;            // per-lane run sums of a chunk
  double* sv2 = sv1 + kDiagT;
  int* skey = reinterpret_cast<int*>(sv2 + kDiagT);  // key on a run's last lane, else -1
  int* ftail = skey + kDiagT;                        // each wave's first run end
  const double fkx = (double)c * ph.dkx, fkxs = fkx * fkx;
  const bool edge = (c == 0 || c == n / 2);
  const size_t cq = (size_t)c * d.kyp;
  int blo, bhi;
  column_shells(ph, rc, c, blo, bhi);
  if (bhi - blo + 1 > nbl) bhi = blo + nbl - 1;  // never: nbl is the maximum
  for (int b = blo + tid; b <= bhi; b += kDiagT) {
    bin1[b - blo] = 0.0;
    bin2[b - blo] = 0.0;
  }
  __syncthreads();
  const int P = rc.p, Ky = rc.ky();
  for (int half = 0; half < 2; ++half) {
    const int r0 = half ? P + 1 : 0, r1 = half ? Ky : P + 1;  // rows [r0, r1)
    for (int base = r0; base < r1; base += kDiagT) {
      const int r = base + tid;
      int key = -1;
      double v1 = 0.0, v2 = 0.0;
      if (r < r1 && !(c == 0 && r == 0)) {  // diagnostics.f90:28
        const int p = fft_of(r, rc, n);
        const double fky = (double)kyi_of(p, n) * ph.dky;
        const double k2 = fkxs + fky * fky;
        const double k = sqrt(k2);
        const int ib = (int)(k / ph.dkx) + 1;  // diagnostics.f90:34
        if (ib < nbin && ib >= blo && ib <= bhi) {
          key = ib;
          const double ka = kalpha(ph, k2);
          const double2 zv = d.z[cq + r];
          if (which == 0) {
            const double fac = edge ? 0.5 : 1.0;  // diagnostics.f90:29-30
            const double am = hypot(zv.x, zv.y);
            const double a2 = am * am;
            v1 = fac * a2 / ka;  // sp, diagnostics.f90:36
            v2 = fac * a2;       // sz, :37
          } else {
            const double fac = edge ? 1.0 : 2.0;  // diagnostics.f90:89-90
            const double2 nv = d.znl[cq + r];
            const double fz = zv.x * nv.x + zv.y * nv.y;  // :93
            v1 = fac * fz;          // flzm, :98
            v2 = fac * (fz / ka);   // flem, :95,99
          }
        }
      }
      const bool last = wave_runs(key, v1, v2);
      const unsigned long long lasts = __ballot(last);
      skey[tid] = last ? key : -1;
      sv1[tid] = v1;
      sv2[tid] = v2;
      if (lane == 0) ftail[wave] = (wave << 6) + __builtin_ctzll(lasts);
      __syncthreads();
      if (last && key >= 0) {
        // A wave's first run continues the previous wave's last run when the
        // keys agree: that run belongs to the lane that ends its first piece.
        const bool cont = wave > 0 && tid == ftail[wave] && skey[(wave << 6) - 1] == key;
        if (!cont) {
          double a1 = v1, a2 = v2;
          if (lane == 63) {
            for (int w2 = wave + 1; w2 < kDiagT / 64; ++w2) {
              const int t = ftail[w2];
              if (skey[t] != key) break;
              a1 = a1 + sv1[t];
              a2 = a2 + sv2[t];
              if ((t & 63) != 63) break;
            }
          }
          bin1[key - blo] = bin1[key - blo] + a1;
          bin2[key - blo] = bin2[key - blo] + a2;
        }
      }
      __syncthreads();
    }
  }
  for (int b = blo + tid; b <= bhi; b += kDiagT) {
    d.part1[(size_t)c * nbin + (b - 1)] = bin1[b - blo];
    d.part2[(size_t)c * nbin + (b - 1)] = bin2[b - blo];
  }
}

// Sum of a wavefront's values in a fixed tree order; lane 0 holds the total.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = v + __shfl_down(v, off);
  return v;
}

// One wavefront per shell: the columns' bins of shell b, over the columns in
// a fixed order.
__global__ void __launch_bounds__(kDiagT) shell_sum_k(DiagArgs d, int which) {
  const Phys& ph = d.ph;
  const Rect rc = d.rc;
  const int n = ph.n, nbin = n / 2;
  const int b = (int)((blockIdx.x * kDiagT + threadIdx.x) >> 6) + 1;  // shell 1..NBIN
  const int lane = threadIdx.x & 63;
  if (b > nbin) return;
  double s1 = 0.0, s2 = 0.0;
  for (int c = lane; c <= rc.kx; c += 64) {
    int lo, hi;
    column_shells(ph, rc, c, lo, hi);
    if (b >= lo && b <= hi) {
      s1 = s1 + d.part1[(size_t)c * nbin + (b - 1)];
      s2 = s2 + d.part2[(size_t)c * nbin + (b - 1)];
    }
  }
  s1 = wave_sum(s1);
  s2 = wave_sum(s2);
  if (lane != 0) return;
  if (which == 0) {  // diagnostics.f90:42-45
    d.out1[b - 1] = s1 / ph.dkx;
    d.out2[b - 1] = s2 / ph.dkx;
  } else {
    d.part1[(size_t)(rc.kx + 1) * nbin + (b - 1)] = s1;  // flzm
    d.part2[(size_t)(rc.kx + 1) * nbin + (b - 1)] = s2;  // flem
  }
}

// flz(i) = sum(flzm(i:NBIN)) in the reference's order (diagnostics.f90:104-109):
// each thread sums its own suffix in increasing q, from LDS (workgroup g
// stages flzm and flem from its first i on).
__global__ void __launch_bounds__(kDiagT) suffix_k(DiagArgs d) {
  extern __shared__ double ssm[];
  const int nbin = d.ph.n / 2;
  const int i0 = blockIdx.x * kDiagT, len = nbin - i0;
  const double* flzm = d.part1 + (size_t)(d.rc.kx + 1) * nbin;
  const double* flem = d.part2 + (size_t)(d.rc.kx + 1) * nbin;
  double* lz = ssm;
  double* le = ssm + len;
  for (int t = threadIdx.x; t < len; t += kDiagT) {
    lz[t] = flzm[i0 + t];
    le[t] = flem[i0 + t];
  }
  __syncthreads();
  const int i = i0 + (int)threadIdx.x;
  if (i >= nbin) return;
  double sz = 0.0, se = 0.0;
  int q = i - i0;
  // sixteen terms loaded together, then added one after the other in q order
  for (; q + 16 <= len; q += 16) {
    double tz[16], te[16];
#pragma unroll
    for (int t = 0; t < 16; ++t) {
      tz[t] = lz[q + t];
      te[t] = le[q + t];
    }
#pragma unroll
    for (int t = 0; t < 16; ++t) {
      sz = sz + tz[t];
      se = se + te[t];
    }
  }
  for (; q < len; ++q) {
    sz = sz + lz[q];
    se = se + le[q];
  }
  d.out1[i] = se;  // fle
  d.out2[i] = sz;  // flz
}

// k2**a for the dissipation exponents of glob: repeated products for the
// small whole exponents the runs use (k^8 hyperviscosity is a = 4, linear
// drag a = 0), pow otherwise.
__device__ __forceinline__ double pow_k2(double k2, double a) {
  if (a == 0.0) return 1.0;
  if (a == 1.0) return k2;
  if (a == 2.0) return k2 * k2;
  if (a == 3.0) return (k2 * k2) * k2;
  if (a == 4.0) {
    const double t = k2 * k2;
    return t * t;
  }
  return pow(k2, a);
}

// src/diagnostics.f90:147-163: one workgroup per column; each thread sums
// its rows in order, then a fixed wavefront tree and the four waves in
// order.  dise_* use k2**alpnu / ka for k2**(alpnu - alpv/2).
__global__ void __launch_bounds__(kDiagT) glob_col_k(DiagArgs d) {
  const Phys& ph = d.ph;
  const Rect rc = d.rc;
  const int n = ph.n, c = blockIdx.x, Ky = rc.ky();
  const double fkx = (double)c * ph.dkx, fkxs = fkx * fkx;
  const double fac = (c == 0 || c == n / 2) ? 0.5 : 1.0;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int r = threadIdx.x; r < Ky; r += kDiagT) {
    if (c == 0 && r == 0) continue;
    const int p = fft_of(r, rc, n);
    const double fky = (double)kyi_of(p, n) * ph.dky;
    const double k2 = fkxs + fky * fky;
    const double ka = kalpha(ph, k2);
    const double2 zv = d.z[(size_t)c * d.zcs + (size_t)r * d.zrs];
    const double am = hypot(zv.x, zv.y);
    const double scra = am * am;
    const double dnu = 2.0 * fac * ph.nu * pow_k2(k2, ph.alpnu) * scra;
    const double dmu = 2.0 * fac * ph.mu * pow_k2(k2, ph.alpmu) * scra;
    acc[0] = acc[0] + fac * scra / ka;  // ee
    acc[1] = acc[1] + fac * scra;       // zz
    acc[2] = acc[2] + dnu;              // disz_nu
    acc[3] = acc[3] + dmu;              // disz_mu
    acc[4] = acc[4] + dnu / ka;         // dise_nu
    acc[5] = acc[5] + dmu / ka;         // dise_mu
  }
  __shared__ double red[6][kDiagT / 64];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    const double v = wave_sum(acc[q]);
    if (lane == 0) red[q][wave] = v;
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    const int q = threadIdx.x;
    double v = red[q][0];
    for (int w = 1; w < kDiagT / 64; ++w) v = v + red[q][w];
    d.gpart[(size_t)c * 6 + q] = v;
  }
}

// One wavefront per quantity, over the columns in a fixed order.
__global__ void __launch_bounds__(6 * 64) glob_sum_k(DiagArgs d) {
  const int q = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double s = 0.0;
  for (int c = lane; c <= d.rc.kx; c += 64) s = s + d.gpart[(size_t)c * 6 + q];
  s = wave_sum(s);
  if (lane == 0) d.gout[q] = s;
}

__global__ void shell_index_k(Phys ph, int* ib) {
  const int n = ph.n, nh = n / 2 + 1;
  const size_t tot = (size_t)nh * n;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < tot;
       e += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(e % nh), p = (int)(e / nh);
    const double fkx = (double)i * ph.dkx;
    ib[e] = (e == 0) ? 0 : shell_of(ph, fkx * fkx, p);
  }
}

static inline unsigned grid_for(size_t tot, int block) {
  size_t g = (tot + block - 1) / block;
  return (unsigned)(g > 65536 ? 65536 : (g == 0 ? 1 : g));
}

#define GT_LOGN_CASE(V, ...) \
  case V: {                     \
    constexpr int L = V;        \
    __VA_ARGS__;                \
  } break;
#define GT_LOGN_SWITCH(LOGN_VAR, ...)                                     \
  switch (LOGN_VAR) {                                                     \
    GT_LOGN_CASE(4, __VA_ARGS__) GT_LOGN_CASE(5, __VA_ARGS__)             \
    GT_LOGN_CASE(6, __VA_ARGS__) GT_LOGN_CASE(7, __VA_ARGS__)             \
    GT_LOGN_CASE(8, __VA_ARGS__) GT_LOGN_CASE(9, __VA_ARGS__)             \
    GT_LOGN_CASE(10, __VA_ARGS__) GT_LOGN_CASE(11, __VA_ARGS__)           \
    GT_LOGN_CASE(12, __VA_ARGS__) GT_LOGN_CASE(13, __VA_ARGS__)           \
    default: return hipErrorInvalidValue;                                 \
  }

hipError_t launch_c_pass(int logn, int mode, const PassArgs& a, hipStream_t s) {
  GT_LOGN_SWITCH(logn, return c_pass_launch<L>(mode, a, s));
  return hipSuccess;
}

hipError_t launch_r_pass(int logn, const PassArgs& a, hipStream_t s) {
  GT_LOGN_SWITCH(logn, return r_pass_launch<L>(a, s));
  return hipSuccess;
}

hipError_t launch_j_pass(int logn, int mode, const AdjArgs& a, hipStream_t s) {
  GT_LOGN_SWITCH(logn, return j_pass_launch<L>(mode, a, s));
  return hipSuccess;
}

hipError_t launch_jr_pass(int logn, const AdjArgs& a, hipStream_t s) {
  GT_LOGN_SWITCH(logn, return jr_pass_launch<L>(a, s));
  return hipSuccess;
}

hipError_t launch_ref_fft(int logn, int dir, double* z, const double2* tw,
                          hipStream_t s) {
  GT_LOGN_SWITCH(logn, return ref_fft_launch<L>(dir, z, tw, s));
  return hipSuccess;
}

hipError_t launch_produ(int n, const double* z, double* u, double* v,
                        hipStream_t s) {
  produ_k<<<grid_for(((size_t)n + 2) * n, 256), 256, 0, s>>>(n, z, u, v);
  return hipGetLastError();
}

hipError_t launch_trunc_ref(const Phys& ph, double* z, hipStream_t s) {
  trunc_ref_k<<<grid_for((size_t)(ph.n / 2 + 1) * ph.n, 256), 256, 0, s>>>(
      ph, reinterpret_cast<double2*>(z));
  return hipGetLastError();
}

hipError_t launch_trunc_cmp(const Phys& ph, const Rect& rc, int kyp, double2* z,
                            hipStream_t s) {
  trunc_cmp_k<<<grid_for((size_t)rc.kc() * rc.ky(), 256), 256, 0, s>>>(ph, rc, kyp, z);
  return hipGetLastError();
}

hipError_t launch_gather(const Phys& ph, const Rect& rc, int kyp, const double* ref,
                         double2* cmp, int apply_trunc, int* outside_flag,
                         hipStream_t s) {
  const double2* rc2 = reinterpret_cast<const double2*>(ref);
  gather_k<<<grid_for((size_t)rc.kc() * rc.ky(), 256), 256, 0, s>>>(
      ph, rc, kyp, rc2, cmp, apply_trunc);
  if (outside_flag)
    outside_k<<<grid_for((size_t)(ph.n / 2 + 1) * ph.n, 256), 256, 0, s>>>(
        ph, rc, rc2, outside_flag);
  return hipGetLastError();
}

hipError_t launch_scatter(const Phys& ph, const Rect& rc, int kyp, const double2* cmp,
                          double* ref, hipStream_t s) {
  scatter_k<<<grid_for((size_t)(ph.n / 2 + 1) * ph.n, 256), 256, 0, s>>>(
      ph, rc, kyp, cmp, reinterpret_cast<double2*>(ref));
  return hipGetLastError();
}

hipError_t launch_forcing(const Phys& ph, int nf, const ForcedMode* modes,
                          const double2* phases, double amp, double2* z,
                          hipStream_t s) {
  (void)ph;
  if (nf <= 0) return hipSuccess;
  forcing_k<<<(nf + 255) / 256, 256, 0, s>>>(nf, modes, phases, amp, z);
  return hipGetLastError();
}

hipError_t launch_ens_c_pass(int ty, int sched, int mode, const PassArgs* descs,
                             const uint32_t* items, int nitems, int grid,
                             hipStream_t s) {
  if (nitems <= 0) return hipSuccess;
  if (ty == 256) return sched ? ens_c_launch<256, 1>(mode, descs, items, nitems, grid, s)
                              : ens_c_launch<256, 0>(mode, descs, items, nitems, grid, s);
  if (ty == 512) return sched ? ens_c_launch<512, 1>(mode, descs, items, nitems, grid, s)
                              : ens_c_launch<512, 0>(mode, descs, items, nitems, grid, s);
  return hipErrorInvalidValue;
}

hipError_t launch_ens_r_pass(int tx, int ty, int sched, const PassArgs* descs,
                             const uint32_t* items, int nitems, int grid,
                             hipStream_t s) {
  if (nitems <= 0) return hipSuccess;
  if (tx == 256) return sched ? ens_r_launch<256, 1>(ty, descs, items, nitems, grid, s)
                              : ens_r_launch<256, 0>(ty, descs, items, nitems, grid, s);
  if (tx == 512) return sched ? ens_r_launch<512, 1>(ty, descs, items, nitems, grid, s)
                              : ens_r_launch<512, 0>(ty, descs, items, nitems, grid, s);
  return hipErrorInvalidValue;
}

hipError_t launch_ka_table(const Phys& ph, const Rect& rc, int kyp, double* ka,
                           hipStream_t s) {
  const size_t tot = (size_t)rc.kc() * rc.ky();
  ka_table_k<<<grid_for(tot, 256), 256, 0, s>>>(ph, rc, kyp, ka);
  return hipGetLastError();
}

hipError_t launch_ens_forcing(const EnsForce* f, int nmem, int total,
                              const int* step, hipStream_t s) {
  if (total <= 0) return hipSuccess;
  ens_forcing_k<<<(total + 255) / 256, 256, 0, s>>>(f, nmem, total, step);
  return hipGetLastError();
}

hipError_t launch_ens_tick(int* step, hipStream_t s) {
  ens_tick_k<<<1, 1, 0, s>>>(step);
  return hipGetLastError();
}

hipError_t launch_spectrum(const DiagArgs& d, int which, hipStream_t s) {
  const int nbin = d.ph.n / 2;
  int nbl = 1;  // the most shells of one column
  for (int c = 0; c <= d.rc.kx; ++c) {
    int lo, hi;
    column_shells(d.ph, d.rc, c, lo, hi);
    nbl = hi - lo + 1 > nbl ? hi - lo + 1 : nbl;
  }
  // LDS: two sums per shell of a column, the chunk's run sums and keys, and
  // one slot per wave
  const size_t lds = sizeof(double) * (2 * (size_t)nbl + 2 * kDiagT) +
                     sizeof(int) * (kDiagT + kDiagT / 64);
  static bool attr = false;  // once per process: more than 64 KB of LDS at N = 8192
  if (!attr && lds > 48 * 1024) {
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&shell_bin_k),
                                             hipFuncAttributeMaxDynamicSharedMemorySize,
                                             (int)lds);
    if (e != hipSuccess) return e;
    attr = true;
  }
  shell_bin_k<<<d.rc.kc(), kDiagT, lds, s>>>(d, which, nbl);
  shell_sum_k<<<(nbin * 64 + kDiagT - 1) / kDiagT, kDiagT, 0, s>>>(d, which);
  if (which == 1) {
    const size_t slds = 2 * sizeof(double) * (size_t)nbin;  // the first workgroup's
    static bool sattr = false;
    if (!sattr && slds > 48 * 1024) {
      const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&suffix_k),
                                               hipFuncAttributeMaxDynamicSharedMemorySize,
                                               (int)slds);
      if (e != hipSuccess) return e;
      sattr = true;
    }
    suffix_k<<<(nbin + kDiagT - 1) / kDiagT, kDiagT, slds, s>>>(d);
  }
  return hipGetLastError();
}

hipError_t launch_glob(const DiagArgs& d, hipStream_t s) {
  glob_col_k<<<d.rc.kc(), kDiagT, 0, s>>>(d);
  glob_sum_k<<<1, 6 * 64, 0, s>>>(d);
  return hipGetLastError();
}

// Partial sums of the allocation-free glob: six per kx column (at most
// N/2 + 1 = 4097 columns, N <= 8192) and the six totals, 196,704 bytes per
// GPU for the life of the process.  Host code never uses the array's name
// as an address: glob_scratch resolves the device address once per device.
constexpr size_t kGlobScratch = 6 * (8192 / 2 + 1) + 6;
static __device__ __align__(256) double scratch[kGlobScratch];

hipError_t glob_scratch(double** ptr, size_t* count) {
  constexpr int kMaxDev = 64;
  static std::mutex mu;
  static double* addr[kMaxDev] = {};
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDev) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(mu);
  if (addr[dev] == nullptr) {
    void* a = nullptr;
    e = hipGetSymbolAddress(&a, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    addr[dev] = static_cast<double*>(a);
  }
  *ptr = addr[dev];
  *count = kGlobScratch;
  return hipSuccess;
}

hipError_t launch_shell_index(const Phys& ph, int* ib, hipStream_t s) {
  shell_index_k<<<grid_for((size_t)(ph.n / 2 + 1) * ph.n, 256), 256, 0, s>>>(ph, ib);
  return hipGetLastError();
}

#endif  // GT_INST_COMMON

}  // namespace gt
