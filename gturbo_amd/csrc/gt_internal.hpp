// gt_internal.hpp -- types shared by the HIP kernels and the host runtime.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace gt {

// Kept-mode rectangle of the half-spectrum in the compact device layout:
// kx = 0..kx, ky = 0..p and -q..-1.  With the 2/3 rule on, p = q = M and
// kx = M (M = largest index inside the truncation circle along an axis);
// without it the rectangle is the whole half-spectrum (p = N/2, q = N/2-1).
struct Rect {
  int kx, p, q;
  __host__ __device__ int ky() const { return p + 1 + q; }
  __host__ __device__ int kc() const { return kx + 1; }
};

// Device-side scalars of commphy/commk (src/modules.f90:30-36,55-65).
struct Phys {
  int n;
  double dkx, dky;
  double ahalf;  // alpv/2 (nlt.f90:27)
  // how ka = k2**(alpv/2) is evaluated: 1 -> k2 (alpv = 2), 2 -> sqrt(k2)
  // (alpv = 1), 3 -> sqrt(sqrt(k2)) (alpv = 1/2), 0 -> pow(k2, alpv/2)
  int akind;
  double dt, dt2, dt3, dt6;
  double rtrunc2, fkxsmax, fkysmax;
  int truncate;
  double inv_n;  // 1/N (fft_dir.f90:19 divides by NY then NX)
  double inv_n2; // 1/N^2, exact for N = 2^k
  double nu, mu, alpnu, alpmu, alpv;
};

// One forcing update of a kx column for the fused step boundary: the mode
// (or Hermitian partner, conj = 1) in compact row `row`, phase k of the
// step's draws, weights ff (src/forcing.f90:25-31).
struct FEntry {
  int row, k, conj;
  double ff1, ff2;
};

struct EnsForce;

// Arguments of the column (y) and row (x) passes of the nonlinear term.
struct PassArgs {
  Rect rc;
  int kyp;     // pitch of a compact column (double2)
  int ntiles;  // column tiles of the y-pass buffers
  Phys ph;
  const double2* tw;  // exp(-2 pi i m/N), m < N
  double2* y1;        // 3 fields after the inverse y transform
  double2* y2;        // 2 fields after the forward x transform
  double2* z;         // stage state (compact)
  double2* znew;      // RK4 accumulator (compact)
  const double* emk;  // integrating factor (compact, real)
  double2* out;       // nonlinear term (compact), NLT mode
  const double* ka;   // k2**(alpv/2) (compact), batched passes with akind 0
  double2* sst;       // if set: the stage input formed by this y pass is
                      // also stored here (compact), for the adjoint's recompute
  // batched ensemble passes only: x-pass packing (0 = one row per line,
  // 1 = row pairs)
  int xvar;
  // fused step boundary (C_RK4_4S, C_RK2_2S) with forcing: column c's updates
  // are fce[fcs[c] .. fcs[c+1]-1], with the step's phases fph (cos, sin) and
  // amplitude famp = sqrt(dt)*famp; fcs = nullptr: no forcing
  const int* fcs;
  const FEntry* fce;
  const double2* fph;
  double famp;
  // batched ensembles: the phases and amplitude come from the member's
  // forcing descriptor instead (fens->phases[*fstep * fens->nf ..],
  // fens->amp), at the step counter the graph advances on the device
  const EnsForce* fens;
  const int* fstep;
};

// Arguments of the adjoint passes (transposed nonlinear term inside the
// reversed RK stages).
struct AdjArgs {
  Rect rc;
  int kyp;
  int ntiles;
  Phys ph;
  const double2* tw;
  double2* y1;           // 5 fields after the inverse y transforms:
                         // lambda_x, lambda_y, q, u, v
  double2* y2;           // 3 fields after the forward x transforms:
                         // u lambda_x + v lambda_y, q lambda_x, q lambda_y
  double2* lam;          // adjoint state: lambda' in, lambda out (J_END)
  double2* lacc;         // the step's adjoint accumulator
  const double* emk;
  const double2* snext;  // stage state of the next transposed term
};

enum JMode : int {
  J_START4 = 0,  // RK4: lacc = E^2 mu, a = dt/6 mu; inverse part for s4
  J_RK4_3 = 1,   // a_s4 in: lacc += E^2 a_s4, a = dt/3 E mu + dt E a_s4 (s3)
  J_RK4_2 = 2,   // a_s3 in: lacc += E a_s3, a = dt/3 E mu + dt/2 a_s3 (s2)
  J_RK4_1 = 3,   // a_s2 in: lacc += E a_s2, a = dt/6 E^2 mu + dt/2 E a_s2 (s1)
  J_END = 4,     // a_s1 in: lambda = lacc + a_s1
  J_START2 = 5,  // RK2: lacc = E^2 mu, a = dt E mu (s2)
  J_RK2_1 = 6,   // a_s2 in: lacc += E a_s2, a = dt/2 E a_s2 (s1)
};

enum CMode : int {
  C_START = 0,  // inverse part only: stage input = z
  C_NLT = 1,    // forward part only: out = N(q)
  C_RK4_1 = 2,
  C_RK4_2 = 3,
  C_RK4_3 = 4,
  C_RK4_4 = 5,
  C_RK2_1 = 6,
  C_RK2_2 = 7,
  // step boundary: the last stage of a step (C_RK4_4 / C_RK2_2), the step's
  // forcing, and the inverse part of the next step (C_START), in one pass
  C_RK4_4S = 8,
  C_RK2_2S = 9,
};

// Column-tile width of the y pass (BT), y rows per workgroup of the x pass,
// and row pairs per workgroup of the reference-layout transforms, per
// log2 N.  A line is N/8 threads and (N + N/16) double2 of LDS.  Two
// columns (rows) per workgroup up to N = 4096: a 4-row block of two columns
// is one 128-byte line, and small grids get as many workgroups as there
// are column pairs (wider tiles left 256^2 with six y workgroups).
__host__ __device__ constexpr int bt_of(int logn) { return logn <= 12 ? 2 : 1; }
__host__ __device__ constexpr int rows_of(int logn) { return logn <= 12 ? 2 : 1; }
__host__ __device__ constexpr int rp_of(int logn) {
  return logn <= 6 ? (1 << (logn - 1)) : logn == 7 ? 32 : logn == 8 ? 16
       : logn == 9 ? 8 : logn == 10 ? 4 : logn == 11 ? 2 : 1;
}

// Layout of the fields between the passes: rows in groups of kYR; group g
// of a field holds, tile after tile, a block of kYR rows x BT columns.  The
// y pass writes whole blocks (full cache lines); the x pass reads a row
// pair's pieces of consecutive tiles from one DRAM page.  Measured at
// 4096^2 (BT = 2): kYR = 4 beats whole-column tiles, 8 and 2 (2 leaves the
// y pass writing half lines).
constexpr int kYR = 4;
__host__ __device__ inline unsigned yrow_off(int y, int ntiles, int bt) {
  return (unsigned)(y / kYR) * (unsigned)(ntiles * kYR * bt) + (unsigned)(y % kYR * bt);
}
__host__ __device__ inline unsigned ycol_off(int c, int bt) {
  return (unsigned)(c / bt) * (unsigned)(kYR * bt) + (unsigned)(c % bt);
}

// N = 4096: the inverse y pass writes its three fields (x pass input) in
// blocks of 2 rows x kY1C columns, one 128-byte line each, so that the x
// pass, which transforms row pairs, reads whole lines.  Within a block the
// two rows of a column are adjacent: the y pass, whose lanes hold
// consecutive rows of one column, writes 32 contiguous bytes per lane pair,
// and a two-column tile half of each line.  Measured at 4096^2 against the
// kYR x BT tiles: x pass -10%, y stages +9%, step -1%.  A field holds
// y1_cols(kc) columns; element (row y, column c).
constexpr int kY1C = 4;
__host__ __device__ inline int y1_cols(int kc) { return (kc + kY1C - 1) / kY1C * kY1C; }
__host__ __device__ inline unsigned y1_off(int y, int c, int cols) {
  return (unsigned)(y >> 1) * (unsigned)(2 * cols) + (unsigned)(c / kY1C) * (2 * kY1C) +
         (unsigned)(c % kY1C * 2 + (y & 1));
}

// Compact row r of FFT index p (ky in FFT order), or -1 outside the rect.
__host__ __device__ inline int row_of(int p, const Rect& rc, int n) {
  if (p <= rc.p) return p;
  if (p >= n - rc.q) return p - (n - rc.ky());
  return -1;
}
// FFT index of compact row r.
__host__ __device__ inline int fft_of(int r, const Rect& rc, int n) {
  return r <= rc.p ? r : n - rc.ky() + r;
}
// Signed ky index of FFT index p (src/defk.f90:25-30).
__host__ __device__ inline int kyi_of(int p, int n) {
  return p <= n / 2 ? p : p - n;
}

struct DiagArgs {
  Rect rc;
  int kyp;
  Phys ph;
  const double2* z;
  // glob only: mode (column c, row r) of z at c * zcs + r * zrs -- (kyp, 1)
  // on the compact layout, (1, N/2+1) on the reference layout (full rect)
  int zcs, zrs;
  const double2* znl;
  double* part1;  // (kx+1) x NBIN partial sums
  double* part2;
  double* out1;   // NBIN
  double* out2;
  double* gpart;  // (kx+1) x 6
  double* gout;   // 6
};

// Forced mode in the compact layout (src/iniforcing.f90 order).
struct ForcedMode {
  int idx;      // compact index of (ikxf, ikyf)
  int pidx;     // compact index of the Hermitian partner, or -1
  double ff1, ff2;
};

// Forced modes of one ensemble member for the batched forcing launch.
struct EnsForce {
  const ForcedMode* modes;
  const double2* phases;  // [step][nf] of the current chunk of steps
  double2* z;
  double amp;             // sqrt(dt)*famp
  int nf;
  int first;  // index of the member's first mode in the flattened launch
};

// Sizes the batched ensemble passes are compiled for (log2 N) and their
// workgroup sizes (threads; a workgroup holds T*8/N lines of a size).
constexpr int kEnsMinLogn = 7, kEnsMaxLogn = 11;
__host__ __device__ constexpr int ens_lines(int t, int logn) { return t * 8 >> logn; }

// Host-side launchers (gturbo_kernels.hip).
hipError_t launch_c_pass(int logn, int mode, const PassArgs& a, hipStream_t s);
hipError_t launch_r_pass(int logn, const PassArgs& a, hipStream_t s);
hipError_t launch_ref_fft(int logn, int dir, double* z, const double2* tw,
                          hipStream_t s);
hipError_t launch_j_pass(int logn, int mode, const AdjArgs& a, hipStream_t s);
hipError_t launch_jr_pass(int logn, const AdjArgs& a, hipStream_t s);
hipError_t launch_produ(int n, const double* z, double* u, double* v,
                        hipStream_t s);
hipError_t launch_trunc_ref(const Phys& ph, double* z, hipStream_t s);
hipError_t launch_trunc_cmp(const Phys& ph, const Rect& rc, int kyp,
                            double2* z, hipStream_t s);
hipError_t launch_gather(const Phys& ph, const Rect& rc, int kyp,
                         const double* ref, double2* cmp, int apply_trunc,
                         int* outside_flag, hipStream_t s);
hipError_t launch_scatter(const Phys& ph, const Rect& rc, int kyp,
                          const double2* cmp, double* ref, hipStream_t s);
hipError_t launch_forcing(const Phys& ph, int nf, const ForcedMode* modes,
                          const double2* phases, double amp, double2* z,
                          hipStream_t s);
hipError_t launch_spectrum(const DiagArgs& d, int which, hipStream_t s);
hipError_t launch_glob(const DiagArgs& d, hipStream_t s);
// Device address of the static partial-sum array of the allocation-free glob
// (gt_glob_noalloc) on the current device, resolved once per device, and its
// size in doubles.
hipError_t glob_scratch(double** ptr, size_t* count);
hipError_t launch_shell_index(const Phys& ph, int* ib, hipStream_t s);
// Batched ensemble passes: items are (member << 16 | tile or row group),
// grid workgroups loop over them; sched is the radix order (fft_reg SCHED).
hipError_t launch_ens_c_pass(int ty, int sched, int mode, const PassArgs* descs,
                             const uint32_t* items, int nitems, int grid,
                             hipStream_t s);
hipError_t launch_ens_r_pass(int tx, int ty, int sched, const PassArgs* descs,
                             const uint32_t* items, int nitems, int grid,
                             hipStream_t s);
hipError_t launch_ens_forcing(const EnsForce* f, int nmem, int total,
                              const int* step, hipStream_t s);
hipError_t launch_ens_tick(int* step, hipStream_t s);
hipError_t launch_ka_table(const Phys& ph, const Rect& rc, int kyp, double* ka,
                           hipStream_t s);

}  // namespace gt
