/*
 * gturbo_oracle.c -- CPU restatement of the reference's pseudospectral time
 * step, written in plain C (OpenMP over the same loops the reference
 * parallelises with OpenACC).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library:
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 * Nothing in gturbo_amd/ links, loads or calls it.
 *
 * Every function follows one reference subroutine and cites it (paths are
 * relative to the reference checkout).  The arithmetic keeps the reference's
 * operand order; the build uses -ffp-contract=off so that no multiply-add is
 * contracted behind the reader's back.
 *
 * The 2D transform is the one piece whose reference implementation (cuFFT,
 * NVIDIA HPC SDK nvhpc/23.11, src/jobscript:18) is closed and absent.  It is
 * restated here from the definition of cuFFT's D2Z/Z2D in-place transforms:
 * forward e^{-i k x}, inverse e^{+i k x} unnormalised, real rows padded to
 * N+2 doubles (src/main.f90:27,32; src/fft_dir.f90:12; src/fft_inv.f90:12).
 * It is pinned against the DFT definition and numpy.fft in tests.
 */
#include "gturbo_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NFORC 2048 /* src/modules.f90:41 */

struct or_ctx {
  int n, nh; /* NX = NY, NX2P1 */
  /* commphy (src/modules.f90:30-36) */
  double xlx, xly, dkx, dky, nu, mu, alpnu, alpmu, alpv;
  double dt, dt2, dt3, dt6, sdt, inpz, inpe;
  /* commforc (src/modules.f90:39-46) */
  double famp, k1f, k2f;
  int nf;
  int32_t ikxf[NFORC], ikyf[NFORC];
  double ff[2 * NFORC];
  /* commsim (src/modules.f90:49-52) */
  int npas, iout, imix, split, rko;
  /* commk (src/modules.f90:55-65) */
  double *emk, *emkx, *emky, *fkxs, *fkx, *fkys, *fky;
  double *ka; /* k2**(alpv/2) of src/nlt.f90:26-27, taken once per mode */
  int32_t *jc; /* 1-based, as jc(j) */
  double rtrunc, fkxsmax, fkysmax;
  int truncate;
  /* transform tables */
  double *tw; /* e^{-2 pi i k/n}, k < n/2 */
  int *rev;   /* bit reversal of 0..n-1 */
};

static char g_err[256];
const char *or_last_error(void) { return g_err; }

void or_set_threads(int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
}

/* ------------------------------------------------------------------ */
/* 1D complex FFTs (iterative radix-2), the oracle's own transform.     */
/* ------------------------------------------------------------------ */

/* NB independent length-n transforms at once, real and imaginary parts
 * apart: element j of transform b is re[j * NB + b] + i im[j * NB + b],
 * with im = re + n * NB.  Each transform sees exactly the operations of a
 * lone radix-2 transform (bit reversal, then butterflies with
 * e^{sign 2 pi i k/len}); the NB-wide inner loops only let the compiler
 * vectorise across transforms, so the result of one transform does not
 * depend on NB or on which other transforms share the block. */
#define FFT_NB 4
typedef double v4d __attribute__((vector_size(4 * sizeof(double))));
static inline v4d ld4(const double *p) {
  v4d v;
  memcpy(&v, p, sizeof v);
  return v;
}
static inline void st4(double *p, v4d v) { memcpy(p, &v, sizeof v); }

static void fft4(const or_ctx *c, double *re, int sign) {
  const int n = c->n;
  double *im = re + (size_t)n * FFT_NB;
  for (int i = 0; i < n; i++) {
    int r = c->rev[i];
    if (r > i) {
      v4d t = ld4(re + i * FFT_NB);
      st4(re + i * FFT_NB, ld4(re + r * FFT_NB));
      st4(re + r * FFT_NB, t);
      t = ld4(im + i * FFT_NB);
      st4(im + i * FFT_NB, ld4(im + r * FFT_NB));
      st4(im + r * FFT_NB, t);
    }
  }
  for (int len = 2; len <= n; len <<= 1) {
    int half = len >> 1, step = n / len;
    for (int s = 0; s < n; s += len) {
      for (int k = 0; k < half; k++) {
        const double wr1 = c->tw[2 * (k * step)];
        const double wi1 = sign * c->tw[2 * (k * step) + 1];
        const v4d wr = {wr1, wr1, wr1, wr1}, wi = {wi1, wi1, wi1, wi1};
        double *pr = re + (size_t)(s + k) * FFT_NB, *pi = im + (size_t)(s + k) * FFT_NB;
        double *qr = re + (size_t)(s + k + half) * FFT_NB;
        double *qi = im + (size_t)(s + k + half) * FFT_NB;
        const v4d ar = ld4(pr), ai = ld4(pi), br = ld4(qr), bi = ld4(qi);
        /* each lane: the radix-2 butterfly of one transform */
        const v4d xr = br * wr - bi * wi;
        const v4d xi = br * wi + bi * wr;
        st4(qr, ar - xr);
        st4(qi, ai - xi);
        st4(pr, ar + xr);
        st4(pi, ai + xi);
      }
    }
  }
}

/* Block element access: line b, position j. */
#define BRE(a, j, b) (a)[(size_t)(j) * FFT_NB + (b)]
#define BIM(a, j, b) (a)[(size_t)c->n * FFT_NB + (size_t)(j) * FFT_NB + (b)]

static int setup_fft(or_ctx *c) {
  const int n = c->n;
  int lg = 0;
  while ((1 << lg) < n) lg++;
  if ((1 << lg) != n) return -1;
  c->tw = (double *)malloc(sizeof(double) * n);
  c->rev = (int *)malloc(sizeof(int) * n);
  const long double twopi = 6.283185307179586476925286766559005768L;
  for (int k = 0; k < n / 2; k++) {
    long double ang = twopi * (long double)k / (long double)n;
    c->tw[2 * k] = (double)cosl(ang);
    c->tw[2 * k + 1] = -(double)sinl(ang);
  }
  for (int i = 0; i < n; i++) {
    int r = 0;
    for (int b = 0; b < lg; b++)
      if (i & (1 << b)) r |= 1 << (lg - 1 - b);
    c->rev[i] = r;
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* inieuler + defk + iniforcing                                        */
/* ------------------------------------------------------------------ */

/* src/defk.f90:1-75 */
static void defk(or_ctx *c) {
  const int n = c->n, nh = c->nh;
  c->jc[0] = 1; /* defk.f90:12-15 */
  for (int j = 2; j <= n; j++) c->jc[j - 1] = n + 2 - j;
  c->dkx = 2.0 * 3.14159265358979323846 / c->xlx; /* defk.f90:18-19 */
  c->dky = 2.0 * 3.14159265358979323846 / c->xly;
  for (int i = 1; i <= nh; i++) c->fkx[i - 1] = (double)(i - 1) * c->dkx;
  for (int j = 1; j <= n / 2 + 1; j++) c->fky[j - 1] = (double)(j - 1) * c->dky;
  for (int j = n / 2 + 2; j <= n; j++)
    c->fky[j - 1] = (double)(j - n - 1) * c->dky; /* defk.f90:28-30 */
  for (int i = 0; i < nh; i++) c->fkxs[i] = c->fkx[i] * c->fkx[i];
  for (int j = 0; j < n; j++) c->fkys[j] = c->fky[j] * c->fky[j];
  /* defk.f90:43-50 */
  c->emk[0] = 0.0;
  for (int j = 0; j < n; j++)
    for (int i = 0; i < nh; i++) {
      if (i == 0 && j == 0) continue;
      double fk2 = c->fkxs[i] + c->fkys[j];
      double scra = c->nu * pow(fk2, c->alpnu) + c->mu * pow(fk2, c->alpmu);
      c->emk[i + (size_t)nh * j] = exp(-0.5 * c->dt * scra);
    }
  for (int j = 0; j < n; j++) { /* defk.f90:58-62 */
    double fk2 = c->fkys[j];
    double scra = c->nu * pow(fk2, c->alpnu) + c->mu;
    c->emky[j] = exp(-0.5 * c->dt * scra);
  }
  for (int i = 0; i < nh; i++) { /* defk.f90:64-68 */
    double fk2 = c->fkxs[i];
    double scra = c->nu * pow(fk2, c->alpnu);
    c->emkx[i] = exp(-0.5 * c->dt * scra);
  }
  /* nlt.f90:26-27: ka = k2**(alpv/2) depends on the mode only; the same
   * pow call, made once here instead of once per nlt. */
  const double e = c->alpv / 2.0;
  c->ka[0] = 1.0; /* the (1,1) mode is skipped (nlt.f90:25) */
  for (int j = 0; j < n; j++)
    for (int i = 0; i < nh; i++)
      if (i || j) c->ka[i + (size_t)nh * j] = pow(c->fkxs[i] + c->fkys[j], e);
  c->fkxsmax = c->fkxs[nh - 1]; /* defk.f90:71-73 */
  c->fkysmax = c->fkys[n / 2];
  c->rtrunc = c->rtrunc * c->rtrunc;
}

/* One candidate of src/iniforcing.f90:17-29 (the three loops share it). */
static int force_candidate(or_ctx *c, int i, int j) {
  double k2 = c->fkxs[i - 1] + c->fkys[j - 1];
  double k2n = c->fkxs[i - 1] / c->fkxsmax + c->fkys[j - 1] / c->fkysmax;
  if (k2n <= c->rtrunc) {
    if ((k2 <= c->k2f * c->k2f) && (k2 > c->k1f * c->k1f)) {
      c->nf++;
      if (c->nf > NFORC) {
        snprintf(g_err, sizeof g_err, "NFORC too small !");
        return -1;
      }
      c->ikxf[c->nf - 1] = i;
      c->ikyf[c->nf - 1] = j;
      c->ff[2 * (c->nf - 1)] = 1.0;
      c->ff[2 * (c->nf - 1) + 1] = 1.0;
      c->inpz = c->inpz + c->famp * c->famp;
      c->inpe = c->inpe + c->famp * c->famp / pow(k2, c->alpv / 2.0);
    }
  }
  return 0;
}

/* src/iniforcing.f90:1-80 -- the order of the forced modes matters, since
 * each one consumes the next draw of rann in forcing(). */
static int iniforcing(or_ctx *c) {
  const int n = c->n, nh = c->nh;
  c->nf = 0;
  c->inpe = 0.0;
  c->inpz = 0.0;
  for (int j = 2; j <= n / 2 + 1; j++)
    if (force_candidate(c, 1, j)) return -1;
  for (int i = 2; i <= n / 2; i++)
    for (int j = 1; j <= n; j++)
      if (force_candidate(c, i, j)) return -1;
  for (int j = 1; j <= n / 2 + 1; j++)
    if (force_candidate(c, nh, j)) return -1;
  return 0;
}

or_ctx *or_create(int n, const double *p) {
  if (n < 4 || (n & (n - 1))) {
    snprintf(g_err, sizeof g_err, "N must be a power of two >= 4 (got %d)", n);
    return NULL;
  }
  or_ctx *c = (or_ctx *)calloc(1, sizeof(or_ctx));
  c->n = n;
  c->nh = n / 2 + 1;
  /* src/inieuler.f90:14-19 */
  c->xlx = p[0];
  c->xly = p[1];
  c->nu = p[2];
  c->alpnu = p[3];
  c->mu = p[4];
  c->alpmu = p[5];
  c->alpv = p[6];
  c->dt = p[7];
  c->rko = (int)p[8];
  c->npas = (int)p[9];
  c->iout = (int)p[10];
  c->imix = (int)p[11];
  c->rtrunc = p[12];
  c->truncate = (int)p[13];
  c->famp = p[14];
  c->k1f = p[15];
  c->k2f = p[16];
  /* src/inieuler.f90:22-36 */
  c->split = 0;
  if (c->alpnu < 1.0005 && c->alpnu > 0.9995 && c->alpmu < 0.0005 &&
      c->alpmu > -0.0005)
    c->split = 1;
  if (c->rko == 4) c->split = 0;
  if (c->rko != 2 && c->rko != 4) c->rko = 2; /* inieuler.f90:63-66 */
  /* src/inieuler.f90:78-81 */
  c->sdt = sqrt(c->dt);
  c->dt2 = c->dt / 2.0;
  c->dt3 = c->dt / 3.0;
  c->dt6 = c->dt / 6.0;
  const int nh = c->nh;
  c->emk = (double *)malloc(sizeof(double) * (size_t)nh * n);
  c->ka = (double *)malloc(sizeof(double) * (size_t)nh * n);
  c->emkx = (double *)malloc(sizeof(double) * nh);
  c->emky = (double *)malloc(sizeof(double) * n);
  c->fkx = (double *)malloc(sizeof(double) * nh);
  c->fkxs = (double *)malloc(sizeof(double) * nh);
  c->fky = (double *)malloc(sizeof(double) * n);
  c->fkys = (double *)malloc(sizeof(double) * n);
  c->jc = (int32_t *)malloc(sizeof(int32_t) * n);
  defk(c);
  if (iniforcing(c) || setup_fft(c)) {
    or_destroy(c);
    return NULL;
  }
  return c;
}

void or_destroy(or_ctx *c) {
  if (!c) return;
  free(c->emk);
  free(c->ka);
  free(c->emkx);
  free(c->emky);
  free(c->fkx);
  free(c->fkxs);
  free(c->fky);
  free(c->fkys);
  free(c->jc);
  free(c->tw);
  free(c->rev);
  free(c);
}

void or_get_scalars(const or_ctx *c, double *out, int32_t *iout) {
  double v[] = {c->dkx,    c->dky,     c->dt2,     c->dt3,  c->dt6, c->sdt,
                c->rtrunc, c->fkxsmax, c->fkysmax, c->inpz, c->inpe};
  memcpy(out, v, sizeof v);
  iout[0] = c->nf;
  iout[1] = c->split;
  iout[2] = c->rko;
  iout[3] = c->truncate;
}

void or_get_defk(const or_ctx *c, double *fkx, double *fky, double *fkxs,
                 double *fkys, int32_t *jc, double *emk, double *emkx,
                 double *emky) {
  const int n = c->n, nh = c->nh;
  memcpy(fkx, c->fkx, sizeof(double) * nh);
  memcpy(fkxs, c->fkxs, sizeof(double) * nh);
  memcpy(fky, c->fky, sizeof(double) * n);
  memcpy(fkys, c->fkys, sizeof(double) * n);
  memcpy(jc, c->jc, sizeof(int32_t) * n);
  memcpy(emk, c->emk, sizeof(double) * (size_t)nh * n);
  memcpy(emkx, c->emkx, sizeof(double) * nh);
  memcpy(emky, c->emky, sizeof(double) * n);
}

void or_get_forcing_modes(const or_ctx *c, int32_t *ikxf, int32_t *ikyf,
                          double *ff) {
  memcpy(ikxf, c->ikxf, sizeof(int32_t) * c->nf);
  memcpy(ikyf, c->ikyf, sizeof(int32_t) * c->nf);
  memcpy(ff, c->ff, sizeof(double) * 2 * c->nf);
}

/* ------------------------------------------------------------------ */
/* Transforms: src/fft_dir.f90, src/fft_inv.f90                         */
/* ------------------------------------------------------------------ */

/* In-place D2Z over rows then columns, then z(:,j)=z(:,j)/NY/NX
 * (src/fft_dir.f90:12,17-21). */
void or_fft_dir(const or_ctx *c, double *z) {
  const int n = c->n, nh = c->nh;
  const size_t ld = (size_t)n + 2;
#pragma omp parallel
  {
    double *a = (double *)malloc(sizeof(double) * 2 * n * FFT_NB);
#pragma omp for schedule(static)
    for (int j0 = 0; j0 < n; j0 += FFT_NB) { /* rows, FFT_NB at a time */
      const int nb = n - j0 < FFT_NB ? n - j0 : FFT_NB;
      if (nb < FFT_NB) memset(a, 0, sizeof(double) * 2 * n * FFT_NB);
      for (int b = 0; b < nb; b++) {
        const double *row = z + ld * (j0 + b);
        for (int x = 0; x < n; x++) {
          BRE(a, x, b) = row[x];
          BIM(a, x, b) = 0.0;
        }
      }
      fft4(c, a, 1);
      for (int b = 0; b < nb; b++) {
        double *row = z + ld * (j0 + b);
        for (int k = 0; k < nh; k++) {
          row[2 * k] = BRE(a, k, b);
          row[2 * k + 1] = BIM(a, k, b);
        }
      }
    }
#pragma omp for schedule(static)
    for (int i0 = 0; i0 < nh; i0 += FFT_NB) { /* columns, FFT_NB at a time */
      const int nb = nh - i0 < FFT_NB ? nh - i0 : FFT_NB;
      if (nb < FFT_NB) memset(a, 0, sizeof(double) * 2 * n * FFT_NB);
      for (int j = 0; j < n; j++)
        for (int b = 0; b < nb; b++) {
          BRE(a, j, b) = z[2 * (i0 + b + (size_t)nh * j)];
          BIM(a, j, b) = z[2 * (i0 + b + (size_t)nh * j) + 1];
        }
      fft4(c, a, 1);
      for (int j = 0; j < n; j++)
        for (int b = 0; b < nb; b++) {
          z[2 * (i0 + b + (size_t)nh * j)] = BRE(a, j, b);
          z[2 * (i0 + b + (size_t)nh * j) + 1] = BIM(a, j, b);
        }
    }
    free(a);
#pragma omp for schedule(static)
    for (int j = 0; j < n; j++)
      for (size_t i = 0; i < ld; i++)
        z[i + ld * j] = z[i + ld * j] / (double)n / (double)n;
  }
}

/* In-place Z2D, unnormalised (src/fft_inv.f90:12): inverse over columns,
 * then complex-to-real over rows.  The imaginary parts of the kx = 0 and
 * kx = N/2 entries are not used by a complex-to-real transform; the two
 * padding doubles of each row are left zero. */
void or_fft_inv(const or_ctx *c, double *z) {
  const int n = c->n, nh = c->nh;
  const size_t ld = (size_t)n + 2;
#pragma omp parallel
  {
    double *a = (double *)malloc(sizeof(double) * 2 * n * FFT_NB);
#pragma omp for schedule(static)
    for (int i0 = 0; i0 < nh; i0 += FFT_NB) { /* columns, FFT_NB at a time */
      const int nb = nh - i0 < FFT_NB ? nh - i0 : FFT_NB;
      if (nb < FFT_NB) memset(a, 0, sizeof(double) * 2 * n * FFT_NB);
      for (int j = 0; j < n; j++)
        for (int b = 0; b < nb; b++) {
          BRE(a, j, b) = z[2 * (i0 + b + (size_t)nh * j)];
          BIM(a, j, b) = z[2 * (i0 + b + (size_t)nh * j) + 1];
        }
      fft4(c, a, -1);
      for (int j = 0; j < n; j++)
        for (int b = 0; b < nb; b++) {
          z[2 * (i0 + b + (size_t)nh * j)] = BRE(a, j, b);
          z[2 * (i0 + b + (size_t)nh * j) + 1] = BIM(a, j, b);
        }
    }
#pragma omp for schedule(static)
    for (int j0 = 0; j0 < n; j0 += FFT_NB) { /* rows, FFT_NB at a time */
      const int nb = n - j0 < FFT_NB ? n - j0 : FFT_NB;
      if (nb < FFT_NB) memset(a, 0, sizeof(double) * 2 * n * FFT_NB);
      for (int b = 0; b < nb; b++) {
        const double *row = z + ld * (j0 + b);
        /* Hermitian extension of the half spectrum */
        BRE(a, 0, b) = row[0];
        BIM(a, 0, b) = 0.0;
        BRE(a, (n / 2), b) = row[2 * (n / 2)];
        BIM(a, (n / 2), b) = 0.0;
        for (int k = 1; k < n / 2; k++) {
          BRE(a, k, b) = row[2 * k];
          BIM(a, k, b) = row[2 * k + 1];
          BRE(a, (n - k), b) = row[2 * k];
          BIM(a, (n - k), b) = -row[2 * k + 1];
        }
      }
      fft4(c, a, -1);
      for (int b = 0; b < nb; b++) {
        double *row = z + ld * (j0 + b);
        for (int x = 0; x < n; x++) row[x] = BRE(a, x, b);
        row[n] = 0.0;
        row[n + 1] = 0.0;
      }
    }
    free(a);
  }
}

/* ------------------------------------------------------------------ */
/* Physics operators                                                    */
/* ------------------------------------------------------------------ */

/* src/produ.f90:11-23 */
void or_produ(const or_ctx *c, const double *z, double *u, double *v) {
  const int n = c->n;
  const size_t ld = (size_t)n + 2;
#pragma omp parallel for schedule(static)
  for (int j = 0; j < n; j++)
    for (size_t i = 0; i < ld; i++) {
      size_t p = i + ld * j;
      if (i < (size_t)n) {
        u[p] = z[p] * u[p];
        v[p] = z[p] * v[p];
      } else {
        u[p] = 0.0;
        v[p] = 0.0;
      }
    }
}

/* src/trunc.f90:10-25 */
void or_trunc(const or_ctx *c, double *z) {
  const int n = c->n, nh = c->nh;
#pragma omp parallel for schedule(static)
  for (int j = 0; j < n; j++)
    for (int i = 0; i < nh; i++) {
      double k1 = c->fkys[j] / c->fkysmax;
      double k2 = k1 + c->fkxs[i] / c->fkxsmax;
      if (k2 > c->rtrunc) {
        z[2 * (i + (size_t)nh * j)] = 0.0;
        z[2 * (i + (size_t)nh * j) + 1] = 0.0;
      }
    }
  z[0] = 0.0;
  z[1] = 0.0;
}

/* src/nlt.f90:1-65.  On return z holds the nonlinear term; u and v hold the
 * normalised transforms of the products, as in the reference. */
void or_nlt(const or_ctx *c, double *z, double *u, double *v) {
  const int n = c->n, nh = c->nh;
  const size_t ns = ((size_t)n + 2) * n;
  memset(u, 0, sizeof(double) * ns); /* nlt.f90:17-20 */
  memset(v, 0, sizeof(double) * ns);
#pragma omp parallel for schedule(static)
  for (int j = 0; j < n; j++) /* nlt.f90:22-36 */
    for (int i = 0; i < nh; i++) {
      if (i == 0 && j == 0) continue;
      double ka = c->ka[i + (size_t)nh * j]; /* k2**(alpv/2) */
      double kx = c->fkx[i] / ka;
      double ky = c->fky[j] / ka;
      size_t p = 2 * (i + (size_t)nh * j);
      u[p] = -ky * z[p + 1];
      u[p + 1] = +ky * z[p];
      v[p] = +kx * z[p + 1];
      v[p + 1] = -kx * z[p];
    }
  or_fft_inv(c, z); /* nlt.f90:41-43 */
  or_fft_inv(c, v);
  or_fft_inv(c, u);
  or_produ(c, z, u, v); /* nlt.f90:46 */
  or_fft_dir(c, u);     /* nlt.f90:50-51 */
  or_fft_dir(c, v);
#pragma omp parallel for schedule(static)
  for (int j = 0; j < n; j++) /* nlt.f90:54-61 */
    for (int i = 0; i < nh; i++) {
      size_t p = 2 * (i + (size_t)nh * j);
      z[p] = +c->fkx[i] * u[p + 1] + c->fky[j] * v[p + 1];
      z[p + 1] = -c->fkx[i] * u[p] - c->fky[j] * v[p];
    }
  if (c->truncate == 1) or_trunc(c, z); /* nlt.f90:63 */
}

/* src/rkstep.f90:2-76 (integrating-factor RK2, split or unsplit). */
void or_step(const or_ctx *c, double *z, double *znl, double *u, double *v) {
  const int n = c->n, nh = c->nh;
  const size_t nc = (size_t)nh * n;
  memcpy(znl, z, sizeof(double) * 2 * nc); /* rkstep.f90:11-13 */
  or_nlt(c, znl, u, v);
#pragma omp parallel for schedule(static)
  for (int j = 0; j < n; j++)
    for (int i = 0; i < nh; i++) {
      size_t q = i + (size_t)nh * j, p = 2 * q;
      double e = c->split ? c->emkx[i] * c->emky[j] : c->emk[q];
      znl[p] = e * (z[p] + c->dt2 * znl[p]); /* rkstep.f90:24 / :34 */
      znl[p + 1] = e * (z[p + 1] + c->dt2 * znl[p + 1]);
    }
  znl[0] = 0.0; /* rkstep.f90:41-43 */
  znl[1] = 0.0;
  or_nlt(c, znl, u, v);
#pragma omp parallel for schedule(static)
  for (int j = 0; j < n; j++)
    for (int i = 0; i < nh; i++) {
      size_t q = i + (size_t)nh * j, p = 2 * q;
      double e = c->split ? c->emkx[i] * c->emky[j] : c->emk[q];
      z[p] = e * (e * z[p] + c->dt * znl[p]); /* rkstep.f90:54 / :64 */
      z[p + 1] = e * (e * z[p + 1] + c->dt * znl[p + 1]);
    }
  z[0] = 0.0; /* rkstep.f90:71-73 */
  z[1] = 0.0;
}

/* src/rkstep.f90:82-131 (Lawson RK4). */
void or_step4(const or_ctx *c, double *z, double *znl, double *znew,
              double *u, double *v) {
  const size_t nc = (size_t)c->nh * c->n;
  const double *E = c->emk;
  const double dt = c->dt, dt2 = c->dt2, dt3 = c->dt3, dt6 = c->dt6;
  memcpy(znl, z, sizeof(double) * 2 * nc); /* rkstep.f90:91-94 */
  or_nlt(c, znl, u, v);
#pragma omp parallel for schedule(static)
  for (size_t q = 0; q < nc; q++)
    for (int r = 0; r < 2; r++) {
      size_t p = 2 * q + r;
      znew[p] = E[q] * E[q] * (z[p] + dt6 * znl[p]); /* :95-97 */
      znl[p] = E[q] * (z[p] + dt2 * znl[p]);         /* :100-102 */
    }
  or_nlt(c, znl, u, v);
#pragma omp parallel for schedule(static)
  for (size_t q = 0; q < nc; q++)
    for (int r = 0; r < 2; r++) {
      size_t p = 2 * q + r;
      znew[p] = znew[p] + dt3 * E[q] * znl[p]; /* :104-106 */
      znl[p] = E[q] * z[p] + dt2 * znl[p];     /* :109-111 */
    }
  or_nlt(c, znl, u, v);
#pragma omp parallel for schedule(static)
  for (size_t q = 0; q < nc; q++)
    for (int r = 0; r < 2; r++) {
      size_t p = 2 * q + r;
      znew[p] = znew[p] + dt3 * E[q] * znl[p];    /* :113-115 */
      znl[p] = E[q] * (E[q] * z[p] + dt * znl[p]); /* :118-120 */
    }
  or_nlt(c, znl, u, v);
#pragma omp parallel for schedule(static)
  for (size_t p = 0; p < 2 * nc; p++) z[p] = znew[p] + dt6 * znl[p]; /* :122-124 */
  z[0] = 0.0; /* :126-128 */
  z[1] = 0.0;
}

/* ------------------------------------------------------------------ */
/* Tangent-linear model of the step (the reference has none): the exact  */
/* linearisation of or_nlt / or_step / or_step4, used to check the      */
/* library's adjoint (a dot-product test) and pinned by finite          */
/* differences of the functions above.                                  */
/* ------------------------------------------------------------------ */

/* The spectral velocity of src/nlt.f90:22-36: u = i ky z/ka, v = -i kx z/ka. */
static void velocity(const or_ctx *c, const double *z, double *u, double *v) {
  const int n = c->n, nh = c->nh;
  const size_t ns = ((size_t)n + 2) * n;
  memset(u, 0, sizeof(double) * ns);
  memset(v, 0, sizeof(double) * ns);
#pragma omp parallel for schedule(static)
  for (int j = 0; j < n; j++)
    for (int i = 0; i < nh; i++) {
      if (i == 0 && j == 0) continue;
      double ka = c->ka[i + (size_t)nh * j];
      double kx = c->fkx[i] / ka;
      double ky = c->fky[j] / ka;
      size_t p = 2 * (i + (size_t)nh * j);
      u[p] = -ky * z[p + 1];
      u[p + 1] = +ky * z[p];
      v[p] = +kx * z[p + 1];
      v[p + 1] = -kx * z[p];
    }
}

/* out = N'(z) dz = -ik.(dq u + q du)^, the derivative of or_nlt at z. */
void or_nlt_tl(const or_ctx *c, const double *z, const double *dz, double *out) {
  const int n = c->n, nh = c->nh;
  const size_t ns = ((size_t)n + 2) * n;
  double *q = malloc(sizeof(double) * ns), *u = malloc(sizeof(double) * ns);
  double *v = malloc(sizeof(double) * ns), *dq = malloc(sizeof(double) * ns);
  double *du = malloc(sizeof(double) * ns), *dv = malloc(sizeof(double) * ns);
  memcpy(q, z, sizeof(double) * ns);
  memcpy(dq, dz, sizeof(double) * ns);
  velocity(c, z, u, v);
  velocity(c, dz, du, dv);
  double *f[6] = {q, u, v, dq, du, dv};
  for (int k = 0; k < 6; k++) or_fft_inv(c, f[k]);
#pragma omp parallel for schedule(static)
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n + 2; i++) {
      size_t p = (size_t)j * (n + 2) + i;
      double a = 0.0, b = 0.0;
      if (i < n) {
        a = dq[p] * u[p] + q[p] * du[p];
        b = dq[p] * v[p] + q[p] * dv[p];
      }
      u[p] = a;
      v[p] = b;
    }
  or_fft_dir(c, u);
  or_fft_dir(c, v);
#pragma omp parallel for schedule(static)
  for (int j = 0; j < n; j++)
    for (int i = 0; i < nh; i++) {
      size_t p = 2 * (i + (size_t)nh * j);
      out[p] = +c->fkx[i] * u[p + 1] + c->fky[j] * v[p + 1];
      out[p + 1] = -c->fkx[i] * u[p] - c->fky[j] * v[p];
    }
  if (c->truncate == 1) or_trunc(c, out);
  for (int k = 0; k < 6; k++) free(f[k]);
}

/* N(s) into k (s kept) and N'(s) ds into dk. */
static void nlt_pair(const or_ctx *c, const double *s, const double *ds, double *k,
                     double *dk, double *u, double *v) {
  const size_t ns = ((size_t)c->n + 2) * c->n;
  memcpy(k, s, sizeof(double) * ns);
  or_nlt(c, k, u, v);
  or_nlt_tl(c, s, ds, dk);
}

/* One or_step4 of z together with its tangent dz (z as or_step4 leaves it). */
void or_step4_tl(const or_ctx *c, double *z, double *dz) {
  const size_t nc = (size_t)c->nh * c->n, ns = 2 * nc;
  const double *E = c->emk;
  const double dt = c->dt, dt2 = c->dt2, dt3 = c->dt3, dt6 = c->dt6;
  double *s = malloc(sizeof(double) * ns), *ds = malloc(sizeof(double) * ns);
  double *k = malloc(sizeof(double) * ns), *dk = malloc(sizeof(double) * ns);
  double *zn = malloc(sizeof(double) * ns), *dzn = malloc(sizeof(double) * ns);
  double *u = malloc(sizeof(double) * ns), *v = malloc(sizeof(double) * ns);
  nlt_pair(c, z, dz, k, dk, u, v);
  for (size_t q = 0; q < nc; q++)
    for (int r = 0; r < 2; r++) {
      size_t p = 2 * q + r;
      zn[p] = E[q] * E[q] * (z[p] + dt6 * k[p]);
      dzn[p] = E[q] * E[q] * (dz[p] + dt6 * dk[p]);
      s[p] = E[q] * (z[p] + dt2 * k[p]);
      ds[p] = E[q] * (dz[p] + dt2 * dk[p]);
    }
  nlt_pair(c, s, ds, k, dk, u, v);
  for (size_t q = 0; q < nc; q++)
    for (int r = 0; r < 2; r++) {
      size_t p = 2 * q + r;
      zn[p] = zn[p] + dt3 * E[q] * k[p];
      dzn[p] = dzn[p] + dt3 * E[q] * dk[p];
      s[p] = E[q] * z[p] + dt2 * k[p];
      ds[p] = E[q] * dz[p] + dt2 * dk[p];
    }
  nlt_pair(c, s, ds, k, dk, u, v);
  for (size_t q = 0; q < nc; q++)
    for (int r = 0; r < 2; r++) {
      size_t p = 2 * q + r;
      zn[p] = zn[p] + dt3 * E[q] * k[p];
      dzn[p] = dzn[p] + dt3 * E[q] * dk[p];
      s[p] = E[q] * (E[q] * z[p] + dt * k[p]);
      ds[p] = E[q] * (E[q] * dz[p] + dt * dk[p]);
    }
  nlt_pair(c, s, ds, k, dk, u, v);
  for (size_t p = 0; p < ns; p++) {
    z[p] = zn[p] + dt6 * k[p];
    dz[p] = dzn[p] + dt6 * dk[p];
  }
  z[0] = z[1] = dz[0] = dz[1] = 0.0;
  free(s); free(ds); free(k); free(dk); free(zn); free(dzn); free(u); free(v);
}

/* One or_step of z together with its tangent dz. */
void or_step_tl(const or_ctx *c, double *z, double *dz) {
  const int n = c->n, nh = c->nh;
  const size_t nc = (size_t)nh * n, ns = 2 * nc;
  double *s = malloc(sizeof(double) * ns), *ds = malloc(sizeof(double) * ns);
  double *k = malloc(sizeof(double) * ns), *dk = malloc(sizeof(double) * ns);
  double *u = malloc(sizeof(double) * ns), *v = malloc(sizeof(double) * ns);
  nlt_pair(c, z, dz, k, dk, u, v);
  for (int j = 0; j < n; j++)
    for (int i = 0; i < nh; i++) {
      size_t q = i + (size_t)nh * j, p = 2 * q;
      double e = c->split ? c->emkx[i] * c->emky[j] : c->emk[q];
      for (int r = 0; r < 2; r++) {
        s[p + r] = e * (z[p + r] + c->dt2 * k[p + r]);
        ds[p + r] = e * (dz[p + r] + c->dt2 * dk[p + r]);
      }
    }
  s[0] = s[1] = ds[0] = ds[1] = 0.0;
  nlt_pair(c, s, ds, k, dk, u, v);
  for (int j = 0; j < n; j++)
    for (int i = 0; i < nh; i++) {
      size_t q = i + (size_t)nh * j, p = 2 * q;
      double e = c->split ? c->emkx[i] * c->emky[j] : c->emk[q];
      for (int r = 0; r < 2; r++) {
        z[p + r] = e * (e * z[p + r] + c->dt * k[p + r]);
        dz[p + r] = e * (e * dz[p + r] + c->dt * dk[p + r]);
      }
    }
  z[0] = z[1] = dz[0] = dz[1] = 0.0;
  free(s); free(ds); free(k); free(dk); free(u); free(v);
}

/* ------------------------------------------------------------------ */
/* Forcing: src/forcing.f90                                             */
/* ------------------------------------------------------------------ */

/* src/forcing.f90:40-51.  The product 69069*irand+1 wraps in 32 bits, the
 * mask keeps 31 bits, and the quotient is returned as a default REAL. */
float or_rann(int32_t *seed) {
  uint32_t s = (uint32_t)(*seed);
  s = (69069u * s + 1u) & 0x7FFFFFFFu;
  *seed = (int32_t)s;
  return (float)((double)s / (double)0x7FFFFFFF);
}

/* te = pi2*rann with the truncated literal of src/forcing.f90:10.  cos and
 * sin are taken in extended precision and rounded once (the correctly
 * rounded values), which makes the phase table independent of how the
 * double-precision libm is reached: the reference evaluates them with the
 * device cos/sin of its compiler, so its own phases may differ from these
 * by one unit in the last place. */
static void phase(int32_t *seed, double *cc, double *cs) {
  const double pi2 = 2.0 * 3.14159265358979;
  double te = pi2 * (double)or_rann(seed);
  *cc = (double)cosl((long double)te);
  *cs = (double)sinl((long double)te);
}

void or_forcing_phases(const or_ctx *c, int32_t *seed, int nsteps,
                       double *cs) {
  for (int s = 0; s < nsteps; s++)
    for (int k = 0; k < c->nf; k++) {
      size_t p = 2 * ((size_t)s * c->nf + k);
      phase(seed, &cs[p], &cs[p + 1]);
    }
}

/* src/forcing.f90:16-33, serial over the forced modes in table order. */
void or_forcing(const or_ctx *c, double *z, int32_t *seed) {
  const int nh = c->nh;
  for (int k = 0; k < c->nf; k++) {
    int i = c->ikxf[k], j = c->ikyf[k]; /* 1-based */
    double cc, cs;
    phase(seed, &cc, &cs);
    size_t p = 2 * ((size_t)(i - 1) + (size_t)nh * (j - 1));
    z[p] = z[p] + c->sdt * c->famp * c->ff[2 * k] * cc;
    z[p + 1] = z[p + 1] + c->sdt * c->famp * c->ff[2 * k + 1] * cs;
    if (i == 1 || i == nh) {
      size_t q = 2 * ((size_t)(i - 1) + (size_t)nh * (c->jc[j - 1] - 1));
      z[q] = z[q] + c->sdt * c->famp * c->ff[2 * k] * cc;
      z[q + 1] = z[q + 1] - c->sdt * c->famp * c->ff[2 * k + 1] * cs;
    }
  }
}

/* ------------------------------------------------------------------ */
/* Diagnostics: src/diagnostics.f90 (serial, so free of the reference's  */
/* unprotected GPU accumulations)                                        */
/* ------------------------------------------------------------------ */

static int shell_of(const or_ctx *c, int i, int j) {
  double k2 = c->fkxs[i] + c->fkys[j];
  double k = sqrt(k2);
  return (int)(k / c->dkx) + 1; /* diagnostics.f90:34 (dk = dkx) */
}

void or_shell_index(const or_ctx *c, int32_t *ib) {
  const int n = c->n, nh = c->nh;
  for (int j = 0; j < n; j++)
    for (int i = 0; i < nh; i++)
      ib[i + (size_t)nh * j] = (i == 0 && j == 0) ? 0 : shell_of(c, i, j);
}

/* src/diagnostics.f90:5-56; sp and sz have NBIN = N/2 entries. */
void or_spectrum(const or_ctx *c, const double *z, double *sp, double *sz) {
  const int n = c->n, nh = c->nh, nbin = n / 2;
  const double dk = c->dkx;
  for (int b = 0; b < nbin; b++) sp[b] = sz[b] = 0.0;
  for (int j = 0; j < n; j++)
    for (int i = 0; i < nh; i++) {
      if (i == 0 && j == 0) continue;
      double fac = 1.0;
      if (i == 0 || i == nh - 1) fac = 0.5;
      double k2 = c->fkxs[i] + c->fkys[j];
      double k = sqrt(k2);
      double ka = pow(k2, c->alpv / 2.0);
      int ib = (int)(k / dk) + 1;
      if (ib < nbin) {
        size_t p = 2 * (i + (size_t)nh * j);
        double a = hypot(z[p], z[p + 1]);
        double a2 = a * a;
        sp[ib - 1] = sp[ib - 1] + fac * a2 / ka;
        sz[ib - 1] = sz[ib - 1] + fac * a2;
      }
    }
  for (int b = 0; b < nbin; b++) { /* diagnostics.f90:42-45 */
    sp[b] = sp[b] / dk;
    sz[b] = sz[b] / dk;
  }
}

/* src/diagnostics.f90:62-119 */
void or_fluxes(const or_ctx *c, const double *z, const double *znl,
               double *fle, double *flz) {
  const int n = c->n, nh = c->nh, nbin = n / 2;
  const double dk = c->dkx;
  double *flem = (double *)calloc(nbin, sizeof(double));
  double *flzm = (double *)calloc(nbin, sizeof(double));
  for (int j = 0; j < n; j++)
    for (int i = 0; i < nh; i++) {
      if (i == 0 && j == 0) continue;
      double fac = 2.0;
      if (i == 0 || i == nh - 1) fac = 1.0;
      double k2 = c->fkxs[i] + c->fkys[j];
      double k = sqrt(k2);
      size_t p = 2 * (i + (size_t)nh * j);
      double fz = z[p] * znl[p] + z[p + 1] * znl[p + 1];
      double ka = pow(k2, c->alpv / 2.0);
      double fe = fz / ka;
      int ib = (int)(k / dk) + 1;
      if (ib < nbin) {
        flzm[ib - 1] = flzm[ib - 1] + fac * fz;
        flem[ib - 1] = flem[ib - 1] + fac * fe;
      }
    }
  for (int b = 0; b < nbin; b++) { /* diagnostics.f90:104-109 */
    double sz = 0.0, se = 0.0;
    for (int q = b; q < nbin; q++) {
      sz += flzm[q];
      se += flem[q];
    }
    flz[b] = sz;
    fle[b] = se;
  }
  free(flem);
  free(flzm);
}

/* src/diagnostics.f90:125-179 */
void or_glob(const or_ctx *c, const double *z, double *out) {
  const int n = c->n, nh = c->nh;
  double ee = 0, zz = 0, dise_nu = 0, dise_mu = 0, disz_nu = 0, disz_mu = 0;
  for (int j = 0; j < n; j++)
    for (int i = 0; i < nh; i++) {
      double fac = 1.0;
      if (i == 0 || i == nh - 1) fac = 0.5;
      if (i == 0 && j == 0) continue;
      double k2 = c->fkxs[i] + c->fkys[j];
      double ka = pow(k2, c->alpv / 2.0);
      size_t p = 2 * (i + (size_t)nh * j);
      double a = hypot(z[p], z[p + 1]);
      double scra = a * a;
      ee = ee + fac * scra / ka;
      zz = zz + fac * scra;
      disz_nu = disz_nu + 2.0 * fac * c->nu * pow(k2, c->alpnu) * scra;
      disz_mu = disz_mu + 2.0 * fac * c->mu * pow(k2, c->alpmu) * scra;
      dise_nu = dise_nu +
                2.0 * fac * c->nu * pow(k2, c->alpnu - c->alpv / 2.0) * scra;
      dise_mu = dise_mu +
                2.0 * fac * c->mu * pow(k2, c->alpmu - c->alpv / 2.0) * scra;
    }
  out[0] = ee;
  out[1] = zz;
  out[2] = disz_nu;
  out[3] = disz_mu;
  out[4] = dise_nu;
  out[5] = dise_mu;
}

/* ------------------------------------------------------------------ */
/* Driver: src/main.f90:89-143 without file output                      */
/* ------------------------------------------------------------------ */

void or_run(const or_ctx *c, double *z, int32_t *seed, int npas, double *t,
            double *diag) {
  const size_t ns = ((size_t)c->n + 2) * c->n;
  const int nbin = c->n / 2;
  double *znl = (double *)calloc(ns, sizeof(double));
  double *znew = (double *)calloc(ns, sizeof(double));
  double *u = (double *)calloc(ns, sizeof(double));
  double *v = (double *)calloc(ns, sizeof(double));
  or_fft_dir(c, z); /* main.f90:89-92 */
  if (c->truncate == 1) or_trunc(c, z);
  size_t nd = 0;
  for (int ipas = 1; ipas <= npas; ipas++) {
    *t = *t + c->dt; /* main.f90:98 */
    if (c->rko == 4)
      or_step4(c, z, znl, znew, u, v);
    else
      or_step(c, z, znl, u, v);
    if (c->imix > 0 && ipas % c->imix == 0 && diag) { /* main.f90:106-117 */
      memcpy(znl, z, sizeof(double) * ns);
      or_nlt(c, znl, u, v);
      double *d = diag + nd * (4 * (size_t)nbin + 6);
      or_spectrum(c, z, d, d + nbin);
      or_fluxes(c, z, znl, d + 2 * nbin, d + 3 * nbin);
      or_glob(c, z, d + 4 * nbin);
      nd++;
    }
    or_forcing(c, z, seed); /* main.f90:136 */
  }
  or_fft_inv(c, z); /* main.f90:143 */
  free(znl);
  free(znew);
  free(u);
  free(v);
}
