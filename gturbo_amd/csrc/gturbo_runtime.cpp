// gturbo_runtime.cpp -- host side of libgturbo_hip.so: contexts, the present
// table of device copies, host tables and the C entry points of gturbo.h.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "gt_runtime.hpp"

namespace gtr {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

}  // namespace gtr

namespace gtr {

int ilog2(int n) {
  int l = 0;
  while ((1 << l) < n) ++l;
  return l;
}

// Largest axis index kept by the mask of src/trunc.f90:13-15, evaluated with
// the reference's own floating-point expression.
int kept_axis(const gt_ctx* c, bool xaxis) {
  int m = 0;
  for (int i = 0; i <= c->n / 2; ++i) {
    double f = (double)i * (xaxis ? c->dkx : c->dky);
    double v = xaxis ? (f * f) / c->fkxsmax : (f * f) / c->fkysmax;
    if (!(v > c->rtrunc)) m = i;
  }
  return m;
}

int rect_kind(const gt_ctx* c, const Rect& r) {
  return (r.kx == c->rtr.kx && r.p == c->rtr.p && r.q == c->rtr.q) ? 0 : 1;
}

int kyp_of(const Rect& r) { return (r.ky() + 7) & ~7; }

hipEvent_t get_event(gt_ctx* c) {
  if (!c->ev_pool.empty()) {
    hipEvent_t e = c->ev_pool.back();
    c->ev_pool.pop_back();
    return e;
  }
  // Timing only: without the system-scope fence a record does not write
  // back and invalidate the caches between the kernels it separates (which
  // would slow the kernels being timed).  gt_timing_read synchronizes the
  // stream before it reads the events.
  hipEvent_t e = nullptr;
  if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return nullptr;
  return e;
}

Timer::Timer(gt_ctx* c_, int cls_) : c(c_), cls(cls_) {
  if (c->timing && (c->timing_mask >> cls & 1u)) {
    a = get_event(c);
    if (a) (void)hipEventRecord(a, c->stream);
  }
}

Timer::~Timer() {
  if (a) {
    hipEvent_t b = get_event(c);
    if (b) {
      (void)hipEventRecord(b, c->stream);
      c->tl.push_back({cls, a, b});
    }
  }
}

// ---- host tables ----------------------------------------------------------

// src/defk.f90:1-75
void defk(gt_ctx* c) {
  const int n = c->n, nh = c->nh;
  c->jc.assign(n, 0);
  c->jc[0] = 1;
  for (int j = 2; j <= n; ++j) c->jc[j - 1] = n + 2 - j;
  c->dkx = 2.0 * 3.14159265358979323846 / c->xlx;
  c->dky = 2.0 * 3.14159265358979323846 / c->xly;
  c->fkx.assign(nh, 0.0);
  c->fkxs.assign(nh, 0.0);
  c->fky.assign(n, 0.0);
  c->fkys.assign(n, 0.0);
  for (int i = 0; i < nh; ++i) c->fkx[i] = (double)i * c->dkx;
  for (int j = 0; j < n; ++j) c->fky[j] = (double)gt::kyi_of(j, n) * c->dky;
  for (int i = 0; i < nh; ++i) c->fkxs[i] = c->fkx[i] * c->fkx[i];
  for (int j = 0; j < n; ++j) c->fkys[j] = c->fky[j] * c->fky[j];
  c->emk.assign((size_t)nh * n, 0.0);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < nh; ++i) {
      if (i == 0 && j == 0) continue;
      double fk2 = c->fkxs[i] + c->fkys[j];
      double scra = c->nu * std::pow(fk2, c->alpnu) + c->mu * std::pow(fk2, c->alpmu);
      c->emk[i + (size_t)nh * j] = std::exp(-0.5 * c->dt * scra);
    }
  c->emky.assign(n, 0.0);
  for (int j = 0; j < n; ++j) {
    double scra = c->nu * std::pow(c->fkys[j], c->alpnu) + c->mu;
    c->emky[j] = std::exp(-0.5 * c->dt * scra);
  }
  c->emkx.assign(nh, 0.0);
  for (int i = 0; i < nh; ++i) {
    double scra = c->nu * std::pow(c->fkxs[i], c->alpnu);
    c->emkx[i] = std::exp(-0.5 * c->dt * scra);
  }
  c->fkxsmax = c->fkxs[nh - 1];
  c->fkysmax = c->fkys[n / 2];
  c->rtrunc = c->rtrunc * c->rtrunc;
}

// src/iniforcing.f90:1-80 (mode order matters: one LCG draw per mode)
int iniforcing(gt_ctx* c) {
  c->nf = 0;
  c->inpz = 0.0;
  c->inpe = 0.0;
  c->ikxf.clear();
  c->ikyf.clear();
  c->ff.clear();
  auto cand = [&](int i, int j) -> int {
    double k2 = c->fkxs[i - 1] + c->fkys[j - 1];
    double k2n = c->fkxs[i - 1] / c->fkxsmax + c->fkys[j - 1] / c->fkysmax;
    if (k2n <= c->rtrunc && k2 <= c->k2f * c->k2f && k2 > c->k1f * c->k1f) {
      if (++c->nf > kNFORC) return fail(GT_ERR_NFORC, "NFORC too small !");
      c->ikxf.push_back(i);
      c->ikyf.push_back(j);
      c->ff.push_back(1.0);
      c->ff.push_back(1.0);
      c->inpz = c->inpz + c->famp * c->famp;
      c->inpe = c->inpe + c->famp * c->famp / std::pow(k2, c->alpv / 2.0);
    }
    return GT_OK;
  };
  const int n = c->n;
  for (int j = 2; j <= n / 2 + 1; ++j) CHK(cand(1, j));
  for (int i = 2; i <= n / 2; ++i)
    for (int j = 1; j <= n; ++j) CHK(cand(i, j));
  for (int j = 1; j <= n / 2 + 1; ++j) CHK(cand(c->nh, j));
  return GT_OK;
}

int upload_twiddles(gt_ctx* c) {
  const int n = c->n;
  std::vector<double2> w(n);
  const long double twopi = 6.283185307179586476925286766559005768L;
  for (int m = 0; m < n; ++m) {
    long double ang = twopi * (long double)m / (long double)n;
    w[m] = make_double2((double)cosl(ang), -(double)sinl(ang));
  }
  HIPCHK(hipMalloc(&c->tw, sizeof(double2) * n));
  HIPCHK(hipMemcpy(c->tw, w.data(), sizeof(double2) * n, hipMemcpyHostToDevice));
  return GT_OK;
}

// Integrating-factor table and forced-mode table for one rectangle.
int rect_tables(gt_ctx* c, const Rect& r, RectTables** out) {
  RectTables& t = c->tab[rect_kind(c, r)];
  *out = &t;
  if (t.emk) return GT_OK;
  t.rc = r;
  t.kyp = kyp_of(r);
  const int n = c->n, nh = c->nh;
  std::vector<double> e((size_t)r.kc() * t.kyp, 0.0);
  for (int col = 0; col < r.kc(); ++col)
    for (int row = 0; row < r.ky(); ++row) {
      const int p = gt::fft_of(row, r, n);
      // split: emkx(i)*emky(j) is the factor of src/rkstep.f90:24,54
      e[(size_t)col * t.kyp + row] =
          c->split ? c->emkx[col] * c->emky[p] : c->emk[col + (size_t)nh * p];
    }
  HIPCHK(hipMalloc(&t.emk, sizeof(double) * e.size()));
  HIPCHK(hipMemcpy(t.emk, e.data(), sizeof(double) * e.size(), hipMemcpyHostToDevice));
  std::vector<ForcedMode> fm(c->nf > 0 ? c->nf : 1);
  for (int k = 0; k < c->nf; ++k) {
    const int col = c->ikxf[k] - 1, p = c->ikyf[k] - 1;
    const int row = gt::row_of(p, r, n);
    if (col > r.kx || row < 0)
      return fail(GT_ERR_ARG, "forced mode (%d,%d) outside the kept rectangle", col, p);
    ForcedMode f;
    f.idx = col * t.kyp + row;
    f.pidx = -1;
    f.ff1 = c->ff[2 * k];
    f.ff2 = c->ff[2 * k + 1];
    if (col == 0 || col == nh - 1) {  // forcing.f90:28-31
      const int pp = c->jc[p] - 1, prow = gt::row_of(pp, r, n);
      if (prow < 0) return fail(GT_ERR_ARG, "forced partner outside the rectangle");
      f.pidx = col * t.kyp + prow;
    }
    fm[k] = f;
  }
  HIPCHK(hipMalloc(&t.modes, sizeof(ForcedMode) * fm.size()));
  HIPCHK(hipMemcpy(t.modes, fm.data(), sizeof(ForcedMode) * fm.size(),
                   hipMemcpyHostToDevice));
  // The same updates by column for the fused step boundary, each column's in
  // table order (a mode, then its partner).
  std::vector<std::vector<gt::FEntry>> bycol(r.kc());
  for (int k = 0; k < c->nf; ++k) {
    const int col = c->ikxf[k] - 1;
    const ForcedMode& f = fm[k];
    bycol[col].push_back(gt::FEntry{f.idx - col * t.kyp, k, 0, f.ff1, f.ff2});
    if (f.pidx >= 0) bycol[col].push_back(gt::FEntry{f.pidx - col * t.kyp, k, 1, f.ff1, f.ff2});
  }
  std::vector<int> fcs(r.kc() + 1, 0);
  std::vector<gt::FEntry> fce;
  for (int col = 0; col < r.kc(); ++col) {
    fcs[col] = (int)fce.size();
    fce.insert(fce.end(), bycol[col].begin(), bycol[col].end());
  }
  fcs[r.kc()] = (int)fce.size();
  if (fce.empty()) fce.push_back(gt::FEntry{-1, 0, 0, 0.0, 0.0});
  HIPCHK(hipMalloc(&t.fcs, sizeof(int) * fcs.size()));
  HIPCHK(hipMemcpy(t.fcs, fcs.data(), sizeof(int) * fcs.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMalloc(&t.fce, sizeof(gt::FEntry) * fce.size()));
  HIPCHK(hipMemcpy(t.fce, fce.data(), sizeof(gt::FEntry) * fce.size(), hipMemcpyHostToDevice));
  return GT_OK;
}

int ensure_ybuf(gt_ctx* c, const Rect& r) {
  const int bt = gt::bt_of(c->logn);
  const size_t ntiles = (size_t)(r.kc() + bt - 1) / bt;
  // columns of a field: the tiles' (y1, y2), or the y1 blocks' at N = 4096
  const size_t cols = std::max(ntiles * bt, (size_t)gt::y1_cols(r.kc()));
  const size_t need = cols * (size_t)c->n;
  if (need <= c->ycap) return GT_OK;
  if (c->y1) (void)hipFree(c->y1);
  if (c->y2) (void)hipFree(c->y2);
  c->y1 = c->y2 = nullptr;
  c->ycap = 0;
  HIPCHK(hipMalloc(&c->y1, sizeof(double2) * need * 3));
  HIPCHK(hipMalloc(&c->y2, sizeof(double2) * need * 2));
  HIPCHK(hipMemsetAsync(c->y1, 0, sizeof(double2) * need * 3, c->stream));
  HIPCHK(hipMemsetAsync(c->y2, 0, sizeof(double2) * need * 2, c->stream));
  c->ycap = need;
  return GT_OK;
}

// ---- present table --------------------------------------------------------

int lookup(gt_ctx* c, const void* host, Mirror** m) {
  auto it = c->present.find(host);
  if (it == c->present.end())
    return fail(GT_ERR_NOT_PRESENT, "array %p is not present on the device", host);
  const size_t need = ((size_t)c->n + 2) * c->n * sizeof(double);
  if (it->second.bytes < need)
    return fail(GT_ERR_ARG, "array %p holds %zu bytes, a field needs %zu", host,
                it->second.bytes, need);
  *m = &it->second;
  return GT_OK;
}

int alloc_cmp(gt_ctx*, Mirror& m, const Rect& r) {
  const int kyp = kyp_of(r);
  const size_t need = (size_t)r.kc() * kyp;
  if (need > m.cmp_cap) {
    if (m.dcmp) (void)hipFree(m.dcmp);
    m.dcmp = nullptr;
    m.cmp_cap = 0;
    HIPCHK(hipMalloc(&m.dcmp, sizeof(double2) * need));
    m.cmp_cap = need;
  }
  m.rect = r;
  m.kyp = kyp;
  return GT_OK;
}

int ensure_ref(gt_ctx* c, Mirror& m) {
  if (m.ref_valid) return GT_OK;
  if (!m.cmp_valid) return fail(GT_ERR_ARG, "array %p has no valid device data", m.host);
  Timer t(c, GT_K_OTHER);
  HIPCHK(gt::launch_scatter(c->ph, m.rect, m.kyp, m.dcmp, m.dref, c->stream));
  m.ref_valid = true;
  return GT_OK;
}

// Compact copy of a spectral field.  With the 2/3 rule on, the truncated
// rectangle is used unless the field has nonzero modes outside it.
int ensure_cmp(gt_ctx* c, Mirror& m) {
  if (m.cmp_valid) return GT_OK;
  Rect r = c->rtr;
  if (c->truncate == 1 && (r.kx != c->rfull.kx || r.p != c->rfull.p || r.q != c->rfull.q)) {
    HIPCHK(hipMemsetAsync(c->flag, 0, sizeof(int), c->stream));
    CHK(alloc_cmp(c, m, r));
    {
      Timer t(c, GT_K_OTHER);
      HIPCHK(gt::launch_gather(c->ph, r, m.kyp, m.dref, m.dcmp, 0, c->flag, c->stream));
    }
    int flag = 0;
    HIPCHK(hipMemcpyAsync(&flag, c->flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (!flag) {
      m.cmp_valid = true;
      return GT_OK;
    }
  }
  r = c->rfull;
  CHK(alloc_cmp(c, m, r));
  Timer t(c, GT_K_OTHER);
  HIPCHK(gt::launch_gather(c->ph, r, m.kyp, m.dref, m.dcmp, 0, nullptr, c->stream));
  m.cmp_valid = true;
  return GT_OK;
}

void set_cmp_only(Mirror& m) {
  m.cmp_valid = true;
  m.ref_valid = false;
}
void set_ref_only(Mirror& m) {
  m.ref_valid = true;
  m.cmp_valid = false;
}

gt::PassArgs pass_args(gt_ctx* c, Mirror& z, RectTables* t) {
  gt::PassArgs a{};
  a.rc = z.rect;
  a.kyp = z.kyp;
  const int bt = gt::bt_of(c->logn);
  a.ntiles = (z.rect.kc() + bt - 1) / bt;
  a.ph = c->ph;
  a.tw = c->tw;
  a.y1 = c->y1;
  a.y2 = c->y2;
  a.z = z.dcmp;
  a.emk = t ? t->emk : nullptr;
  return a;
}

int c_pass(gt_ctx* c, int mode, const gt::PassArgs& a) {
  const int cls = mode == gt::C_START ? GT_K_YSTART
                : (mode == gt::C_RK4_4 || mode == gt::C_RK2_2 || mode == gt::C_NLT)
                    ? GT_K_YFINAL
                : (mode == gt::C_RK4_4S || mode == gt::C_RK2_2S) ? GT_K_YBOUND : GT_K_YSTAGE;
  Timer t(c, cls);
  HIPCHK(gt::launch_c_pass(c->logn, mode, a, c->stream));
  return GT_OK;
}

int r_pass(gt_ctx* c, const gt::PassArgs& a) {
  Timer t(c, GT_K_XPASS);
  HIPCHK(gt::launch_r_pass(c->logn, a, c->stream));
  return GT_OK;
}

// Forcing phases of one call: nf draws of rann (src/forcing.f90:40-51),
// te = pi2*rann with pi2 = 2*3.14159265358979 (src/forcing.f90:10), and
// (cos te, sin te) evaluated in extended precision and rounded once: the
// correctly rounded values, identical whichever compiler built the caller.
void forcing_phases(int nf, int32_t* seed, double2* out) {
  const double pi2 = 2.0 * 3.14159265358979;
  uint32_t s = (uint32_t)*seed;
  for (int k = 0; k < nf; ++k) {
    s = (69069u * s + 1u) & 0x7FFFFFFFu;
    const float r = (float)((double)s / (double)0x7FFFFFFF);
    const double te = pi2 * (double)r;
    out[k] = make_double2((double)cosl((long double)te), (double)sinl((long double)te));
  }
  *seed = (int32_t)s;
}

int ensure_diag(gt_ctx* c, const Rect& r) {
  const size_t nbin = c->n / 2;
  const size_t need = 2 * (size_t)(r.kc() + 1) * nbin + 2 * nbin + (size_t)r.kc() * 6 + 6;
  if (need <= c->dg_cap) return GT_OK;
  if (c->dg) (void)hipFree(c->dg);
  c->dg = nullptr;
  c->dg_cap = 0;
  HIPCHK(hipMalloc(&c->dg, sizeof(double) * need));
  c->dg_cap = need;
  return GT_OK;
}

gt::DiagArgs diag_args(gt_ctx* c, const Mirror& z) {
  gt::DiagArgs d{};
  d.rc = z.rect;
  d.kyp = z.kyp;
  d.ph = c->ph;
  d.z = z.dcmp;
  d.zcs = z.kyp;
  d.zrs = 1;
  const size_t nbin = c->n / 2, kc = z.rect.kc();
  d.part1 = c->dg;
  d.part2 = d.part1 + (kc + 1) * nbin;
  d.out1 = d.part2 + (kc + 1) * nbin;
  d.out2 = d.out1 + nbin;
  d.gpart = d.out2 + nbin;
  d.gout = d.gpart + kc * 6;
  return d;
}

// Bring two compact fields onto the same rectangle.
int same_rect(gt_ctx* c, Mirror& a, Mirror& b) {
  if (a.rect.kx == b.rect.kx && a.rect.p == b.rect.p && a.rect.q == b.rect.q) return GT_OK;
  for (Mirror* m : {&a, &b}) {
    if (rect_kind(c, m->rect) == 0) {
      CHK(ensure_ref(c, *m));
      m->cmp_valid = false;
      CHK(alloc_cmp(c, *m, c->rfull));
      Timer t(c, GT_K_OTHER);
      HIPCHK(gt::launch_gather(c->ph, c->rfull, m->kyp, m->dref, m->dcmp, 0, nullptr,
                               c->stream));
      m->cmp_valid = true;
    }
  }
  return GT_OK;
}

// src/forcing.f90 on a compact field: the phases of this call (host LCG)
// go through a ring of pinned slots to the device.
// The next forcing draws (src/forcing.f90:16-24) into a free slot of the
// phase ring, copied to the device on the stream; the slot is marked used
// once its reader is launched (slot_read).
int forcing_upload(gt_ctx* c, int32_t* seed, unsigned* slot_out) {
  const unsigned slot = c->ph_next++ % kRing;
  if (c->ph_ev_used[slot]) HIPCHK(hipEventSynchronize(c->ph_ev[slot]));
  double2* hp = c->ph_host + (size_t)slot * c->nf;
  double2* dp = c->ph_dev + (size_t)slot * c->nf;
  forcing_phases(c->nf, seed, hp);
  HIPCHK(hipMemcpyAsync(dp, hp, sizeof(double2) * c->nf, hipMemcpyHostToDevice, c->stream));
  *slot_out = slot;
  return GT_OK;
}
int slot_read(gt_ctx* c, unsigned slot) {
  HIPCHK(hipEventRecord(c->ph_ev[slot], c->stream));
  c->ph_ev_used[slot] = true;
  return GT_OK;
}
int forcing_slot(gt_ctx* c, RectTables* t, double2* z, unsigned slot) {
  double2* dp = c->ph_dev + (size_t)slot * c->nf;
  {
    Timer tt(c, GT_K_FORCING);
    HIPCHK(gt::launch_forcing(c->ph, c->nf, t->modes, dp, c->sdt * c->famp, z, c->stream));
  }
  return slot_read(c, slot);
}
int forcing_on(gt_ctx* c, RectTables* t, double2* z, int32_t* seed) {
  if (c->nf == 0) return GT_OK;
  unsigned slot;
  CHK(forcing_upload(c, seed, &slot));
  return forcing_slot(c, t, z, slot);
}

// ---- the deferred end of a step ----------------------------------------------

int flush_pending(gt_ctx* c) {
  Pending& p = c->pend;
  if (p.kind == PEND_NONE) return GT_OK;
  const Pending q = p;
  p = Pending{};
  CHK(c_pass(c, q.kind == PEND_RK4 ? gt::C_RK4_4 : gt::C_RK2_2, q.a));
  if (q.forced) CHK(forcing_slot(c, q.t, q.a.z, q.slot));
  return GT_OK;
}

// Whether the deferred end can be fused into a step of the same kind on
// the same arrays (state z, and for RK4 the accumulator of the deferred
// step, which its last stage reads).
static bool pending_fuses(const gt_ctx* c, int kind, const Mirror* m, const Mirror* mn) {
  const Pending& p = c->pend;
  return p.kind == kind && p.z == m && (kind == PEND_RK2 || p.znew == mn) &&
         p.a.rc.kx == m->rect.kx && p.a.rc.p == m->rect.p && p.a.rc.q == m->rect.q &&
         p.a.z == m->dcmp && (kind == PEND_RK2 || p.a.znew == mn->dcmp);
}

// First pass of a step: the fused boundary of the deferred step when it
// fuses (its last stage, its forcing, this step's inverse part), else the
// inverse part alone.
static int step_start(gt_ctx* c, int kind, Mirror* m, Mirror* mn, const gt::PassArgs& a) {
  if (pending_fuses(c, kind, m, mn)) {
    const Pending q = c->pend;
    c->pend = Pending{};
    gt::PassArgs b = q.a;
    if (q.forced) {
      b.fcs = q.t->fcs;
      b.fce = q.t->fce;
      b.fph = c->ph_dev + (size_t)q.slot * c->nf;
      b.famp = c->sdt * c->famp;
    }
    CHK(c_pass(c, kind == PEND_RK4 ? gt::C_RK4_4S : gt::C_RK2_2S, b));
    if (q.forced) CHK(slot_read(c, q.slot));
    return GT_OK;
  }
  CHK(flush_pending(c));
  return c_pass(c, gt::C_START, a);
}

}  // namespace gtr

// ===========================================================================
// C ABI
// ===========================================================================

using namespace gtr;

extern "C" {

const char* gt_last_error(void) { return g_err.c_str(); }

int gt_create(int32_t n, const double* p, gt_ctx** out) {
  if (!out || !p) return fail(GT_ERR_ARG, "null argument");
  *out = nullptr;
  if (n < 16 || n > 8192 || (n & (n - 1)))
    return fail(GT_ERR_ARG, "N must be a power of two in [16, 8192] (got %d)", n);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(GT_ERR_NOGPU, "no HIP device available");
  gt_ctx* c = new gt_ctx();
  c->n = n;
  c->nh = n / 2 + 1;
  c->logn = ilog2(n);
  // src/inieuler.f90:14-19
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
  // src/inieuler.f90:22-36,59-66
  c->split = (c->alpnu < 1.0005 && c->alpnu > 0.9995 && c->alpmu < 0.0005 &&
              c->alpmu > -0.0005) ? 1 : 0;
  if (c->rko == 4) c->split = 0;
  if (c->rko != 2 && c->rko != 4) c->rko = 2;
  // src/inieuler.f90:78-81
  c->sdt = std::sqrt(c->dt);
  c->dt2 = c->dt / 2.0;
  c->dt3 = c->dt / 3.0;
  c->dt6 = c->dt / 6.0;
  defk(c);
  int rc = iniforcing(c);
  if (rc != GT_OK) {
    delete c;
    return rc;
  }
  Phys& ph = c->ph;
  ph.n = n;
  ph.dkx = c->dkx;
  ph.dky = c->dky;
  ph.ahalf = c->alpv / 2.0;
  ph.akind = ph.ahalf == 1.0 ? 1 : ph.ahalf == 0.5 ? 2 : ph.ahalf == 0.25 ? 3 : 0;
  ph.dt = c->dt;
  ph.dt2 = c->dt2;
  ph.dt3 = c->dt3;
  ph.dt6 = c->dt6;
  ph.rtrunc2 = c->rtrunc;
  ph.fkxsmax = c->fkxsmax;
  ph.fkysmax = c->fkysmax;
  ph.truncate = c->truncate;
  ph.inv_n = 1.0 / (double)n;
  ph.inv_n2 = 1.0 / ((double)n * (double)n);
  ph.nu = c->nu;
  ph.mu = c->mu;
  ph.alpnu = c->alpnu;
  ph.alpmu = c->alpmu;
  ph.alpv = c->alpv;
  c->rfull = Rect{n / 2, n / 2, n / 2 - 1};
  if (c->truncate == 1) {
    const int mx = kept_axis(c, true), my = kept_axis(c, false);
    c->rtr = Rect{mx, my < n / 2 ? my : n / 2, my < n / 2 ? my : n / 2 - 1};
  } else {
    c->rtr = c->rfull;
  }
#define CREATE_CHK(call)                  \
  do {                                    \
    int r_ = (call);                      \
    if (r_ != GT_OK) {                    \
      std::string keep = g_err;           \
      gt_destroy(c);                      \
      g_err = keep;                       \
      return r_;                          \
    }                                     \
  } while (0)
#define CREATE_HIP(call)                                                   \
  do {                                                                     \
    hipError_t e_ = (call);                                                \
    if (e_ != hipSuccess) {                                                \
      gt_destroy(c);                                                       \
      return fail(GT_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_));     \
    }                                                                      \
  } while (0)
  CREATE_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  CREATE_CHK(upload_twiddles(c));
  CREATE_HIP(hipMalloc(&c->flag, sizeof(int)));
  // the allocation-free glob's partial sums on this device (resolved once)
  CREATE_HIP(gt::glob_scratch(&c->glob_scratch, &c->glob_scratch_n));
  const int nfa = c->nf > 0 ? c->nf : 1;
  CREATE_HIP(hipHostMalloc(&c->ph_host, sizeof(double2) * nfa * kRing, hipHostMallocDefault));
  CREATE_HIP(hipMalloc(&c->ph_dev, sizeof(double2) * nfa * kRing));
  for (int i = 0; i < kRing; ++i) CREATE_HIP(hipEventCreateWithFlags(&c->ph_ev[i], hipEventDisableTiming));
  *out = c;
  return GT_OK;
}

int gt_destroy(gt_ctx* c) {
  if (!c) return GT_OK;
  c->pend = Pending{};  // a deferred step end is dropped with the context
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  adj_free(c);
  for (auto& kv : c->present) {
    Mirror& m = kv.second;
    if (m.owned && m.dref) (void)hipFree(m.dref);
    if (m.dcmp) (void)hipFree(m.dcmp);
  }
  for (auto& t : c->tab) {
    if (t.emk) (void)hipFree(t.emk);
    if (t.modes) (void)hipFree(t.modes);
    if (t.fcs) (void)hipFree(t.fcs);
    if (t.fce) (void)hipFree(t.fce);
  }
  if (c->tw) (void)hipFree(c->tw);
  if (c->y1) (void)hipFree(c->y1);
  if (c->y2) (void)hipFree(c->y2);
  if (c->flag) (void)hipFree(c->flag);
  if (c->dg) (void)hipFree(c->dg);
  if (c->ph_host) (void)hipHostFree(c->ph_host);
  if (c->ph_dev) (void)hipFree(c->ph_dev);
  for (int i = 0; i < kRing; ++i)
    if (c->ph_ev[i]) (void)hipEventDestroy(c->ph_ev[i]);
  for (auto& t : c->tl) {
    (void)hipEventDestroy(t.a);
    (void)hipEventDestroy(t.b);
  }
  for (auto e : c->ev_pool) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return GT_OK;
}

int gt_info(const gt_ctx* c, int32_t* ii, double* dd) {
  if (!c) return fail(GT_ERR_ARG, "null context");
  if (ii) {
    ii[0] = c->nf;
    ii[1] = c->split;
    ii[2] = c->rko;
    ii[3] = c->truncate;
    ii[4] = c->rtr.kx;
    ii[5] = c->rtr.p;
    ii[6] = c->rtr.q;
  }
  if (dd) {
    double v[] = {c->inpz, c->inpe, c->dt2, c->dt3, c->dt6, c->sdt,
                  c->dkx,  c->dky,  c->rtrunc, c->fkxsmax, c->fkysmax};
    std::memcpy(dd, v, sizeof v);
  }
  return GT_OK;
}

int gt_get_defk(const gt_ctx* c, double* fkx, double* fky, double* fkxs, double* fkys,
                int32_t* jc, double* emk, double* emkx, double* emky) {
  if (!c) return fail(GT_ERR_ARG, "null context");
  const size_t n = c->n, nh = c->nh;
  if (fkx) std::memcpy(fkx, c->fkx.data(), sizeof(double) * nh);
  if (fky) std::memcpy(fky, c->fky.data(), sizeof(double) * n);
  if (fkxs) std::memcpy(fkxs, c->fkxs.data(), sizeof(double) * nh);
  if (fkys) std::memcpy(fkys, c->fkys.data(), sizeof(double) * n);
  if (jc) std::memcpy(jc, c->jc.data(), sizeof(int32_t) * n);
  if (emk) std::memcpy(emk, c->emk.data(), sizeof(double) * nh * n);
  if (emkx) std::memcpy(emkx, c->emkx.data(), sizeof(double) * nh);
  if (emky) std::memcpy(emky, c->emky.data(), sizeof(double) * n);
  return GT_OK;
}

int gt_get_forcing_modes(const gt_ctx* c, int32_t* ikxf, int32_t* ikyf, double* ff) {
  if (!c) return fail(GT_ERR_ARG, "null context");
  if (c->nf == 0) return GT_OK;
  if (ikxf) std::memcpy(ikxf, c->ikxf.data(), sizeof(int32_t) * c->nf);
  if (ikyf) std::memcpy(ikyf, c->ikyf.data(), sizeof(int32_t) * c->nf);
  if (ff) std::memcpy(ff, c->ff.data(), sizeof(double) * 2 * c->nf);
  return GT_OK;
}

int gt_plan_create(gt_ctx* c, int32_t kind, int32_t* plan) {
  if (!c || !plan) return fail(GT_ERR_ARG, "null argument");
  if (kind != GT_PLAN_D2Z && kind != GT_PLAN_Z2D)
    return fail(GT_ERR_ARG, "unknown plan kind %d", kind);
  *plan = kind;
  return GT_OK;
}

int gt_mem_info(size_t* fr, size_t* tot) {
  size_t f = 0, t = 0;
  HIPCHK(hipMemGetInfo(&f, &t));
  if (fr) *fr = f;
  if (tot) *tot = t;
  return GT_OK;
}

int gt_enter_data(gt_ctx* c, void* host, int64_t bytes, int32_t copyin) {
  GT_RANGE();
  if (!c || !host || bytes <= 0) return fail(GT_ERR_ARG, "bad enter_data arguments");
  CHK(flush_pending(c));
  if (c->present.count(host)) return fail(GT_ERR_ARG, "array %p already present", host);
  Mirror m;
  m.host = host;
  m.bytes = (size_t)bytes;
  HIPCHK(hipMalloc(&m.dref, m.bytes));
  if (copyin) {
    HIPCHK(hipMemcpyAsync(m.dref, host, m.bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  } else {
    HIPCHK(hipMemsetAsync(m.dref, 0, m.bytes, c->stream));
  }
  set_ref_only(m);
  c->present[host] = m;
  return GT_OK;
}

int gt_map_data(gt_ctx* c, void* key, void* dev, int64_t bytes) {
  if (!c || !key || !dev || bytes <= 0) return fail(GT_ERR_ARG, "bad map_data arguments");
  CHK(flush_pending(c));
  if (c->present.count(key)) return fail(GT_ERR_ARG, "key %p already present", key);
  Mirror m;
  m.host = nullptr;
  m.bytes = (size_t)bytes;
  m.dref = static_cast<double*>(dev);
  m.owned = false;
  set_ref_only(m);
  c->present[key] = m;
  return GT_OK;
}

static int drop(gt_ctx* c, const void* key) {
  auto it = c->present.find(key);
  Mirror& m = it->second;
  HIPCHK(hipStreamSynchronize(c->stream));
  if (m.owned && m.dref) HIPCHK(hipFree(m.dref));
  if (m.dcmp) HIPCHK(hipFree(m.dcmp));
  c->present.erase(it);
  return GT_OK;
}

int gt_unmap_data(gt_ctx* c, void* key) {
  if (!c) return fail(GT_ERR_ARG, "null context");
  CHK(flush_pending(c));
  auto it = c->present.find(key);
  if (it == c->present.end()) return fail(GT_ERR_NOT_PRESENT, "key %p not present", key);
  if (it->second.owned) return fail(GT_ERR_ARG, "key %p was not mapped", key);
  if (!it->second.ref_valid) CHK(ensure_ref(c, it->second));
  return drop(c, key);
}

int gt_exit_data(gt_ctx* c, void* host, int32_t copyout) {
  GT_RANGE();
  if (!c) return fail(GT_ERR_ARG, "null context");
  CHK(flush_pending(c));
  auto it = c->present.find(host);
  if (it == c->present.end()) return fail(GT_ERR_NOT_PRESENT, "array %p not present", host);
  Mirror& m = it->second;
  if (copyout) {
    if (!m.owned) return fail(GT_ERR_ARG, "mapped array %p has no host copy", host);
    if (m.bytes >= ((size_t)c->n + 2) * c->n * sizeof(double)) CHK(ensure_ref(c, m));
    HIPCHK(hipMemcpyAsync(host, m.dref, m.bytes, hipMemcpyDeviceToHost, c->stream));
  }
  return drop(c, host);
}

int gt_update_host(gt_ctx* c, void* host) {
  GT_RANGE();
  if (!c) return fail(GT_ERR_ARG, "null context");
  CHK(flush_pending(c));
  auto it = c->present.find(host);
  if (it == c->present.end()) return fail(GT_ERR_NOT_PRESENT, "array %p not present", host);
  Mirror& m = it->second;
  if (m.bytes >= ((size_t)c->n + 2) * c->n * sizeof(double)) CHK(ensure_ref(c, m));
  if (m.owned) {
    HIPCHK(hipMemcpyAsync(host, m.dref, m.bytes, hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return GT_OK;
}

int gt_sync_ref(gt_ctx* c, void* key) { return gt_update_host(c, key); }

int gt_update_device(gt_ctx* c, void* host) {
  GT_RANGE();
  if (!c) return fail(GT_ERR_ARG, "null context");
  CHK(flush_pending(c));
  auto it = c->present.find(host);
  if (it == c->present.end()) return fail(GT_ERR_NOT_PRESENT, "array %p not present", host);
  Mirror& m = it->second;
  if (m.owned) {
    HIPCHK(hipMemcpyAsync(m.dref, host, m.bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  set_ref_only(m);
  return GT_OK;
}

int gt_copy(gt_ctx* c, void* dst, const void* src) {
  GT_RANGE();
  if (!c) return fail(GT_ERR_ARG, "null context");
  CHK(flush_pending(c));
  Mirror *d, *s;
  CHK(lookup(c, dst, &d));
  CHK(lookup(c, src, &s));
  if (d == s) return GT_OK;
  Timer t(c, GT_K_OTHER);
  if (s->cmp_valid) {
    CHK(alloc_cmp(c, *d, s->rect));
    HIPCHK(hipMemcpyAsync(d->dcmp, s->dcmp, sizeof(double2) * (size_t)s->rect.kc() * s->kyp,
                          hipMemcpyDeviceToDevice, c->stream));
    d->cmp_valid = true;
    d->ref_valid = false;
    if (s->ref_valid) {
      HIPCHK(hipMemcpyAsync(d->dref, s->dref, ((size_t)c->n + 2) * c->n * sizeof(double),
                            hipMemcpyDeviceToDevice, c->stream));
      d->ref_valid = true;
    }
  } else {
    HIPCHK(hipMemcpyAsync(d->dref, s->dref, ((size_t)c->n + 2) * c->n * sizeof(double),
                          hipMemcpyDeviceToDevice, c->stream));
    set_ref_only(*d);
  }
  return GT_OK;
}

int gt_synchronize(gt_ctx* c) {
  if (!c) return fail(GT_ERR_ARG, "null context");
  CHK(flush_pending(c));
  HIPCHK(hipStreamSynchronize(c->stream));
  return GT_OK;
}

// ---- transforms and operators ----------------------------------------------

int gt_fft_dir(gt_ctx* c, double* z, int32_t plan) {
  GT_RANGE();
  if (!c) return fail(GT_ERR_ARG, "null context");
  CHK(flush_pending(c));
  if (plan != GT_PLAN_D2Z) return fail(GT_ERR_ARG, "fft_dir needs a D2Z plan (got %d)", plan);
  Mirror* m;
  CHK(lookup(c, z, &m));
  CHK(ensure_ref(c, *m));
  Timer t(c, GT_K_OTHER);
  HIPCHK(gt::launch_ref_fft(c->logn, -1, m->dref, c->tw, c->stream));
  set_ref_only(*m);
  return GT_OK;
}

int gt_fft_inv(gt_ctx* c, double* z, int32_t plan) {
  GT_RANGE();
  if (!c) return fail(GT_ERR_ARG, "null context");
  CHK(flush_pending(c));
  if (plan != GT_PLAN_Z2D) return fail(GT_ERR_ARG, "fft_inv needs a Z2D plan (got %d)", plan);
  Mirror* m;
  CHK(lookup(c, z, &m));
  CHK(ensure_ref(c, *m));
  Timer t(c, GT_K_OTHER);
  HIPCHK(gt::launch_ref_fft(c->logn, +1, m->dref, c->tw, c->stream));
  set_ref_only(*m);
  return GT_OK;
}

int gt_produ(gt_ctx* c, const double* z, double* u, double* v) {
  GT_RANGE();
  if (!c) return fail(GT_ERR_ARG, "null context");
  CHK(flush_pending(c));
  Mirror *mz, *mu, *mv;
  CHK(lookup(c, z, &mz));
  CHK(lookup(c, u, &mu));
  CHK(lookup(c, v, &mv));
  CHK(ensure_ref(c, *mz));
  CHK(ensure_ref(c, *mu));
  CHK(ensure_ref(c, *mv));
  Timer t(c, GT_K_OTHER);
  HIPCHK(gt::launch_produ(c->n, mz->dref, mu->dref, mv->dref, c->stream));
  set_ref_only(*mu);
  set_ref_only(*mv);
  return GT_OK;
}

int gt_trunc(gt_ctx* c, double* z) {
  GT_RANGE();
  if (!c) return fail(GT_ERR_ARG, "null context");
  CHK(flush_pending(c));
  Mirror* m;
  CHK(lookup(c, z, &m));
  Timer t(c, GT_K_OTHER);
  if (m->cmp_valid) {
    HIPCHK(gt::launch_trunc_cmp(c->ph, m->rect, m->kyp, m->dcmp, c->stream));
    set_cmp_only(*m);
    return GT_OK;
  }
  // trunc applies the mask whatever the truncate flag: the result fits the
  // truncated rectangle when the rule is on, else the full one.
  const Rect r = c->truncate == 1 ? c->rtr : c->rfull;
  CHK(alloc_cmp(c, *m, r));
  HIPCHK(gt::launch_gather(c->ph, r, m->kyp, m->dref, m->dcmp, 1, nullptr, c->stream));
  set_cmp_only(*m);
  return GT_OK;
}

int gt_nlt(gt_ctx* c, double* z, double* u, double* v) {
  GT_RANGE();
  if (!c) return fail(GT_ERR_ARG, "null context");
  CHK(flush_pending(c));
  Mirror *m, *mu, *mv;
  CHK(lookup(c, z, &m));
  CHK(lookup(c, u, &mu));
  CHK(lookup(c, v, &mv));
  CHK(ensure_cmp(c, *m));
  CHK(ensure_ybuf(c, m->rect));
  gt::PassArgs a = pass_args(c, *m, nullptr);
  a.out = m->dcmp;
  CHK(c_pass(c, gt::C_START, a));
  CHK(r_pass(c, a));
  CHK(c_pass(c, gt::C_NLT, a));
  set_cmp_only(*m);
  return GT_OK;
}

int gt_step(gt_ctx* c, double* z, double* znl, double* u, double* v) {
  GT_RANGE();
  if (!c) return fail(GT_ERR_ARG, "null context");
  Mirror *m, *m1, *mu, *mv;
  CHK(lookup(c, z, &m));
  CHK(lookup(c, znl, &m1));
  CHK(lookup(c, u, &mu));
  CHK(lookup(c, v, &mv));
  if (!pending_fuses(c, PEND_RK2, m, nullptr)) CHK(flush_pending(c));
  CHK(ensure_cmp(c, *m));
  CHK(ensure_ybuf(c, m->rect));
  RectTables* t;
  CHK(rect_tables(c, m->rect, &t));
  gt::PassArgs a = pass_args(c, *m, t);
  CHK(step_start(c, PEND_RK2, m, nullptr, a));
  CHK(r_pass(c, a));
  CHK(c_pass(c, gt::C_RK2_1, a));
  CHK(r_pass(c, a));
  // the last stage (C_RK2_2) is deferred: the next step fuses it with its
  // forcing and its first pass when nothing else comes between
  c->pend = Pending{PEND_RK2, m, nullptr, t, a, false, 0};
  set_cmp_only(*m);
  return GT_OK;
}

int gt_step4(gt_ctx* c, double* z, double* znl, double* znew, double* u, double* v) {
  GT_RANGE();
  if (!c) return fail(GT_ERR_ARG, "null context");
  Mirror *m, *m1, *mn, *mu, *mv;
  CHK(lookup(c, z, &m));
  CHK(lookup(c, znl, &m1));
  CHK(lookup(c, znew, &mn));
  CHK(lookup(c, u, &mu));
  CHK(lookup(c, v, &mv));
  if (mn == m) return fail(GT_ERR_ARG, "step4: z and znew must differ");
  if (!pending_fuses(c, PEND_RK4, m, mn)) CHK(flush_pending(c));
  CHK(ensure_cmp(c, *m));
  CHK(ensure_ybuf(c, m->rect));
  RectTables* t;
  CHK(rect_tables(c, m->rect, &t));
  CHK(alloc_cmp(c, *mn, m->rect));
  gt::PassArgs a = pass_args(c, *m, t);
  a.znew = mn->dcmp;
  CHK(step_start(c, PEND_RK4, m, mn, a));
  CHK(r_pass(c, a));
  CHK(c_pass(c, gt::C_RK4_1, a));
  CHK(r_pass(c, a));
  CHK(c_pass(c, gt::C_RK4_2, a));
  CHK(r_pass(c, a));
  CHK(c_pass(c, gt::C_RK4_3, a));
  CHK(r_pass(c, a));
  // the last stage (C_RK4_4) is deferred: the next step fuses it with its
  // forcing and its first pass when nothing else comes between
  c->pend = Pending{PEND_RK4, m, mn, t, a, false, 0};
  set_cmp_only(*m);
  set_cmp_only(*mn);
  return GT_OK;
}

int gt_forcing(gt_ctx* c, double* z, int32_t* seed) {
  GT_RANGE();
  if (!c || !seed) return fail(GT_ERR_ARG, "null argument");
  Mirror* m;
  CHK(lookup(c, z, &m));
  if (c->nf == 0) return GT_OK;
  Pending& p = c->pend;
  if (p.kind != PEND_NONE && p.z == m && !p.forced) {
    // after a deferred step end: draw now, apply with that end
    CHK(forcing_upload(c, seed, &p.slot));
    p.forced = true;
    return GT_OK;
  }
  CHK(flush_pending(c));
  CHK(ensure_cmp(c, *m));
  RectTables* t;
  CHK(rect_tables(c, m->rect, &t));
  CHK(forcing_on(c, t, m->dcmp, seed));
  set_cmp_only(*m);
  return GT_OK;
}

// ---- diagnostics ------------------------------------------------------------

int gt_spectrum(gt_ctx* c, const double* z, double* sp, double* sz) {
  GT_RANGE();
  if (!c || !sp || !sz) return fail(GT_ERR_ARG, "null argument");
  CHK(flush_pending(c));
  Mirror* m;
  CHK(lookup(c, z, &m));
  CHK(ensure_cmp(c, *m));
  CHK(ensure_diag(c, m->rect));
  gt::DiagArgs d = diag_args(c, *m);
  {
    Timer t(c, GT_K_OTHER);
    HIPCHK(gt::launch_spectrum(d, 0, c->stream));
  }
  const size_t nbin = c->n / 2;
  HIPCHK(hipMemcpyAsync(sp, d.out1, sizeof(double) * nbin, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(sz, d.out2, sizeof(double) * nbin, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return GT_OK;
}

int gt_fluxes(gt_ctx* c, const double* z, const double* znl, double* fle, double* flz) {
  GT_RANGE();
  if (!c || !fle || !flz) return fail(GT_ERR_ARG, "null argument");
  CHK(flush_pending(c));
  Mirror *m, *mn;
  CHK(lookup(c, z, &m));
  CHK(lookup(c, znl, &mn));
  CHK(ensure_cmp(c, *m));
  CHK(ensure_cmp(c, *mn));
  CHK(same_rect(c, *m, *mn));
  CHK(ensure_diag(c, m->rect));
  gt::DiagArgs d = diag_args(c, *m);
  d.znl = mn->dcmp;
  {
    Timer t(c, GT_K_OTHER);
    HIPCHK(gt::launch_spectrum(d, 1, c->stream));
  }
  const size_t nbin = c->n / 2;
  HIPCHK(hipMemcpyAsync(fle, d.out1, sizeof(double) * nbin, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(flz, d.out2, sizeof(double) * nbin, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return GT_OK;
}

int gt_glob(gt_ctx* c, const double* z, double* out) {
  GT_RANGE();
  if (!c || !out) return fail(GT_ERR_ARG, "null argument");
  CHK(flush_pending(c));
  Mirror* m;
  CHK(lookup(c, z, &m));
  CHK(ensure_cmp(c, *m));
  CHK(ensure_diag(c, m->rect));
  gt::DiagArgs d = diag_args(c, *m);
  {
    Timer t(c, GT_K_OTHER);
    HIPCHK(gt::launch_glob(d, c->stream));
  }
  HIPCHK(hipMemcpyAsync(out, d.gout, sizeof(double) * 6, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return GT_OK;
}

int gt_glob_noalloc(gt_ctx* c, const double* z, double* out_dev) {
  // no profiler range here: nothing but kernel launches runs on this path
  if (!c || !out_dev) return fail(GT_ERR_ARG, "null argument");
  CHK(flush_pending(c));  // a step's deferred end (launches only)
  Mirror* m;
  CHK(lookup(c, z, &m));
  gt::DiagArgs d{};
  d.ph = c->ph;
  if (m->cmp_valid) {  // the compact copy, as gt_glob reads it
    d.rc = m->rect;
    d.kyp = m->kyp;
    d.z = m->dcmp;
    d.zcs = m->kyp;
    d.zrs = 1;
  } else if (m->ref_valid) {  // the reference layout: the whole half-spectrum
    d.rc = c->rfull;
    d.kyp = 0;
    d.z = reinterpret_cast<const double2*>(m->dref);
    d.zcs = 1;
    d.zrs = c->nh;
  } else {
    return fail(GT_ERR_ARG, "array %p has no valid device data", z);
  }
  const size_t need = 6 * (size_t)d.rc.kc() + 6;
  if (c->glob_scratch == nullptr || need > c->glob_scratch_n)
    return fail(GT_ERR_ARG, "gt_glob_noalloc: %zu partial sums exceed the static array of %zu",
                need, c->glob_scratch_n);
  // every entry the launch reads it writes first: nothing to clear
  d.gpart = c->glob_scratch;
  d.gout = out_dev;
  HIPCHK(gt::launch_glob(d, c->stream));
  return GT_OK;
}

int gt_shell_index(gt_ctx* c, int32_t* ib) {
  if (!c || !ib) return fail(GT_ERR_ARG, "null argument");
  const size_t tot = (size_t)c->nh * c->n;
  int* d = nullptr;
  HIPCHK(hipMalloc(&d, sizeof(int) * tot));
  hipError_t e = gt::launch_shell_index(c->ph, d, c->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(ib, d, sizeof(int) * tot, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(GT_ERR_HIP, "shell_index: %s", hipGetErrorString(e));
  return GT_OK;
}

// ---- timing -------------------------------------------------------------------

int gt_timing_enable(gt_ctx* c, int32_t on) {
  if (!c) return fail(GT_ERR_ARG, "null context");
  c->timing = on != 0;
  return GT_OK;
}

int gt_timing_classes(gt_ctx* c, int32_t mask) {
  if (!c) return fail(GT_ERR_ARG, "null context");
  c->timing_mask = (unsigned)mask;
  return GT_OK;
}

int gt_timing_read(gt_ctx* c, double* ms, int32_t* count) {
  if (!c) return fail(GT_ERR_ARG, "null context");
  CHK(flush_pending(c));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int i = 0; i < GT_K_NCLASS; ++i) {
    if (ms) ms[i] = 0.0;
    if (count) count[i] = 0;
  }
  for (auto& t : c->tl) {
    float e = 0.f;
    HIPCHK(hipEventElapsedTime(&e, t.a, t.b));
    if (ms) ms[t.cls] += e;
    if (count) count[t.cls] += 1;
    c->ev_pool.push_back(t.a);
    c->ev_pool.push_back(t.b);
  }
  c->tl.clear();
  return GT_OK;
}

}  // extern "C"
