// gturbo_adjoint.cpp -- gradients with respect to the initial field
// (gt_adj_forward / gt_adj_backward of gturbo.h).
//
// Forward: the input frame goes through fft_dir and trunc into a compact
// state, the steps run with the single-run passes, and the state is kept
// every kchk steps (with its forcing seed).  Backward: the adjoint of the
// final field enters as trunc(fft_dir(gT)); for each checkpoint segment,
// last first, the segment is recomputed with the y passes storing every RK
// stage input, then stepped backwards through the adjoint passes (the
// transposed nonlinear term inside the reversed stages); the forcing adds a
// state-independent term and has no adjoint.  The gradient leaves as
// fft_inv of the adjoint state.  With fft_dir = F (1/N^2 included) and
// fft_inv = G, the adjoint of w -> P F w is G P and that of z -> G z is
// N^2 F; the two factors N^2 cancel, so the state carried backwards is the
// adjoint in the Parseval inner product divided by N^2.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "gt_runtime.hpp"

using namespace gtr;

struct AdjState {
  int nsteps = 0, kchk = 0, nck = 0, nstage = 0;
  bool ready = false;
  Rect rc{0, 0, 0};
  int kyp = 0;
  size_t nc = 0;                 // double2 per compact field
  std::vector<double2*> ck;      // state at steps 0, kchk, 2 kchk, ...
  std::vector<int32_t> ckseed;   // forcing seed at those steps
  double2* z = nullptr;          // working state
  double2* znew = nullptr;
  double2* stages = nullptr;     // kchk x nstage stage inputs of a segment
  double2* lam = nullptr;        // adjoint state
  double2* lacc = nullptr;
  double2* y1 = nullptr;         // 5 fields between the adjoint passes
  double2* y2 = nullptr;         // 3 fields
  double* tmp = nullptr;         // one reference-layout field
};

namespace gtr {

void adj_free(gt_ctx* c) {
  AdjState* s = c->adj;
  if (!s) return;
  for (double2* p : s->ck)
    if (p) (void)hipFree(p);
  for (void* p : {(void*)s->z, (void*)s->znew, (void*)s->stages, (void*)s->lam,
                  (void*)s->lacc, (void*)s->y1, (void*)s->y2, (void*)s->tmp})
    if (p) (void)hipFree(p);
  delete s;
  c->adj = nullptr;
}

}  // namespace gtr

namespace {

double2* stage_ptr(AdjState* s, int slot, int stage) {
  return s->stages + ((size_t)slot * s->nstage + stage) * s->nc;
}

// Buffers for nsteps steps with a checkpoint every kchk steps.
int adj_alloc(gt_ctx* c, int nsteps, int kchk) {
  if (!c->adj) c->adj = new AdjState();
  AdjState* s = c->adj;
  const int nck = (nsteps + kchk - 1) / kchk;
  const int nstage = c->rko == 4 ? 4 : 2;
  const size_t nc = (size_t)c->rtr.kc() * kyp_of(c->rtr);
  if (s->nc != nc || s->kchk != kchk || s->nstage != nstage) {
    adj_free(c);
    c->adj = new AdjState();
    s = c->adj;
    s->rc = c->rtr;
    s->kyp = kyp_of(c->rtr);
    s->nc = nc;
    s->kchk = kchk;
    s->nstage = nstage;
    const int bt = gt::bt_of(c->logn);
    const size_t ycap = (size_t)((s->rc.kc() + bt - 1) / bt) * bt * c->n;
    HIPCHK(hipMalloc(&s->z, sizeof(double2) * nc));
    HIPCHK(hipMalloc(&s->znew, sizeof(double2) * nc));
    HIPCHK(hipMalloc(&s->stages, sizeof(double2) * nc * nstage * kchk));
    HIPCHK(hipMalloc(&s->lam, sizeof(double2) * nc));
    HIPCHK(hipMalloc(&s->lacc, sizeof(double2) * nc));
    HIPCHK(hipMalloc(&s->y1, sizeof(double2) * ycap * 5));
    HIPCHK(hipMalloc(&s->y2, sizeof(double2) * ycap * 3));
    HIPCHK(hipMalloc(&s->tmp, sizeof(double) * ((size_t)c->n + 2) * c->n));
    for (double2* p : {s->z, s->znew, s->lam, s->lacc})
      HIPCHK(hipMemsetAsync(p, 0, sizeof(double2) * nc, c->stream));
    HIPCHK(hipMemsetAsync(s->stages, 0, sizeof(double2) * nc * nstage * kchk, c->stream));
    HIPCHK(hipMemsetAsync(s->y1, 0, sizeof(double2) * ycap * 5, c->stream));
    HIPCHK(hipMemsetAsync(s->y2, 0, sizeof(double2) * ycap * 3, c->stream));
  }
  while ((int)s->ck.size() < nck) {
    double2* p = nullptr;
    HIPCHK(hipMalloc(&p, sizeof(double2) * nc));
    s->ck.push_back(p);
  }
  s->ckseed.assign(nck, 0);
  s->nck = nck;
  s->nsteps = nsteps;
  return GT_OK;
}

// Compact state from a physical field: fft_dir and trunc (src/main.f90:89-92).
int from_physical(gt_ctx* c, const double* dref, double2* out) {
  AdjState* s = c->adj;
  const size_t bytes = sizeof(double) * ((size_t)c->n + 2) * c->n;
  HIPCHK(hipMemcpyAsync(s->tmp, dref, bytes, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(gt::launch_ref_fft(c->logn, -1, s->tmp, c->tw, c->stream));
  HIPCHK(gt::launch_gather(c->ph, s->rc, s->kyp, s->tmp, out, 1, nullptr, c->stream));
  return GT_OK;
}

// Physical field of a compact state: fft_inv (src/main.f90:143).
int to_physical(gt_ctx* c, const double2* in, double* dref) {
  AdjState* s = c->adj;
  HIPCHK(gt::launch_scatter(c->ph, s->rc, s->kyp, in, dref, c->stream));
  HIPCHK(gt::launch_ref_fft(c->logn, +1, dref, c->tw, c->stream));
  return GT_OK;
}

// One step of src/rkstep.f90 on the working state; with slot >= 0 every
// stage input is kept in the stage store.
int fwd_step(gt_ctx* c, RectTables* t, int slot) {
  AdjState* s = c->adj;
  CHK(ensure_ybuf(c, s->rc));
  gt::PassArgs a{};
  a.rc = s->rc;
  a.kyp = s->kyp;
  const int bt = gt::bt_of(c->logn);
  a.ntiles = (s->rc.kc() + bt - 1) / bt;
  a.ph = c->ph;
  a.tw = c->tw;
  a.y1 = c->y1;
  a.y2 = c->y2;
  a.z = s->z;
  a.znew = s->znew;
  a.emk = t->emk;
  auto st = [&](int k) { return slot >= 0 ? stage_ptr(s, slot, k) : nullptr; };
  if (c->rko == 4) {
    a.sst = st(0);
    CHK(c_pass(c, gt::C_START, a));
    CHK(r_pass(c, a));
    a.sst = st(1);
    CHK(c_pass(c, gt::C_RK4_1, a));
    CHK(r_pass(c, a));
    a.sst = st(2);
    CHK(c_pass(c, gt::C_RK4_2, a));
    CHK(r_pass(c, a));
    a.sst = st(3);
    CHK(c_pass(c, gt::C_RK4_3, a));
    CHK(r_pass(c, a));
    a.sst = nullptr;
    CHK(c_pass(c, gt::C_RK4_4, a));
  } else {
    a.sst = st(0);
    CHK(c_pass(c, gt::C_START, a));
    CHK(r_pass(c, a));
    a.sst = st(1);
    CHK(c_pass(c, gt::C_RK2_1, a));
    CHK(r_pass(c, a));
    a.sst = nullptr;
    CHK(c_pass(c, gt::C_RK2_2, a));
  }
  return GT_OK;
}

// The adjoint of one step: lam (adjoint of the step's output) -> lam
// (adjoint of its input), with the step's stage inputs in slot.
int adj_step(gt_ctx* c, RectTables* t, int slot) {
  AdjState* s = c->adj;
  gt::AdjArgs j{};
  j.rc = s->rc;
  j.kyp = s->kyp;
  const int bt = gt::bt_of(c->logn);
  j.ntiles = (s->rc.kc() + bt - 1) / bt;
  j.ph = c->ph;
  j.tw = c->tw;
  j.y1 = s->y1;
  j.y2 = s->y2;
  j.lam = s->lam;
  j.lacc = s->lacc;
  j.emk = t->emk;
  auto yp = [&](int mode, int stage) -> int {
    j.snext = stage >= 0 ? stage_ptr(s, slot, stage) : nullptr;
    Timer tm(c, mode == gt::J_START4 || mode == gt::J_START2 ? GT_K_YSTART
                : mode == gt::J_END                        ? GT_K_YFINAL
                                                           : GT_K_YSTAGE);
    HIPCHK(gt::launch_j_pass(c->logn, mode, j, c->stream));
    return GT_OK;
  };
  auto xp = [&]() -> int {
    Timer tm(c, GT_K_XPASS);
    HIPCHK(gt::launch_jr_pass(c->logn, j, c->stream));
    return GT_OK;
  };
  if (c->rko == 4) {
    CHK(yp(gt::J_START4, 3));
    CHK(xp());
    CHK(yp(gt::J_RK4_3, 2));
    CHK(xp());
    CHK(yp(gt::J_RK4_2, 1));
    CHK(xp());
    CHK(yp(gt::J_RK4_1, 0));
    CHK(xp());
    CHK(yp(gt::J_END, -1));
  } else {
    CHK(yp(gt::J_START2, 1));
    CHK(xp());
    CHK(yp(gt::J_RK2_1, 0));
    CHK(xp());
    CHK(yp(gt::J_END, -1));
  }
  return GT_OK;
}

int adj_check(gt_ctx* c) {
  if (!c) return fail(GT_ERR_ARG, "null context");
  const int lim = 3 * (c->n / 8);
  if (c->truncate != 1 || c->rtr.kx >= lim || c->rtr.p >= lim || c->rtr.q >= lim)
    return fail(GT_ERR_ARG, "the adjoint needs the 2/3 rule (truncate = 1)");
  return GT_OK;
}

}  // namespace

extern "C" {

int gt_adj_forward(gt_ctx* c, const double* w0, double* wT, int32_t nsteps, int32_t kchk,
                   int32_t* seed) {
  GT_RANGE();
  CHK(adj_check(c));
  CHK(flush_pending(c));
  if (!seed) return fail(GT_ERR_ARG, "null seed");
  if (nsteps < 1 || kchk < 1 || kchk > nsteps)
    return fail(GT_ERR_ARG, "need 1 <= kchk <= nsteps (got %d, %d)", kchk, nsteps);
  Mirror *m0, *mT;
  CHK(lookup(c, w0, &m0));
  CHK(lookup(c, wT, &mT));
  CHK(ensure_ref(c, *m0));
  CHK(adj_alloc(c, nsteps, kchk));
  AdjState* s = c->adj;
  s->ready = false;
  RectTables* t;
  CHK(rect_tables(c, s->rc, &t));
  CHK(from_physical(c, m0->dref, s->z));
  for (int n = 0; n < nsteps; ++n) {
    if (n % kchk == 0) {
      HIPCHK(hipMemcpyAsync(s->ck[n / kchk], s->z, sizeof(double2) * s->nc,
                            hipMemcpyDeviceToDevice, c->stream));
      s->ckseed[n / kchk] = *seed;
    }
    CHK(fwd_step(c, t, -1));
    CHK(forcing_on(c, t, s->z, seed));  // src/main.f90:136
  }
  CHK(to_physical(c, s->z, mT->dref));
  set_ref_only(*mT);
  s->ready = true;
  return GT_OK;
}

int gt_adj_backward(gt_ctx* c, const double* gT, double* g0) {
  GT_RANGE();
  CHK(adj_check(c));
  CHK(flush_pending(c));
  AdjState* s = c->adj;
  if (!s || !s->ready) return fail(GT_ERR_ARG, "gt_adj_backward needs a gt_adj_forward first");
  Mirror *mT, *m0;
  CHK(lookup(c, gT, &mT));
  CHK(lookup(c, g0, &m0));
  CHK(ensure_ref(c, *mT));
  RectTables* t;
  CHK(rect_tables(c, s->rc, &t));
  CHK(from_physical(c, mT->dref, s->lam));
  for (int seg = s->nck - 1; seg >= 0; --seg) {
    const int n0 = seg * s->kchk, n1 = std::min(s->nsteps, n0 + s->kchk);
    HIPCHK(hipMemcpyAsync(s->z, s->ck[seg], sizeof(double2) * s->nc, hipMemcpyDeviceToDevice,
                          c->stream));
    int32_t seed = s->ckseed[seg];
    for (int n = n0; n < n1; ++n) {
      CHK(fwd_step(c, t, n - n0));
      if (n + 1 < n1) CHK(forcing_on(c, t, s->z, &seed));
    }
    for (int n = n1 - 1; n >= n0; --n) CHK(adj_step(c, t, n - n0));
  }
  CHK(to_physical(c, s->lam, m0->dref));
  set_ref_only(*m0);
  return GT_OK;
}

}  // extern "C"
