// gturbo_ensemble.cpp -- batched ensembles of forced runs (gt_ens_* of
// gturbo.h).
//
// Every member keeps its own compact fields (z, znew), y-pass buffers and
// tables; a device array of per-member pass arguments and two work lists
// (y tiles, x row groups, largest N first) let one launch of a batched pass
// kernel cover every member.  The RK steps of the whole ensemble -- five y
// passes and four x passes each for RK4, then the forcing of every member
// and the step counter, with a step's last y pass, its forcing and the next
// step's first y pass fused into one inside a run of steps -- are captured
// once in a HIP graph (a few steps per graph) and replayed.  The forcing
// phases are drawn on the host, member by member in the reference's order,
// for chunks of steps; a chunk is drawn while the GPU runs the previous one
// (two buffers).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "gt_runtime.hpp"

using namespace gtr;

namespace {

constexpr int kChunk = 256;  // RK steps of forcing phases a buffer holds
constexpr int kPipe = 32;    // RK steps per upload: the host draws the next
                             // chunk while the GPU runs this one

struct EnsMember {
  gt_ctx* ctx = nullptr;  // tables: defk, iniforcing, twiddles, emk
  RectTables* tab = nullptr;
  Rect rc{0, 0, 0};
  int kyp = 0, logn = 0;
  double2* z = nullptr;
  double2* znew = nullptr;
  double2* y1 = nullptr;
  double2* y2 = nullptr;
  double* ka = nullptr;  // k2**(alpv/2), compact, when alpv != 2
  size_t ycap = 0;       // double2 per y field
  int32_t seed = 481929;
  int first = 0;         // first forced mode in the flattened forcing launch
};

}  // namespace

struct gt_ens {
  int nmem = 0, rko = 4, ncu = 0;
  std::vector<EnsMember> mem;
  hipStream_t stream = nullptr;
  int32_t cfg[GT_ENS_NCFG] = {};
  gt::PassArgs* descs = nullptr;
  uint32_t* items = nullptr;  // y items, then x items
  size_t items_cap = 0;
  int ny = 0, nx = 0;
  gt::EnsForce* fdesc = nullptr;
  int ftotal = 0;
  int* step = nullptr;  // chunk step counter read by the forcing launch
  int* flag = nullptr;
  double2* ph_dev[2] = {};
  double2* ph_host[2] = {};
  gt::EnsForce* fd_host[2] = {};
  hipEvent_t ph_ev[2] = {};
  bool ph_used[2] = {};
  unsigned nchunk = 0;
  bool pending = false;  // the last step's forcing is held back (gt_ens_advance)
  int* zero_host = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  hipEvent_t t0 = nullptr, t1 = nullptr;
};

namespace {

bool valid_cfg(const int32_t* c) {
  return (c[0] == 256 || c[0] == 512) && (c[1] == 256 || c[1] == 512) &&
         (c[2] == 0 || c[2] == 1 || c[2] == 2 || c[2] == 4) && (c[3] == 0 || c[3] == 1) &&
         c[5] >= 1 && c[5] <= kChunk;
}

void drop_graph(gt_ens* e) {
  if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
  if (e->graph) (void)hipGraphDestroy(e->graph);
  e->gexec = nullptr;
  e->graph = nullptr;
}

int ntiles_of(const EnsMember& m, int ty) {
  const int bt = gt::ens_lines(ty, m.logn);
  return (m.rc.kc() + bt - 1) / bt;
}

// Pass arguments and work lists of a schedule; uploaded, graph dropped.
int apply(gt_ens* e, const int32_t* cfg) {
  if (!valid_cfg(cfg)) return fail(GT_ERR_ARG, "invalid ensemble schedule");
  std::vector<gt::PassArgs> d(e->nmem);
  std::vector<int> order(e->nmem);
  for (int m = 0; m < e->nmem; ++m) order[m] = m;
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return e->mem[a].logn > e->mem[b].logn; });
  std::vector<uint32_t> yi, xi;
  for (int m = 0; m < e->nmem; ++m) {
    const EnsMember& mb = e->mem[m];
    gt::PassArgs a{};
    a.rc = mb.rc;
    a.kyp = mb.kyp;
    a.ntiles = ntiles_of(mb, cfg[0]);
    a.ph = mb.ctx->ph;
    a.tw = mb.ctx->tw;
    a.y1 = mb.y1;
    a.y2 = mb.y2;
    a.z = mb.z;
    a.znew = mb.znew;
    a.emk = mb.tab->emk;
    a.out = nullptr;
    a.ka = mb.ka;
    a.xvar = (cfg[4] >> mb.logn) & 1;
    // the fused step boundary forces the member's columns with the phases
    // of the step the device counter is at
    a.fcs = mb.tab->fcs;
    a.fce = mb.tab->fce;
    a.fens = e->fdesc + m;
    a.fstep = e->step;
    d[m] = a;
  }
  for (int m : order) {
    const EnsMember& mb = e->mem[m];
    const int nt = d[m].ntiles;
    for (int t = 0; t < nt; ++t) yi.push_back(((uint32_t)m << 16) | (uint32_t)t);
    const int rows = gt::ens_lines(cfg[1], mb.logn) * (d[m].xvar ? 2 : 1);
    const int ng = (1 << mb.logn) / rows;
    for (int g = 0; g < ng; ++g) xi.push_back(((uint32_t)m << 16) | (uint32_t)g);
  }
  const size_t need = yi.size() + xi.size();
  if (need > e->items_cap) {
    if (e->items) (void)hipFree(e->items);
    e->items = nullptr;
    e->items_cap = 0;
    HIPCHK(hipMalloc(&e->items, sizeof(uint32_t) * need));
    e->items_cap = need;
  }
  HIPCHK(hipStreamSynchronize(e->stream));
  HIPCHK(hipMemcpy(e->items, yi.data(), sizeof(uint32_t) * yi.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(e->items + yi.size(), xi.data(), sizeof(uint32_t) * xi.size(),
                   hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(e->descs, d.data(), sizeof(gt::PassArgs) * d.size(), hipMemcpyHostToDevice));
  e->ny = (int)yi.size();
  e->nx = (int)xi.size();
  std::memcpy(e->cfg, cfg, sizeof(e->cfg));
  drop_graph(e);
  return GT_OK;
}

int grid_of(const gt_ens* e, int nitems, int threads) {
  if (e->cfg[2] == 0) return nitems;
  const int g = e->ncu * e->cfg[2] * 256 / threads;  // 4 SIMDs x waves, 64-lane waves
  return std::max(1, std::min(nitems, g));
}

// n RK steps of every member on the ensemble stream; with forcing, each
// step's forcing and the step counter follow it.  Between two forced steps
// the last stage, the forcing and the next step's first pass run as one y
// pass (C_RK4_4S / C_RK2_2S), as gt_step4 / gt_step fuse a driver's step,
// forcing, step: one launch fewer and the state read and written once.
int launch_steps(gt_ens* e, int n, bool forcing) {
  const int ty = e->cfg[0], tx = e->cfg[1], sc = e->cfg[3];
  const uint32_t* yi = e->items;
  const uint32_t* xi = e->items + e->ny;
  const int gy = grid_of(e, e->ny, ty), gx = grid_of(e, e->nx, tx);
  auto cp = [&](int mode) -> int {
    HIPCHK(gt::launch_ens_c_pass(ty, sc, mode, e->descs, yi, e->ny, gy, e->stream));
    return GT_OK;
  };
  auto rp = [&]() -> int {
    HIPCHK(gt::launch_ens_r_pass(tx, ty, sc, e->descs, xi, e->nx, gx, e->stream));
    return GT_OK;
  };
  bool started = false;  // this step's first pass ran fused with the last one's end
  for (int s = 0; s < n; ++s) {
    const bool fuse = forcing && s + 1 < n;  // this step's end with the next's start
    if (!started) CHK(cp(gt::C_START));
    CHK(rp());
    if (e->rko == 4) {  // src/rkstep.f90:82-131
      CHK(cp(gt::C_RK4_1));
      CHK(rp());
      CHK(cp(gt::C_RK4_2));
      CHK(rp());
      CHK(cp(gt::C_RK4_3));
      CHK(rp());
      CHK(cp(fuse ? gt::C_RK4_4S : gt::C_RK4_4));
    } else {  // src/rkstep.f90:2-76
      CHK(cp(gt::C_RK2_1));
      CHK(rp());
      CHK(cp(fuse ? gt::C_RK2_2S : gt::C_RK2_2));
    }
    if (forcing) {  // src/main.f90:136
      if (!fuse)
        HIPCHK(gt::launch_ens_forcing(e->fdesc, e->nmem, e->ftotal, e->step, e->stream));
      HIPCHK(gt::launch_ens_tick(e->step, e->stream));
    }
    started = fuse;
  }
  return GT_OK;
}

int capture(gt_ens* e) {
  drop_graph(e);
  HIPCHK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
  int rc = GT_OK;
  rc = launch_steps(e, e->cfg[5], true);
  hipGraph_t g = nullptr;
  const hipError_t ee = hipStreamEndCapture(e->stream, &g);
  if (rc != GT_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (ee != hipSuccess)
    return fail(GT_ERR_HIP, "graph capture: %s", hipGetErrorString(ee));
  e->graph = g;
  HIPCHK(hipGraphInstantiate(&e->gexec, e->graph, nullptr, nullptr, 0));
  return GT_OK;
}

// The held-back forcing of gt_ens_advance, on the step counter and phase
// buffer it was drawn with.
int apply_pending(gt_ens* e) {
  if (!e->pending) return GT_OK;
  HIPCHK(gt::launch_ens_forcing(e->fdesc, e->nmem, e->ftotal, e->step, e->stream));
  HIPCHK(gt::launch_ens_tick(e->step, e->stream));
  e->pending = false;
  return GT_OK;
}

// Forcing phases of `steps` steps for every member into host buffer b
// (member m's block: [step][mode] at first_m * steps), drawn in parallel
// over members, each in the reference's order; the seeds advance.
void draw_phases(gt_ens* e, int b, int steps) {
  const int nm = e->nmem;
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const int nt = std::min<int>(nm, (int)std::min(16u, hw));
  auto work = [&](int t) {
    for (int m = t; m < nm; m += nt) {
      EnsMember& mb = e->mem[m];
      const int nf = mb.ctx->nf;
      double2* out = e->ph_host[b] + (size_t)mb.first * steps;
      for (int s = 0; s < steps; ++s) forcing_phases(nf, &mb.seed, out + (size_t)s * nf);
    }
  };
  std::vector<std::thread> th;
  for (int t = 1; t < nt; ++t) th.emplace_back(work, t);
  work(0);
  for (auto& x : th) x.join();
  for (int m = 0; m < nm; ++m) {
    EnsMember& mb = e->mem[m];
    gt::EnsForce f;
    f.modes = mb.tab->modes;
    f.phases = e->ph_dev[b] + (size_t)mb.first * steps;
    f.z = mb.z;
    f.amp = mb.ctx->sdt * mb.ctx->famp;
    f.nf = mb.ctx->nf;
    f.first = mb.first;
    e->fd_host[b][m] = f;
  }
}

int check_member(gt_ens* e, int32_t m) {
  if (!e) return fail(GT_ERR_ARG, "null ensemble");
  if (m < 0 || m >= e->nmem) return fail(GT_ERR_ARG, "member %d out of range", m);
  return GT_OK;
}

int time_steps(gt_ens* e, int steps, const std::vector<double2*>& backup, double* ms) {
  double best = 1e30;
  for (int rep = 0; rep < 3; ++rep) {
    for (int m = 0; m < e->nmem; ++m) {
      const EnsMember& mb = e->mem[m];
      HIPCHK(hipMemcpyAsync(mb.z, backup[m], sizeof(double2) * mb.rc.kc() * mb.kyp,
                            hipMemcpyDeviceToDevice, e->stream));
    }
    CHK(launch_steps(e, 1, false));
    HIPCHK(hipEventRecord(e->t0, e->stream));
    CHK(launch_steps(e, steps, false));
    HIPCHK(hipEventRecord(e->t1, e->stream));
    HIPCHK(hipEventSynchronize(e->t1));
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, e->t0, e->t1));
    best = std::min(best, (double)t / steps);
  }
  *ms = best;
  return GT_OK;
}

}  // namespace

extern "C" {

int gt_ens_destroy(gt_ens* e) {
  if (!e) return GT_OK;
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  drop_graph(e);
  for (auto& m : e->mem) {
    for (void* p : {(void*)m.z, (void*)m.znew, (void*)m.y1, (void*)m.y2, (void*)m.ka})
      if (p) (void)hipFree(p);
    if (m.ctx) gt_destroy(m.ctx);
  }
  for (void* p : {(void*)e->descs, (void*)e->items, (void*)e->fdesc, (void*)e->step,
                  (void*)e->flag, (void*)e->ph_dev[0], (void*)e->ph_dev[1]})
    if (p) (void)hipFree(p);
  for (void* p : {(void*)e->ph_host[0], (void*)e->ph_host[1], (void*)e->fd_host[0],
                  (void*)e->fd_host[1], (void*)e->zero_host})
    if (p) (void)hipHostFree(p);
  for (hipEvent_t v : {e->ph_ev[0], e->ph_ev[1], e->t0, e->t1})
    if (v) (void)hipEventDestroy(v);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
  return GT_OK;
}

int gt_ens_create(int32_t nmem, const int32_t* n, const double* params, gt_ens** out) {
  if (!out || !n || !params) return fail(GT_ERR_ARG, "null argument");
  *out = nullptr;
  if (nmem < 1 || nmem > 65535) return fail(GT_ERR_ARG, "1 to 65535 members (got %d)", nmem);
  for (int m = 0; m < nmem; ++m) {
    const int l = ilog2(n[m]);
    if ((1 << l) != n[m] || l < gt::kEnsMinLogn || l > gt::kEnsMaxLogn)
      return fail(GT_ERR_ARG, "member %d: N must be a power of two in [%d, %d] (got %d)", m,
                  1 << gt::kEnsMinLogn, 1 << gt::kEnsMaxLogn, n[m]);
  }
  gt_ens* e = new gt_ens();
  e->nmem = nmem;
  e->mem.resize(nmem);
#define ENS_CHK(call)                    \
  do {                                   \
    int r_ = (call);                     \
    if (r_ != GT_OK) {                   \
      std::string keep = g_err;          \
      gt_ens_destroy(e);                 \
      g_err = keep;                      \
      return r_;                         \
    }                                    \
  } while (0)
#define ENS_HIP(call)                                                      \
  do {                                                                     \
    hipError_t e_ = (call);                                                \
    if (e_ != hipSuccess) {                                                \
      gt_ens_destroy(e);                                                   \
      return fail(GT_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_));     \
    }                                                                      \
  } while (0)
  int dev = 0;
  ENS_HIP(hipGetDevice(&dev));
  ENS_HIP(hipDeviceGetAttribute(&e->ncu, hipDeviceAttributeMultiprocessorCount, dev));
  ENS_HIP(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
  int first = 0;
  for (int m = 0; m < nmem; ++m) {
    EnsMember& mb = e->mem[m];
    ENS_CHK(gt_create(n[m], params + 17 * (size_t)m, &mb.ctx));
    gt_ctx* c = mb.ctx;
    if (m == 0) e->rko = c->rko;
    if (c->rko != e->rko)
      ENS_CHK(fail(GT_ERR_ARG, "member %d: RK order %d differs from member 0 (%d)", m, c->rko,
                   e->rko));
    mb.logn = c->logn;
    mb.rc = c->rtr;
    if (c->truncate != 1 || mb.rc.kx >= 3 * (c->n / 8) || mb.rc.p >= 3 * (c->n / 8) ||
        mb.rc.q >= 3 * (c->n / 8))
      ENS_CHK(fail(GT_ERR_ARG, "member %d: ensemble members need the 2/3 rule", m));
    mb.kyp = kyp_of(mb.rc);
    ENS_CHK(rect_tables(c, mb.rc, &mb.tab));
    const size_t nc = (size_t)mb.rc.kc() * mb.kyp;
    ENS_HIP(hipMalloc(&mb.z, sizeof(double2) * nc));
    ENS_HIP(hipMalloc(&mb.znew, sizeof(double2) * nc));
    ENS_HIP(hipMemsetAsync(mb.z, 0, sizeof(double2) * nc, e->stream));
    ENS_HIP(hipMemsetAsync(mb.znew, 0, sizeof(double2) * nc, e->stream));
    for (int ty : {256, 512}) {
      const int bt = gt::ens_lines(ty, mb.logn);
      mb.ycap = std::max(mb.ycap, (size_t)ntiles_of(mb, ty) * bt * c->n);
    }
    ENS_HIP(hipMalloc(&mb.y1, sizeof(double2) * mb.ycap * 3));
    ENS_HIP(hipMalloc(&mb.y2, sizeof(double2) * mb.ycap * 2));
    ENS_HIP(hipMemsetAsync(mb.y1, 0, sizeof(double2) * mb.ycap * 3, e->stream));
    ENS_HIP(hipMemsetAsync(mb.y2, 0, sizeof(double2) * mb.ycap * 2, e->stream));
    if (c->ph.akind != 1) {  // ka table (alpv != 2), built as the y pass builds ka
      ENS_HIP(hipMalloc(&mb.ka, sizeof(double) * nc));
      ENS_HIP(hipMemsetAsync(mb.ka, 0, sizeof(double) * nc, e->stream));
      ENS_HIP(gt::launch_ka_table(c->ph, mb.rc, mb.kyp, mb.ka, e->stream));
    }
    mb.first = first;
    first += c->nf;
  }
  e->ftotal = first;
  ENS_HIP(hipMalloc(&e->descs, sizeof(gt::PassArgs) * nmem));
  ENS_HIP(hipMalloc(&e->fdesc, sizeof(gt::EnsForce) * nmem));
  ENS_HIP(hipMalloc(&e->step, sizeof(int)));
  ENS_HIP(hipMalloc(&e->flag, sizeof(int)));
  ENS_HIP(hipHostMalloc(&e->zero_host, sizeof(int), hipHostMallocDefault));
  *e->zero_host = 0;
  const size_t nph = (size_t)std::max(1, e->ftotal) * kChunk;
  for (int b = 0; b < 2; ++b) {
    ENS_HIP(hipMalloc(&e->ph_dev[b], sizeof(double2) * nph));
    ENS_HIP(hipHostMalloc(&e->ph_host[b], sizeof(double2) * nph, hipHostMallocDefault));
    ENS_HIP(hipHostMalloc(&e->fd_host[b], sizeof(gt::EnsForce) * nmem, hipHostMallocDefault));
    ENS_HIP(hipEventCreateWithFlags(&e->ph_ev[b], hipEventDisableTiming));
  }
  ENS_HIP(hipEventCreate(&e->t0));
  ENS_HIP(hipEventCreate(&e->t1));
  // default schedule: 512-thread workgroups, one workgroup per work item,
  // radix 8 first, row pairs in the x pass from N = 2048, 10 steps a graph
  int32_t cfg[GT_ENS_NCFG] = {512, 512, 0, 0, 1 << 11, 10};
  ENS_CHK(apply(e, cfg));
  ENS_HIP(hipStreamSynchronize(e->stream));
  *out = e;
  return GT_OK;
#undef ENS_CHK
#undef ENS_HIP
}

int gt_ens_set_field(gt_ens* e, int32_t m, const double* z, int32_t space) {
  GT_RANGE();
  CHK(check_member(e, m));
  if (!z) return fail(GT_ERR_ARG, "null field");
  if (space != GT_SPACE_PHYSICAL && space != GT_SPACE_SPECTRAL)
    return fail(GT_ERR_ARG, "unknown space %d", space);
  EnsMember& mb = e->mem[m];
  gt_ctx* c = mb.ctx;
  const size_t bytes = ((size_t)c->n + 2) * c->n * sizeof(double);
  double* tmp = nullptr;
  HIPCHK(hipMalloc(&tmp, bytes));
  int flag = 0;
  hipError_t err = hipMemcpyAsync(tmp, z, bytes, hipMemcpyHostToDevice, e->stream);
  if (err == hipSuccess && space == GT_SPACE_PHYSICAL)
    err = gt::launch_ref_fft(c->logn, -1, tmp, c->tw, e->stream);  // fft_dir
  if (err == hipSuccess) err = hipMemsetAsync(e->flag, 0, sizeof(int), e->stream);
  if (err == hipSuccess)  // trunc (physical input) or the rectangle as is
    err = gt::launch_gather(c->ph, mb.rc, mb.kyp, tmp, mb.z, space == GT_SPACE_PHYSICAL,
                            space == GT_SPACE_SPECTRAL ? e->flag : nullptr, e->stream);
  if (err == hipSuccess)
    err = hipMemcpyAsync(&flag, e->flag, sizeof(int), hipMemcpyDeviceToHost, e->stream);
  if (err == hipSuccess) err = hipStreamSynchronize(e->stream);
  (void)hipFree(tmp);
  if (err != hipSuccess) return fail(GT_ERR_HIP, "set_field: %s", hipGetErrorString(err));
  if (flag) return fail(GT_ERR_ARG, "member %d: spectrum has modes outside the 2/3 rectangle", m);
  return GT_OK;
}

int gt_ens_get_field(gt_ens* e, int32_t m, double* z, int32_t space) {
  GT_RANGE();
  CHK(check_member(e, m));
  if (!z) return fail(GT_ERR_ARG, "null field");
  if (space != GT_SPACE_PHYSICAL && space != GT_SPACE_SPECTRAL)
    return fail(GT_ERR_ARG, "unknown space %d", space);
  EnsMember& mb = e->mem[m];
  gt_ctx* c = mb.ctx;
  const size_t bytes = ((size_t)c->n + 2) * c->n * sizeof(double);
  double* tmp = nullptr;
  HIPCHK(hipMalloc(&tmp, bytes));
  hipError_t err = gt::launch_scatter(c->ph, mb.rc, mb.kyp, mb.z, tmp, e->stream);
  if (err == hipSuccess && space == GT_SPACE_PHYSICAL)
    err = gt::launch_ref_fft(c->logn, +1, tmp, c->tw, e->stream);  // fft_inv
  if (err == hipSuccess) err = hipMemcpyAsync(z, tmp, bytes, hipMemcpyDeviceToHost, e->stream);
  if (err == hipSuccess) err = hipStreamSynchronize(e->stream);
  (void)hipFree(tmp);
  if (err != hipSuccess) return fail(GT_ERR_HIP, "get_field: %s", hipGetErrorString(err));
  return GT_OK;
}

int gt_ens_set_seed(gt_ens* e, int32_t m, int32_t seed) {
  CHK(check_member(e, m));
  e->mem[m].seed = seed;
  return GT_OK;
}

int gt_ens_get_seed(gt_ens* e, int32_t m, int32_t* seed) {
  CHK(check_member(e, m));
  if (!seed) return fail(GT_ERR_ARG, "null seed");
  *seed = e->mem[m].seed;
  return GT_OK;
}

int gt_ens_schedule(gt_ens* e, int32_t* cfg, int32_t set) {
  if (!e || !cfg) return fail(GT_ERR_ARG, "null argument");
  if (!set) {
    std::memcpy(cfg, e->cfg, sizeof(e->cfg));
    return GT_OK;
  }
  return apply(e, cfg);
}

int gt_ens_run(gt_ens* e, int32_t nsteps) {
  GT_RANGE();
  return gt_ens_advance(e, nsteps, 0);
}

int gt_ens_advance(gt_ens* e, int32_t nsteps, int32_t hold) {
  GT_RANGE();
  if (!e) return fail(GT_ERR_ARG, "null ensemble");
  if (nsteps < 0) return fail(GT_ERR_ARG, "negative step count");
  CHK(apply_pending(e));
  if (nsteps == 0) return GT_OK;
  if (!e->gexec) CHK(capture(e));
  const int gs = e->cfg[5];
  const int cap = std::max(gs, kPipe / gs * gs);
  for (int left = nsteps; left > 0;) {
    const int ch = std::min(left, cap);
    // the held step is the chunk's last: its phases are drawn and uploaded
    // with the chunk, and the step counter points at them when it is applied
    const int full = (hold && ch == left) ? ch - 1 : ch;
    const int b = (int)(e->nchunk++ % 2);
    if (e->ph_used[b]) HIPCHK(hipEventSynchronize(e->ph_ev[b]));  // host buffer free
    draw_phases(e, b, ch);
    HIPCHK(hipMemcpyAsync(e->ph_dev[b], e->ph_host[b],
                          sizeof(double2) * (size_t)e->ftotal * ch, hipMemcpyHostToDevice,
                          e->stream));
    HIPCHK(hipMemcpyAsync(e->fdesc, e->fd_host[b], sizeof(gt::EnsForce) * e->nmem,
                          hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipMemcpyAsync(e->step, e->zero_host, sizeof(int), hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipEventRecord(e->ph_ev[b], e->stream));
    e->ph_used[b] = true;
    for (int i = 0; i < full / gs; ++i) HIPCHK(hipGraphLaunch(e->gexec, e->stream));
    CHK(launch_steps(e, full % gs, true));
    if (full < ch) {
      CHK(launch_steps(e, 1, false));
      e->pending = true;
    }
    left -= ch;
  }
  return GT_OK;
}

int gt_ens_synchronize(gt_ens* e) {
  if (!e) return fail(GT_ERR_ARG, "null ensemble");
  HIPCHK(hipStreamSynchronize(e->stream));
  return GT_OK;
}

int gt_ens_tune(gt_ens* e, int32_t steps, int32_t* cfg_out, double* ms_out) {
  GT_RANGE();
  if (!e) return fail(GT_ERR_ARG, "null ensemble");
  if (steps < 1) return fail(GT_ERR_ARG, "steps must be positive");
  CHK(apply_pending(e));
  HIPCHK(hipStreamSynchronize(e->stream));
  std::vector<double2*> backup(e->nmem, nullptr);
  int rc = GT_OK;
  for (int m = 0; m < e->nmem && rc == GT_OK; ++m) {
    const EnsMember& mb = e->mem[m];
    const size_t bytes = sizeof(double2) * mb.rc.kc() * mb.kyp;
    if (hipMalloc(&backup[m], bytes) != hipSuccess ||
        hipMemcpyAsync(backup[m], mb.z, bytes, hipMemcpyDeviceToDevice, e->stream) != hipSuccess)
      rc = fail(GT_ERR_HIP, "tune: field backup failed");
  }
  int32_t best[GT_ENS_NCFG];
  std::memcpy(best, e->cfg, sizeof(best));
  double tbest = 0.0;
  if (rc == GT_OK) rc = apply(e, best);
  if (rc == GT_OK) rc = time_steps(e, steps, backup, &tbest);
  // coordinate descent: workgroup sizes, grid, radix order, then the x-pass
  // packing of each size present
  auto trial = [&](int idx, int32_t val, int32_t mask) -> int {
    int32_t c[GT_ENS_NCFG];
    std::memcpy(c, best, sizeof(c));
    if (mask) c[idx] = (c[idx] & ~mask) | (val ? mask : 0);
    else c[idx] = val;
    if (std::memcmp(c, best, sizeof(c)) == 0) return GT_OK;
    CHK(apply(e, c));
    double t = 0.0;
    CHK(time_steps(e, steps, backup, &t));
    if (t < tbest * 0.99) {
      tbest = t;
      std::memcpy(best, c, sizeof(best));
    }
    return GT_OK;
  };
  for (int v : {256, 512}) if (rc == GT_OK) rc = trial(0, v, 0);
  for (int v : {256, 512}) if (rc == GT_OK) rc = trial(1, v, 0);
  for (int v : {0, 1, 2, 4}) if (rc == GT_OK) rc = trial(2, v, 0);
  for (int v : {0, 1}) if (rc == GT_OK) rc = trial(3, v, 0);
  for (int l = gt::kEnsMinLogn; l <= gt::kEnsMaxLogn && rc == GT_OK; ++l) {
    bool present = false;
    for (const auto& mb : e->mem) present |= mb.logn == l;
    if (present)
      for (int v : {0, 1}) if (rc == GT_OK) rc = trial(4, v, 1 << l);
  }
  // restore the fields, keep the best schedule
  for (int m = 0; m < e->nmem; ++m) {
    const EnsMember& mb = e->mem[m];
    if (backup[m] && rc == GT_OK &&
        hipMemcpyAsync(mb.z, backup[m], sizeof(double2) * mb.rc.kc() * mb.kyp,
                       hipMemcpyDeviceToDevice, e->stream) != hipSuccess)
      rc = fail(GT_ERR_HIP, "tune: field restore failed");
  }
  if (hipStreamSynchronize(e->stream) != hipSuccess && rc == GT_OK)
    rc = fail(GT_ERR_HIP, "tune: synchronize failed");
  for (double2* p : backup)
    if (p) (void)hipFree(p);
  if (rc != GT_OK) return rc;
  CHK(apply(e, best));
  if (cfg_out) std::memcpy(cfg_out, best, sizeof(best));
  if (ms_out) *ms_out = tbest;
  return GT_OK;
}

}  // extern "C"
