// gt_runtime.hpp -- host-side state shared by the runtime translation units
// (single-run context, batched ensemble, adjoint): the context, the present
// table entries and the helpers that act on them.
#pragma once
#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <string>
#include <unordered_map>
#include <vector>

#include "gt_internal.hpp"
#include "gturbo.h"

using gt::ForcedMode;
using gt::Phys;
using gt::Rect;

namespace gtr {

extern thread_local std::string g_err;

int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));

#define HIPCHK(call)                                                         \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess)                                                    \
      return fail(GT_ERR_HIP, "%s: %s (%s:%d)", #call, hipGetErrorString(e_), \
                  __FILE__, __LINE__);                                       \
  } while (0)

#define CHK(call)            \
  do {                       \
    int r_ = (call);         \
    if (r_ != GT_OK) return r_; \
  } while (0)

constexpr int kNFORC = 2048;  // src/modules.f90:41
constexpr int kRing = 64;     // forcing-phase slots in flight

struct Mirror {
  void* host = nullptr;
  size_t bytes = 0;
  double* dref = nullptr;   // reference layout
  bool owned = true;
  double2* dcmp = nullptr;  // compact layout
  size_t cmp_cap = 0;       // double2 capacity of dcmp
  Rect rect{0, 0, 0};
  int kyp = 0;
  bool ref_valid = true;
  bool cmp_valid = false;
};

struct RectTables {
  Rect rc{-1, -1, -1};
  int kyp = 0;
  double* emk = nullptr;
  ForcedMode* modes = nullptr;
  // the forced modes by kx column (fused step boundary): column c's updates
  // are fce[fcs[c] .. fcs[c+1]-1]
  int* fcs = nullptr;
  gt::FEntry* fce = nullptr;
};

// The last stage of a step left for the next step to fuse with its first
// pass (the driver's step, forcing, step sequence): kind GT_PEND_RK4 or
// GT_PEND_RK2, the state and accumulator arrays and the pass arguments; with
// forcing = true the step's forcing draws are already in slot `slot`.
enum : int { PEND_NONE = 0, PEND_RK4 = 1, PEND_RK2 = 2 };
struct Pending {
  int kind = PEND_NONE;
  Mirror* z = nullptr;
  Mirror* znew = nullptr;
  RectTables* t = nullptr;
  gt::PassArgs a{};
  bool forced = false;
  unsigned slot = 0;
};

struct TimedLaunch {
  int cls;
  hipEvent_t a, b;
};

}  // namespace gtr

struct AdjState;  // gturbo_adjoint.cpp

struct gt_ctx {
  int n = 0, logn = 0, nh = 0;
  // commphy / commsim / commforc / commk (src/modules.f90)
  double xlx, xly, dkx, dky, nu, mu, alpnu, alpmu, alpv;
  double dt, dt2, dt3, dt6, sdt, inpz = 0, inpe = 0;
  double famp, k1f, k2f, rtrunc, fkxsmax, fkysmax;
  int npas, iout, imix, split, rko, truncate;
  int nf = 0;
  std::vector<int32_t> ikxf, ikyf, jc;
  std::vector<double> ff, fkx, fky, fkxs, fkys, emk, emkx, emky;

  Phys ph{};
  Rect rtr{0, 0, 0}, rfull{0, 0, 0};

  hipStream_t stream = nullptr;
  double2* tw = nullptr;
  gtr::RectTables tab[2];  // [0] truncated rect, [1] full rect
  double2* y1 = nullptr;
  double2* y2 = nullptr;
  size_t ycap = 0;  // double2 per field
  int* flag = nullptr;

  // forcing phases
  double2* ph_host = nullptr;  // pinned, kRing * nf
  double2* ph_dev = nullptr;
  hipEvent_t ph_ev[gtr::kRing];
  bool ph_ev_used[gtr::kRing] = {};
  unsigned ph_next = 0;

  // diagnostics
  double* dg = nullptr;
  size_t dg_cap = 0;
  // device address of the static partial sums of gt_glob_noalloc on this
  // context's device, and their count in doubles
  double* glob_scratch = nullptr;
  size_t glob_scratch_n = 0;

  // timing
  bool timing = false;
  unsigned timing_mask = ~0u;  // classes timed (bit 1 << GT_K_*)
  std::vector<gtr::TimedLaunch> tl;
  std::vector<hipEvent_t> ev_pool;

  std::unordered_map<const void*, gtr::Mirror> present;

  AdjState* adj = nullptr;  // checkpoints and buffers of gt_adj_*

  gtr::Pending pend;  // deferred end of the last step (fused into the next)
};

namespace gtr {

int ilog2(int n);
int kyp_of(const Rect& r);
int rect_kind(const gt_ctx* c, const Rect& r);
int rect_tables(gt_ctx* c, const Rect& r, RectTables** out);
int ensure_ybuf(gt_ctx* c, const Rect& r);
int lookup(gt_ctx* c, const void* host, Mirror** m);
int alloc_cmp(gt_ctx* c, Mirror& m, const Rect& r);
int ensure_ref(gt_ctx* c, Mirror& m);
int ensure_cmp(gt_ctx* c, Mirror& m);
void set_cmp_only(Mirror& m);
void set_ref_only(Mirror& m);
gt::PassArgs pass_args(gt_ctx* c, Mirror& z, RectTables* t);
int c_pass(gt_ctx* c, int mode, const gt::PassArgs& a);
int r_pass(gt_ctx* c, const gt::PassArgs& a);
void forcing_phases(int nf, int32_t* seed, double2* out);
int forcing_on(gt_ctx* c, RectTables* t, double2* z, int32_t* seed);
void adj_free(gt_ctx* c);  // gturbo_adjoint.cpp
// Runs the deferred end of the last step (its last stage and forcing), if
// any: every entry point that reads or writes device data calls it first.
int flush_pending(gt_ctx* c);

// roctx range over one C entry point (rocprofv3 --marker-trace shows them;
// without a profiler attached they cost a call each).
struct Range {
  explicit Range(const char* name) { roctxRangePushA(name); }
  ~Range() { roctxRangePop(); }
  Range(const Range&) = delete;
  Range& operator=(const Range&) = delete;
};
#define GT_RANGE() gtr::Range gt_range_(__func__)

// Timer of one launch class on the context's stream (GT_K_* of gturbo.h).
struct Timer {
  gt_ctx* c;
  int cls;
  hipEvent_t a = nullptr;
  Timer(gt_ctx* c_, int cls_);
  ~Timer();
};

}  // namespace gtr
