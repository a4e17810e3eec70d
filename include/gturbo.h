/*
 * gturbo.h -- C ABI of libgturbo_hip.so, the MI355X (gfx950) implementation
 * of Gturbo's pseudospectral time step.
 *
 * Every entry point replaces one subroutine or data clause of the reference
 * (paths relative to the reference checkout) and keeps its argument meaning:
 * arrays are passed by the address of the caller's host array, exactly as
 * Fortran passes them to an external subroutine with an implicit interface,
 * and the library finds the device copy in a present table keyed by that
 * address -- the role OpenACC's present table plays in the reference
 * (src/main.f90:67-83).  A field array holds (N+2)*N doubles: the Fortran
 * array z(N+2,N) in physical space, or complex z(N/2+1,N) in spectral space
 * (src/modules.f90:4-9).
 *
 * Device copies of spectral fields may be kept in an internal layout (the
 * 2/3-truncated rectangle, ky fastest); they are converted back to the
 * reference layout by gt_update_host / gt_exit_data / gt_sync_ref and by the
 * physical-space entry points (gt_fft_inv, gt_produ), never silently lost.
 *
 * Every function returns 0 on success or a GT_ERR_* code; gt_last_error()
 * gives the message.  A context is used by one host thread; all device work
 * of a context runs in order on the context's own HIP stream.
 */
#ifndef GTURBO_H
#define GTURBO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GT_OK 0
#define GT_ERR_HIP 1         /* a HIP runtime call failed */
#define GT_ERR_ARG 2         /* invalid argument or size */
#define GT_ERR_NOT_PRESENT 3 /* array has no device copy (gt_enter_data) */
#define GT_ERR_NFORC 4       /* more than NFORC forced modes */
#define GT_ERR_NOGPU 5       /* no usable gfx950 device */

#define GT_PLAN_D2Z 1 /* cufftPlan2D(plan_dir,NY,NY,CUFFT_D2Z), main.f90:32 */
#define GT_PLAN_Z2D 2 /* cufftPlan2D(plan_inv,NY,NY,CUFFT_Z2D), main.f90:27 */

typedef struct gt_ctx gt_ctx;

/* --- set-up ------------------------------------------------------------ */

/* Replaces src/inieuler.f90:13-88 (split and RK-order rules, dt/2, dt/3,
 * dt/6, sqrt(dt)) with src/defk.f90 and src/iniforcing.f90, plus the
 * streams of src/main.f90:28,33.  params holds the params.dat values in file
 * order: xlx xly nu alpnu mu alpmu alpv dt rko npas iout imix rtrunc
 * truncate famp k1f k2f.  n is the compile-time NX=NY of src/modules.f90:4,
 * a power of two in [16, 8192]. */
int gt_create(int32_t n, const double *params, gt_ctx **ctx);
/* Replaces src/main.f90:185-188 (plans and streams destroyed). */
int gt_destroy(gt_ctx *ctx);
const char *gt_last_error(void);

/* Scalars after set-up.  iinfo: nf split rko truncate kx p q ;
 * dinfo: inpz inpe dt2 dt3 dt6 sdt dkx dky rtrunc^2 fkxsmax fkysmax. */
int gt_info(const gt_ctx *ctx, int32_t *iinfo, double *dinfo);
/* Host copies of the src/defk.f90 tables (jc is 1-based, emk is
 * (N/2+1)*N in Fortran order). */
int gt_get_defk(const gt_ctx *ctx, double *fkx, double *fky, double *fkxs,
                double *fkys, int32_t *jc, double *emk, double *emkx,
                double *emky);
/* The src/iniforcing.f90 table (1-based ikxf, ikyf; ff(2,nf)). */
int gt_get_forcing_modes(const gt_ctx *ctx, int32_t *ikxf, int32_t *ikyf,
                         double *ff);
/* Replaces cufftPlan2D (src/main.f90:27,32): kind GT_PLAN_D2Z/Z2D. */
int gt_plan_create(gt_ctx *ctx, int32_t kind, int32_t *plan);
/* Replaces show_mem's cudaMemGetInfo (src/trunc.f90:32-49). */
int gt_mem_info(size_t *free_bytes, size_t *total_bytes);

/* --- data clauses (src/main.f90:67-83,112,125,142-145,164-179) --------- */

/* !$acc enter data copyin(a) / create(a): copyin = 1 copies the host array. */
int gt_enter_data(gt_ctx *ctx, void *host, int64_t bytes, int32_t copyin);
/* !$acc exit data copyout(a) / delete(a) finalize: copyout = 1 copies back. */
int gt_exit_data(gt_ctx *ctx, void *host, int32_t copyout);
/* !$acc update host(a) / update device(a). */
int gt_update_host(gt_ctx *ctx, void *host);
int gt_update_device(gt_ctx *ctx, void *host);
/* acc_map_data: attach caller-owned device memory (e.g. a torch tensor) as
 * the device copy of key; gt_sync_ref makes that memory current in the
 * reference layout. */
int gt_map_data(gt_ctx *ctx, void *key, void *dev, int64_t bytes);
int gt_unmap_data(gt_ctx *ctx, void *key);
int gt_sync_ref(gt_ctx *ctx, void *key);
/* Device-side array copy dst = src, as !$acc kernels present(z,znl);
 * znl=z (src/main.f90:107-109,120-122). */
int gt_copy(gt_ctx *ctx, void *dst, const void *src);
/* Waits for the context's stream. */
int gt_synchronize(gt_ctx *ctx);

/* --- the hot path ------------------------------------------------------ */

/* src/fft_dir.f90:1-23: in-place D2Z of z, then z/NY/NX. */
int gt_fft_dir(gt_ctx *ctx, double *z, int32_t plan);
/* src/fft_inv.f90:1-15: in-place Z2D, unnormalised. */
int gt_fft_inv(gt_ctx *ctx, double *z, int32_t plan);
/* src/produ.f90:1-25: u = z*u, v = z*v; padding rows zeroed. */
int gt_produ(gt_ctx *ctx, const double *z, double *u, double *v);
/* src/trunc.f90:1-28: 2/3 dealiasing mask and z(1,1) = 0. */
int gt_trunc(gt_ctx *ctx, double *z);
/* src/nlt.f90:1-65: z <- N(z) = -ik.(q u)^, truncated when truncate = 1.
 * u and v are the reference's work arrays; their contents afterwards are
 * not specified. */
int gt_nlt(gt_ctx *ctx, double *z, double *u, double *v);
/* src/rkstep.f90:2-76: integrating-factor RK2 (split or not).  znl, u, v
 * are work arrays. */
int gt_step(gt_ctx *ctx, double *z, double *znl, double *u, double *v);
/* src/rkstep.f90:82-131: Lawson RK4.  znew receives the stage accumulator
 * (src/rkstep.f90:96-114); znl, u, v are work arrays.
 *
 * gt_step and gt_step4 leave their last stage (the update of z) queued in
 * the context: when the next call on the context is gt_forcing(z) and then
 * the same step on the same arrays -- the driver's loop, src/main.f90:97-136
 * -- that stage, the forcing and the next step's first pass run as one pass
 * over the fields.  Any other entry point runs the queued work first, so
 * every result a caller can observe is the same as with the three calls run
 * one after the other. */
int gt_step4(gt_ctx *ctx, double *z, double *znl, double *znew, double *u,
             double *v);
/* src/forcing.f90:1-51: adds one draw of the ring forcing and advances the
 * LCG seed (the seed is kept on the host; the draws are replayed there). */
int gt_forcing(gt_ctx *ctx, double *z, int32_t *seed);

/* --- diagnostics (src/diagnostics.f90) --------------------------------- */

/* spectrum (:5-56): sp, sz of NBIN = N/2 shells, already divided by dk. */
int gt_spectrum(gt_ctx *ctx, const double *z, double *sp, double *sz);
/* fluxes (:62-119): fle, flz (suffix sums over shells). */
int gt_fluxes(gt_ctx *ctx, const double *z, const double *znl, double *fle,
              double *flz);
/* glob (:125-179): out = ee zz disz_nu disz_mu dise_nu dise_mu. */
int gt_glob(gt_ctx *ctx, const double *z, double *out);
/* No counterpart in the reference: gt_glob for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The same six sums, written to out_dev
 * (6 doubles of device memory) asynchronously on the context's stream
 * (gt_synchronize waits for them); nothing is allocated, copied to the host
 * or synchronised.  The partial sums live in a static device array of
 * 6*(N/2+1)+6 doubles for N <= 8192 (196,704 bytes per GPU, held until the
 * process ends), shared by every context on that GPU: calls that overlap on
 * one GPU race.  z must be present; its current device copy (compact or
 * reference layout) is read as it is. */
int gt_glob_noalloc(gt_ctx *ctx, const double *z, double *out_dev);
/* Shell index int(k/dk)+1 of every half-spectrum mode as the device computes
 * it (diagnostics.f90:31-34), in Fortran order; 0 for the mean mode. */
int gt_shell_index(gt_ctx *ctx, int32_t *ib);

/* --- measurement ------------------------------------------------------- */

/* Kernel classes timed with HIP events on the context's stream. */
#define GT_K_YSTART 0 /* y pass: inverse part only */
#define GT_K_YSTAGE 1 /* y pass: forward + stage update + inverse */
#define GT_K_YFINAL 2 /* y pass: forward + final update */
#define GT_K_XPASS 3  /* x pass */
#define GT_K_FORCING 4
#define GT_K_OTHER 5
#define GT_K_YBOUND 6 /* y pass: a step's last stage + its forcing + the next
                         step's inverse part (see gt_step4) */
#define GT_K_NCLASS 7
int gt_timing_enable(gt_ctx *ctx, int32_t on);
/* Which classes gt_timing_enable times: bit (1 << GT_K_*) set = timed
 * (default: all).  Each timed launch records two events on the stream;
 * timing only the class of interest keeps the others' launches back to
 * back. */
int gt_timing_classes(gt_ctx *ctx, int32_t mask);
/* Sums of event-timed kernel durations per class since the last read (ms)
 * and launch counts; resets the sums. */
int gt_timing_read(gt_ctx *ctx, double *ms, int32_t *count);

/* --- adjoint -------------------------------------------------------------
 * Gradients with respect to the initial field, for initial-condition
 * inversion.  The reference has no adjoint: these entries differentiate the
 * run of src/main.f90:89-136 -- fft_dir and trunc of the input frame, then
 * nsteps x (step or step4, forcing), then fft_inv -- through hand-written
 * transposes of src/nlt.f90 (transposed transforms, adjoint products and
 * alpha operator) and of the RK stages of src/rkstep.f90 (adjoint
 * integrating factors).  Needs the 2/3 rule (truncate = 1).  Fields are
 * present arrays of (N+2)*N doubles in physical space. */

/* Runs the forward model from w0 for nsteps steps (forcing from *seed, which
 * advances), writes the final field to wT and keeps a checkpoint of the
 * state every kchk steps on the device for gt_adj_backward. */
int gt_adj_forward(gt_ctx *ctx, const double *w0, double *wT, int32_t nsteps,
                   int32_t kchk, int32_t *seed);
/* Given gT = dL/dwT, the gradient g0 = dL/dw0 for the run of the last
 * gt_adj_forward: each checkpoint segment is recomputed with its stage
 * states kept, then stepped backwards. */
int gt_adj_backward(gt_ctx *ctx, const double *gT, double *g0);

/* --- batched ensembles ---------------------------------------------------
 * The reference runs an ensemble as copies of one run directory, each
 * submitted as its own job (src/reset.sh:12-77, src/jobscript:14), each job
 * the main loop of src/main.f90:97-137.  Here the members -- each with its
 * own params.dat values, N (128 to 2048), field and forcing seed -- advance
 * together: every pass of the RK step is one launch for all members, and the
 * whole step (passes, then the forcing of every member, src/main.f90:99-136
 * without the diagnostics) is captured in a HIP graph and replayed.  All
 * members use the same RK order and the 2/3 rule (truncate = 1). */
typedef struct gt_ens gt_ens;

/* Field spaces of gt_ens_set_field / gt_ens_get_field. */
#define GT_SPACE_PHYSICAL 0 /* z(N+2,N) in physical space (src/read.f90) */
#define GT_SPACE_SPECTRAL 1 /* complex z(N/2+1,N), reference layout */

/* Schedule of the batched passes (gt_ens_schedule, gt_ens_tune):
 * [0] threads per y-pass workgroup (256 or 512; a workgroup holds that many
 *     times 8/N kx-columns), [1] threads per x-pass workgroup (256 or 512),
 * [2] waves per SIMD of the persistent grid (1, 2 or 4; 0 = one workgroup
 *     per work item), [3] radix order of the transforms (0 = radix 8 first,
 *     1 = the radix-2/4 pass first), [4] bit L set: members with N = 2^L
 *     use the row-pair x pass, [5] RK steps per captured graph (1-256). */
#define GT_ENS_NCFG 6

/* nmem members; n[m] = N of member m; params holds nmem rows of the 17
 * params.dat values of gt_create.  Members start from a zero field and the
 * forcing seed 481929 (src/seed.0). */
int gt_ens_create(int32_t nmem, const int32_t *n, const double *params, gt_ens **ens);
int gt_ens_destroy(gt_ens *ens);
/* Member m's field from a host array of (N+2)*N doubles.  GT_SPACE_PHYSICAL
 * applies fft_dir and trunc as src/main.f90:89-92 does to the input frame;
 * GT_SPACE_SPECTRAL takes the spectrum as is (it must vanish outside the
 * 2/3 rectangle). */
int gt_ens_set_field(gt_ens *ens, int32_t m, const double *z, int32_t space);
/* Member m's field into a host array of (N+2)*N doubles (GT_SPACE_PHYSICAL:
 * after fft_inv, as src/main.f90:143 before writef). */
int gt_ens_get_field(gt_ens *ens, int32_t m, double *z, int32_t space);
/* The forcing LCG seed of member m (src/forcing.f90:40-51). */
int gt_ens_set_seed(gt_ens *ens, int32_t m, int32_t seed);
int gt_ens_get_seed(gt_ens *ens, int32_t m, int32_t *seed);
/* set = 0: read the schedule into cfg; set = 1: use cfg. */
int gt_ens_schedule(gt_ens *ens, int32_t *cfg, int32_t set);
/* Autotuner: times candidate schedules over `steps` RK steps of this
 * ensemble (fields and seeds are left as they were), keeps the fastest and
 * returns it in cfg with its time per RK step (ms, forcing excluded). */
int gt_ens_tune(gt_ens *ens, int32_t steps, int32_t *cfg, double *ms_per_step);
/* nsteps x (RK step of every member, then its forcing).  Asynchronous. */
int gt_ens_run(gt_ens *ens, int32_t nsteps);
/* As gt_ens_run; with hold != 0 the forcing of the last of the nsteps steps
 * is held back, so that the fields show the state after that step's RK
 * update and before its forcing -- where the reference's driver takes its
 * diagnostics and snapshots (src/main.f90:97-136).  A held forcing is applied
 * first by the next gt_ens_advance / gt_ens_run (nsteps = 0 applies it
 * alone) or gt_ens_tune. */
int gt_ens_advance(gt_ens *ens, int32_t nsteps, int32_t hold);
int gt_ens_synchronize(gt_ens *ens);

#ifdef __cplusplus
}
#endif
#endif
