/*
 * tsphere.h — C ABI of the MI355X (gfx950) spectral IMEX library.
 *
 * This is the drop-in boundary for the reference's hot path: the spherical-
 * harmonic transform plan of src/sht/sht.ts and the element-wise-kernel /
 * step-recording backend of src/mgpu/{plan,wgsl,model}.ts.  Everything is
 * plain C: integer status codes (TSPHERE_OK = 0, never abort), opaque
 * handles, caller-owned device buffers, plan-owned tables and scratch, and
 * no implicit synchronisation except in the functions documented as blocking
 * (the host<->device reads, as in the reference, whose only await is the
 * readback, src/mgpu/model.ts:261).
 *
 * Threading: one host thread drives a plan; all work of a plan is ordered on
 * the stream passed to each call.  Calls that overlap on one plan from
 * different streams race on its scratch (the reference's shared fmBuf has the
 * same rule, src/sht/sht.ts:291-298).
 *
 * Spectral layout (bit-exact with src/sht/layout.ts:1-32): complex Q_lm for
 * m >= 0 stored as interleaved float pairs [re, im], m-major (for m = 0..mmax,
 * for l = m..lmax), index lm = m(lmax+1) - m(m-1)/2 + (l-m).  Spatial layout:
 * float spat[ilat*nphi + iphi], ilat north -> south (Gauss nodes decreasing),
 * phi_j = 2 pi j / nphi.  Orthonormal harmonics with Condon-Shortley phase.
 */
#ifndef TSPHERE_H
#define TSPHERE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
#define TSPHERE_OK 0
#define TSPHERE_ERR_INVALID 1     /* bad argument or configuration            */
#define TSPHERE_ERR_HIP 2         /* HIP runtime error (message has details)  */
#define TSPHERE_ERR_COMPILE 3     /* hipRTC compile failure (log in message)  */
#define TSPHERE_ERR_UNSUPPORTED 4 /* valid request the library cannot serve    */
#define TSPHERE_ERR_NODEVICE 5    /* no GPU visible                            */

/* Library version, as 10000*major + 100*minor + patch. */
int tsphere_version(void);
/* Text of the last error raised on the calling thread ("" if none). */
const char *tsphere_last_error(void);
/* Number of visible GPUs (0 without a device; never an error). */
int tsphere_device_count(int *count);

/* ------------------------------------------------------- device plumbing */
typedef void *tsphere_stream; /* a hipStream_t; NULL = the null stream */
typedef struct tsphere_event_t *tsphere_event;

int tsphere_stream_create(tsphere_stream *out); /* default priority only */
int tsphere_stream_destroy(tsphere_stream s);
int tsphere_stream_sync(tsphere_stream s);      /* blocking */
int tsphere_device_sync(void);                  /* blocking */

int tsphere_malloc(size_t bytes, void **out);
int tsphere_free(void *ptr);
/* Host -> device, enqueued on s (the host buffer must stay alive until s
 * reaches the copy; tsphere_stream_sync to be sure). */
int tsphere_memcpy_htod(void *dst, const void *src, size_t bytes, tsphere_stream s);
/* Device -> host: blocking (waits for s, then copies). */
int tsphere_memcpy_dtoh(void *dst, const void *src, size_t bytes, tsphere_stream s);
int tsphere_memcpy_dtod(void *dst, const void *src, size_t bytes, tsphere_stream s);
/* dst[i] = value for i < count, by a kernel (safe inside graph capture). */
int tsphere_fill_f32(float *dst, float value, long long count, tsphere_stream s);

int tsphere_event_create(tsphere_event *out);
int tsphere_event_destroy(tsphere_event e);
int tsphere_event_record(tsphere_event e, tsphere_stream s);
/* Make later work on `s` wait for `e` (no host block).  Inside a capture this
 * forks/joins streams into one graph (the reference has one queue and needs
 * no counterpart). */
int tsphere_stream_wait(tsphere_stream s, tsphere_event e);
/* Milliseconds between two recorded events (blocking on `end`). */
int tsphere_event_elapsed(tsphere_event start, tsphere_event end, float *ms);

/* -------------------------------------------------- the transform plan
 * Replaces ShtPlan (src/sht/sht.ts:116-469).  The plan computes its tables in
 * f64 on the host (Gauss nodes/weights gauss.ts:11-51, recurrence
 * coefficients coeffs.ts:28-56, the seed ytilde_m^m = amm sin^m(theta) with a
 * scaled exponent, the twiddle table sht.ts:200-205) and keeps them on the
 * device as f32, plus a scratch Fourier buffer for `max_fields` fields.
 */
#define TSPHERE_FOURIER_AUTO 0 /* FFT when nphi is a power of two >= 8, else DFT */
#define TSPHERE_FOURIER_FFT 1
#define TSPHERE_FOURIER_DFT 2

#define TSPHERE_LEG_FP32 0   /* Legendre stage on the FP32 VALU                */
#define TSPHERE_LEG_BF16X3 1 /* Legendre stage on MFMA, split-BF16 x3, FP32 acc */
#define TSPHERE_LEG_F64REC 2 /* FP64 recurrence, FP32 sums on the VALU          */

typedef struct {
  int lmax, mmax, nlat, nphi; /* validated as layout.ts:34-43               */
  int fourier;                /* TSPHERE_FOURIER_*  (ShtOptions.fourier)     */
  int legendre;               /* TSPHERE_LEG_*                               */
  int max_fields;             /* fields per batched call (>= 1)             */
} tsphere_sht_config;

typedef struct {
  int lmax, mmax, nlat, nphi;
  long long nlm;      /* layout.ts:24-27 */
  int fourier;        /* resolved: TSPHERE_FOURIER_FFT or _DFT (ShtPlan.fourierMode) */
  int legendre;
  int leg_lat;        /* latitudes one Legendre lane walks: ceil(nlat/2) (ShtPlan.legLat) */
  int max_fields;
  size_t scratch_bytes; /* plan-owned device memory: tables + Fourier scratch */
} tsphere_plan_info;

typedef struct tsphere_plan_t *tsphere_plan;

/* ShtPlan.create (sht.ts:166-183): validates (TSPHERE_ERR_INVALID with the
 * reference's message), resolves the Fourier mode (an explicit FFT request
 * that cannot be met is TSPHERE_ERR_INVALID, sht.ts:173-178). */
int tsphere_plan_create(const tsphere_sht_config *cfg, tsphere_plan *out);
int tsphere_plan_destroy(tsphere_plan p); /* ShtPlan.destroy (sht.ts:463-468) */
int tsphere_plan_query(tsphere_plan p, tsphere_plan_info *info);
/* ShtPlan.theta / cosTheta / gaussWeights (sht.ts:122-125): f64, nlat each;
 * any pointer may be NULL. */
int tsphere_plan_grid(tsphere_plan p, double *theta, double *cos_theta, double *weights);

/* The plan's f32 device tables, computed on the host without a device (for
 * inspection and tests).  Any pointer may be NULL.  Sizes: ab 2*nlm,
 * amm mmax+1, ctstw 3*nlat ([cos | sin | w*2pi/nphi], sht.ts:194-199),
 * seed_y0/seed_y1/seed_l (mmax+1)*ceil(nlat/2): for m and latitude pair j,
 * the recurrence starts at l = seed_l (the first l where |ytilde_l^m| >
 * 1e-15, rounded down to m + 4k; INT_MAX if none up to lmax) from the f64
 * values ytilde_l = seed_y0, ytilde_{l+1} = seed_y1; trig 2*nphi. */
int tsphere_plan_tables(const tsphere_sht_config *cfg, float *ab, float *amm, float *ctstw,
                        float *seed_y0, float *seed_y1, int *seed_l, float *trig);

/* synth / analys on caller-owned device buffers (createSynthBinding +
 * encodeSynthInto, sht.ts:299-338; createAnalysBinding + encodeAnalysInto,
 * sht.ts:312-353).  qlm: 2*nlm floats; spat: nlat*nphi floats.  Enqueued on
 * s, no synchronisation. */
int tsphere_synth(tsphere_plan p, const float *qlm, float *spat, tsphere_stream s);
int tsphere_analys(tsphere_plan p, const float *spat, float *qlm, tsphere_stream s);

/* Batched: nfield <= max_fields fields, field f at qlm + f*qlm_stride and
 * spat + f*spat_stride (strides in floats).  One launch per stage for all
 * fields, so the Legendre recurrence is shared by every field. */
int tsphere_synth_batch(tsphere_plan p, int nfield, const float *qlm, long long qlm_stride,
                        float *spat, long long spat_stride, tsphere_stream s);
int tsphere_analys_batch(tsphere_plan p, int nfield, const float *spat, long long spat_stride,
                         float *qlm, long long qlm_stride, tsphere_stream s);

/* tsphere_synth_batch for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * allocation).  No counterpart in the reference.  The same launches, with the
 * Fourier-space intermediate in the library's static device scratch (4 Mi
 * complex values = 32 MiB per GPU, held until the process ends) instead of
 * the plan's; the call allocates nothing.  Limit: nfield*(mmax+1)*nlat <=
 * 4194304 (e.g. 4 fields at lmax 1023 on 1024 latitudes), else
 * TSPHERE_ERR_INVALID; nfield is not bounded by max_fields.  The plan's
 * device must be current.  Every call on one GPU shares its scratch: calls
 * that overlap on one GPU (different streams, or a graph replaying beside a
 * direct call) race; the scratch is fully overwritten by each call, so
 * nothing needs clearing between calls. */
int tsphere_synth_batch_noalloc(tsphere_plan p, int nfield, const float *qlm,
                                long long qlm_stride, float *spat, long long spat_stride,
                                tsphere_stream s);

/* Batched over a pointer list: field f reads qlm[f] / spat[f], nfield <= 8
 * (and <= max_fields).  This is how a step that transforms several separate
 * buffers -- synth(U) and synth(V) of a two-species model, whose two
 * encodeSynthInto calls the reference records back to back (plan.ts:464-469)
 * -- runs as one launch per stage. */
#define TSPHERE_MAX_MULTI 8
int tsphere_synth_multi(tsphere_plan p, int nfield, const float *const *qlm, float *const *spat,
                        tsphere_stream s);
int tsphere_analys_multi(tsphere_plan p, int nfield, const float *const *spat, float *const *qlm,
                         tsphere_stream s);

/* Single stages (ShtPlan.encodeStage, sht.ts:368-401; what
 * scripts/diagnose-sht.ts reads to tell the halves apart).  fm is the
 * Fourier-space intermediate, complex [(f*nlat + ilat)*(mmax+1) + m] as
 * interleaved floats: row-major per latitude, the transpose of the
 * reference's fmBuf ([m*nlat + ilat], sht.ts:291-298), so that the FFT moves
 * contiguous rows. */
int tsphere_leg_synth(tsphere_plan p, int nfield, const float *qlm, long long qlm_stride,
                      float *fm, tsphere_stream s);
int tsphere_fourier_synth(tsphere_plan p, int nfield, const float *fm, float *spat,
                          long long spat_stride, tsphere_stream s);
int tsphere_fourier_analys(tsphere_plan p, int nfield, const float *spat,
                           long long spat_stride, float *fm, tsphere_stream s);
int tsphere_leg_analys(tsphere_plan p, int nfield, const float *fm, float *qlm,
                       long long qlm_stride, tsphere_stream s);

/* The two Legendre stages over a pointer list of spectral fields (field f at
 * qlm[f]) and a caller-owned Fourier scratch fm laid out as above, nfield
 * fields.  With tsphere_plan_fourier_table and tsphere_kernel_launch_grid
 * they let a generated grid kernel take the place of the Fourier stages and
 * the element-wise statements between them (the planner's grid fusion of the
 * reference's encodeSynthInto -> element-wise dispatches -> encodeAnalysInto
 * sequence, plan.ts:436-478); no reference counterpart of their own. */
int tsphere_leg_synth_multi(tsphere_plan p, int nfield, const float *const *qlm, float *fm,
                            tsphere_stream s);
int tsphere_leg_analys_multi(tsphere_plan p, int nfield, const float *fm, float *const *qlm,
                             tsphere_stream s);
/* The same for nptr * per fields addressed as tsphere_synth_fields (field f
 * at qlm[f / per] + (f % per) * qlm_stride floats; Fourier-scratch field f):
 * an ensemble's grid ops. */
int tsphere_leg_synth_fields(tsphere_plan p, int nptr, int per, const float *const *qlm,
                             long long qlm_stride, float *fm, tsphere_stream s);
int tsphere_leg_analys_fields(tsphere_plan p, int nptr, int per, const float *fm,
                              float *const *qlm, long long qlm_stride, tsphere_stream s);
/* The implicit step folded into the next synthesis: state field f is
 * replaced by X_new = (X + s1[f] A_f) / (1 + s2[f] lam) per real coefficient
 * (A_f = analysed[f], rounded as written, as tsphere_leg_spectral_imex), and
 * the new state synthesised into fm (as tsphere_leg_synth_multi): the
 * implicit-solve dispatch and the next step's encodeSynthInto (plan.ts:
 * 436-478) as one launch.  nfield <= 4; split-BF16 plans up to 512
 * latitude pairs (TSPHERE_ERR_INVALID above). */
int tsphere_leg_imex_synth(tsphere_plan p, int nfield, const float *const *analysed,
                           float *const *state, const float *lam, const float *s1,
                           const float *s2, float *fm, tsphere_stream s);
/* tsphere_leg_imex_synth for a batch (an ensemble's runs): field f = pointer
 * f / per, copy f % per at qlm_stride floats (as tsphere_leg_synth_fields),
 * the same layout for analysed and state; coef is a DEVICE array of 2 nptr
 * per floats, (s1, s2) of field f at coef[2 f], coef[2 f + 1], so a graph
 * that captured this call takes new scalars from an upload.  Any field
 * count up to max_fields; split-BF16 plans up to 512 latitude pairs.
 * No reference counterpart of its own (it fuses the implicit-solve dispatch
 * and encodeSynthInto of plan.ts:436-478 for a batch of runs). */
int tsphere_leg_imex_synth_fields(tsphere_plan p, int nptr, int per,
                                  const float *const *analysed, float *const *state,
                                  long long qlm_stride, const float *lam, const float *coef,
                                  float *fm, tsphere_stream s);
/* The implicit step behind the analysis, for a batch: the Legendre analysis
 * of the scratch fm (as tsphere_leg_analys_fields), then state field f
 * replaced by X_new = (X + s1 A_f) / (1 + s2 lam) per real coefficient, A_f
 * its analysed values rounded as written, (s1, s2) = coef[2 f], coef[2 f + 1]
 * (device array, as tsphere_leg_imex_synth_fields); field f = pointer f / per,
 * copy f % per at qlm_stride floats for analysed and state alike.  With the
 * 4x4-block or per-parity product analysis tiles (impl 3, 5) of the FP32 and
 * FP64-recurrence forms the step runs in the analysis kernel and analysed is
 * not written; other tiles write analysed and run the step as a second
 * kernel (analysed and state must not overlap).  Any field
 * count up to max_fields.  No reference
 * counterpart of its own (it fuses encodeAnalysFrom and the implicit-solve
 * dispatch of plan.ts:436-478 for a batch of runs). */
int tsphere_leg_analys_imex_fields(tsphere_plan p, int nptr, int per, const float *fm,
                                   float *const *analysed, float *const *state,
                                   long long qlm_stride, const float *lam, const float *coef,
                                   tsphere_stream s);
/* The spectral op: per m-row, the Legendre analysis of the scratch fm
 * (nfield fields, as tsphere_leg_analys_multi), the diagonal implicit step
 *   X_new = (X + s1[f] A_f) / (1 + s2[f] lam)     per real coefficient
 * (bench/shtns/spec.h:166-169; the last lines of the models' .m files,
 * Un = (U + dt*analys(..))
 * ./ (1 + (dt*D)*lam), rounded as written) of state field species[f] in
 * place, and the Legendre synthesis of the new state into fm (field slot =
 * state index), ready for the next step's grid kernel: the analysis of step
 * n and the synthesis of step n+1 share one generation of the recurrence.
 * A(l,m) below every lane's significance point is 0, as in the analysis.
 * FP32 and FP64-recurrence plans up to 1024 latitude pairs, split-BF16
 * plans up to 512 (TSPHERE_ERR_INVALID above); nfield <= 4.
 * No reference counterpart of its own (it fuses encodeAnalysInto, the
 * implicit-solve dispatch and encodeSynthInto, plan.ts:436-478). */
int tsphere_leg_spectral_imex(tsphere_plan p, int nfield, float *fm, float *const *state,
                              const float *lam, const float *s1, const float *s2,
                              const int *species, tsphere_stream s);
/* The plan's device twiddle table, which the row FFT of a generated grid
 * kernel must read to round as the plan's Fourier stages do: the nphi
 * entries (cos, sin)(2 pi k / nphi) (2*nphi floats), followed, for FFT
 * plans, by the row FFT's radix-4 pass tables (copies of those entries,
 * fft_row.h pass_table_size(log2(nphi/2)) float pairs). */
int tsphere_plan_fourier_table(tsphere_plan p, const void **trig);

/* ------------------------------------------- generated element-wise kernels
 * Replaces the WGSL pipeline creation of src/mgpu/plan.ts:145-170 for the
 * kernels buildKernel (src/mgpu/wgsl.ts:321-372) emits: HIP C++ source is
 * compiled for gfx950 by hipRTC at model-compile time.  A compile failure is
 * TSPHERE_ERR_COMPILE with the hipRTC log in tsphere_last_error(). */
typedef struct tsphere_module_t *tsphere_module;
typedef struct tsphere_kernel_t *tsphere_kernel;

/* options: extra hipRTC options separated by '\n' (may be NULL). */
int tsphere_module_compile(const char *source, const char *options, tsphere_module *out);
int tsphere_module_destroy(tsphere_module m);
int tsphere_module_kernel(tsphere_module m, const char *name, tsphere_kernel *out);
int tsphere_kernel_destroy(tsphere_kernel k); /* the handle only; the module owns the code */
/* Compile-only check (no device needed): 0 if `source` compiles for gfx950. */
int tsphere_rtc_check(const char *source, const char *options);

/* Launch a generated element-wise kernel whose signature is
 *   (long long count, float *out, const float *in0.., float s0..)
 * over `count` elements: nptr device pointers (out first), nscalar floats
 * passed BY VALUE (changing a parameter never recompiles, as the reference's
 * params buffer, plan.ts:410-420). */
int tsphere_kernel_launch(tsphere_kernel k, long long count, int nptr, void *const *ptrs,
                          int nscalar, const float *scalars, tsphere_stream s);

/* Launch a generated kernel on an explicit grid (gx x gy workgroups of
 * `block` threads) with arguments (pointers..., floats...): the grid kernels,
 * one workgroup per latitude row (or rows). */
int tsphere_kernel_launch_grid(tsphere_kernel k, unsigned gx, unsigned gy, unsigned block,
                               int nptr, void *const *ptrs, int nscalar, const float *scalars,
                               tsphere_stream s);

/* Ensembles: nptr * per fields in one batched launch per stage, field f at
 * ptr[f / per] + (f % per) * stride floats -- e.g. two species of 64 runs
 * each.  (Batched form of encodeSynthInto/encodeAnalysInto, sht.ts:326-353.) */
int tsphere_synth_fields(tsphere_plan p, int nptr, int per, const float *const *qlm,
                         long long qlm_stride, float *const *spat, long long spat_stride,
                         tsphere_stream s);
int tsphere_analys_fields(tsphere_plan p, int nptr, int per, const float *const *spat,
                          long long spat_stride, float *const *qlm, long long qlm_stride,
                          tsphere_stream s);

/* Transposes of the transforms, the vector-Jacobian products of reverse-mode
 * gradients (no reference counterpart: the reference has no adjoint).  With
 * Q the 2*nlm real coefficients and g the nlat*nphi grid values,
 *   synth_adjoint:  qbar = S^T gbar   (S: Q -> grid, the map tsphere_synth applies)
 *   analys_adjoint: gbar = A^T qbar   (A: grid -> Q, the map tsphere_analys applies)
 * exactly, up to FP32 rounding; fields addressed as tsphere_synth_fields.
 * Outputs are overwritten (use tsphere_accumulate to sum contributions). */
int tsphere_synth_adjoint(tsphere_plan p, int nptr, int per, const float *const *gbar,
                          long long g_stride, float *const *qbar, long long q_stride,
                          tsphere_stream s);
int tsphere_analys_adjoint(tsphere_plan p, int nptr, int per, const float *const *qbar,
                           long long q_stride, float *const *gbar, long long g_stride,
                           tsphere_stream s);
int tsphere_accumulate(float *dst, const float *src, long long n, tsphere_stream s); /* dst += src */

/* ------------------------------------------------------- tile shapes
 * The Legendre stage's workgroup shape (no reference counterpart: WebGPU's
 * workgroup sizes are fixed in sht.ts:70-89).  Stage 0 = synthesis,
 * 1 = analysis; fields left 0 use the built-in choice.
 *   FP32 and FP64-recurrence forms:  impl 1 = contraction on the VALU,
 *               2 = on the FP32 matrix cores as 16x16x4 tiles
 *               (v_mfma_f32_16x16x4_f32), 3 = on the FP32 matrix cores as 16
 *               4x4 blocks per wave (v_mfma_f32_4x4x1_16b_f32: two fields'
 *               re/im fill the 4 columns); the recurrence stays on the VALU;
 *               nf fields per workgroup (1, 2, 4; 8 for impl 3 with k 1 or
 *               2: batched transforms), nwaves waves (1..4; 1..8 for impl 3),
 *               k latitude pairs per lane (1, 2, 4, 8; matrix cores 1, 2, 4),
 *               64*nwaves*k >= pairs; rows = m-rows per workgroup for impl 1
 *               and 3 (2: the pair (m, mmax - m), balanced; 1: one row, twice
 *               the workgroups).  impl 5 = per-parity matrix products on
 *               16x16x4 tiles whose columns are (field, re/im) of 8 fields,
 *               for batched transforms: nf 8 or 16, k 1, rows 1 (the
 *               8-field synthesis also 2: consecutive rows m, m + 1 per
 *               workgroup), a wave per 64 latitude pairs (up to 512 pairs;
 *               beyond, impl 3).
 *   Split-BF16 form:  impl 4 = 16 4x4x4 blocks per wave
 *               (v_mfma_f32_4x4x4_16b_bf16: two fields' re/im fill the 4
 *               columns; the built-in choice up to 512 latitude pairs) with
 *               nf, nwaves and rows as impl 3 and k = 1; impl 0 with ctc =
 *               16-column tiles of v_mfma_f32_16x16x32_bf16 per workgroup (1,
 *               2; 4 for synthesis; the built-in choice above 512 pairs). */
typedef struct {
  int nf, nwaves, k, ctc, impl, rows;
} tsphere_leg_tile;

int tsphere_plan_set_tile(tsphere_plan p, int stage, const tsphere_leg_tile *tile); /* NULL = built-in */
int tsphere_plan_get_tile(tsphere_plan p, int stage, int nfield, tsphere_leg_tile *out);
/* Autotune: time every candidate tile of both stages for `nfield` fields
 * (`reps` launches each, on `s`, on plan scratch and a temporary buffer),
 * keep the fastest in the plan; best[2] / best_ms[2] may be NULL.  Blocks. */
int tsphere_plan_tune(tsphere_plan p, int nfield, int reps, tsphere_stream s,
                      tsphere_leg_tile *best, float *best_ms);

/* ------------------------------------------------------- step recording
 * The reference re-records its op list every step (plan.ts:436-478); here a
 * step is captured once into a HIP graph and replayed.  Capture what the
 * stream does between begin and end (kernels, transforms, copies). */
typedef struct tsphere_graph_t *tsphere_graph;

int tsphere_graph_begin(tsphere_stream s);
int tsphere_graph_end(tsphere_stream s, tsphere_graph *out);
int tsphere_graph_launch(tsphere_graph g, tsphere_stream s);
int tsphere_graph_destroy(tsphere_graph g);

#ifdef __cplusplus
}
#endif
#endif /* TSPHERE_H */
