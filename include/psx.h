/* psx.h -- C ABI of the MI355X pseudo-spectral solver (libpsx.so).
 *
 * Drop-in boundary for PySpectral's serial hot path (PySpectral/src).  All state
 * arrays are device pointers to complex128 values stored as interleaved
 * (re, im) doubles in the reference's per-field Fourier layout
 * (N1, N2, N3/2+1), C order, k1/k2 in FFT order, k3 >= 0
 * (src/Classes.py:19-25,565-567), with the three velocity components stacked:
 * shape (3, N1, N2, N3/2+1).  The reference interleaves the components into one
 * (3N1, 3N2, 3(N3/2+1)) array Q (src/Classes.py:30-34); the Python layer maps
 * that view onto this compact one.
 *
 * Every call is asynchronous on the context's HIP stream and returns 0 on
 * success or a negative PSX_E* code; psx_last_error() gives the message.
 * PSX_E_NONFINITE is a status, not a failure: the call's work was enqueued as
 * usual (the reference keeps stepping a NaN state, src/PySpec_3d.py:121-141), but
 * a CFL or diagnostics reduction that has completed met a NaN or an infinite value.
 * Device memory is allocated by the create calls, by psx_sgs_dns (first call),
 * psx_set_tuning / psx_autotune (chunked-pipeline buffers, a temporary state) and
 * psx_mlp_grad (activation store, when the batch size changes); by nothing else.
 * A context is not thread-safe; separate contexts are independent.  With cfl > 0 the
 * DNS, Smagorinsky and learned-closure contexts also hold a second library stream and
 * two 3-field state-sized buffers: psx_step runs its CFL reduction there, overlapping
 * the first stage's transforms, and joins it before the first stage update.
 */
#ifndef PSX_H
#define PSX_H

#ifdef __cplusplus
extern "C" {
#endif

#define PSX_OK 0
#define PSX_E_ARG (-1)
#define PSX_E_HIP (-2)
#define PSX_E_OOM (-3)
#define PSX_E_UNSUPPORTED (-4)
#define PSX_E_NONFINITE (-5)

/* turb_model values (src/Classes.py:28,77,101,129,192) */
#define PSX_MODEL_DNS 0          /* computeRHS_NOSGS (src/RHSfunctions.py:3-50)        */
#define PSX_MODEL_SMAGORINSKY 1  /* computeRHS_SMAG  (src/RHSfunctions.py:146-237)     */
#define PSX_MODEL_TMODEL 2       /* computeRHS_tmodel (src/RHSfunctions.py:240-363)    */
#define PSX_MODEL_STATIC_TAU 3   /* computeRHS_staumodel (src/RHSfunctions.py:366-497) */
#define PSX_MODEL_DYNAMIC_TAU 4  /* computeRHS_Dtaumodel (src/RHSfunctions.py:592-878) */
#define PSX_MODEL_DSMAG 5        /* computeRHS_DSMAG, turb_model 7 (src/RHSfunctions.py:1663-1862) */
#define PSX_MODEL_LEARNED 6      /* learned closure (no counterpart in the reference): the
                                    Smagorinsky-form RHS with the stress of an attached
                                    network, see psx_set_closure                        */

typedef struct psx_params {
  double nu;              /* viscosity (main.nu)                                     */
  double L1, L2, L3;      /* box lengths (gridclass L1..L3, src/PySpec_3d.py:61-72)   */
  double dx, dy, dz;      /* case grid spacings (src/Classes.py:562-564): CFL and Delta */
  double cfl;             /* > 0: CFL-limited dt; <= 0: fixed dt = -cfl (src/Classes.py:687-698) */
  double Cs;              /* Smagorinsky constant, 0.16 in the reference (src/RHSfunctions.py:197) */
  double Ct;              /* t-model constant (src/Classes.py:103-108)                */
  double tau_a, tau_b;    /* static tau-model tau = tau_a + tau_b t (src/Classes.py:194-195) */
  int kc;                 /* sharp cutoff of Gf (src/Classes.py:584-588)              */
  int rotate;             /* Coriolis terms on/off (src/RHSfunctions.py:33-50), every model */
  double Om1, Om2, Om3;   /* rotation vector                                         */
  double force;           /* linear forcing f = force * u of the velocity as the model
                             uses it (no counterpart in the reference; every model)   */
} psx_params;

typedef struct psx_context psx_context;
typedef struct psx_mlp psx_mlp; /* the learned closure's network, see psx_mlp_create */

/* Replaces FFTclass + gridclass + the turb_model dispatch of variables
 * (src/Classes.py:590-625, 548-588, 28-229): plans, twiddle tables and pipeline
 * scratch for an N1 x N2 x N3 grid.  stream is a hipStream_t (NULL = default). */
int psx_create(psx_context** ctx, int n1, int n2, int n3, int model, const psx_params* params, int device,
               void* stream);

/* No counterpart in the reference (which runs one grid per process): `members`
 * independent instances of the same grid and model in one context, for ensembles.
 * Every state argument then holds the members one after another,
 * (members, 3, N1, N2, N3/2+1); dt_dev arguments hold one dt per member; nu gives
 * each member's viscosity (NULL: params->nu for all).  Each kernel launch covers all
 * members.  DNS, Smagorinsky and learned closure; psx_spectrum, psx_filter and psx_sgs_dns take
 * single-member contexts; psx_stats writes 3 values per member. */
int psx_create_batch(psx_context** ctx, int n1, int n2, int n3, int model, const psx_params* params, int members,
                     const double* nu, int device, void* stream);

/* Ensembles of uneven grid sizes (no counterpart in the reference): count members with
 * grid dims[3*i..3*i+2] and viscosity nu[i]; the grid spacings are L/N and kc = N1/2
 * for each member, the other parameters come from common.  Members of one size share
 * a batch context.  The ensemble state is one array of psx_ensemble_size complex
 * values; member i starts at psx_ensemble_offset(i) and its dt / stats entries are at
 * slot psx_ensemble_slot(i).  psx_ensemble_step advances every member by one
 * advanceQ_RK4 with its own CFL dt; psx_ensemble_step_graph captures the whole
 * ensemble step as one hipGraph on the ensemble's (non-default) stream and replays it. */
typedef struct psx_ensemble psx_ensemble;
int psx_ensemble_create(psx_ensemble** ens, int count, const int* dims, const double* nu, int model,
                        const psx_params* common, int device, void* stream);
int psx_ensemble_destroy(psx_ensemble* ens);
const char* psx_ensemble_last_error(const psx_ensemble* ens);
long long psx_ensemble_size(const psx_ensemble* ens);
long long psx_ensemble_offset(const psx_ensemble* ens, int member);
int psx_ensemble_slot(const psx_ensemble* ens, int member);
int psx_ensemble_set_stream(psx_ensemble* ens, void* stream);
int psx_ensemble_step(psx_ensemble* ens, double* q, double* qs, double* dt_dev, double t);
int psx_ensemble_step_graph(psx_ensemble* ens, double* q, double* qs, double* dt_dev);
int psx_ensemble_stats(psx_ensemble* ens, const double* q, double* out_dev);
/* Number of size classes (distinct member grids, in order of first appearance). */
int psx_ensemble_classes(const psx_ensemble* ens);
/* Kernel shapes for every size class (no counterpart in the reference): coordinate
 * descent on the time of the whole ensemble step, the classes running concurrently as
 * psx_ensemble_step runs them -- for each class, the tile widths, z-pass variant, block
 * orders and stress-sum placement of psx_set_tuning (tune[0..3], [7..10]) are tried in
 * turn, each candidate timed on reps steps from a copy of q (the ensemble state; q is
 * not changed), and the fastest kept (a change must gain 1%).  Allocates two ensemble
 * states for the trials.  tune_out (may be NULL) receives PSX_TUNE_N values per class,
 * class-major.  The captured ensemble step is re-captured by its next call. */
int psx_ensemble_autotune(psx_ensemble* ens, const double* q, int reps, int* tune_out);
/* psx_status over every member: waits for the ensemble's stream. */
int psx_ensemble_status(psx_ensemble* ens);
/* psx_set_closure for every member of a PSX_MODEL_LEARNED ensemble (one network). */
int psx_ensemble_set_closure(psx_ensemble* ens, psx_mlp* net, double bound, int flags);
int psx_destroy(psx_context* ctx);
int psx_set_stream(psx_context* ctx, void* stream);
/* Replaces the parameters; a step captured by psx_step_graph is re-captured by its
 * next call (as after psx_set_tuning, psx_autotune and psx_set_closure). */
int psx_set_params(psx_context* ctx, const psx_params* params);
const char* psx_last_error(const psx_context* ctx);

/* Non-finite detection (no counterpart in the reference, whose driver prints NaN
 * statistics, src/PySpec_3d.py:129-141).  The CFL reduction of psx_step /
 * psx_step_graph / psx_compute_dt (whose dt then is NaN for a NaN velocity and 0 for
 * an infinite one, as numpy's amax and minimum give in src/Classes.py:687-698) and
 * psx_stats raise a sticky per-context flag when they meet a NaN or an infinite
 * value.  Entry points that run them return PSX_E_NONFINITE once the flag is set by
 * work that has completed.  psx_status waits for the context's stream and returns
 * PSX_E_NONFINITE if the flag is set (PSX_OK otherwise); psx_clear_status waits and
 * clears it. */
int psx_status(psx_context* ctx);
int psx_clear_status(psx_context* ctx);
/* Bytes of device scratch the context holds. */
long long psx_scratch_bytes(const psx_context* ctx);

/* main.computeRHS(main, grid, myFFT, utilities) (src/RHSfunctions.py:3,146,240,366,592):
 * rhs = RHS(uhat) at step-start time t.  uhat_model (nullable) receives the
 * velocity as the model leaves it in main.uhat (Nyquist entries zeroed for the
 * models that truncate, src/RHSfunctions.py:10-12).  w0 (nullable) receives the
 * subgrid term main.w0_{u,v,w} (closures only). */
int psx_rhs(psx_context* ctx, const double* uhat, double* rhs, double* uhat_model, double* w0, double t);

/* One stage of advanceQ_RK4 (src/PySpec_3d.py:102-104): out = q0 + dt*c*RHS(uin),
 * dt read from the device scalar *dt_dev.  out may alias uin or q0. */
int psx_rk_stage(psx_context* ctx, const double* q0, const double* uin, double* out, const double* dt_dev,
                 double c, double t, int zero_nyquist_out);

/* utilitiesClass.compute_dt (src/Classes.py:687-698) on uhat, result in *dt_dev. */
int psx_compute_dt(psx_context* ctx, const double* uhat, double* dt_dev);

/* advanceQ_RK4 (src/PySpec_3d.py:96-105): dt from the stale velocity qs (the
 * reference's main.uhat), then the four low-storage stages.  On return q holds
 * the new state and qs the stage-4 input as the model left it (main.uhat after
 * the step).  Before the first step qs must hold the initial velocity. */
int psx_step(psx_context* ctx, double* q, double* qs, double* dt_dev, double t);

/* advanceQ_RK4 followed by the driver's t += dt (src/PySpec_3d.py:124-125) with the
 * step-start time held on the device (*t_dev, advanced by dt), as one hipGraph.
 * The first call for a given (q, qs, dt_dev, t_dev, stream) captures the step on
 * the context's stream, which must not be the default stream; later calls replay
 * it until the parameters, tuning or closure change (then the next call captures
 * again).  No counterpart in the reference (whose loop is host Python); for callers
 * that run many steps without host work in between. */
int psx_step_graph(psx_context* ctx, double* q, double* qs, double* dt_dev, double* t_dev);

/* utilitiesClass.computeEnergy / computeEnstrophy / computeEnergy_resolved
 * (src/Classes.py:652-676,733-744): out_dev[0..2] = E, Omega, E_resolved. */
int psx_stats(psx_context* ctx, const double* uhat, double* out_dev);

/* No counterpart in the reference: psx_stats for callers that cannot allocate device
 * memory (code under hipGraph capture, serving loops, harnesses that forbid hipMalloc),
 * without a context.  uhat holds `members` states (members, 3, N1, N2, N3/2+1) of an
 * N1 x N2 x N3 grid with box lengths L1..L3 and sharp cutoff kc; out_dev[3m..3m+2]
 * receives member m's E, Omega, E_resolved, bitwise as psx_stats gives them.  Runs on
 * `stream` (a hipStream_t, NULL = default) of the current device and calls no allocator:
 * its reduction partials live in a static device array of 1024 x 3 doubles (24 KiB),
 * which each GPU holds until the process ends.  Any grid size is accepted (the partials
 * are capped at 1024 blocks; every block writes its own slot before the final sum reads
 * it, so nothing needs clearing).  Calls that overlap on one GPU share that array and
 * race: order them on one stream.  Returns PSX_OK, PSX_E_ARG or PSX_E_HIP; there is no
 * context, hence no message and no non-finite flag (a NaN state gives NaN values). */
int psx_stats_noalloc(const double* uhat, int members, int n1, int n2, int n3, double L1, double L2, double L3,
                      int kc, double* out_dev, void* stream);

/* computeSpectrum / computeSpectrum_resolved (src/Classes.py:762-797,825-844):
 * spec_dev[(kmax+1) x 3], kmax from psx_spectrum_bins(). */
int psx_spectrum_bins(const psx_context* ctx);
int psx_spectrum(psx_context* ctx, const double* uhat, double* spec_dev, int resolved);

/* Filters of a state-layout field (3, N1, N2, N3/2+1).  PSX_FILTER_SHARP is
 * gridclass.Gf / utilitiesClass.myFilter (src/Classes.py:584-588, 1271-1279): keep
 * |i1|, |i2| < kc and i3 < kc (signed indices, kc from psx_params).  PSX_FILTER_BOX
 * (no counterpart in the reference) is the top-hat of width delta on every axis,
 * transfer function prod_i sin(k_i delta/2) / (k_i delta/2). */
#define PSX_FILTER_SHARP 0
#define PSX_FILTER_BOX 1
int psx_filter(psx_context* ctx, const double* uhat, double* out, int filter, double delta);

/* utilitiesClass.computeSGS_DNS (src/Classes.py:861-929): with ubar = G uhat,
 * tau_hat_ij = G FFT(u_i u_j) - FFT(ubar_i ubar_j) (3/2-rule products) and
 * w0_i = -i k_j tau_ij - k_i (k_m k_j tau_mj)/k^2.  w0_out: (3, N1, N2, N3/2+1) or
 * NULL; tau_out: (6, N1, N2, N3/2+1), order 11 22 33 12 13 23, or NULL.  Works on
 * a context of any model; its work buffers are allocated on the first call. */
int psx_sgs_dns(psx_context* ctx, const double* uhat, double* w0_out, double* tau_out, int filter, double delta);

/* The dynamic tau-model's coefficient from the last RHS evaluation (main.tau,
 * src/RHSfunctions.py:857-861), copied to host memory (synchronizes the stream):
 * out[0] tau as computed, out[1] tau clamped at 0 (the value used), out[2..7] the
 * reductions LE, PLQLE_filt, filt_PLQLE as (re, im) pairs.  For the dynamic
 * Smagorinsky model out[0] = out[1] = Cs^2 (src/RHSfunctions.py:1818). */
int psx_tau(psx_context* ctx, double* out_host8);

/* Kernel shapes (no counterpart in the reference): tune[0..2] = tile widths (k3 columns
 * per block: 4, 8 or 16, and 24 or 32 up to length 96; 0 = the built-in default: 16 up
 * to length 96, or 24, else 8, where 16 does not divide the column count but that width
 * does; 8 up to 384; 4 beyond) of the x-inverse, y-inverse and y-forward passes, tune[3] = z-pass variant (0: one wavefront per position where the
 * z length is a multiple of 64, 1: the block-wide Stockham kernel), tune[4] = number of
 * x-row chunks (0 or 1: none; at most 8) in which the y-inverse, z and y-forward passes
 * run alternately on the context stream and a second library stream, so that their
 * HBM- and compute-bound phases overlap; tune[5] = 1 runs the chunks in order on the
 * context stream instead (0 or 2: two streams), so that a chunk's intermediates can be
 * re-read from the last-level cache; tune[6] = 0 evaluates the learned closure's network
 * inside the z pass where the padded z length is 64 * {1,2,3,4,6,8,12}, 1 always as its
 * own kernel between a gradient and a products z pass; tune[7..9] = block order of the
 * x-inverse, y-inverse and y-forward passes (0: each XCD runs a contiguous range of
 * blocks, 1: dispatch order, 2: each outer row's column tiles on one XCD, consecutive
 * rows round-robin over the XCDs); tune[10] = where the Smagorinsky-form stresses are combined
 * into the three k1-polynomial coefficients the projection needs (0: the y-forward sums
 * them and hands the epilogue 9 fields, 1: the epilogue reads all 12 and sums them).
 * psx_autotune times every candidate of each
 * knob on reps RHS evaluations of uhat with HIP events, keeps the fastest (knob by
 * knob), and returns the choice in tune_out (may be NULL); it allocates one temporary
 * state, and the chunked pipeline's buffers when it tries chunks, as psx_set_tuning
 * does when it sets them.  Every model: the t-model family tunes the tiles and block
 * orders of its 2x-grid pencil passes (tile widths are checked against those lengths), the
 * learned stage does not chunk, chunking is for the DNS, Smagorinsky and DSMAG stages. */
#define PSX_TUNE_N 11
int psx_set_tuning(psx_context* ctx, const int* tune);
int psx_get_tuning(const psx_context* ctx, int* tune);
int psx_autotune(psx_context* ctx, const double* uhat, int reps, int* tune_out);

/* Learned subgrid closure network (no counterpart in the reference): a pointwise MLP
 * 9 -> 64 -> 64 -> 64 -> 6 with tanh hidden layers mapping filtered velocity-gradient
 * components to the six SGS stresses, BF16 operands with FP32 accumulation on the matrix
 * cores, FP32 master weights.  Parameters (PSX_MLP_NPARAM floats, row-major):
 * W1[64][9] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[6][64] b4[6].  x: (n, 9) and
 * y, target: (n, 6) float32 device arrays; inputs and targets are normalised inside
 * with psx_mlp_set_norm's means and scales (inputs as (x - mean) * (1/scale), the
 * reciprocal rounded to float32).  psx_mlp_grad computes the MSE loss (into
 * loss_dev, may be NULL) and the parameter gradient (psx_mlp_gradient copies it out);
 * psx_mlp_train_step adds one Adam step (beta 0.9/0.999, eps 1e-8).  Training batches
 * are multiples of 16 points. */
#define PSX_MLP_NPARAM 9350
int psx_mlp_create(psx_mlp** net, int device, void* stream, unsigned long long seed);
int psx_mlp_destroy(psx_mlp* net);
const char* psx_mlp_last_error(const psx_mlp* net);
int psx_mlp_set_stream(psx_mlp* net, void* stream);
int psx_mlp_get_params(psx_mlp* net, float* host_params);
int psx_mlp_set_params(psx_mlp* net, const float* host_params);
int psx_mlp_set_norm(psx_mlp* net, const float* x_mean, const float* x_scale, const float* y_mean,
                     const float* y_scale);
int psx_mlp_forward(psx_mlp* net, const float* x, float* y, long long n);
/* The closure as the learned LES evaluates it: stresses tau (n, 6) from gradients g (n, 9)
 * (du_i/dx_j at column 3i+j), the Smagorinsky stress -2 c2 |S| S_ij outside the input
 * range [-bound, bound] and, with PSX_CLOSURE_CLIP_BACKSCATTER in flags, where the
 * network's stress is backscatter (see psx_set_closure). */
int psx_mlp_closure(psx_mlp* net, const float* g, float* tau, long long n, double c2, double bound, int flags);
int psx_mlp_grad(psx_mlp* net, const float* x, const float* target, long long n, float* loss_dev);
int psx_mlp_gradient(psx_mlp* net, float* grad_dev);
int psx_mlp_train_step(psx_mlp* net, const float* x, const float* target, long long n, float lr, float* loss_dev);

/* Learned-closure LES (no counterpart in the reference).  psx_set_closure attaches a
 * network to a PSX_MODEL_LEARNED context (the context does not own it; it must outlive
 * the context's use).  Each RHS evaluation forms the nine velocity gradients
 * du_i/dx_j (row-major (i, j), physical units of u = irfftn(uhat) sqrt(Nt)) at every
 * point of the 3/2-padded grid, maps them through the network to the six stresses
 * (order 11 22 33 12 13 23) and uses them in place of the Smagorinsky stress of
 * computeRHS_SMAG (src/RHSfunctions.py:146-237): w0 and the RHS follow
 * src/RHSfunctions.py:215-237.  A point whose normalised inputs leave [-bound, bound]
 * takes the Smagorinsky stress -2 (Cs Delta)^2 |S| S_ij of its gradients (bound =
 * INFINITY: never; a negative bound: everywhere).  With flags PSX_CLOSURE_CLIP_BACKSCATTER
 * a point where the network's stress is backscatter (tau_ij S_ij > 0: negative subgrid
 * dissipation) takes the Smagorinsky stress too.  Re-attach after changing the
 * network's normalisation: a captured step (psx_step_graph) holds it by value. */
#define PSX_CLOSURE_CLIP_BACKSCATTER 1
int psx_set_closure(psx_context* ctx, psx_mlp* net, double bound, int flags);

/* Grid-space fields (no counterpart as one call in the reference, which calls
 * myFFT.ifftT_obj, src/Classes.py:596-607): out (nfields, N1, N2, N3) =
 * irfftn(fhat) sqrt(N1 N2 N3) of nfields spectral fields fhat (nfields, N1, N2, N3/2+1),
 * the inverse of uhat = rfftn(u)/sqrt(Nt).  Single-member contexts of the DNS,
 * Smagorinsky, DSMAG and learned models. */
int psx_physical(psx_context* ctx, const double* fhat, int nfields, double* out);

/* The inverse of psx_physical (no counterpart as one call in the reference, which calls
 * myFFT.fft_obj on padded grids, src/Classes.py:608-613, and FFTFromFile's plans in
 * post-processing, src/postProcessingFunctions.py:67-108): fhat (nfields, N1, N2, N3/2+1)
 * = rfftn(phys) / sqrt(N1 N2 N3) of nfields real grid fields phys (nfields, N1, N2, N3),
 * every spectral entry kept.  Single-member contexts of any model. */
int psx_spectral(psx_context* ctx, const double* phys, int nfields, double* fhat);

/* The nine velocity gradients du_i/dx_j = irfftn(i k_j uhat_i) sqrt(Nt) at the grid
 * points, as float32 out[(x N2 + y) N3 + z][3 i + j] (the learned closure's input
 * layout).  Single-member contexts, as psx_physical. */
int psx_gradients(psx_context* ctx, const double* uhat, float* out);

/* Spectral restriction from a finer grid (s1, s2, s3) to the context's grid: the
 * signed modes |k1| < N1/2, |k2| < N2/2, 0 <= k3 < N3/2 of nfields fields
 * src (nfields, s1, s2, s3/2+1), times sqrt(N1 N2 N3 / (s1 s2 s3)) so that grid values
 * are kept; the context grid's Nyquist entries are zero.  For filtered-DNS training
 * data on an LES grid. */
int psx_restrict(psx_context* ctx, const double* src, int s1, int s2, int s3, int nfields, double* out);

/* Instrumentation (no counterpart in the reference, which only prints wall time,
 * src/PySpec_3d.py:136-145): bracket every kernel launch with HIP events on the
 * context stream between begin and end.  psx_timing_end synchronizes the
 * stream and returns, per kernel class, the summed milliseconds and launch
 * counts in ms[PSX_NKERNEL], count[PSX_NKERNEL]. */
#define PSX_KERNEL_X_INVERSE 0
#define PSX_KERNEL_Y_INVERSE 1
#define PSX_KERNEL_Z_PASS 2
#define PSX_KERNEL_Y_FORWARD 3
#define PSX_KERNEL_X_FORWARD_EPILOGUE 4
#define PSX_KERNEL_CFL_X 5
#define PSX_KERNEL_CFL_Y 6
#define PSX_KERNEL_CFL_Z 7
#define PSX_KERNEL_OTHER 8
#define PSX_KERNEL_CLOSURE 9  /* the learned closure's network */
#define PSX_NKERNEL 10
int psx_timing_begin(psx_context* ctx);
int psx_timing_end(psx_context* ctx, double* ms, int* count);

/* Streaming bandwidth probe (instrumentation, no counterpart in the reference): mode 0
 * copies n complex128 values src -> dst with the workgroups striding through the array,
 * mode 2 with each workgroup copying contiguous 16 KiB chunks, mode 3 as mode 2 with
 * non-temporal stores; mode 1 reads them and writes one partial sum per workgroup into
 * dst[0 .. blocks).  `blocks` workgroups of 256 lanes, four 16-byte loads in flight per
 * lane.  Asynchronous on `stream`. */
int psx_stream_probe(const double* src, double* dst, long long n, int mode, int blocks, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* PSX_H */
