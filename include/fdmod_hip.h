/*
 * fdmod_hip.h -- C-ABI of libfdmod_hip, the MI355X (gfx950) implementation of the
 * TTI-FDModeling rotated-staggered-grid (RSG) 2-D TTI elastic modelling loop.
 *
 * The reference has no library boundary for this path: the loop is inline in each
 * program's main().  These entry points replace the following reference code, so
 * that the three host programs (coordinate rotation, tensor rotation, Bond
 * transformation) keep their getpar surface, model files and SU I/O and hand the
 * work between "model read" and "trace write" to the GPU:
 *
 *   fdmod_create         grid sizing, BT:296-340 (N=5, ealThick+=N, Nz, Nx)
 *   fdmod_set_model      padding BT:452-457 (padModel2d, L/absorbc/padmod.c:33-58),
 *                        stability BT:472 (L/fdmod/stable.c:217-242),
 *                        MEAL BT:474-475 (L/absorbc/meal.c:51-201, 450-491),
 *                        Thomsen->Cij(theta) BT:493-526 / TR:483-503 / CR:497-514,
 *                        invrho / trig averaging BT:527-542 / CR:516-535,
 *                        Bond free-surface fix-up BT:543-568
 *   fdmod_set_source     source geometry after read2dsource (L/fdmod/obsinfo.c:34-87)
 *   fdmod_set_receivers  receiver geometry after read2dreceiver (obsinfo.c:175-230)
 *   fdmod_set_snapshots  snapshot cadence, BT:916-919 (L/fdmod/isoelastic.c:240-289)
 *   fdmod_set_energy     energy diagnostic switch, BT:926-952
 *   fdmod_run            the time loop BT:641-954 (TR:604-935, CR:601-926)
 *   fdmod_get_traces     Receiver2D.recv after the loop, handed to traces2d_output
 *                        (L/fdmod/traces.c:146-196)
 *   fdmod_get_array      the side-effect dumps BT:476-491, 569-596 and the fields
 *
 * Citations: BT/TR/CR = src/main/sufdmod2dTTIelastic/sufdmod2dTTIelastic_{bond_
 * transformation,tensor_rotation,coordinate_rotation}.c of the reference.
 *
 * Conventions.  All arrays cross the boundary as host pointers in the reference's
 * layout (CWP alloc2float(n1=nz, n2=nx): nx records of nz floats, iz fastest).  The
 * library copies inputs into device memory it owns and never keeps a host pointer
 * past a call.  Every entry point returns 0 or a negative FDMOD_E* code; the message
 * of the last failure is available from fdmod_last_error().  Argument validation
 * (including the stability check) runs on the host before any device work.  One
 * context belongs to one host thread; calls on one context are not re-entrant.
 */
#ifndef FDMOD_HIP_H
#define FDMOD_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define FDMOD_OK 0
#define FDMOD_EINVAL (-1)   /* bad argument (sizes, flags, NULL pointers)           */
#define FDMOD_EUNSTABLE (-2) /* dt larger than the stability limit (stable.c:236)   */
#define FDMOD_ERANGE (-3)   /* source/receiver outside the model (obsinfo.c:69-80)  */
#define FDMOD_EDEVICE (-4)  /* HIP runtime failure                                  */
#define FDMOD_ENOMEM (-5)   /* device or host allocation failed                     */
#define FDMOD_ESTATE (-6)   /* call out of order (e.g. run before set_model)        */

/* rotation forms: the three programs under src/main/sufdmod2dTTIelastic/ */
#define FDMOD_FORM_COORD 0  /* sufdmod2dTTIelastic_coordinate_rotation */
#define FDMOD_FORM_TENSOR 1 /* sufdmod2dTTIelastic_tensor_rotation     */
#define FDMOD_FORM_BOND 2   /* sufdmod2dTTIelastic_bond_transformation */

/* trace components, the flag values of traces2d_output (traces.c:166-190) */
#define FDMOD_TRACE_VX 1
#define FDMOD_TRACE_VZ 3
#define FDMOD_TRACE_P 4

/* array ids for fdmod_get_array / fdmod_set_array (padded Nx*Nz, iz fastest) */
enum fdmod_array {
  FDMOD_A_VP = 0, FDMOD_A_VS, FDMOD_A_RHO, FDMOD_A_EPS, FDMOD_A_DELTA, FDMOD_A_THETA,
  FDMOD_A_INVRHO, FDMOD_A_C11, FDMOD_A_C13, FDMOD_A_C15, FDMOD_A_C33, FDMOD_A_C35, FDMOD_A_C55,
  FDMOD_A_SIN2, FDMOD_A_COS2, FDMOD_A_SIN2T, FDMOD_A_SIN, FDMOD_A_COS, FDMOD_A_SINV, FDMOD_A_COSV,
  FDMOD_A_D, FDMOD_A_BETA, FDMOD_A_COEF1, FDMOD_A_COEF2,
  FDMOD_A_VX, FDMOD_A_VZ, FDMOD_A_SXX, FDMOD_A_SZZ, FDMOD_A_SXZ, FDMOD_A_TXX, FDMOD_A_TZZ, FDMOD_A_TXZ,
  FDMOD_A_COUNT
};

/* getpar parameters of the three programs that shape the run (BT:134-249). */
typedef struct fdmod_params {
  int form;        /* FDMOD_FORM_*                                                 */
  int nz, nx, nt;  /* model samples (1st/2nd dimension) and time samples           */
  float dx, dz, dt;
  int ealThick;    /* absorbing layers as given by the user; N=5 is added inside   */
  float eal_alpha; /* -1 selects the library default 2.5 (meal.c:74-77)            */
  int freesurface; /* 0/1                                                          */
  float fpeak;     /* MEAL reference frequency                                     */
  float scalar;    /* applied to theta (degrees)                                   */
  int mode;        /* 1: it = 0..nt-1; 2: it = nt-1..0 (BT:651)                    */
} fdmod_params;

typedef struct fdmod_ctx fdmod_ctx;

/* Snapshot sink: called on the host at each snapshot time with one cropped field
 * (kind "vx", "vz" or "p"; n1 = samples per record (z), n2 = records (x)).  The data
 * pointer is valid for the duration of the call.  Return nonzero to abort the run. */
typedef int (*fdmod_snapshot_fn)(void *user, int it, const char *kind, const float *data, int n1, int n2);

int fdmod_create(const fdmod_params *p, fdmod_ctx **out);
void fdmod_destroy(fdmod_ctx *ctx);
const char *fdmod_last_error(const fdmod_ctx *ctx);
const char *fdmod_strerror(int code);

/* out[0..6] = Nz, Nx, ealThick (effective), izStart, izEnd, ixStart, ixEnd */
int fdmod_dims(const fdmod_ctx *ctx, int *out7);

/* Host-only stability check of staggered_grid_stable2d (L/fdmod/stable.c:217-242,
 * called BT:472) on an unpadded nz*nx vp: FDMOD_OK or FDMOD_EUNSTABLE, with the
 * limit in *dt_max.  Touches no device, so callers can validate before creating a
 * context; fdmod_set_model performs the same check. */
int fdmod_check_stability(const fdmod_params *p, const float *vp, double *dt_max);

/* Unpadded nz*nx arrays in file order; theta in degrees.  vs and rho must be given
 * (the host programs apply vp2vs / vp2densityGardner, L/rockmod/modparconvert.c,
 * when the files are absent).  On FDMOD_EUNSTABLE *dt_max holds the limit. */
int fdmod_set_model(fdmod_ctx *ctx, const float *vp, const float *vs, const float *rho, const float *eps,
                    const float *delta, const float *theta, double *dt_max);

/* stype 1..8 (BT:41-56); positions in cells (0..nx / 0..nz inclusive, obsinfo.c:69-80);
 * src is [ntrace][nt] (Source2D.src). */
int fdmod_set_source(fdmod_ctx *ctx, int stype, int ntrace, const int *posx, const int *posz, const float *src);

/* seistype 0..3 (BT:57-61); ntsample = (nt-1)/mt + 1 (obsinfo.c:199). */
int fdmod_set_receivers(fdmod_ctx *ctx, int seistype, int mt, int ntrace, const int *posx, const int *posz);

/* wsave 0..3, savew 0/1, snap_sample >= 1 (BT:79-92); fn may be NULL when wsave == 0. */
int fdmod_set_snapshots(fdmod_ctx *ctx, int wsave, int savew, int snap_sample, fdmod_snapshot_fn fn, void *user);

int fdmod_set_energy(fdmod_ctx *ctx, int on);

/* Run loop iterations k0 .. k0+count-1 (it = k or nt-1-k).  The wavefield persists
 * between calls, so a run may be split.  *seconds (may be NULL) receives the device
 * time of the call, measured with HIP events around the loop. */
int fdmod_run(fdmod_ctx *ctx, int k0, int count, double *seconds);

/* Reset the wavefield, traces and energy to zero (model and geometry are kept). */
int fdmod_reset(fdmod_ctx *ctx);

/* Traces in Receiver2D.recv layout [ntrace][ntsample]. */
int fdmod_get_traces(fdmod_ctx *ctx, int component, float *out);

/* Energy series E, Ek, Ep of length nt (BT:946-950). */
int fdmod_get_energy(fdmod_ctx *ctx, float *E, float *Ek, float *Ep);

/* Padded arrays (Nx*Nz floats, iz fastest): model after MEAL, parameters, fields. */
int fdmod_get_array(fdmod_ctx *ctx, int which, float *out);
int fdmod_set_array(fdmod_ctx *ctx, int which, const float *in);

/* ---- Batched shots (no counterpart in the reference, which runs one shot per program
 * invocation, BT:239-243 shotID).  A batch holds nshots independent shots that share
 * every parameter of *p except nx (shot s has nx[s] columns: its aperture crop).  Each
 * shot is set up exactly like a single-shot context; the time loop then advances all
 * shots together -- one launch per sweep covers every shot -- with K steps captured
 * once in a hipGraph and replayed.  Snapshots and the energy series are single-shot
 * features.  Results per shot are identical to running that shot alone. */
typedef struct fdmod_batch fdmod_batch;
int fdmod_batch_create(const fdmod_params *p, int nshots, const int *nx, fdmod_batch **out);
void fdmod_batch_destroy(fdmod_batch *b);
const char *fdmod_batch_last_error(const fdmod_batch *b);
int fdmod_batch_set_model(fdmod_batch *b, int shot, const float *vp, const float *vs, const float *rho,
                          const float *eps, const float *delta, const float *theta, double *dt_max);
int fdmod_batch_set_source(fdmod_batch *b, int shot, int stype, int ntrace, const int *posx, const int *posz,
                           const float *src);
int fdmod_batch_set_receivers(fdmod_batch *b, int shot, int seistype, int mt, int ntrace, const int *posx,
                              const int *posz);
/* The shot's own context, for per-shot queries (fdmod_dims, fdmod_get_array, ...). */
fdmod_ctx *fdmod_batch_shot(fdmod_batch *b, int shot);
/* Autotuner: time the candidate segment lengths of both sweeps on this batch and keep
 * the fastest (reps launches per candidate, after a warm-up); resets the wavefields.
 * segLV 0 (the default, and a tuner candidate) gives each shot group the shortest
 * velocity segment whose blocks all run in one generation on the device. */
int fdmod_batch_tune(fdmod_batch *b, int reps, int *segLV, int *segLS);
int fdmod_batch_set_tiles(fdmod_batch *b, int segLV, int segLS);
int fdmod_batch_run(fdmod_batch *b, int k0, int count, double *seconds);
int fdmod_batch_reset(fdmod_batch *b);
int fdmod_batch_get_traces(fdmod_batch *b, int shot, int component, float *out);

/* ---- Adjoint and misfit gradient (no counterpart in the reference, which only runs
 * the forward loop; all three forms).  The adjoint is the exact transpose of the loop
 * iteration BT:651-954 as implemented by fdmod_run (sweeps, velocity/stress sources,
 * free-surface imaging, receiver sampling), with the absorbing layers (MEAL D, Beta,
 * coef1, coef2 of the last fdmod_set_model) held fixed.
 *
 * fdmod_gradient: J = 1/2 sum_c sum_r sum_s (d_c[r][s] - obs_c[r][s])^2 over the
 * recorded components c whose obs_c is given (vx and vz go together), with d from a
 * run from rest; grad receives dJ/d(vp, epsilon, delta, theta in degrees) as
 * [4][nx][nz] doubles in the model-file layout.  The forward wavefield is rebuilt by
 * checkpointed recomputation with ncheckpoints stored states (<= 0: ceil(sqrt(nt))).
 * Imaging sums accumulate in FP64.  Afterwards the traces are those of the run from
 * rest; the wavefield is not.  vs and rho are not differentiated. */
int fdmod_gradient(fdmod_ctx *ctx, const float *obs_vx, const float *obs_vz, const float *obs_p, int ncheckpoints,
                   double *misfit, double *grad);

/* Storage of the wavefields the gradient keeps: checkpointed states and the strain
 * E(v_{k+1}) = (vxdx, vzdz, vzdx + vxdz) the stress sweep computes at every step,
 * which is all the imaging needs.  FP16 stores each field as FP16 with one
 * power-of-two scale per column and 64-row block (stencil arithmetic stays FP32).
 * history selects how the backward pass gets the forward wavefield:
 *   FDMOD_HIST_CHECKPOINTS   ncheckpoints states, each segment recomputed with its strain;
 *   FDMOD_HIST_FULL          the strain of every step (3*nt fields, no recomputation);
 *   FDMOD_HIST_RECONSTRUCT   boundary-saving reconstruction: the forward saves the state on
 *                            a band 5 points wide around the undamped interior and the
 *                            strain outside it, every step; the backward pass runs the loop
 *                            backwards in time over the interior (dt negated), putting the
 *                            band back after each sweep.  ncheckpoints > 0 also stores that
 *                            many full states and restarts the reversed run from each (0:
 *                            none).  FP32 only (storage must be FDMOD_STORE_F32).
 *   FDMOD_HIST_RECONSTRUCT_INTERIOR  as _RECONSTRUCT without keeping the absorbing layers'
 *                            strain (only the band, and the surface row under a free
 *                            surface, are saved): the gradient leaves out the absorbing
 *                            layers' imaging, which the transpose of the edge padding adds
 *                            only to the model's first and last columns, its bottom row
 *                            and, without a free surface, its top row; every other cell is
 *                            that of _RECONSTRUCT.
 * fdmod_gradient is {ncheckpoints, FDMOD_STORE_F32, FDMOD_HIST_CHECKPOINTS}. */
#define FDMOD_HIST_CHECKPOINTS 0
#define FDMOD_HIST_FULL 1
#define FDMOD_HIST_RECONSTRUCT 2
#define FDMOD_HIST_RECONSTRUCT_INTERIOR 3
#define FDMOD_STORE_F32 0
#define FDMOD_STORE_F16 1      /* strain FP16, checkpoints FP32 */
#define FDMOD_STORE_F16_ALL 2  /* strain and checkpoints FP16: restarts from rounded
                                  states; on a 1126 x 2000, nt = 4000 shot within 3.3e-4 of
                                  the FP32 path, except the tensor form's theta (2.1e-3) */
typedef struct fdmod_grad_opts {
  int ncheckpoints; /* <= 0: ceil(sqrt(nt))                    */
  int storage;      /* FDMOD_STORE_F32, _F16 or _F16_ALL        */
  int history;      /* FDMOD_HIST_CHECKPOINTS, _FULL, _RECONSTRUCT or _RECONSTRUCT_INTERIOR */
} fdmod_grad_opts;
int fdmod_gradient_opts(fdmod_ctx *ctx, const float *obs_vx, const float *obs_vz, const float *obs_p,
                        const fdmod_grad_opts *opts, double *misfit, double *grad);

/* Misfit gradient of every shot of the batch, each as fdmod_gradient_opts
 * would compute it, with the shots of a group advancing together on their own streams
 * (groups are as large as free device memory allows for the chosen storage).  obs_* are
 * arrays of nshots pointers (the array itself NULL = component absent).  With x0, shot s
 * covers model columns [x0[s], x0[s]+nx[s]) of an nx_full-column model and grad is the
 * FP64 sum over shots, [4][nx_full][nz]; with x0 NULL, grad holds the per-shot gradients
 * one after another ([shot][4][nx[s]][nz]).  *misfit is the sum of the shots' misfits. */
int fdmod_batch_gradient(fdmod_batch *b, const float *const *obs_vx, const float *const *obs_vz,
                         const float *const *obs_p, const fdmod_grad_opts *opts, const int *x0, int nx_full,
                         double *misfit, double *grad);
/* Replace the model keeping the absorbing layers of the last fdmod_set_model (the
 * model the gradient is defined with); same array conventions as fdmod_set_model. */
int fdmod_update_model(fdmod_ctx *ctx, const float *vp, const float *vs, const float *rho, const float *eps,
                       const float *delta, const float *theta);
/* F^T y, the adjoint of the modelling operator F (source time functions -> traces)
 * applied to data-space series y ([ntrace][ntsample] per component, NULL = absent):
 * out [ntrace_src][nt] doubles. */
int fdmod_adjoint_source(fdmod_ctx *ctx, const float *y_vx, const float *y_vz, const float *y_p, double *out);
/* Dot-product test on the device: *lhs = <F s, y> from a run from rest with the
 * current source s, *rhs = <s, F^T y>; both reductions in FP64 on the GPU. */
int fdmod_adjoint_dot(fdmod_ctx *ctx, const float *y_vx, const float *y_vz, const float *y_p, double *lhs,
                      double *rhs);

/* Measurement hook (no counterpart in the reference): runs min(reps, nt) loop
 * iterations from iteration 0, as fdmod_run does, and returns the mean duration of
 * the velocity sweep and of the stress sweep (with the loop's fused source and
 * receiver work; any velocity-source launch counts with the stress sweep) from HIP
 * events recorded on the context's stream between the launches.  No snapshots are
 * written.  The wavefield and the recorded traces advance; call fdmod_reset() before
 * reusing the context. */
int fdmod_time_sweeps(fdmod_ctx *ctx, int reps, double *ms_v, double *ms_s);

/* ---- Allocation-free run (no counterpart in the reference; for callers that cannot
 * allocate device memory: code under HIP graph capture, serving loops that must not call
 * hipMalloc).  Runs loop iterations k0 .. k0+count-1 as fdmod_run does and writes the
 * energy diagnostic of each (BT:926-952; the values fdmod_set_energy + fdmod_run +
 * fdmod_get_energy give) to E/Ek/Ep[j] for iteration k0 + j.  It calls no hipMalloc or
 * hipFree: its device scratch (the energy kernels' partial sums and the series of the
 * call) is one static array per GPU, held until the process ends.  At most
 * FDMOD_NOALLOC_MAX_STEPS iterations per call (longer runs: several calls); snapshots are
 * refused (wsave must be 0); the context's own energy series is not advanced.  Calls that
 * overlap on one GPU share that array and race. */
#define FDMOD_NOALLOC_MAX_STEPS 16384
int fdmod_run_energy_noalloc(fdmod_ctx *ctx, int k0, int count, float *E, float *Ek, float *Ep, double *seconds);

#ifdef __cplusplus
}
#endif
#endif /* FDMOD_HIP_H */
