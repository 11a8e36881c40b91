/*
 * fwi2d_hip.h -- C ABI of the MI355X (gfx950) elastic propagation engine.
 *
 * Four levels, all plain C (pointers, sizes, status codes):
 *
 *  1. Operator level: a drop-in for the reference's non-linear elastic operators
 *     nl_we_op_e / nl_we_op_vti (src/include/we_op.hpp:54-86, 150-176): the
 *     apply_forward / apply_jacobianT pair that the reference's objectives
 *     (nlls_fwi, nlls_fwi_eco: src/include/optimization.hpp:351, 452, 656, 890)
 *     and WE_MODELING.x (src/main/we_modeling.cpp:73) call.
 *
 *  1a. Objective and solver on top of it: nlls_fwi / nlls_fwi_eco over all shots in
 *     two batched phases (fwi2d_nlls_*), and the reference's l-BFGS with the regular
 *     Wolfe line search over an objective callback (fwi2d_lbfgs).
 *
 *  2. Engine level: batched device propagation that replaces the per-shot device
 *     hook nl_we_op_e::propagate_gpu (src/include/we_op.hpp:82, 174; body in
 *     src/cuda/cudaWeOp.cpp:10-876).  Many shots advance together; the adjoint
 *     recomputes the forward wavefield from device-resident checkpoints instead of
 *     reading a host-stored full wavefield.
 *
 * Conventions: arrays are [slowest]...[fastest]; models [np][nx][nz] (z fastest),
 * traces [component][trace][time].  Status 0 is success; on failure the call
 * returns non-zero and fwi2d_last_error() describes the error (the reference throws
 * std::logic_error from successCheck, src/cpp/misc.cpp:4-9).  Calls are synchronous
 * at return.  One engine or operator per host thread.
 */
#ifndef FWI2D_HIP_H
#define FWI2D_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------- */
/* Engine level                                                               */
/* ------------------------------------------------------------------------- */

/* Propagation grid and boundary set-up; the fields mean what the same-named param
   fields mean in the reference (src/include/param.hpp:10-41). */
typedef struct {
    int nx, nz;            /* model axes 2 and 1 */
    double dx, dz;
    int nt;                /* number of propagation time steps (par.nt) */
    double dt;             /* propagation time step (par.dt) */
    int sub;               /* imaging/snapshot sub-sampling (par.sub); 0 = forward only */
    int bc_top, bc_bottom, bc_left, bc_right;              /* 1 free surface, 2 absorbing */
    int taper_top, taper_bottom, taper_left, taper_right;  /* cos^2 taper widths (cells) */
    double taper_strength;
    int vti;               /* 0: model (lambda, mu, rho); 1: (lambda, mu, rho, c13, eps) */
    double gdamp_tmin, gdamp_tmax, gdamp_power;            /* gradient time window */
} fwi2d_grid;

typedef struct fwi2d_engine fwi2d_engine;

/* Create an engine on HIP device `device`.  No reference counterpart: replaces the
   per-call set-up of propagate_gpu (src/cuda/cudaWeOp.cpp:12-58). */
int fwi2d_engine_create(const fwi2d_grid* grid, int device, fwi2d_engine** out);

/* Upload the converted model [np][nx][nz] (np = 3 iso, 5 vti): lambda, mu, rho[, c13, eps],
   as produced by convert_model (src/cpp/we_op.cpp:476-494, 1441-1462).  Invalidates
   checkpoints of a previous forward run. */
int fwi2d_engine_set_model(fwi2d_engine* e, const float* model);

/* Acquisition: nshots sources at sxz[s] = (x, z) and nrec[s] point receivers per shot
   at rxz (all shots concatenated, (x, z) pairs), grid origin (ox, oz).  The point
   injector/extractor weights are those of delta_m3 (src/cpp/injector.cpp:3-64). */
int fwi2d_engine_set_shots(fwi2d_engine* e, int nshots, const double* sxz, const int* nrec,
                           const double* rxz, double ox, double oz);

/* fwi2d_engine_set_shots with the reference's other injectors (par.mt, par.gl):
   moment_tensor != 0: the sources are moment tensors injected with ddelta_m3
   (src/cpp/injector.cpp:114-239; we_op.cpp:640-654) and the source traces of
   fwi2d_engine_forward are [3][nshots][nt] = (mxx, mzz, mxz);
   gauge_length > 0: the receivers are DAS channels extracted and injected with dipole_m3
   (injector.cpp:241-378) at x -/+ gl/2 cos(dip), z -/+ gl/2 sin(dip), rdip [nrec_total]
   their dips (rad); each receiver trace pair [2][r][nt] then holds the differences of
   the two displacement components between the gauge ends (before dipole_to_strain). */
int fwi2d_engine_set_acquisition(fwi2d_engine* e, int nshots, const double* sxz, const int* nrec, const double* rxz,
                                 double ox, double oz, int moment_tensor, double gauge_length, const double* rdip);

/* Forward propagation of all shots (propagate(adj=false), src/cpp/we_op.cpp:564-902).
   src: [2][nshots][nt] source traces at the propagation step (x then z force);
   rcv: [2][nrec_total][nt] displacement traces (output).  keep_checkpoints != 0 keeps
   the device checkpoints needed by fwi2d_engine_adjoint.  on_device != 0 means src and
   rcv are device pointers (inputs already resident in HBM). */
int fwi2d_engine_forward(fwi2d_engine* e, const float* src, float* rcv, int keep_checkpoints, int on_device);

/* Adjoint propagation with imaging (propagate(adj=true) + compute_gradients,
   src/cpp/we_op.cpp:505-562, 1478-1557).  adj: [2][nrec_total][nt] adjoint source
   traces in adjoint time order (already time-reversed and scaled, as prepared by
   apply_jacobianT, src/cpp/we_op.cpp:1112-1127); grad: [np][nx][nz] raw gradient
   (before H quadrature and chain rule) summed over shots, overwritten.  Needs the
   checkpoints of a forward run on the same model. */
int fwi2d_engine_adjoint(fwi2d_engine* e, const float* adj, float* grad, int on_device);

/* FP64 reduction on the device: out = sum_i a_i * b_i * w_i (b, w may be NULL; b NULL
   means b = a).  Misfit, norms and line-search dot products (the data_t reductions of
   src/include/vecReg.hpp:257-325).  Pointers are device pointers. */
int fwi2d_engine_dot_f64(fwi2d_engine* e, const float* a, const float* b, const float* w, long long n,
                         double* out);

/* fwi2d_engine_dot_f64 without an engine and without any device allocation.  No
   counterpart in the reference: for callers that cannot allocate device memory (code
   under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  Same
   kernels and block count as fwi2d_engine_dot_f64, so the result is bit-identical to it.
   The block partials and the result live in a static device array of the library
   ((1024 + 1) doubles, 8 KB per device, held until the process ends), whose address is
   looked up once per device on the first call there; nothing else is allocated.  Runs
   on the null stream of `device` (after earlier work on the library's streams) and
   returns synchronously; the caller's current device is restored.  a, b, w are device
   pointers (b NULL means b = a, w NULL means w = 1); out is a host pointer.  n must lie
   in [0, FWI2D_DOT_NOALLOC_MAX_N].  Every call on one GPU uses the same array: calls
   that overlap on one GPU (from several host threads) race and must be serialised by
   the caller. */
#define FWI2D_DOT_NOALLOC_MAX_N (1LL << 40)
int fwi2d_dot_f64_noalloc(int device, const float* a, const float* b, const float* w, long long n, double* out);

/* Tuning and test options (no reference counterpart): "tile_x" (16 | 32 cells per tile
   along x), "shots_per_block" (1..64 shots advanced in turn by one step workgroup of an
   interior tile), "shots_per_block_edge" (the same for tiles at the grid edge),
   "shots_per_block_edge_img" (the same for edge tiles of adjoint steps with the imaging
   fused in; 0, the default: "shots_per_block_edge"),
   "edge_stream" (1: edge-band launches run on a second stream beside the core launch,
   enqueued before it; 2: the same, enqueued after it; 0: after it on the same stream),
   "shots_per_block_img" (shots per imaging workgroup; 0: 4, or "shots_per_block" when
   fused),
   "image_fused" (1: the imaging condition runs inside the adjoint step kernels, by parts
   in time (with sub = 1 and the imaging lag; otherwise as 0); 0: in its own kernel after
   each adjoint step and its point sources; -1, the default: 1 on isotropic grids, 0 on
   VTI grids; the stat reads the choice in effect),
   "image_buffers" (1, the default, or 2 LDS window buffers in the imaging kernel; with
   2 the next shot's windows load during this shot's imaging),
   "image_stream" (1, the default: the imaging kernel of adjoint step it runs on a second
   stream beside step it+1; 0: on the engine stream; results are identical),
   "image_sbp" (1, the default for sub = 1: the density term of the imaging condition is
   summed by parts in time, one snapshot per imaging launch and four rotating adjoint
   levels; 0: the reference's direct form with three snapshots per launch; the two agree
   to rounding),
   "autotune" (1: at the next forward run, time a few steps of every tile_x,
   shots-per-workgroup and edge_stream choice and keep the fastest, edge_stream 2 winning
   ties within 1.5% (it replays the gradient faster); forward results are unchanged),
   "checkpoint_format" (0 FP32; 1 BF16, 2 FP16: checkpoint levels stored in 16 bits with
   a max-abs scale per level and shot, which halves their memory; the gradient then
   carries the error of the recomputation from rounded levels),
   "stiffness_gradients" (VTI, 1: fwi2d_engine_adjoint returns raw gradients w.r.t.
   c11, c13, c33, c55, rho instead of lambda, mu, rho, 0, 0),
   "segment" (snapshots per checkpoint segment, 0 = about sqrt(2 nt/sub), or with
   "checkpoint_budget" the longest that fits it), "checkpoint_budget" (MiB of device
   memory for the checkpoint levels and the segment's snapshot slots; 0, the default: the
   sqrt scheme's minimum.  Longer segments cost the same recomputation per step, and the
   last segment, which the forward run keeps whole, is not recomputed: a budget of about
   nt/sub / 4 levels saves a quarter of the recomputation), "lag" (1 the
   reference's one-step imaging lag, 0 none), "timing" (1 time the forward step loop with
   HIP events), "graph" (1: capture each forward / adjoint time loop in a hipGraph the
   first time it runs in a configuration and replay it afterwards).  Statistics: "steps", "recompute_steps",
   "image_calls", "segment", "nseg", "tile_x", "shots_per_block", "shots_per_block_edge", "shots_per_block_edge_img", "edge_stream",
   "image_fused", "image_buffers", "image_stream", "image_sbp", "ntiles", "interior_tiles", "graph_captures", "graph_replays", "nrec",
   "nshots", "fwd_loop_ns", "fwd_loop_steps", "adj_ns" (with the "timing" option: the last
   adjoint run's device time, recomputation included); -1 for an unknown key. */
int fwi2d_engine_set_option(fwi2d_engine* e, const char* key, long long value);
long long fwi2d_engine_get_stat(fwi2d_engine* e, const char* key);

/* Timings of the last autotuning run, "tile_x=16 spb=4 spb_edge=1: 512.3 us; ..."
   ("" before any); valid until the next call on the engine. */
const char* fwi2d_engine_tune_report(fwi2d_engine* e);

void fwi2d_engine_destroy(fwi2d_engine* e);

/* Message of the last failed call on this thread ("" if none). */
const char* fwi2d_last_error(void);

/* ------------------------------------------------------------------------- */
/* Acoustic engine: nl_we_op_a::propagate (src/cpp/we_op.cpp:1937-2135) for all */
/* shots at once (the reference runs this operator on the CPU only)            */
/* ------------------------------------------------------------------------- */
typedef struct fwi2d_acoustic fwi2d_acoustic;

/* The grid as for the elastic engine (vti ignored); the boundaries are 1 (free surface,
   the Dirichlet SAT asat_dirichlet_*), 2 (absorbing, asat_absorbing_*) or 3 (rigid,
   asat_neumann_*); free_surface_stiffness is par.free_surface_stiffness (raised to 1 as
   the reference does for this operator, we_op.cpp:16; the SAT parameter is
   0.2508560249 / stiffness, we_op.cpp:1962). */
int fwi2d_acoustic_create(const fwi2d_grid* grid, double free_surface_stiffness, int device, fwi2d_acoustic** out);

/* The converted model [2][nx][nz]: K = rho vp^2 and 1/rho (nl_we_op_a::convert_model,
   we_op.cpp:1863-1881).  Invalidates snapshots of a previous forward run. */
int fwi2d_acoustic_set_model(fwi2d_acoustic* e, const float* model);

/* Point sources and receivers with delta_m3 footprints (src/cpp/injector.cpp:3-64), as
   fwi2d_engine_set_shots. */
int fwi2d_acoustic_set_shots(fwi2d_acoustic* e, int nshots, const double* sxz, const int* nrec, const double* rxz,
                             double ox, double oz);

/* propagate(adj = false) of every shot: source traces src [nshots][nt], receiver traces
   rcv [nrec total][nt]; keep_snapshots: keep what fwi2d_acoustic_adjoint needs of the
   1 + nt/sub snapshots per shot (we_op.cpp:923) on the device: all of them when they fit
   the "snapshot_budget" option, otherwise checkpoints of segments of about sqrt(2 nsnap)
   snapshots, from which the adjoint run recomputes them (one more forward run's steps;
   the source traces are copied for it).  on_device: src and rcv are device pointers. */
int fwi2d_acoustic_forward(fwi2d_acoustic* e, const float* src, float* rcv, int keep_snapshots, int on_device);

/* propagate(adj = true) with compute_gradients (we_op.cpp:1883-1935): adjoint sources adj
   [nrec total][nt] injected at the receivers; grad [2][nx][nz] receives the raw K and
   1/rho gradients summed over the shots (before the H quadrature and the chain rule of
   apply_jacobianT, we_op.cpp:1192-1230). */
int fwi2d_acoustic_adjoint(fwi2d_acoustic* e, const float* adj, float* grad, int on_device);

/* Options (no counterpart in the reference, which keeps every snapshot on the host):
   "snapshot_budget" (MiB, default 131072): the largest snapshot history kept whole; 0
   always checkpoints; "image_fused" (1: the adjoint step kernel images the snapshots as
   it steps; 0, the default and faster: a separate imaging kernel before each imaged step;
   the gradients are identical).
   Statistics: "image_fused", "recompute_steps" (total), "whole_snapshots" (1 or 0,
   the last forward run's choice), "segment", "nseg", "snapshot_budget",
   "shots_per_block" (of the step kernel); -1 for an unknown key. */
int fwi2d_acoustic_set_option(fwi2d_acoustic* e, const char* key, long long value);
long long fwi2d_acoustic_get_stat(fwi2d_acoustic* e, const char* key);

void fwi2d_acoustic_destroy(fwi2d_acoustic* e);

/* ------------------------------------------------------------------------- */
/* Operator level (drop-in for nl_we_op_e / nl_we_op_vti / nl_we_op_a)         */
/* ------------------------------------------------------------------------- */

/* Operator description: the reference's param fields used on this path
   (src/include/param.hpp:10-82) plus the model domain and source time axis.
   Point forces or moment tensors (mt), point receivers or DAS channels (gl).  With
   nparam = 2 the operator is the acoustic nl_we_op_a (model (vp, rho); one source and
   one data component, allsrc [ns][nt_data] and data [nrec][nt_data]; mt and gl ignored,
   as the reference's analyzeGeometry does, we_op.cpp:13-20). */
typedef struct {
    int nx, nz; double dx, dz, ox, oz;            /* model domain axes 2, 1 */
    int nt_data; double dt_data;                  /* source/data time axis (origin 0) */
    int nparam;                                   /* 2 acoustic (vp,rho), 3 iso (vp,vs,rho) or
                                                     5 vti (+delta,eps) */
    int ns; const double* sxz;                    /* shots (x, z) */
    const int* nrec; const double* rxz;           /* receivers per shot, (x, z) concatenated */
    double courant, dt;                           /* par.courant, par.dt (<0 auto) */
    int sub;                                      /* par.sub (<0 auto, 0 forward only) */
    int bc_top, bc_bottom, bc_left, bc_right;
    int taper_top, taper_bottom, taper_left, taper_right; double taper_strength;
    int seismotype;                               /* 0 displacement, 1 particle velocity */
    int resampling;                               /* 0 sinc, 1 linear */
    int sinc_half_length;
    double vpmin, vpmax, vsmin, vsmax, rhomin, rhomax, deltamin, deltamax, epsilonmin, epsilonmax;
    int soft_clip;
    double gdamp_tmin, gdamp_tmax, gdamp_power;
    int device;
    int mt;                                       /* par.mt: moment-tensor sources, allsrc [3][ns][nt_data]
                                                     = (mxx, mzz, mxz) (analyzeWavelet, we_op.cpp:216-238) */
    double gl;                                    /* par.gl > 0: DAS receivers of this gauge length (km);
                                                     data [2][nrec][nt_data] with the strain (rate) in
                                                     component 0 and zeros in 1 (dipole_to_strain,
                                                     misc.cpp:18-43) */
    const double* rdip;                           /* par.rdip: receiver dips (rad), nrec_total values,
                                                     read when gl > 0 */
    double free_surface_stiffness;                /* par.free_surface_stiffness of the acoustic
                                                     free surface (raised to 1 as the reference
                                                     does; 0 reads as 1) */
} fwi2d_op_desc;

typedef struct fwi2d_op fwi2d_op;

/* Constructor of nl_we_op_e / nl_we_op_vti (we_op.hpp:58-66) and nl_we_op_a
   (we_op.hpp:238-245, desc->nparam = 2): allsrc is the source after analyzeWavelet,
   [2][ns][nt_data] (point forces), [3][ns][nt_data] (moment tensors, desc->mt) or
   [1][ns][nt_data] (acoustic). */
int fwi2d_op_create(const fwi2d_op_desc* desc, const float* allsrc, fwi2d_op** out);

/* nl_we_op::apply_forward (we_op.hpp:48; we_op.cpp:904-1055):
   pdat = add*pdat + F(pmod); pmod [nparam][nx][nz] physical model (km/s, g/cc),
   pdat [2][nrec_total][nt_data] ([1][nrec_total][nt_data] for the acoustic operator).
   Keeps device checkpoints for apply_jacobianT. */
int fwi2d_op_apply_forward(fwi2d_op* op, int add, const float* pmod, float* pdat);

/* nl_we_op::apply_jacobianT (we_op.hpp:49; we_op.cpp:1057-1239):
   pmod = add*pmod + J(pmod0)^T pdat, gradient w.r.t. (vp, vs, rho[, 0, 0]), or (vp, rho)
   for the acoustic operator. */
int fwi2d_op_apply_jacobianT(fwi2d_op* op, int add, float* pmod, const float* pmod0, const float* pdat);

/* Access to the engine of an operator (batch options, statistics); NULL before the
   operator's first apply_forward / apply_jacobianT and for the acoustic operator. */
fwi2d_engine* fwi2d_op_engine(fwi2d_op* op);

/* Propagation time axis chosen by analyzeModel for the last call. */
int fwi2d_op_time_axis(fwi2d_op* op, int* nt, double* dt, int* sub);

void fwi2d_op_destroy(fwi2d_op* op);

/* ------------------------------------------------------------------------- */
/* Least-squares objective on the operator (nlls_fwi / nlls_fwi_eco)           */
/* ------------------------------------------------------------------------- */

typedef struct fwi2d_nlls fwi2d_nlls;

/* Options of the objective (bit flags):
   FWI2D_NLLS_NORMALIZE_OBJ: nlls_fwi_eco's par.normalize_obj_func: functional and gradient
     divided by dnorm = dt_data * |P d|^2 (1 below 1e-16); without it dnorm = 1 (nlls_fwi);
   FWI2D_NLLS_INTEGRATE: par.integrate / nlls_fwi's integrate: data and synthetics
     integrated in time (integral, src/include/operator.hpp:650-689);
   FWI2D_NLLS_TRACE_NORMALIZE: par.normalize / nlls_fwi's normalize: every trace divided
     by its L2 norm (ttnormalize, src/cpp/misc.cpp:97-110). */
#define FWI2D_NLLS_NORMALIZE_OBJ 1
#define FWI2D_NLLS_INTEGRATE 2
#define FWI2D_NLLS_TRACE_NORMALIZE 4

/* Constructors of nlls_fwi (src/include/optimization.hpp:309-344) and nlls_fwi_eco
   (optimization.hpp:482-590) without envelope, filter, interferometry, double difference
   or source rescaling.  d: observed data [2][nrec_total][nt_data]; w: residual weights
   of the same shape, or NULL; gmask: gradient mask [nparam][nx][nz], or NULL; flags: the
   options above.  The data are processed once as the constructors do: P d = normalise(W
   integrate(d)).  Negative weights or mask values are rejected as in the reference.
   The objective keeps a copy of d, w and gmask, and uses op until it is destroyed. */
int fwi2d_nlls_create(fwi2d_op* op, const float* d, const float* w, const float* gmask, int flags,
                      fwi2d_nlls** out);

/* par.filter_phase (src/include/param.hpp:58): how the objective's 1-D filter is prepared
   before use (optimization.hpp:526-534).  DEFAULT: as given, convolved centred; ZERO:
   zero_phase (src/cpp/operator.cpp:2197-2250), centred; MINIMUM: minimum_phase
   (operator.cpp:2252-2320, eps 1e-7), convolved causally (not centred).  The zero and
   minimum phases pad an even-length filter with one zero sample. */
#define FWI2D_FILTER_DEFAULT 0
#define FWI2D_FILTER_ZERO 1
#define FWI2D_FILTER_MINIMUM 2

/* nlls_fwi_eco's constructor with its filter argument (optimization.hpp:483, 526-539):
   as fwi2d_nlls_create, plus a 1-D filter of nf samples on the data time axis (or NULL)
   prepared by filter_phase.  Observed data and synthetics are convolved with it in time
   (conv1dnd, src/include/operator.hpp:1035-1057; src/cpp/operator.cpp:2092-2137) after
   the weights and before the trace normalisation, and the adjoint source with its
   adjoint (optimization.hpp:688-696, 846-854). */
int fwi2d_nlls_create_filtered(fwi2d_op* op, const float* d, const float* w, const float* gmask, int flags,
                               const float* filter, int nf, int filter_phase, fwi2d_nlls** out);

/* The filter preparation alone (host only, no device): out receives the prepared filter
   in double precision (room for nf + 1 values), *nout its length. */
int fwi2d_filter_prepare(const float* filter, int nf, int filter_phase, double* out, int* nout);

/* conv1dnd apply_forward / apply_adjoint (operator.cpp:2092-2137) alone, host only, in
   double precision: out = C in (adjoint: C^T in) for ntr traces of nt samples, with the
   filter's centred (shift floor((nf + 1)/2) - 1) or causal convolution, on up to 16
   host threads; in and out must not overlap. */
int fwi2d_filter_apply(const double* filter, int nf, int centered, int adjoint, long long ntr, int nt,
                       const double* in, double* out);

/* res + getFunc + grad at the physical model m (nlls_fwi: optimization.hpp:346-455;
   nlls_fwi_eco: compute_res_and_grad, getFunc, grad, optimization.hpp:592-951), over all
   shots in two batched phases (forward, then adjoint):
     r = P F(m) - P d,
     func = 1/2 dt (|r|^2 - 1/2 sum over traces of r[0]^2 + r[nt-1]^2) / dnorm  (FP64),
     g = gmask * J(m)^T P'(F(m))^T Ht r / dnorm     [nparam][nx][nz],
   with P' the Jacobian of the processing (integral adjoint, weights, the trace
   normalisation's I/|u| - u u^T/|u|^3). 
   func or g may be NULL (g NULL: residual and functional only, no adjoint). */
int fwi2d_nlls_evaluate(fwi2d_nlls* f, const float* m, double* func, float* g);

/* Residual r of the last evaluation ([2][nrec_total][nt_data], owned by f), or NULL. */
const float* fwi2d_nlls_residual(fwi2d_nlls* f);

/* The normalisation dnorm of the functional and gradient (1 for nlls_fwi). */
double fwi2d_nlls_dnorm(fwi2d_nlls* f);

void fwi2d_nlls_destroy(fwi2d_nlls* f);

/* ------------------------------------------------------------------------- */
/* Solver: l-BFGS with the regular Wolfe line search                         */
/* ------------------------------------------------------------------------- */

/* Objective callback: the functional *f and the gradient g (n values) at model m;
   returns 0 on success.  The least-squares objective itself is one:
   fwi2d_nlls_objective with ctx = the fwi2d_nlls handle. */
typedef int (*fwi2d_objective_fn)(void* ctx, const float* m, double* f, float* g);
int fwi2d_nlls_objective(void* ctx, const float* m, double* f, float* g);

/* Parameters of lbfgs (src/include/nlsolver.hpp:308-324) and regular_wolfe
   (nlsolver.hpp:85-114); the reference's defaults are memory 5, c1 1e-4, c2 0.9, a0 1,
   a1 0, step_max 1e6, version 0. */
typedef struct {
    int niter;          /* iterations */
    int max_trial;      /* trial steps per line search */
    double threshold;   /* stop when the relative decrease of the functional falls below it */
    int memory;         /* stored (s, y) pairs */
    double c1, c2;      /* sufficient decrease and curvature constants, 0 < c1 < c2 < 1 */
    double a0, a1;      /* first step a0 / |g|, or a1 |m| / |g| if larger */
    double step_max;    /* largest step length */
    int version;        /* 0: next first step stp * df_prev / df; 1: 2 (f - f_prev) / df, at most 1 */
    const float* H0;    /* initial diagonal inverse Hessian (n values > 0), or NULL for 1 */
} fwi2d_lbfgs_options;

/* What the reference prints about a run at verbose=1 (nlsolver.cpp:1262-1346 and the
   trial blocks of regular_wolfe::lineSearch / zoom, nlsolver.cpp:236-311, 350-420):
   the evaluation counts, as lbfgs::run counts them (a function evaluation per objective
   call, a gradient evaluation for the starting model and each trial whose gradient the
   search reads), and one row per line-search trial. */
typedef struct {
    int feval, geval;   /* output: function / gradient evaluations */
    int ntrials;        /* output: trials made (rows beyond trials_cap are not stored) */
    int trials_cap;     /* input: rows available in trials */
    double* trials;     /* input: [trials_cap][5] or NULL; each row: iteration (from 0), trial
                           (from 1), step length, function value, gradient norm (NaN when the
                           trial's gradient was not read: "Last gradient norm" in the reference) */
} fwi2d_lbfgs_report;

/* lbfgs::run (src/cpp/nlsolver.cpp:1207-1415) with regular_wolfe::lineSearch / zoom
   (nlsolver.cpp:184-421) on model m (n values, updated in place).  Every evaluation
   returns functional and gradient together, as nlls_fwi_eco's res() does.
   func_history (niter + 1 values, or NULL) receives the functional before the first
   iteration and after each accepted step; *iterations the accepted steps; report (or
   NULL) the evaluation counts and the trials.  Stops, as the reference, after niter
   iterations, a relative decrease below threshold, a functional or gradient norm below
   1e-10, a failed line search (the model is then the last trial) or a search direction
   that is not a descent direction. */
int fwi2d_lbfgs(fwi2d_objective_fn fn, void* ctx, long long n, float* m, const fwi2d_lbfgs_options* opt,
                double* func_history, int* iterations, fwi2d_lbfgs_report* report);

/* ------------------------------------------------------------------------- */
/* On-disk formats (host only; no device)                                     */
/* ------------------------------------------------------------------------- */

/* read<T>(file, format) / write<T>(vec, file, format) (src/include/IO.hpp:432-443):
   FWI2D_FORMAT_SEP is a SEPlib header with its binary (sepRead / sepWrite, IO.hpp:14-176),
   FWI2D_FORMAT_BINARY the plain descriptor of binRead / binWrite (IO.hpp:180-430). */
#define FWI2D_FORMAT_SEP 0
#define FWI2D_FORMAT_BINARY 1

/* A regular hypercube of up to 6 axes (n, o, d, label; axis 1 fastest) and the binary it
   describes. */
typedef struct {
    int ndim;            /* axes kept: trailing unit axes are dropped, down to ndim0 */
    long long n[6];
    double o[6], d[6];
    char label[6][64];
    int esize;           /* bytes per element in the file: 4 or 8 */
    int xdr;             /* 1: SEPlib data_format "xdr_*" (big-endian 4-byte words); 0: native */
    char in[4096];       /* path of the binary */
} fwi2d_hypercube;

/* The header of `path`: n1..n6, o1..o6, d1..d6, label1..label6 (defaults 1, 0, 1, ""),
   esize (default 4), in=, and for SEPlib data_format (default native).  A later
   definition of a key overrides an earlier one (SEPlib header history); the binary
   format skips lines starting with '#'.  A relative in= path that does not exist from the
   working directory is looked up next to the header. */
int fwi2d_read_header(const char* path, int format, int ndim0, fwi2d_hypercube* h);

/* The n1 * ... * n6 values of the binary of h, converted to single precision (data_t);
   the binary's size must equal n1 * ... * n6 * esize (IO.hpp:404-406). */
int fwi2d_read_data(const fwi2d_hypercube* h, float* data);

/* Writes the header `path` and its binary <datapath>/<name of path>@ (datapath NULL or
   "": $DATAPATH, else the header's directory) with the ndim axes of h and n1 * ... * nN
   single-precision values: SEPlib as data_format="xdr_float" (srite's default), the
   binary format in native byte order (binWrite).  h->esize, h->xdr and h->in are ignored. */
int fwi2d_write(const char* path, int format, const fwi2d_hypercube* h, const float* data, const char* datapath);

/* readCoord (src/cpp/param.cpp:4-77): the source-receiver file (lines "id sx sz rx rz
   [dip]", '#' comments; consecutive lines with the same integer id form one shot).
   Call with sxz, nrec, rxz NULL to get *ns and *nrec_total; then with arrays of those
   sizes (sxz [ns][2], nrec [ns], rxz [nrec_total][3] = x, z, dip) and *ns, *nrec_total
   holding their capacities.  Values are read in single precision as the reference's
   data_t.  Blank lines are skipped (the reference would repeat the previous line's id). */
int fwi2d_read_coord(const char* path, int* ns, long long* nrec_total, double* sxz, int* nrec, double* rxz);

/* The regular acquisition of readParameters (param.cpp:221-244) in the reference's
   single-precision arithmetic: shot s at (sx0 + s sxinc, sz0 + s szinc), receiver r of
   every shot at (rx0 + r rxinc, rz0 + r rzinc) with dip rdip0 + r rdipinc.
   sxz [ns][2], rxz [ns][nr][3]. */
int fwi2d_regular_geometry(int ns, float sx0, float sz0, float sxinc, float szinc, int nr, float rx0, float rz0,
                           float rxinc, float rzinc, float rdip0, float rdipinc, double* sxz, double* rxz);

#ifdef __cplusplus
}
#endif
#endif /* FWI2D_HIP_H */
