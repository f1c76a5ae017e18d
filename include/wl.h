/*
 * wl.h -- C ABI of the MI355X (gfx950) WaterLily time-step library (libwaterlily_hip.so).
 *
 * Plain pointers and sizes only.  Every host buffer is in the memory layout Julia uses for the
 * corresponding WaterLily array: column-major, x fastest, ghost layers included, the vector
 * component index last (u[x,y,z,i], mu1[x,y,z,i,j]).  A `ccall` can therefore pass
 * `pointer(A)` of the reference's own arrays.
 *
 * All entry points return 0 on success and a negative WL_E* code on failure; the message of the
 * last failure on the calling thread is returned by wl_last_error().  Calls on one handle are not
 * re-entrant; separate handles are independent.  Every call is synchronous for the host outputs
 * it returns.
 *
 * Reference interfaces replaced (WaterLily.jl v1.8.0; paths relative to the WaterLily repository):
 *   wl_flow_create      Flow(N, uBC; Δt, ν, g, λ, T, perdir, exitBC)      src/Flow.jl:133-147
 *   wl_flow_upload /    the fields u, u⁰, f, p, σ, V, μ₀, μ₁ of Flow          src/Flow.jl:114-125
 *     wl_flow_download
 *   wl_flow_measure     measure!(flow, body; t, ϵ) for analytic bodies     src/Body.jl:28-51
 *   wl_flow_pois_update update!(::MultiLevelPoisson)                        src/MultiLevelPoisson.jl:79-86
 *   wl_mom_step         mom_step!(flow, pois) incl. CFL                     src/Flow.jl:156-167,234-244
 *   wl_pressure_force   pressure_force(flow, body)                          src/Metrics.jl:126-133
 *   wl_pois_create      Poisson(x,L,z) / MultiLevelPoisson(x,L,z)          src/Poisson.jl:32-38,
 *                                                                           src/MultiLevelPoisson.jl:61-77
 *   wl_pois_mult        mult!(pois, x)                                      src/Poisson.jl:63-69
 *   wl_pois_solve       solver!(pois; tol, itmx)                            src/Poisson.jl:204-214,
 *                                                                           src/MultiLevelPoisson.jl:108-127
 *   wl_pois_update      update!(pois)                                       src/Poisson.jl:47
 */
#ifndef WL_H
#define WL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ---- */
#define WL_OK 0
#define WL_EINVAL (-1)   /* bad argument: shape, size, enum (the reference's @assert cases) */
#define WL_EHIP (-2)     /* a HIP runtime call failed */
#define WL_ENAN (-3)     /* a NaN/Inf reached the CFL reduction */
#define WL_ENOMEM (-4)   /* device allocation failed */

/* ---- enums ---- */
enum wl_dtype { WL_F32 = 0, WL_F64 = 1 };
/* convective scheme λ(u,c,d): quick (default), vanLeer, cds -- src/Flow.jl:4-6 */
enum wl_scheme { WL_QUICK = 0, WL_VANLEER = 1, WL_CDS = 2 };
/* Flow fields -- src/Flow.jl:114-125 */
enum wl_field { WL_U = 0, WL_U0 = 1, WL_F = 2, WL_P = 3, WL_SIGMA = 4, WL_V = 5, WL_MU0 = 6, WL_MU1 = 7 };
/* Poisson level fields -- src/Poisson.jl:22-31 */
enum wl_pfield { WL_PX = 0, WL_PL = 1, WL_PD = 2, WL_PID = 3, WL_PEPS = 4, WL_PR = 5, WL_PZ = 6 };
/* Pressure solver modes */
enum wl_solver_mode {
  WL_SOLVER_PLAIN = 0, /* MultiLevelPoisson exactly as the reference: FP32 V-cycles, GS-RB at every level */
  WL_SOLVER_MIXED = 1  /* FP32 V-cycles inside an FP64 defect-correction loop */
};
/* Analytic bodies.  A body list is a union (SetBody(min,...), src/Body.jl:91-103). */
enum wl_body_kind {
  WL_BODY_SPHERE = 1,    /* AutoBody((x,t)->√sum(abs2,x.-c)-r); a circle in 2D */
  WL_BODY_ELLIPSOID = 2  /* AutoBody((x,t)->(√sum(abs2,(x.-c)./a)-1)*minimum(a)) */
};

typedef struct {
  int32_t kind;       /* wl_body_kind */
  int32_t reserved;
  double center[3];
  double radius;      /* sphere radius */
  double axes[3];     /* ellipsoid semi-axes */
} wl_body;

/* Flow(N, uBC; ...) arguments, src/Flow.jl:133-134 and src/WaterLily.jl:93-98.
 * The boundary velocity is uniform in space: uBC(i,x,t) = uBC[i] + uBC_rate[i]*t.
 * The body force is uniform in space: g(i,x,t) = g0[i] + g1[i]*t (has_g = 0 means g=nothing). */
typedef struct {
  int32_t D;            /* 2 or 3 */
  int32_t n[3];         /* interior cells per direction (n[2] ignored when D==2) */
  int32_t dtype;        /* wl_dtype */
  int32_t scheme;       /* wl_scheme */
  int32_t perdir_mask;  /* bit j set: direction j+1 periodic */
  int32_t exit_bc;      /* convective exit in direction 1 */
  int32_t has_g;
  int32_t has_uBC_rate;
  double uBC[3];
  double uBC_rate[3];
  double g0[3];
  double g1[3];
  double nu;
  double dt0;
} wl_flow_desc;

typedef struct wl_flow_s* wl_flow_t;
typedef struct wl_pois_s* wl_pois_t;

const char* wl_last_error(void);
int wl_device_count(int* n);

/* ---- Flow: src/Flow.jl ---- */
int wl_flow_create(wl_flow_t* out, const wl_flow_desc* desc);
int wl_flow_destroy(wl_flow_t h);
/* element count of a field (ghosted cells x components) */
int wl_flow_field_size(wl_flow_t h, int field, size_t* count);
int wl_flow_upload(wl_flow_t h, int field, const void* src, size_t nbytes);
int wl_flow_download(wl_flow_t h, int field, void* dst, size_t nbytes);
int wl_flow_measure(wl_flow_t h, const wl_body* bodies, int nbodies, double t, double eps);
int wl_flow_pois_update(wl_flow_t h);
int wl_flow_set_solver(wl_flow_t h, int mode, double tol, int itmx);
/* One mom_step!: dt = Δt[end], t0 = t1 - dt, t1 = sum(Δt).  Returns the CFL time step that the
 * caller pushes onto Δt, and the V-cycle counts of the two projections (pois.n entries). */
int wl_mom_step(wl_flow_t h, double dt, double t0, double t1, double* dt_next, int16_t* n_iter);
/* F = Σ p·nds(body) with the product in T and the sum in Float64; F has D entries. */
int wl_pressure_force(wl_flow_t h, const wl_body* bodies, int nbodies, double t, double* F);
/* pressure_force, viscous_force, pressure_moment and viscous_moment (src/Metrics.jl:126-197) of the
 * flow's p and u about x0 (D entries): each cell's density is formed in T as the reference forms it
 * (moments: loc − x₀ and the cross product in Float64, rounded to T) and the densities are summed
 * in Float64.  Fp, Fv: D entries; Mp, Mv: 1 entry in 2D (the 2-vector cross product), 3 in 3D;
 * any output may be NULL.  Replaces the Julia functions of the same names (Metrics.jl:146-197). */
int wl_flow_metrics(wl_flow_t h, const wl_body* bodies, int nbodies, double t, const double* x0, double* Fp,
                    double* Fv, double* Mp, double* Mv);
/* wl_flow_metrics with no device allocation.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory at call time (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc), where wl_flow_metrics allocates its block-partial buffer
 * on its first call.  Same arguments and results, bit for bit.  The partials go to a static device
 * array of 12 MiB per device (flows of at most 131072 blocks of 256 inside cells, e.g. 512×256×256;
 * larger flows return WL_EINVAL).  All calls on one device share that array: calls that overlap on
 * one device race. */
int wl_flow_metrics_noalloc(wl_flow_t h, const wl_body* bodies, int nbodies, double t, const double* x0,
                            double* Fp, double* Fv, double* Mp, double* Mv);

/* MeanFlow (src/Metrics.jl:199-261): running averages P, U (and UU = u⊗u with uu_stats) on the
 * flow's device.  The time vector t stays with the caller, which forms each update's weight
 * ε = dt/(dt + time(meanflow) + eps(T)) (1 on the first update) as update!(meanflow, flow) does
 * (:229-241).  Fields download in Julia layout: P (Ng), U (Ng×D), UU and τ = uu (Ng×D×D). */
typedef struct wl_meanflow_s* wl_meanflow_t;
enum wl_mfield { WL_MF_P = 0, WL_MF_U = 1, WL_MF_UU = 2, WL_MF_TAU = 3 };
int wl_meanflow_create(wl_meanflow_t* out, wl_flow_t flow, int uu_stats); /* MeanFlow(flow; uu_stats), :208-216 */
int wl_meanflow_destroy(wl_meanflow_t m);
int wl_meanflow_update(wl_meanflow_t m, wl_flow_t flow, double eps);     /* update!(meanflow, flow), :229-241 */
int wl_meanflow_reset(wl_meanflow_t m);                                  /* reset!(meanflow), :222-227 (fields) */
int wl_meanflow_download(wl_meanflow_t m, int field, void* dst, size_t nbytes); /* P, U, UU; τ = uu(m), :243-252 */
int wl_meanflow_copy_to_flow(wl_meanflow_t m, wl_flow_t flow);           /* copy!(flow, meanflow), :254-257 */
/* Put stored averages (P, U, UU) back on the device: load!(meanflow), ext/WaterLilyJLD2Ext.jl:57-70 */
int wl_meanflow_upload(wl_meanflow_t m, int field, const void* src, size_t nbytes);

/* Space- and time-dependent boundary values and forcing (function uBC, g), evaluated by the caller
 * on every cell in the layout of u (Ng…×D) and valid for the next wl_mom_step:
 *   WL_FN_UBC:  uBC(i, loc(i,I), t₁); BC! then runs the reference's loops on it (src/core.jl:200-219);
 *   WL_FN_ACC0: g(i, loc(i,I), t₀) + ∂uBC/∂t(i, loc(i,I), t₀), added by the predictor's
 *               accelerate! (src/Flow.jl:69-73,193);
 *   WL_FN_ACC1: the same at t₁ for the corrector (:208).
 * A field replaces the descriptor's uniform values; src = NULL clears it.  This is how a Julia
 * closure `uBC`/`g` reaches the device: the shim evaluates it with `apply!` on the host. */
enum wl_fn_field { WL_FN_UBC = 0, WL_FN_ACC0 = 1, WL_FN_ACC1 = 2 };
int wl_flow_set_field_fn(wl_flow_t h, int which, const void* src, size_t nbytes);
/* After the caller uploaded u (the initial condition uBC(i,x,0), src/Flow.jl:138): BC!(u, uBC,
 * exitBC, perdir) at t = 0 with the WL_FN_UBC field if set, exitBC!(u,u,0) and u⁰ = u (:140-141). */
int wl_flow_apply_ic(wl_flow_t h);

/* udf!(flow, force!, u, t) (src/Flow.jl:247-257): a user function that mom_step! calls in the
 * predictor (phase 0, t = t₀, the advecting field is u⁰) and in the corrector (phase 1, t = t₁,
 * the advecting field is u), after conv_diff! and before accelerate! and BDIM! (:193,207).  It runs
 * on the calling thread with the stream idle and may download and upload any field of `h`
 * (typically f, which then holds conv_diff!'s result); a nonzero return aborts the step with
 * WL_EINVAL.  fn = NULL removes it.  Replaces the `udf` keyword of mom_step!/sim_step!
 * (src/WaterLily.jl:128-140), which a Julia closure cannot cross the C ABI to provide otherwise. */
typedef int (*wl_udf_fn)(void* user, wl_flow_t h, int phase, double t);
int wl_flow_set_udf(wl_flow_t h, wl_udf_fn fn, void* user);

/* CFL(a) (src/Flow.jl:234-237) of the flow's current u, as mom_step! ends with it: used when a
 * restart replaces u (load!, ext/WaterLilyReadVTKExt.jl:38). */
int wl_flow_cfl(wl_flow_t h, double* dt);

/* Batches: many flows of different sizes stepped together, one kernel launch per stage of the step
 * for all of them.  No counterpart in the reference, which steps one Simulation at a time
 * (src/WaterLily.jl:127-140 is the loop each member follows).  Members are ordinary flow handles
 * (created, measured and pois-updated as usual, still owned by the caller); they share D, dtype and
 * scheme, use the plain solver, and have no periodic directions, no exitBC and no warm-start
 * network attached (checked when the batch is created).  One wl_batch_step is one mom_step! of
 * every member (dt, t0, t1, dt_next: one entry per member; n_iter: two per member) with the same
 * results as wl_mom_step on each. */
typedef struct wl_batch_s* wl_batch_t;
int wl_batch_create(wl_batch_t* out, const wl_flow_t* flows, int nflows);
int wl_batch_destroy(wl_batch_t b);
int wl_batch_step(wl_batch_t b, const double* dt, const double* t0, const double* t1, double* dt_next,
                  int16_t* n_iter);

/* Learned warm start of the pressure solves.  No counterpart in the reference, whose warm start is
 * the previous pressure, x = p·Δt (mom_project!, src/Flow.jl:225).  A network (a bias-free 3D
 * U-Net of 3×3×3 convolutions, 16/32/64/64 channels, bf16 weights and activations, FP32
 * accumulation on the MFMA units) maps the warm start's defect, the warm start and the geometry,
 * on the multigrid's level-1 grid, to a guess of the increment that is added to x before the solver
 * loop (src/MultiLevelPoisson.jl:108-127), which still iterates to the same Σr² tolerance.  The
 * network trains on the device (Adam) from pairs the solver records: (input, solved x − warm
 * start), both on the level-1 grid.
 *   grid: the network grid, the flow's interior halved in each direction (multiples of 8).
 *   weights: canonical FP32, the 9 layers concatenated, each w[tap][ci][co] with
 *     tap = (dz+1)·9 + (dy+1)·3 + (dx+1); (ci, co) = (8,16) (16,16) (16,32) (32,64) (64,64)
 *     (64,32) (32,16) (16,16) (16,8); input channels 3..7 and output channels 1..7 are padding.
 *   wl_net_forward / wl_net_gradient: the network on a given input [V][8] (rounded to bf16), V the
 *     grid's voxel count, index (z·ny + y)·nx + x; out / target: [V]; grad: canonical layout.
 *   wl_net_train: `steps` Adam steps, each on a pair drawn from the recorded ones; returns the mean
 *     loss Σ(e−ē)²/Σ(t−t̄)² before each update.
 *   wl_flow_set_warmstart: attach (net != NULL) or detach a network; flags WL_WARM_*.  A network
 *     must stay alive while attached; the flow must be 3D and its level-1 grid the network's grid.
 *     Either solver mode takes the guess (WL_SOLVER_MIXED adds it to its FP64 solution). */
typedef struct wl_net_s* wl_net_t;
enum wl_warm_flags {
  WL_WARM_GUESS = 1,  /* add the network's guess before the solver loop */
  WL_WARM_RECORD = 2, /* record (input, target) pairs of every projection */
  WL_WARM_STATS = 4,  /* keep Σr² before and after the guess (wl_flow_warm_stats) */
  WL_WARM_PREDICTOR = 8 /* only in the predictor's projection (w = 1, src/Flow.jl:162) */
};
int wl_net_create(wl_net_t* out, const int32_t* grid, int replay_pairs);
int wl_net_destroy(wl_net_t n);
int wl_net_num_weights(wl_net_t n, int64_t* count);
int wl_net_set_weights(wl_net_t n, const float* w, int64_t count);
int wl_net_get_weights(wl_net_t n, float* w, int64_t count);
int wl_net_forward(wl_net_t n, const float* in, float* out);
int wl_net_gradient(wl_net_t n, const float* in, const float* target, double* loss, float* grad);
int wl_net_adam(wl_net_t n, double lr); /* one Adam step with the last gradient */
int wl_net_train(wl_net_t n, int steps, double lr, uint64_t seed, double* mean_loss);
int wl_net_pairs(wl_net_t n, int* count);
int wl_net_get_pair(wl_net_t n, int index, float* in, float* target); /* 0 = oldest kept */
int wl_flow_set_warmstart(wl_flow_t h, wl_net_t net, int flags);
int wl_flow_warm_stats(wl_flow_t h, double* r2, int cap, int* count); /* (before, after) pairs */

/* Tile autotuner, no counterpart in the reference: times `reps` level-0 launches of the GS half
 * sweeps and of the restriction/prolongation kernels for each candidate blocks-in-flight target
 * (1024..16384) and keeps the fastest for this flow (targets[0]: GS, targets[1]: restriction /
 * prolongation; may be NULL).  Only kernels without reductions are tuned, so results do not
 * change; the fields are saved and restored around the timing.  A batch created afterwards uses
 * each member's targets. */
int wl_flow_autotune(wl_flow_t h, int reps, int32_t* targets);
/* Set the two targets directly (0: the default), e.g. to repeat a tuned run without timing. */
int wl_flow_set_tiles(wl_flow_t h, int32_t gs, int32_t rp);

/* Σr² residual history of the last mom_step! (both projections); returns the count written. */
int wl_flow_last_residuals(wl_flow_t h, double* r2, int cap, int* count);

/* Instrumentation, no counterpart in the reference: per-kernel HIP-event timing on the handle's
 * stream.  wl_flow_profile(h,1) starts recording (and clears), wl_flow_profile_read aggregates
 * launches, total milliseconds and the algorithmic bytes per launch by kernel name and level,
 * then clears the records. */
typedef struct {
  char name[40];
  int32_t level;
  int32_t reserved;
  int64_t launches;
  double total_ms;
  double bytes_per_launch;
} wl_kernel_stat;
int wl_flow_profile(wl_flow_t h, int enable);
int wl_flow_profile_read(wl_flow_t h, wl_kernel_stat* out, int cap, int* count);
/* Device self-tests, no counterpart in the reference.  test 0: the number of finite floats x for
 * which the kernels' division by 6 differs in any bit from the IEEE quotient (expected 0 apart
 * from the ranges it delegates to the division itself). */
int wl_self_test(int test, int64_t* result);

/* ---- Poisson / MultiLevelPoisson: src/Poisson.jl, src/MultiLevelPoisson.jl ---- */
/* Ng: ghosted sizes; L: Ng x D coefficients (host, layout as above); x, z start at zero. */
int wl_pois_create(wl_pois_t* out, int D, const int32_t* Ng, int dtype, const void* L, int multilevel,
                   int perdir_mask);
int wl_pois_destroy(wl_pois_t h);
int wl_pois_nlevels(wl_pois_t h, int* n);
int wl_pois_level_dims(wl_pois_t h, int level, int32_t* Ng);
int wl_pois_upload(wl_pois_t h, int level, int pfield, const void* src, size_t nbytes);
int wl_pois_download(wl_pois_t h, int level, int pfield, void* dst, size_t nbytes);
/* z = A x on level 1 (x given on the host; z written to the level's z and to z_out) */
int wl_pois_mult(wl_pois_t h, const void* x, void* z_out, size_t nbytes);
int wl_pois_update(wl_pois_t h);
int wl_pois_solve(wl_pois_t h, double tol, int itmx, int32_t* n_iter);

#ifdef __cplusplus
}
#endif
#endif /* WL_H */
