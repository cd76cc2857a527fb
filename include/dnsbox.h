/*
 * dnsbox.h — C-ABI of the MI355X (gfx950) pseudospectral Navier-Stokes box.
 *
 * Drop-in for the reference's hot path: the rhs / rhsAdj / stepper module
 * procedures of dnsbox (Fortran 90), which operate on module globals.  Each
 * entry point below names the reference routine it replaces (file:line in the
 * reference tree).  Plain pointers, sizes and a POD parameter struct only; no
 * torch or HIP types cross this boundary (streams are passed as void*).
 *
 * Conventions
 *   - Every function returns 0 on success and a negative DNSBOX_E* code on
 *     failure; dnsbox_last_error() gives the message (thread-local).  Nothing
 *     calls stop/abort (the reference stops on allocation failure,
 *     variables.f90:165-168).
 *   - A context owns its device memory (state, scratch, FFT tables).  Host
 *     pointers are borrowed for the duration of the call.
 *   - Work is stream-ordered on the context's stream; calls that return host
 *     data (get_*, diagnostics) synchronise that stream.
 *   - One context per host thread at a time; no MPI (one GPU holds the box).
 *   - Host spectral arrays use the reference layout: Fortran
 *     uhat(Nz, Ny, Nx/2+1) complex(8) (p3dfft STRIDE1, z fastest), one array per
 *     velocity component.  Host physical arrays: Fortran u(Nx, Ny, Nz) real(8).
 *   - Device-pointer entry points (suffix _dev) use the internal layout:
 *     (3, Nz, Ny, Nx/2+1) complex in the context precision, kx fastest; they need
 *     state_precision equal to precision.
 */
#ifndef DNSBOX_H
#define DNSBOX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DNSBOX_E_ARG     (-1)  /* invalid argument / unsupported size       */
#define DNSBOX_E_HIP     (-2)  /* HIP runtime error                         */
#define DNSBOX_E_NOMEM   (-3)  /* device allocation failed                  */
#define DNSBOX_E_STATE   (-4)  /* call not valid in the current state       */

/* diagnostics vector (io.f90:651-658 Ekin.dat columns, plus state_spectrum prints) */
enum {
    DNSBOX_D_EKIN = 0,   /* state_kinetic            state.f90:429-454           */
    DNSBOX_D_DISP,       /* state_dissipation        state.f90:517-544           */
    DNSBOX_D_PROD,       /* 2 Q Ekin                 io.f90:644,654              */
    DNSBOX_D_COURANT,    /* io_Courant               io.f90:664-676              */
    DNSBOX_D_DIVMAX,     /* state_divergence         state.f90:413-427           */
    DNSBOX_D_EZERO,      /* state_spectrum           state.f90:370-374           */
    DNSBOX_D_EONE,       /*                          state.f90:376-381           */
    DNSBOX_D_ESQRT2,     /*                          state.f90:383-388           */
    DNSBOX_D_EBAND,      /* rhs_Eband                rhs.f90:137-168             */
    DNSBOX_D_TIME,       /* time(n+1)                dnsboxheun.f90:128          */
    DNSBOX_D_DT,         /* dt                                                   */
    DNSBOX_NDIAG
};

/* Runtime replacement of the compile-time parameters.f90 constants. */
typedef struct dnsbox_params {
    double nu;          /* kinematic viscosity          parameters.f90:20 */
    double Q;           /* linear forcing multiplier    parameters.f90:21 */
    double Pin;         /* band power input             parameters.f90:23 */
    double kcut;        /* band forcing cut-off         parameters.f90:22 */
    double kolmo_amp;   /* Kolmogorov forcing f = amp sin(kolmo_k y) e_x (0 = off) */
    double deltak;      /* shell width                  parameters.f90:26 */
    double cfl_min;     /* CourantMin                   parameters.f90:33 */
    double cfl_max;     /* CourantMax                   parameters.f90:34 */
    double dt_max;      /* tStepMax                     parameters.f90:35 */
    double dt;          /* initial time step            variables.f90:9   */
    int n;              /* grid size of the axes nx, ny, nz leave at 0: 16, 32, 48, 64, 128, 256,
                           512 (Nx = Ny = Nz = n when all three are 0)  parameters.f90:12-14 */
    int precision;      /* 64: all FP64; 32: FFTs/nonlinear term FP32, IF/CFL/diagnostics FP64 */
    int kolmo_k;        /* Kolmogorov forcing wavenumber (0 = off) */
    int bandlim;        /* band-limited forcing         parameters.f90:52 */
    int spherical;      /* spherical truncation         parameters.f90:50 */
    int parity_mode;    /* 1: reference k-table (Nyquist wavenumber 0, Nyquist modes kept by
                           the mask, variables.f90:229,245,259); 0: Nyquist modes removed */
    int adaptive_dt;    /* 1: CFL controller (tStepFix = .false., parameters.f90:37) */
    int state_precision;/* 0 = precision; 64 with precision 32: FP64 state, FP32 transforms */
    double alpha_x;     /* 2 pi / Lx (0 = 1: the 2 pi box)   parameters.f90:18; held in single  */
    double alpha_y;     /* 2 pi / Ly                          parameters.f90:19; precision, as  */
    double alpha_z;     /* 2 pi / Lz                          parameters.f90:20; the reference's */
                        /* real(kind=4) parameters are (wavenumbers, 2/3 rule, Nspec, Courant dx) */
    int nx, ny, nz;     /* grid Nx, Ny, Nz (0 = n)            parameters.f90:12-14: a cubic grid of
                           the sizes above, or one of the non-cubic grids built into the library
                           (32x48x16, 48x16x32, 32x32x64, 64x64x128, 256x256x128); any other grid
                           is rejected with DNSBOX_E_ARG.  Groups and dnsbox_rhs_dev_noalloc take
                           cubic grids only.  Spectral arrays are uhat(Nz, Ny, Nx/2+1), physical
                           ones u(Nx, Ny, Nz); kolmo_k < Ny/2 */
} dnsbox_params;

typedef struct dnsbox_ctx dnsbox_ctx;

/* library */
const char* dnsbox_last_error(void);
const char* dnsbox_version(void);
int dnsbox_params_default(dnsbox_params* p);               /* parameters.f90 values, n = 32 */

/* context: var_init (variables.f90:99-280) / deallocate (dnsboxheun.f90:197-204) */
int dnsbox_create(const dnsbox_params* p, dnsbox_ctx** out);
int dnsbox_destroy(dnsbox_ctx* ctx);
int dnsbox_set_stream(dnsbox_ctx* ctx, void* hip_stream);
int dnsbox_sync(dnsbox_ctx* ctx);
int dnsbox_nspec(dnsbox_ctx* ctx);                          /* Nspec, variables.f90:112-123 */

/* state transfer (host, reference layout) */
int dnsbox_set_state(dnsbox_ctx* ctx, const double* uhat, const double* vhat, const double* what);
int dnsbox_get_state(dnsbox_ctx* ctx, double* uhat, double* vhat, double* what);
/* state_u2uhat (state.f90:75-81): physical u(Nx,Ny,Nz) -> state (unnormalised r2c) */
int dnsbox_set_state_physical(dnsbox_ctx* ctx, const double* u, const double* v, const double* w);
/* state_uhat2u (state.f90:83-96): state -> physical velocity (scaled by 1/(Nx Ny Nz)) */
int dnsbox_get_state_physical(dnsbox_ctx* ctx, double* u, double* v, double* w);

/* initial fields and field operators */
/* seeded solenoidal field, E(k) ~ k^4 exp(-2 (k/k0)^2), state_kinetic energy = energy
   (replaces state_init_rand, state.f90:211-285) */
int dnsbox_init_random(dnsbox_ctx* ctx, uint64_t seed, int box, double k0, double energy);
/* state_dealias (state.f90:181-209) and/or state_project (state.f90:120-179) */
int dnsbox_project(dnsbox_ctx* ctx, int dealias, int project);

/* rhs_int_fact (rhs.f90:8-16): sets the step size of the integrating factor; the factor
   exp((-nu|k|^2+Q) dt) is evaluated in FP64 on the fly in the stepper epilogue */
int dnsbox_int_fact(dnsbox_ctx* ctx, double dt);
/* rhs_nonlinear (rhs.f90:18-135): nonlin <- N(state) (rotational form, 2/3 mask,
   projection, band / Kolmogorov forcing) */
int dnsbox_rhs_nonlinear(dnsbox_ctx* ctx);
int dnsbox_get_nonlin(dnsbox_ctx* ctx, double* nuhat, double* nvhat, double* nwhat);
/* the time loop body of dnsboxheun.f90:81-157: stepper 1 = IF-Heun predictor-corrector
   (dnsboxheun.f90:84-126), 2 = variable-step IF-AB2 (dnsbox.f90:87-216 with consistent weights:
   u_{n+1} = E_h u_n + h (1 + h/2h') E_h N_n - (h^2/2h') E_{h+h'} N_{n-1}, h' = t_n - t_{n-1};
   the first step after a state change is IF-Heun, dnsbox.f90:87-141), 3 = IF-Heun with the
   state projected and dealiased after every step (dnsboxband.f90:91-156); each with the CFL
   controller set_time_step (dnsboxheun.f90:227-246) when adaptive_dt */
int dnsbox_step(dnsbox_ctx* ctx, int nsteps, int stepper);
/* io_save_stats (io.f90:628-661) + state_spectrum (state.f90:350-411): diag[DNSBOX_NDIAG],
   espec[Nspec] (may be NULL) */
int dnsbox_diagnostics(dnsbox_ctx* ctx, double* diag, double* espec);
int dnsbox_get_time(dnsbox_ctx* ctx, double* t, double* dt);
int dnsbox_set_time(dnsbox_ctx* ctx, double t, double dt);

/* measurement: average duration (ms, HIP events on the context stream) of each kernel pass
   of the step, in the order of DNSBOX_PASS_*: the passes of `reps` IF-Heun steps taken on a
   copy of the state (an event after every launch), and DNSBOX_PASS_ZINV (run once per pipeline
   start) over `reps` back-to-back launches.  State, time and dt of the context are unchanged.
   No reference counterpart (the reference has no timers). */
enum {
    DNSBOX_PASS_ZINV = 0,  /* z-inverse FFT + spectral curl: reads 3 fields, writes 6     */
    DNSBOX_PASS_YINV,      /* y-inverse FFT of 6 fields in place: reads 6, writes 6       */
    DNSBOX_PASS_XCROSS,    /* c2r + (omega x u) + r2c along x: reads 6, writes 3          */
    DNSBOX_PASS_YFWD,      /* y-forward FFT of 3 fields in place: reads 3, writes 3       */
    DNSBOX_PASS_ZPRED,     /* z-forward FFT + projection + IF-Heun predictor + z-inverse/curl of
                              the predicted state: reads 6, writes 9                        */
    DNSBOX_PASS_ZCORR,     /* z-forward FFT + projection + IF-Heun corrector + z-inverse/curl of
                              the new state: reads 6, writes 9                              */
    DNSBOX_NPASS
};
int dnsbox_time_passes(dnsbox_ctx* ctx, int reps, double* ms);
/* algorithmic HBM bytes of one launch of each pass (DNSBOX_PASS_* order) as the next step would
   run it: every element the pass must read or write, once.  No reference counterpart. */
int dnsbox_pass_bytes(dnsbox_ctx* ctx, double* bytes);

/* Transforms pruned to the 2/3-rule cube.  The mask of rhs_nonlinear (rhs.f90:74-88) zeroes
   every mode with |k_i| >= N/3 on some axis, so a state with no energy there keeps none through
   a step.  While that holds (after dnsbox_init_random, dnsbox_project with dealias = 1, every
   step, and dnsbox_set_state of such a field) the IF-Heun step skips the loads and stores of
   those modes; it skips only exact zeros, so the results agree with the unpruned step to
   rounding.  Taking dnsbox_state_ptr or setting a
   physical field ends it until the state is dealiased again.  set_pruning(ctx, 0) turns it off
   (default on); get_pruning reports whether the next step runs pruned.  No reference
   counterpart (p3dfft transforms whole fields). */
int dnsbox_set_pruning(dnsbox_ctx* ctx, int enable);
int dnsbox_get_pruning(dnsbox_ctx* ctx, int* active);

/* Step and RHS adjoints about states confined to the 2/3-rule cube.  set_vjp_cube(ctx, 1) is the
   caller's guarantee that the states the *_vjp_* entry points linearise about (u, and u* where
   given) have no energy outside the kept cube; their passes up to the x pass then read the cube
   only (the cotangent enters masked), and the results agree with the full passes to rounding.
   The guarantee is checked on the device: the first pass of every such VJP also scans its
   state outside the cube (on blocks the pruned pass leaves idle) and raises a flag, and
   set_vjp_cube(ctx, 0) reads the flag back (synchronising the context's stream) and returns
   DNSBOX_E_STATE if a state outside the cube was differentiated since the guarantee was set:
   the adjoints computed in between are then wrong.  States made by pruned steps, which
   get_last_step_pruned reports for the last dnsbox_step_dev(_record) call, satisfy it.
   Default off.  No reference counterpart. */
int dnsbox_set_vjp_cube(dnsbox_ctx* ctx, int confined);
int dnsbox_get_last_step_pruned(dnsbox_ctx* ctx, int* pruned);

/* device-pointer entry points (internal layout, context precision) */
int dnsbox_state_ptr(dnsbox_ctx* ctx, void** dev_state);   /* the context's state buffer (ends
                                                              pruning until the next dealias) */
/* rhs_nonlinear on device buffers: out <- N(in) */
int dnsbox_rhs_dev(dnsbox_ctx* ctx, const void* in, void* out);
/* the time loop of dnsboxheun.f90:81-126 on a caller-owned device state (in place); the passes
   run pruned when a device check finds the buffer empty outside the 2/3-rule cube */
int dnsbox_step_dev(dnsbox_ctx* ctx, void* state, int nsteps);
/* dnsbox_step_dev that also stores the state after step j + 1 at traj + j * (3 fields of the
   internal layout), j < nsteps, and (traj_pred not NULL) stores that step's predicted state
   u* = E (u + dt N(u)) at traj_pred + j * 3 fields: the recomputed trajectory of an adjoint
   checkpoint segment in one call (the step pipeline runs through it).  No reference counterpart. */
int dnsbox_step_dev_record(dnsbox_ctx* ctx, void* state, int nsteps, void* traj, void* traj_pred);
/* exact discrete adjoint of rhs_nonlinear (rhs.f90:18-135): out <- J_N(u)^T lam in the
   Euclidean inner product of the stored coordinates, sum Re(conj(a) b).  No reference
   counterpart: the reference's adjoint (rhsAdj.f90) is the continuous one below. */
int dnsbox_rhs_vjp_dev(dnsbox_ctx* ctx, const void* u, const void* lam, void* out);
/* adjoint of one fixed-dt IF-Heun step (dnsboxheun.f90:84-126) taken from state u:
   lam_out <- (d u_{n+1} / d u_n)^T lam.  Backward step of the T-step map Phi_T. */
int dnsbox_step_vjp_dev(dnsbox_ctx* ctx, const void* u, const void* lam, void* lam_out, double dt);
/* dnsbox_step_vjp_dev with the step's predicted state ustar = E (u + dt N(u)) supplied (as
   recorded by dnsbox_step_dev_record), so it is not recomputed.  No reference counterpart. */
int dnsbox_step_vjp_pred_dev(dnsbox_ctx* ctx, const void* u, const void* ustar, const void* lam, void* lam_out,
                             double dt);
/* dnsbox_step_dev_record (traj and traj_pred both required) that also records, for step j + 1,
   the x-pass inputs of its two right-hand sides, N(u_j) and N(u*_j): u, v, w and the vorticity
   after the z- and y-inverse transforms, six fields of the internal layout each, in fields 3..8
   of the 9-field slots mixed + (2 j) * 9 fields and mixed + (2 j + 1) * 9 fields (mixed holds
   18 nsteps fields).  Needs state_precision = precision.  No reference counterpart: the forward
   half of the backward step of the T-step map (dnsboxheun.f90:84-126). */
int dnsbox_step_dev_record_mixed(dnsbox_ctx* ctx, void* state, int nsteps, void* traj, void* traj_pred,
                                 void* mixed);
/* dnsbox_step_vjp_pred_dev that reads the x-pass inputs at u and at ustar from the record slots
   mixed_u and mixed_ustar that dnsbox_step_dev_record_mixed wrote for that step, instead of
   transforming u and ustar again (fields 0..2 of both slots are overwritten).  pruned_record:
   dnsbox_get_last_step_pruned after the recording; a pruned recording holds the kept kx only and
   needs dnsbox_set_vjp_cube(ctx, 1), else DNSBOX_E_ARG.  No reference counterpart. */
int dnsbox_step_vjp_mixed_dev(dnsbox_ctx* ctx, const void* u, const void* ustar, const void* lam, void* lam_out,
                              double dt, void* mixed_u, void* mixed_ustar, int pruned_record);
/* dnsbox_step_vjp_mixed_dev over the recorded steps of a segment, last step first: for
   j = nsteps-1 .. 0, lam <- (d u_{j+1} / d u_j)^T lam at u[j] with predicted state ustar[j] and the
   record slots mixed[2 j] (the RHS at u[j]) and mixed[2 j + 1] (at ustar[j]).  The result is left in
   lam_out; lam is overwritten (work space).  With pruned passes, each step's last pass also runs
   the first pass of the step before (one launch fewer per step).  No reference counterpart. */
int dnsbox_step_vjp_mixed_seq_dev(dnsbox_ctx* ctx, const void* const* u, const void* const* ustar, void* lam,
                                  void* lam_out, double dt, void* const* mixed, int nsteps, int pruned_record);
/* rhsAdj (rhsAdj.f90:356-381 via rhsNonlinearAdj :261-354, rhsupptilde :177-213, convective
   rhsNonlinear :64-175), with the never-assigned phatadj (:343-348) read as phat (:335-340).
   lin_sign = +1 reproduces the reference's L u~ - P n_adj; lin_sign = -1 gives the descent
   direction -DF^+(u~) = -L u~ - P n_adj of J = 1/2 ||F(u)||^2_{H^-1}. */
int dnsbox_rhs_adj_dev(dnsbox_ctx* ctx, const void* u, void* out, int lin_sign);
/* the adjoint-descent loop body of dnsboxAdj.f90:76-126, nsteps times, in place on a device state:
   G = rhsAdj(u) (:79), u* = u + dtau G (:83-101), u <- u + (dtau G)/2 + (dtau rhsAdj(u*))/2
   (:87-113), state_project (:117).  lin_sign as in dnsbox_rhs_adj_dev (+1: the reference's rhsAdj;
   -1: the descent direction). */
int dnsbox_adj_descent_step_dev(dnsbox_ctx* ctx, void* u, double dtau, int nsteps, int lin_sign);
/* the same on the context's state; its clock advances by dtau per step (time(n+1) = time(n) + dt,
   dnsboxAdj.f90:125) */
int dnsbox_adj_descent_step(dnsbox_ctx* ctx, double dtau, int nsteps, int lin_sign);

/* rhs_nonlinear on device buffers (dnsbox_rhs_dev's computation, out <- N(in), internal layout,
   precision = state_precision) with no context and no device allocation.  No counterpart in the
   reference: it is for callers that cannot allocate device memory (code under HIP graph capture,
   serving loops, harnesses that forbid hipMalloc).  n must be 16, 32, 48 or 64: its transform
   scratch, twiddles and band-energy scalar are one static device array of 12.98 MB per GPU
   (6 FP64 fields of 64^3), kept until the process ends and shared by every call on that GPU,
   so calls that overlap on one GPU race.  Larger n fail with DNSBOX_E_ARG.  Runs on hip_stream
   (NULL: the null stream); the caller synchronises. */
int dnsbox_rhs_dev_noalloc(const dnsbox_params* p, const void* in, void* out, void* hip_stream);

/* Groups: many independent boxes (N in {16, 32, 48, 64, 128}, one precision, each with its
   own parameters) advanced together, one launch per pass for all of them, the step captured
   in a hipGraph.  Batched counterpart of dnsbox_step for adjoint-descent guesses.  No
   reference counterpart (the reference runs one box per MPI job). */
typedef struct dnsbox_group dnsbox_group;
int dnsbox_group_create(const dnsbox_params* params, int nbox, dnsbox_group** out);
int dnsbox_group_destroy(dnsbox_group* g);
int dnsbox_group_size(dnsbox_group* g);
/* borrowed context of box i (set/get state, init, diagnostics, ...; not destroyable) */
int dnsbox_group_box(dnsbox_group* g, int i, dnsbox_ctx** box);
/* nsteps IF-Heun steps of every box of the group (dnsboxheun.f90:84-157 per box) */
int dnsbox_group_step(dnsbox_group* g, int nsteps);
int dnsbox_group_sync(dnsbox_group* g);
/* stream (hipStream_t) the device-buffer group entry points run on, and that
   dnsbox_group_step is ordered with; NULL = the null stream */
int dnsbox_group_set_stream(dnsbox_group* g, void* stream);
/* batched dnsbox_step_dev: states[i] is a device state of box i (internal layout, group
   precision), advanced in place by nsteps IF-Heun steps at box i's step size */
int dnsbox_group_step_dev(dnsbox_group* g, void* const* states, int nsteps);
/* batched dnsbox_step_vjp_dev: lam_out[i] <- (d u_{n+1} / d u_n)^T lam[i] for one IF-Heun
   step of box i from u[i] at step size dt[i] */
int dnsbox_group_step_vjp_dev(dnsbox_group* g, const void* const* u, const void* const* lam,
                              void* const* lam_out, const double* dt);
/* batched dnsbox_step_dev_record: box i's state after step j + 1 is copied to traj[i] + j * 3
   fields and, when traj_pred is not NULL, that step's predicted state u* is stored at
   traj_pred[i] + j * 3 fields (device buffers of the group precision, internal layout).  No
   reference counterpart (it serves the checkpointed adjoint of the T-step map). */
int dnsbox_group_step_dev_record(dnsbox_group* g, void* const* states, int nsteps, void* const* traj,
                                 void* const* traj_pred);
/* batched dnsbox_step_vjp_pred_dev: dnsbox_group_step_vjp_dev with each box's predicted state
   u* = E (u + dt N(u)) given (as recorded by dnsbox_group_step_dev_record).  No reference
   counterpart. */
int dnsbox_group_step_vjp_pred_dev(dnsbox_group* g, const void* const* u, const void* const* ustar,
                                   const void* const* lam, void* const* lam_out, const double* dt);
/* batched dnsbox_step_dev_record_mixed: box i's step j + 1 records the x-pass inputs of its two
   right-hand sides in fields 3..8 of the two 9-field slots at mixed[i] + 18 j fields.  No
   reference counterpart. */
int dnsbox_group_step_dev_record_mixed(dnsbox_group* g, void* const* states, int nsteps, void* const* traj,
                                       void* const* traj_pred, void* const* mixed);
/* batched dnsbox_step_vjp_mixed_dev: mixed[i] is box i's slot pair of the step (18 fields, the RHS
   at u then at u*), pruned_record as dnsbox_group_get_last_step_pruned after the recording (a
   pruned recording needs dnsbox_group_set_vjp_cube).  No reference counterpart. */
int dnsbox_group_step_vjp_mixed_dev(dnsbox_group* g, const void* const* u, const void* const* ustar,
                                    const void* const* lam, void* const* lam_out, const double* dt,
                                    void* const* mixed, int pruned_record);
/* the group forms of dnsbox_set_vjp_cube / dnsbox_get_last_step_pruned (every box's states;
   the last dnsbox_group_step_dev(_record) call).  No reference counterpart. */
int dnsbox_group_set_vjp_cube(dnsbox_group* g, int confined);
int dnsbox_group_get_last_step_pruned(dnsbox_group* g, int* pruned);

#ifdef __cplusplus
}
#endif
#endif /* DNSBOX_H */
