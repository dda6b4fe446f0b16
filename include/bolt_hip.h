/*
 * bolt_hip.h -- C ABI of libbolt_hip.so, the MI355X (gfx950) implementation of
 * Bolt.jl's perturbation hot path:
 *
 *   hierarchy! RHS + boltsolve (KenCarp4 ESDIRK)  ->  source_grid  ->  cltt / plin
 *
 * Every entry point is extern "C", takes plain pointers and sizes, returns an
 * int status (0 = OK, negative = error class, see BOLT_E_*) and leaves a
 * human readable message retrievable with bolt_last_error() (thread local).
 * Functions without the _dev suffix take host arrays (what a Julia ccall with
 * GC.@preserve hands over).  The _dev variants take device pointers and leave
 * their results in device memory.  All entry points return after their
 * kernels have finished; a context serialises its callers.
 *
 * Citations are into the Bolt.jl source tree (src/...).
 */
#ifndef BOLT_HIP_H
#define BOLT_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BOLT_ABI_VERSION 1

/* status codes */
#define BOLT_OK 0
#define BOLT_E_ARG (-1)      /* invalid argument (sizes, NULL pointers, config) */
#define BOLT_E_DEVICE (-2)   /* HIP runtime / device error, or no GPU present */
#define BOLT_E_NOMEM (-3)    /* device allocation failed */
#define BOLT_E_NUMERIC (-4)  /* a mode failed (see per-mode status) */
#define BOLT_E_STATE (-5)    /* unknown cosmology id / context misuse */

/* per-mode status written by the solvers */
#define BOLT_MODE_OK 0
#define BOLT_MODE_MAXSTEPS 1 /* step budget exhausted before x = 0 */
#define BOLT_MODE_DTMIN 2    /* step size underflow */
#define BOLT_MODE_NONFINITE 3/* NaN/Inf in the state or error norm */

/* CosmoParams in Bolt's field order (src/Bolt.jl:56-66).  Sum_m_nu is in
 * natural units (Mpc^-1), as Bolt stores it (0.06*mass_natural by default). */
typedef struct bolt_cosmo_params {
    double h, Omega_r, Omega_b, Omega_c, A, n, Y_p, N_nu, Sum_m_nu;
} bolt_cosmo_params;

/* Spline tables.  Each table is a cubic B-spline with natural ("Line(OnGrid())")
 * end conditions on the uniform grid x_j = x0 + j*dx, j = 0..nx-1, stored as its
 * nx+2 padded coefficients (src/util.jl:11-13).  Table order: */
enum {
    BOLT_T_HC = 0,   /* bg.ℋ      (src/background.jl:121) */
    BOLT_T_HC_P,     /* bg.ℋ′     (spline_∂ₓ(ℋ_))  */
    BOLT_T_HC_PP,    /* bg.ℋ′′    (spline_∂ₓ²(ℋ_)) */
    BOLT_T_ETA,      /* bg.η */
    BOLT_T_ETA_P,    /* bg.η′ */
    BOLT_T_ETA_PP,   /* bg.η′′ */
    BOLT_T_RHO0M,    /* bg.ρ₀ℳ (massive-ν background density) */
    BOLT_T_XE,       /* ih.Xₑ      (src/ionization/recfast.jl:705) */
    BOLT_T_TAU,      /* ih.τ */
    BOLT_T_TAU_P,    /* ih.τ′ */
    BOLT_T_TAU_PP,   /* ih.τ′′ */
    BOLT_T_G,        /* ih.g̃ */
    BOLT_T_G_P,      /* ih.g̃′ */
    BOLT_T_G_PP,     /* ih.g̃′′ */
    BOLT_T_TMAT,     /* ih.Tmat */
    BOLT_T_CSB2,     /* ih.csb² */
    BOLT_NTAB
};

/* scalar entries of the background / ionization history */
enum { BOLT_S_H0 = 0, BOLT_S_ETA0, BOLT_S_RHO_CRIT, BOLT_S_OMEGA_L, BOLT_S_TAU0, BOLT_NSCAL };

typedef struct bolt_tables {
    int nx;             /* number of x nodes (2001 for x_grid = -20:0.01:0) */
    int nq;             /* Gauss-Legendre order of the q quadrature */
    double x0, x1;      /* first and last node; x_j = (x0*(nx-1-j) + x1*j)/(nx-1) */
    double scal[BOLT_NSCAL];
    const double* xq;   /* [nq] Gauss-Legendre nodes on [-1,1], ascending */
    const double* wq;   /* [nq] weights */
    const double* coef; /* [BOLT_NTAB][nx+2] */
} bolt_tables;

/* Hierarchy sizes and solver settings (src/perturbations.jl:7-21, 25-33).
 * n = 2(ell_g+1) + (ell_nu+1) + (ell_mnu+1)*nq + 5. */
typedef struct bolt_hier_cfg {
    int ell_g, ell_nu, ell_mnu, nq;
    double reltol, abstol;
    int max_steps;      /* per mode; 0 -> 100,000 */
    int flags;          /* 0 (FP64 throughout), or BOLT_CFG_FP32_SOLVE [| refinement sweeps << 4] */
} bolt_hier_cfg;

/* cfg.flags.  BOLT_CFG_FP32_SOLVE: the stage equations (I − hγJ)z = r of the
 * KenCarp4 Newton solve (perturbations.jl:25-33) are factored and substituted
 * in FP32, and each stage value is refined against its FP64 residual
 * r − z + hγJz until every correction is below 1e-14 of its component, at
 * most (flags & BOLT_CFG_REFINE_MASK) >> 4 sweeps (0 selects 12); a stage
 * still unconverged after the last sweep is taken as it stands.  The state,
 * stage combinations, error estimate, step control, residuals and spline
 * lookups stay FP64.  All modes of one launch use the same setting.  The
 * step sequence then differs from the FP64 solve's, so outputs differ from
 * it by up to the integration's own error at the given tolerances. */
#define BOLT_CFG_FP32_SOLVE 1
#define BOLT_CFG_REFINE_MASK 0xF0

typedef struct bolt_stats {
    int64_t steps_accepted, steps_rejected, rhs_evals, linear_solves;
    double kernel_ms;   /* device time of the solver launch (hipEvent) */
} bolt_stats;

/* ---------------------------------------------------------------- host --- */
int bolt_abi_version(void);
const char* bolt_last_error(void);
int bolt_state_length(const bolt_hier_cfg* cfg);

/* Background(par; x_grid, nq) + IonizationHistory(RECFAST(bg), par, bg) on the
 * host, FP64 (src/background.jl:104-128, src/ionization/recfast.jl:674-726).
 * coef_out: [BOLT_NTAB][nx+2]; scal_out: [BOLT_NSCAL]; xq_out, wq_out: [nq]. */
int bolt_host_tables(const bolt_cosmo_params* par, double x0, double x1, int nx, int nq,
                     double* coef_out, double* scal_out, double* xq_out, double* wq_out);

/* bolt_host_tables for ncosmo cosmologies on nthreads host threads (<= 0: one
 * per hardware thread), the (cosmology, k) batch's table stage.  Bolt builds
 * Background + IonizationHistory per cosmology on one thread.  Cosmology c's
 * tables go to coef_out + c*BOLT_NTAB*(nx+2) and scal_out + c*BOLT_NSCAL, bitwise
 * what bolt_host_tables writes for it; xq_out, wq_out: [nq], shared.  status[c]
 * is that call's return code; the function returns BOLT_E_NUMERIC (message of
 * the first failure) if any cosmology failed, after filling all the others. */
int bolt_host_tables_batch(const bolt_cosmo_params* pars, int ncosmo, double x0, double x1, int nx, int nq,
                           int nthreads, double* coef_out, double* scal_out, double* xq_out, double* wq_out,
                           int* status);

/* Forward-mode derivative of bolt_host_tables along the parameter direction
 * dpar: with p(t) = par + t*dpar, dcoef_out = d coef/dt and dscal_out =
 * d scal/dt at t = 0 (same layouts as coef_out / scal_out).  This is what
 * ForwardDiff.Dual parameters carry through Background and
 * IonizationHistory(RECFAST) (src/background.jl, src/ionization/recfast.jl;
 * exercised by test/runtests.jl:54-80).  coef_out, scal_out, xq_out and wq_out
 * receive exactly what bolt_host_tables writes.  Branch decisions follow the
 * values (as under ForwardDiff); the RECFAST switch redshifts carry the
 * implicit-function tangent of their root conditions; the RECFAST ODE takes
 * the values' steps. */
int bolt_host_tables_jvp(const bolt_cosmo_params* par, const bolt_cosmo_params* dpar, double x0, double x1,
                         int nx, int nq, double* coef_out, double* scal_out, double* dcoef_out,
                         double* dscal_out, double* xq_out, double* wq_out);

/* The derivative of bolt_host_tables with respect to every bolt_cosmo_params
 * field at once: ForwardDiff's chunked Duals, nine partials in one pass.
 * dcoef_out: [9][BOLT_NTAB][nx+2] and dscal_out: [9][BOLT_NSCAL], with partial p
 * in struct order (h, Omega_r, Omega_b, Omega_c, A, n, Y_p, N_nu, Sum_m_nu).
 * Partial p is bitwise bolt_host_tables_jvp with dpar the p-th unit vector;
 * coef_out, scal_out, xq_out and wq_out receive what bolt_host_tables writes. */
int bolt_host_tables_jacobian(const bolt_cosmo_params* par, double x0, double x1, int nx, int nq,
                              double* coef_out, double* scal_out, double* dcoef_out, double* dscal_out,
                              double* xq_out, double* wq_out);

/* RECFAST X_e(z) before reionization, Xe_RECFAST(recfastsolve(RECFAST(bg; Tnow)))
 * (src/ionization/recfast.jl:413-475).  Tnow <= 0 takes the default derived from
 * Omega_r (recfast.jl:108). */
int bolt_host_recfast_xe(const bolt_cosmo_params* par, double Tnow, const double* z, int nz,
                         double* xe_out);

/* Evaluate table `tab` (or its first/second x-derivative, deriv = 0,1,2) at xs. */
int bolt_host_spline_eval(const bolt_tables* t, int tab, int deriv, const double* xs, int m,
                          double* out);

/* ---------------------------------------------------------------- device - */
typedef struct bolt_ctx bolt_ctx;

int bolt_ctx_create(int device, bolt_ctx** out);
int bolt_ctx_destroy(bolt_ctx* ctx);
int bolt_ctx_sync(bolt_ctx* ctx);

/* bolt_host_tables_batch computed on the GPU of ctx: Background and
 * IonizationHistory(RECFAST) (src/background.jl:85-128, src/ionization/
 * recfast.jl:446-536, 674-726) with one workgroup per cosmology, the host's
 * algorithm and arithmetic; RECFAST's adaptive steps follow the device's math
 * library, so the tables agree with the host's to the ODE tolerance, not
 * bitwise.  Same layouts as bolt_host_tables_batch; status[c] is 0 or
 * BOLT_E_NUMERIC (RECFAST out of steps, or a non-finite entry). */
int bolt_device_tables_batch(bolt_ctx* ctx, const bolt_cosmo_params* pars, int ncosmo, double x0, double x1, int nx,
                             int nq, double* coef_out, double* scal_out, double* xq_out, double* wq_out, int* status);

/* Layout introspection (no reference counterpart; host only, no GPU needed):
 * the solver's per-k-mode LDS workspace for state length n and nq momentum
 * bins (fp32 != 0: as the FP32 solve mode uses it), as `*nregions` [begin, end) byte ranges written to regions[2*i..] (at
 * most cap of them), the byte offset of the slot that absorbs stores of rows
 * past n, the bytes a launch reserves per workgroup, and the farthest byte
 * the solver's branch-free reads past n may touch.  For tests of the layout. */
int bolt_ws_layout(int n, int nq, int fp32, int cap, long long* regions, int* nregions, long long* dump_off,
                   long long* total_bytes, long long* read_reach);

/* Upload one cosmology's parameters and tables; returns its id. */
int bolt_cosmo_upload(bolt_ctx* ctx, const bolt_cosmo_params* par, const bolt_tables* t,
                      int* cosmo_id);
int bolt_cosmo_release(bolt_ctx* ctx, int cosmo_id);

/* ---------------------------------------------- forward-mode derivatives ---
 * Bolt's headline feature: ForwardDiff.Dual parameters through the whole
 * pipeline (README.md; source_grid, cltt and plin are generic over the
 * parameter type: spectra.jl:6-23, 84-96, 163-198; runtests.jl:54-80;
 * examples/plot_deriv_cl.jl:100-125).  Here the derivative along one
 * parameter direction dpar is carried beside the solution on the device.
 *
 * bolt_cosmo_upload_jvp: bolt_cosmo_upload plus the cosmology's tangent along
 * dpar.  dcoef [BOLT_NTAB][nx+2] and dscal [BOLT_NSCAL] are the derivatives of
 * t->coef and t->scal (what bolt_host_tables_jvp writes); the q grid's and
 * every derived constant's derivatives are formed from dpar and dscal.  The id
 * serves every other entry point as well. */
int bolt_cosmo_upload_jvp(bolt_ctx* ctx, const bolt_cosmo_params* par, const bolt_cosmo_params* dpar,
                          const bolt_tables* t, const double* dcoef, const double* dscal, int* cosmo_id);

/* The derivative of hierarchy! at (u, u̇): du̇ = J u̇ + (∂f/∂p)·dpar u, for B modes at x
 * (perturbations.jl:161-271 under Duals).  u, ud, dud: [n][B], Bolt's order.
 * Checks the tangent of every row recipe, table and q weight. */
int bolt_rhs_jvp_batch(bolt_ctx* ctx, int cosmo_id, const bolt_hier_cfg* cfg, const double* k, int B, double x,
                       const double* u, const double* ud, double* dud);
/* initial_conditions (perturbations.jl:274-338) and their derivative; [n][B] */
int bolt_initial_conditions_jvp(bolt_ctx* ctx, int cosmo_id, const bolt_hier_cfg* cfg, const double* k, int B,
                                double* u0, double* ud0);

/* Tangent-linear boltsolve / source_grid: the primal solve and, beside it,
 * u̇ along the same accepted steps (each stage's tangent equation
 * (I − hγJ)Ż = h(J ȧrg + J̇(arg + γZ)) solved with the primal's factors), so the
 * outputs' derivatives are those of the discrete solution the primal
 * computes.  Step control looks at the primal only (ForwardDiff under
 * OrdinaryDiffEq may also weigh the partials in its error norm; both agree
 * with the exact derivative to the integration's tolerance).  Temperature
 * source only; FP64; hierarchies whose workspace plus tangent fit one
 * workgroup's LDS (n = 323 at ℓᵧ = ℓ_ν = 50 does).  S, dS: [nk][nx]. */
int bolt_source_grid_jvp(bolt_ctx* ctx, int cosmo_id, const bolt_hier_cfg* cfg, const double* k, int nk,
                         double* S, double* dS, int* status, bolt_stats* stats);
int bolt_source_grid_jvp_dev(bolt_ctx* ctx, int cosmo_id, const bolt_hier_cfg* cfg, const double* d_k, int nk,
                             double* d_S, double* d_dS, int* d_status, int64_t* d_counters, float* ms);
/* cltt (spectra.jl:84-96) and its derivative along the cosmology's tangent
 * direction, from S and dS = dS/dp (bolt_source_grid_jvp): the derivative of
 * Σ_k T_ℓ² P(k) Δk/k with T_ℓ = Σ_x j_ℓ(k(η₀ − η(x))) S Δx, through S, the
 * Bessel argument (j′_ℓ k(dη₀ − dη)) and P(k) = A (k/0.05)^(n−1); the k and x
 * grids carry no derivative (assume_nondual, spectra.jl:46-47).  Exact j_ℓ
 * (bessel_mode 0). */
int bolt_cltt_jvp(bolt_ctx* ctx, int cosmo_id, const double* S, const double* dS, const double* k_src, int nk_src,
                  const double* kgrid, int nkd, const int* ells, int nell, double* cl, double* dcl);
int bolt_cltt_jvp_dev(bolt_ctx* ctx, int cosmo_id, const double* d_S, const double* d_dS, const double* d_k_src,
                      int nk_src, const double* d_kgrid, int nkd, const int* ells, int nell, double* d_cl,
                      double* d_dcl, float* ms);
/* states and their derivatives at sorted x_save; u_out, ud_out [B][nsave][n] */
int bolt_boltsolve_jvp_batch(bolt_ctx* ctx, int cosmo_id, const bolt_hier_cfg* cfg, const double* k, int B,
                             const double* x_save, int nsave, double* u_out, double* ud_out, int* status,
                             bolt_stats* stats);

/* hierarchy!(du, u, h, x) for B modes at one x (src/perturbations.jl:161-271).
 * u, du: [n][B] (mode fastest).  If mutate != 0 the RSA branch's in-place
 * update of u (perturbations.jl:217-227) is written back into u. */
int bolt_rhs_batch(bolt_ctx* ctx, int cosmo_id, const bolt_hier_cfg* cfg, const double* k, int B,
                   double x, double* u, double* du, int mutate);

/* initial_conditions(x_i, h) for B modes (perturbations.jl:274-338); u0: [n][B]. */
int bolt_initial_conditions(bolt_ctx* ctx, int cosmo_id, const bolt_hier_cfg* cfg,
                            const double* k, int B, double* u0);

/* boltsolve(h; reltol, abstol) evaluated at x_save (perturbations.jl:25-33).
 * u_out: [B][nsave][n]; status: [B]. */
int bolt_boltsolve_batch(bolt_ctx* ctx, int cosmo_id, const bolt_hier_cfg* cfg, const double* k,
                         int B, const double* x_save, int nsave, double* u_out, int* status,
                         bolt_stats* stats);

/* boltsolve_rsa (perturbations.jl:86-111): the solution on every x node with
 * the RSA overwrite after the switch.  out: [B][nx][n]. */
int bolt_boltsolve_rsa_batch(bolt_ctx* ctx, int cosmo_id, const bolt_hier_cfg* cfg,
                             const double* k, int B, double* out, int* status, bolt_stats* stats);

/* source_grid(par, bg, ih, k_grid, BasicNewtonian(); ℓᵧ, reltol) (spectra.jl:6-23).
 * S: [nk][nx], i.e. Julia's grid[nx, nk] column-major. */
int bolt_source_grid(bolt_ctx* ctx, int cosmo_id, const bolt_hier_cfg* cfg, const double* k,
                     int nk, double* S, int* status, bolt_stats* stats);
/* Device-resident source_grid: d_k [nk], d_S [nk][nx], d_status [nk] and
 * d_counters [nk][4] (accepted, rejected, RHS evaluations, linear solves; may be
 * NULL) are device pointers; *ms (may be NULL) receives the solver kernel's
 * device time.  Returns when the kernel has finished. */
int bolt_source_grid_dev(bolt_ctx* ctx, int cosmo_id, const bolt_hier_cfg* cfg,
                         const double* d_k, int nk, double* d_S, int* d_status,
                         int64_t* d_counters, float* ms);

/* source_grid_P(par, bg, ih, k_grid, BasicNewtonian(); ℓᵧ, reltol)
 * (spectra.jl:25-42) with source_function_P (perturbations.jl:385-404).
 * S_P: [nk][nx]; its last-node value is ±Inf (y = 0 there), as in Bolt. */
int bolt_source_grid_P(bolt_ctx* ctx, int cosmo_id, const bolt_hier_cfg* cfg, const double* k,
                       int nk, double* S_P, int* status, bolt_stats* stats);
/* Both grids from ONE solve per k-mode: replaces the pair source_grid +
 * source_grid_P that cltt/clte/clee callers run back to back (runtests.jl:156-157).
 * Either S_T or S_P may be NULL. */
int bolt_source_grid_tp(bolt_ctx* ctx, int cosmo_id, const bolt_hier_cfg* cfg, const double* k,
                        int nk, double* S_T, double* S_P, int* status, bolt_stats* stats);
/* Device-resident form of bolt_source_grid_tp (pointers as in bolt_source_grid_dev). */
int bolt_source_grid_tp_dev(bolt_ctx* ctx, int cosmo_id, const bolt_hier_cfg* cfg,
                            const double* d_k, int nk, double* d_ST, double* d_SP,
                            int* d_status, int64_t* d_counters, float* ms);

/* cltt(ℓ, sf, kgrid, par, bg) for a list of ℓ (spectra.jl:84-96, 132-150).
 * S: [nk_src][nx] source grid on k_src; kgrid: the dense k grid (nkd nodes, the
 * sum runs over its nkd-1 midpoints); bessel_mode 0 = j_ℓ by recurrence. */
int bolt_cltt(bolt_ctx* ctx, int cosmo_id, const double* S, const double* k_src, int nk_src,
              const double* kgrid, int nkd, const int* ells, int nell, int bessel_mode,
              double* cl);
/* Device-resident cltt: S, k_src, kgrid and cl on the device, the ℓ list on the
 * host.  d_steps (may be NULL) receives the executed recurrence steps per
 * dense-k midpoint; *ms the LOS kernel's device time. */
int bolt_cltt_dev(bolt_ctx* ctx, int cosmo_id, const double* d_S, const double* d_k_src, int nk_src,
                  const double* d_kgrid, int nkd, const int* ells, int nell, double* d_cl,
                  int64_t* d_steps, float* ms);

/* bolt_cltt_dev that allocates nothing (no reference counterpart; for callers
 * that cannot allocate: code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments and results, bit for bit,
 * as bolt_cltt_dev; the LOS work space is a library-owned device array of
 * 12.3 MB per GPU (held until the process ends), so ℓ <= 3000 and at most 3001
 * entries in ells; larger requests return BOLT_E_ARG.  Calls that overlap on
 * one GPU (from different contexts) share that array and race. */
int bolt_cltt_dev_noalloc(bolt_ctx* ctx, int cosmo_id, const double* d_S, const double* d_k_src, int nk_src,
                          const double* d_kgrid, int nkd, const int* ells, int nell, double* d_cl,
                          int64_t* d_steps, float* ms);

/* Autotuner (no reference counterpart: Bolt's only tunable is the @qthreads
 * loop over k, spectra.jl:10).  Times every shape of the solver -- rows per
 * lane of its substitution scans, SR = 6..8 -- on source_grid over the host
 * array k[nk] (increasing) with hierarchy cfg, and, when kgrid is not NULL,
 * every shape of the LOS kernel -- x slots per wavefront, 1, 2, 4 -- on cltt
 * for ℓ = 2..lmax over kgrid[nkd] and the source grid just computed.  The
 * fastest shapes are kept for the context's device, that state length and
 * that (nkd, lmax), and used by later calls of every context on the device;
 * the timings come back in *rep.
 * A given shape gives bitwise reproducible results; two shapes agree to
 * roundoff (their sums run in different orders). */
typedef struct bolt_tune_report {
    int solver_sr;      /* chosen rows per lane (6..8) */
    float solver_ms[4]; /* one source_grid launch per SR = 6, 7, 8 in entries 1..3; entry 0 (SR = 5, a shape
                           that is not launched) stays 0 */
    int los_spw;        /* chosen x slots per wavefront (1, 2, 4); 0 when kgrid is NULL */
    float los_ms[3];    /* one LOS launch per SPW = 1, 2, 4 */
} bolt_tune_report;
int bolt_autotune(bolt_ctx* ctx, int cosmo_id, const bolt_hier_cfg* cfg, const double* k, int nk,
                  const double* kgrid, int nkd, int lmax, bolt_tune_report* rep);

/* A whole spectrum as one HIP graph (no reference counterpart: Bolt calls
 * source_grid then cltt, spectra.jl:6-23, 84-96).  Captures, on the context's
 * stream: the solver's work-counter reset (a kernel, not a memset), the
 * persistent source_grid launch over d_k[nk] with hierarchy cfg into
 * d_S[nk][nx] (status d_status[nk], counters d_counters[nk][4] or NULL), and the
 * LOS for C_ℓ^TT at ells[nell] over d_kgrid[nkd] into d_cl[nell], with the
 * solver and LOS shapes in force at creation.  The graph owns its descriptors
 * and LOS work space; the caller's device buffers must outlive it.
 * bolt_graph_launch replays it and waits; each replay recomputes the spectrum
 * bit for bit as source_grid_dev + cltt_dev do. */
typedef struct bolt_graph bolt_graph;
int bolt_graph_spectrum_create(bolt_ctx* ctx, int cosmo_id, const bolt_hier_cfg* cfg, const double* d_k, int nk,
                               const double* d_kgrid, int nkd, const int* ells, int nell, double* d_S,
                               int* d_status, int64_t* d_counters, double* d_cl, bolt_graph** out);
int bolt_graph_launch(bolt_graph* g);
int bolt_graph_destroy(bolt_graph* g);

/* source_grid for many cosmologies at once: the @qthreads loop over k_grid
 * (spectra.jl:10) widened to every (cosmology, k) pair of the batch, run by one
 * persistent launch with per-mode adaptive steps (largest k first).
 * cosmo_ids[c] and cfgs[c] give cosmology c and its hierarchy (each cfg's nq
 * must match its tables; all x grids must have the same length nx); k is
 * [ncosmo][nk]; S_T / S_P (either may be NULL) are [ncosmo][nk][nx]; status is
 * [ncosmo][nk].  Each mode's result is the same as in a single-cosmology call. */
int bolt_source_grid_batch(bolt_ctx* ctx, int ncosmo, const int* cosmo_ids,
                           const bolt_hier_cfg* cfgs, const double* k, int nk, double* S_T,
                           double* S_P, int* status, bolt_stats* stats);
/* Device-resident form (cosmo_ids and cfgs on the host, the rest on the device). */
int bolt_source_grid_batch_dev(bolt_ctx* ctx, int ncosmo, const int* cosmo_ids,
                               const bolt_hier_cfg* cfgs, const double* d_k, int nk, double* d_ST,
                               double* d_SP, int* d_status, int64_t* d_counters, float* ms);

/* cltt, clte and clee for a list of ℓ (spectra.jl:84-130, 132-160) from one
 * LOS pass: T_ℓ of the temperature source S_T and the polarization source S_P
 * share each j_ℓ; E terms carry the spin factor √((ℓ+2)(ℓ+1)ℓ(ℓ−1)).
 * S_T / S_P: [nk_src][nx] or NULL; any of cltt/clte/clee may be NULL, a
 * requested spectrum needs its sources. */
int bolt_cl(bolt_ctx* ctx, int cosmo_id, const double* S_T, const double* S_P, const double* k_src,
            int nk_src, const double* kgrid, int nkd, const int* ells, int nell, int bessel_mode,
            double* cltt, double* clte, double* clee);
/* Device-resident bolt_cl (ℓ list on the host; d_steps and ms as in bolt_cltt_dev). */
int bolt_cl_dev(bolt_ctx* ctx, int cosmo_id, const double* d_ST, const double* d_SP,
                const double* d_k_src, int nk_src, const double* d_kgrid, int nkd, const int* ells,
                int nell, double* d_tt, double* d_te, double* d_ee, int64_t* d_steps, float* ms);

/* plin(k, par, bg, ih, n_q, ℓᵧ, ℓ_ν, ℓ_mν, x, reltol) for nk wavenumbers
 * (spectra.jl:163-198); x must be a grid node or 0. */
int bolt_plin(bolt_ctx* ctx, int cosmo_id, const bolt_hier_cfg* cfg, const double* k, int nk,
              double x, double* pk, int* status, bolt_stats* stats);

/* ------------------------------------------------------------------ serving
 * Continuous batching of spectrum requests on one context (no Bolt.jl
 * counterpart: its callers run source_grid/cltt/plin per request).  Each
 * request brings a cosmology; the library builds its background and RECFAST
 * tables on the host, solves its k-modes and, for C_ℓ, runs the LOS on a free
 * slot (one stream and its own preallocated HBM); requests on different slots
 * run concurrently and their kernels interleave on the device.  A request's
 * result equals that of the same computation run alone. */
enum { BOLT_REQ_PK = 0, BOLT_REQ_CLTT = 1 };
enum { BOLT_REQ_QUEUED = 0, BOLT_REQ_RUNNING = 1, BOLT_REQ_DONE = 2 };

typedef struct {
    bolt_cosmo_params par;
    int kind;            /* BOLT_REQ_PK: plin at x = 0 (spectra.jl:163-198) on nk log-spaced k in
                            [kmin, kmax]; BOLT_REQ_CLTT: source_grid on those k, then cltt for
                            ℓ = 2..ell_max on quadratic_k(0.01H₀, 1000H₀, nkd) (spectra.jl:84-96) */
    int nk;              /* 2..800 */
    double kmin, kmax;   /* Mpc⁻¹ */
    const double* k;     /* optional explicit k grid (nk positive values, increasing); NULL: log-spaced
                            in [kmin, kmax]; caller-owned until the request is done */
    int ell_max;         /* CLTT: 2..2500 */
    int nkd;             /* CLTT: dense k nodes, 2..5000 (Bolt uses 5000) */
    bolt_hier_cfg cfg;   /* hierarchy and tolerances (state length <= 1024, nq <= 48) */
    double* out;         /* caller-owned, kept alive until the request is done: nk values
                            (PK) or ell_max - 1 values (CLTT) */
} bolt_request;

typedef struct {
    int state;           /* BOLT_REQ_QUEUED / RUNNING / DONE */
    int status;          /* when DONE: BOLT_OK or the error code */
    double queue_ms;     /* submit -> start on a slot */
    double host_ms;      /* background + RECFAST tables on the host */
    double device_ms;    /* slots: solver + LOS kernel time; queue server: the C_l or P(k) step after the solve */
    double latency_ms;   /* submit -> done */
    double solve_ms;     /* queue server: publication of the request's modes -> the last of them solved */
    int joined_running;  /* queue server: 1 if the request's modes were published while a solver launch was
                            running and that launch solved all of them */
    char error[160];     /* when DONE with an error: its message */
} bolt_result;

typedef struct {
    long long launches;          /* solver launches so far */
    long long served;            /* requests completed */
    long long joined_running;    /* of those, admitted into a launch that was already running */
    long long arena_bytes;       /* HBM arena the requests' memory comes from */
    long long arena_peak_bytes;  /* its high-water mark */
    int solver_workgroups;       /* workgroups of each solver launch (CUs minus those left for C_l / P(k)) */
} bolt_queue_info;

/* Start serving with up to `slots` concurrent requests, bounded by hbm_gib of
 * device memory (each slot preallocates ~0.13 GiB). */
int bolt_serve_start(bolt_ctx* ctx, int slots, double hbm_gib);
/* Serving from a device work queue, for callers that submit requests while
 * others run (no Bolt counterpart: Bolt's callers run source_grid/cltt/plin
 * per request, spectra.jl:6-23, 84-96, 163-198).  One persistent solver
 * launch takes every request's k-modes: the host uploads them into a work
 * queue in device memory and raises a published count in pinned host memory
 * while the launch runs, and its workgroups claim them as they free up; a
 * launch is drained 5 ms after the server holds no request, and a new one
 * starts when work arrives.  The solver's stream is masked off one XCD (32
 * CUs), on which the C_l / P(k) steps of finished requests run
 * (BOLT_QS_RESERVE_CU: more, in whole XCDs).  Each request's device memory
 * (tables, k grids, outputs) is taken from one HBM arena (hbm_gib less ~30 MB
 * of fixed work space, at most 15 GiB) by its own size and returned when it
 * is done; a request larger than the arena is refused at submit
 * (BOLT_E_NOMEM).  host_threads (0: 4) build the requests' host tables.
 * Requests are the bolt_request of bolt_serve_start; bolt_submit / poll /
 * wait / serve_stop serve both. */
int bolt_serve_start_queue(bolt_ctx* ctx, double hbm_gib, int host_threads);
/* Counters of the running queue server (BOLT_E_STATE if there is none). */
int bolt_serve_queue_info(bolt_ctx* ctx, bolt_queue_info* info);
/* Number of slots of the running server (0 if none). */
int bolt_serve_slots(bolt_ctx* ctx);
/* Queue a request; *id identifies it.  The request struct is copied. */
int bolt_submit(bolt_ctx* ctx, const bolt_request* req, uint64_t* id);
/* Non-blocking state of a request. */
int bolt_poll(bolt_ctx* ctx, uint64_t id, bolt_result* res);
/* Block until the request is done; the result is then dropped from the
 * server's table (a later poll or wait of the id reports an unknown id). */
int bolt_wait(bolt_ctx* ctx, uint64_t id, bolt_result* res);
/* Finish the queued and running requests, then free the slots. */
int bolt_serve_stop(bolt_ctx* ctx);

/* ---- Moments of spherical Bessel functions and the Filon rule
 * (src/bessel/moments.jl, interpolator.jl, integrator.jl).
 * I_m(x) = ∫_0^x t^m j_ν(t) dt. */
typedef struct bolt_moment_table bolt_moment_table;
/* sph_bessel_interpolator(ν, order, x_min, x_max, N) (interpolator.jl:85-111):
 * the moments m = 0..order-1 (order 3 or 4) on LinRange(x_min, x_max, N) as
 * cubic B-splines with Line(OnGrid()) ends, built on the device: nodes below
 * x = 20 by the ₁F₂ series in double-double, from 20 to 40 by exact
 * recurrences down to Si(x) (the reference: Weniger's transformation of ₁F₂ in
 * Double64 below 50), from 40 on by the Lommel asymptotic.
 * ν = 1, 2 or 3 (the reference builds tables for 2 and 3); 4 <= N <= 2^28. */
int bolt_moment_table_create(bolt_ctx* ctx, int nu, int order, double x_min, double x_max, int n,
                             bolt_moment_table** out);
int bolt_moment_table_destroy(bolt_moment_table* tab);
/* MomentTable(x) (interpolator.jl:60-68): spline in [x_min, x_max], the FP64
 * Maclaurin series below, the asymptotic above; out[i*order + m] (host arrays) */
int bolt_moment_table_eval(bolt_moment_table* tab, const double* x, int nx, double* out);
/* the spline coefficients, [order][N + 2] (host array) */
int bolt_moment_table_coefficients(bolt_moment_table* tab, double* out);
/* integrate_sph_bessel_filon(f, f′, f″, k, a, b, itp) (integrator.jl:5-18) for
 * n intervals: ∫_a^b q(t) j_ν(kt) dt of the quadratic q through f, f′, f″ at a */
int bolt_filon(bolt_moment_table* tab, const double* f, const double* fp, const double* fpp, const double* k,
               const double* a, const double* b, int n, double* out);
/* Σ_i of the Filon rule over (t_i, t_{i+1}), f, f′, f″ at t_i, for each of nk
 * wavenumbers: a piecewise-quadratic integrand against j_ν(kt) (the loop form,
 * integrator.jl:20-38) */
int bolt_filon_sum(bolt_moment_table* tab, const double* t, const double* f, const double* fp, const double* fpp,
                   int nt, const double* k, int nk, double* out);
/* the moment functions at n points (x_i, ν_i, m_i): method 0 the Lommel
 * asymptotic (sph_j_moment_asymp, moments.jl:76-86; x > ν), 1 the FP64
 * Maclaurin series (sph_j_moment_maclaurin_₁F₂, moments.jl:97-102), 2 a table
 * node's method: the series in double-double below x = 20, the recurrences from
 * 20 to 40 (integer 0 <= m <= 3, ν <= 3) */
int bolt_sph_moments(bolt_ctx* ctx, int method, const double* x, const int* nu, const double* m, int n, double* out);

#ifdef __cplusplus
}
#endif
#endif /* BOLT_HIP_H */
