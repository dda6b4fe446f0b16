/*
 * CPU FP64 oracle for the perturbation hot path -- TEST INFRASTRUCTURE ONLY.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it,
 * and only as the checker.  The product path never links or calls it.
 *
 * It restates Bolt.jl's src/perturbations.jl and src/spectra.jl in plain C on
 * the same spline tables the device path consumes (bolt_tables layout of
 * include/bolt_hip.h).
 */
#ifndef BOLT_ORACLE_H
#define BOLT_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    /* CosmoParams order (src/Bolt.jl:56-66) */
    double h, Omega_r, Omega_b, Omega_c, A, n, Y_p, N_nu, Sum_m_nu;
} orc_params;

typedef struct {
    int nx, nq;
    double x0, x1;
    double H0, eta0, rho_crit;
    const double* xq;   /* [nq] */
    const double* wq;   /* [nq] */
    const double* coef; /* [16][nx+2] same order as BOLT_T_* */
} orc_tables;

typedef struct {
    int ell_g, ell_nu, ell_mnu, nq;
    double reltol, abstol;
    int max_steps;
    int lu_refine; /* 1: one residual-correction sweep after each stage solve */
} orc_cfg;

typedef struct {
    int64_t accepted, rejected, rhs_evals, factorizations;
    int status; /* 0 ok, 1 max steps, 2 dt underflow, 3 non-finite, 4 bad configuration */
} orc_stats;

/* largest momentum-bin count the oracle's fixed per-bin arrays hold; calls with
 * nq outside [1, ORC_MAXQ] or a cut-off below 2 return status 4 / NaN outputs */
#define ORC_MAXQ 60

int orc_state_length(const orc_cfg* c);
/* KenCarp4 ESDIRK tableau used by the integrator: A [6][6] (diagonal γ), b, b̂, c */
void orc_tableau(double* A, double* b, double* bh, double* c, double* gamma);
double orc_spline(const orc_tables* t, int tab, double x, int deriv);

/* hierarchy! (perturbations.jl:161-271); mutate != 0 applies the RSA overwrite to u */
void orc_rhs(const orc_params* p, const orc_tables* t, const orc_cfg* c, double k, double x,
             double* u, double* du, int mutate);
/* analytic Jacobian of hierarchy! at (k, x), dense row-major n x n */
void orc_jacobian(const orc_params* p, const orc_tables* t, const orc_cfg* c, double k, double x,
                  double* J);
void orc_initial_conditions(const orc_params* p, const orc_tables* t, const orc_cfg* c, double k,
                            double* u0);
double orc_source(const orc_params* p, const orc_tables* t, const orc_cfg* c, double k, double x,
                  const double* u, const double* du);
void orc_rho_sigma(const orc_params* p, const orc_tables* t, int nq, double a, const double* M0,
                   const double* M2, double* rho, double* sigma);
double orc_theta(const orc_params* p, const orc_tables* t, int nq, double a, const double* M1);
/* Ω_ν of plin's non-relativistic massive neutrino (spectra.jl:182-186) */
double orc_omega_nu(const orc_params* p, const orc_tables* t);
/* plin with a deliberately planted mistake, for tests that show a pin can
 * fail: plant bit 1 flips the sign of the 3ℋv/k gauge terms, bit 2 scales
 * νfac by 1.1, bit 4 drops the factor 2 of 2π²/k³, bit 8 doubles θ.
 * plant = 0 is orc_plin. */
double orc_plin_planted(const orc_params* p, const orc_tables* t, const orc_cfg* c, double k, double x,
                        orc_stats* st, int plant);
double orc_omega_nu_planted(const orc_params* p, const orc_tables* t, int plant);

/* boltsolve + dense output at sorted x_save; out [nsave][n] */
int orc_boltsolve(const orc_params* p, const orc_tables* t, const orc_cfg* c, double k,
                  const double* x_save, int nsave, double* out, orc_stats* st);
/* boltsolve + accepted-step log (x_n list) for diagnostics */
int orc_boltsolve_steps(const orc_params* p, const orc_tables* t, const orc_cfg* c, double k,
                        double* xs, int max_xs, int* n_xs, orc_stats* st);
/* boltsolve_rsa (perturbations.jl:86-111); out [nx][n] */
int orc_boltsolve_rsa(const orc_params* p, const orc_tables* t, const orc_cfg* c, double k,
                      double* out, orc_stats* st);
/* source_grid column for one k (spectra.jl:10-18); S [nx] */
/* source_function_P (perturbations.jl:385-404) at x from the (mutated) u */
double orc_source_P(const orc_params* p, const orc_tables* t, const orc_cfg* c, double k, double x,
                    const double* u);

/* one solve; temperature and/or polarization columns (either may be NULL) */
int orc_source_column_tp(const orc_params* p, const orc_tables* t, const orc_cfg* c, double k, double* S_T,
                         double* S_P, orc_stats* st);

int orc_source_column(const orc_params* p, const orc_tables* t, const orc_cfg* c, double k,
                      double* S, orc_stats* st);
/* the temperature column along a fixed step sequence: the accepted steps whose
 * end points orc_boltsolve_steps recorded, each accepted without error control.
 * Smooth in (p, t), so central differences of it are the derivative of the
 * fixed-step discrete solution (the tangent-linear solve's reference). */
int orc_source_column_replay(const orc_params* p, const orc_tables* t, const orc_cfg* c, double k,
                             const double* step_xs, int n_steps, double* S, orc_stats* st);
/* source_grid over nk modes with nthreads OpenMP threads; S [nk][nx] */
int orc_source_grid(const orc_params* p, const orc_tables* t, const orc_cfg* c, const double* k,
                    int nk, double* S, int* status, int nthreads);
/* plin (spectra.jl:163-198) */
double orc_plin(const orc_params* p, const orc_tables* t, const orc_cfg* c, double k, double x,
                orc_stats* st);

/* spherical Bessel j_0..j_lmax at z by normalised downward (Miller) recurrence */
void orc_sph_bessel(int lmax, double z, double* j);
/* natural cubic spline table of j_ell on [0, zmax] with 5001 nodes (spectra.jl:49-58):
 * coef [5003] */
void orc_bessel_table(int ell, double zmax, double* coef);
double orc_bessel_table_eval(const double* coef, double zmax, double z);

/* cltt(ℓ⃗, sf, kgrid, par, bg) (spectra.jl:84-96); bessel_mode 0 = exact j_ℓ,
 * 1 = Bolt's 5001-node spline table.  k-midpoint range [kbeg, kend) of the
 * nkd-1 midpoints (pass 0, nkd-1 for all); partial sums before the 4π factor
 * are returned when partial != 0. */
/* TT, TE, EE from temperature S and polarization S_P source grids (S_P may
 * be NULL; any output may be NULL) */
int orc_cl(const orc_params* p, const orc_tables* t, const double* S, const double* S_P, const double* k_src,
           int nk_src, const double* kgrid, int nkd, const int* ells, int nell, int bessel_mode, double* cl,
           double* clte, double* clee, int nthreads, int kbeg, int kend);

int orc_cltt(const orc_params* p, const orc_tables* t, const double* S, const double* k_src,
             int nk_src, const double* kgrid, int nkd, const int* ells, int nell, int bessel_mode,
             double* cl, int nthreads, int kbeg, int kend);

/* ---- bessel_moments.c: I_m(x) = ∫_0^x t^m j_ν(t) dt (src/bessel/moments.jl),
 * their tables (interpolator.jl) and the Filon rule (integrator.jl) */
double orc_sphj_moment_asymp_prefactor(int nu, double m);
double orc_sphj_moment_asymp(double x, int nu, double m);
double orc_sphj_moment_maclaurin(double x, int nu, double m);
double orc_sphj_moment_maclaurin_q(double x, int nu, double m);
double orc_J_moment_halfint(double x, int n, double alpha, int method);
double orc_linrange(double xmin, double xmax, int n, int i);
typedef struct {
    int nu, order, n;
    double xmin, xmax;
    double* coef; /* [order][n + 2] */
} orc_moment_table;
int orc_moment_table_build(orc_moment_table* T, int nu, int order, double xmin, double xmax, int n, int nthreads);
void orc_moment_table_free(orc_moment_table* T);
void orc_moment_eval(const orc_moment_table* T, double x, double* out);
double orc_filon(double f, double fp, double fpp, double k, double a, double b, const orc_moment_table* T);
void orc_filon_sum(const orc_moment_table* T, const double* t, const double* f, const double* fp, const double* fpp,
                   int nt, const double* k, int nk, double* out, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
