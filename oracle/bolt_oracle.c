/*
 * CPU FP64 oracle for Bolt.jl's perturbation hot path -- TEST INFRASTRUCTURE.
 * Used only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
 * always as the checker; the product (libbolt_hip.so) never links it.
 *
 * Restated from Bolt.jl sources (file:line into src/):
 *   unpack                      perturbations.jl:114-125   -> layout()
 *   ρ_σ, θ                      perturbations.jl:127-158   -> orc_rho_sigma, orc_theta
 *   hierarchy!                  perturbations.jl:161-271   -> orc_rhs
 *   initial_conditions          perturbations.jl:274-338   -> orc_initial_conditions
 *   source_function             perturbations.jl:343-383   -> orc_source
 *   boltsolve (KenCarp4)        perturbations.jl:25-33     -> integrate()
 *   rsa_perts!, boltsolve_rsa   perturbations.jl:35-111    -> orc_boltsolve_rsa
 *   source_grid                 spectra.jl:6-23            -> orc_source_grid
 *   bessel_interpolator         spectra.jl:49-58           -> orc_bessel_table
 *   Tl, cltt                    spectra.jl:70-96           -> orc_cltt
 *   plin                        spectra.jl:163-198         -> orc_plin
 *
 * The ODE integrator is KenCarp4's ESDIRK tableau (Kennedy & Carpenter 2003,
 * ARK4(3)6L[2]SA, implicit part).  OrdinaryDiffEq's controller, initial step,
 * Newton/W-reuse policy and interpolant are not in the reference tree; this
 * file defines them: I-controller (safety 0.9, exponent 1/4, factors
 * [0.2, 10]) on Shampine's smoothed error estimate W⁻¹Σ(b-b̂)z (KenCarp4's
 * smooth_est default), Hairer's starting step, the exact stage solve (hierarchy! is
 * linear in u, so W = I - hγJ at the stage abscissa solves each stage in one
 * Newton iteration), cubic Hermite dense output, and FSAL evaluation of the
 * accepted state (which is where hierarchy!'s RSA mutation of u persists).
 * The linear solve is Gaussian elimination with partial pivoting on W
 * (symmetrically permuted so hierarchy tails come first), skipping exact-zero
 * multipliers and pivot-row entries.
 *
 * What pins each function (tests/test_oracle.py, tests/test_oracle_plin.py):
 *   orc_rhs, orc_jacobian       exact Jacobian/linearity identities; CLASS Φ, δ_b fixture 1e-3
 *   orc_initial_conditions      closed forms (Φ, δ = 3Θ₀, v = −3kΘ₁), super-horizon stasis
 *   integrate (KenCarp4)        tableau order conditions in rational arithmetic; CLASS fixture
 *   orc_source*, orc_cl*        CAMB C_ℓ fixture 1.1e-1; single-shell closed form; TE/EE spin factors
 *   orc_sph_bessel, tables      scipy/mpmath j_ℓ 1e-12; scipy natural spline
 *   orc_rho_sigma, orc_theta    massless closed form 4·(7/8)(π²/15)T_ν⁴; Poisson constraint
 *   orc_plin, orc_omega_nu      CAMB P(k) fixture 5e-3; Poisson constraint 1e-4; Ω_ν h² = Σm_ν/93.14 eV
 * orc_plin_planted / orc_omega_nu_planted exist only so that the tests can
 * show each of those pins failing on a planted mistake.
 */
#include "bolt_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

enum { T_HC = 0, T_HC_P, T_HC_PP, T_ETA, T_ETA_P, T_ETA_PP, T_RHO0M, T_XE, T_TAU, T_TAU_P, T_TAU_PP,
       T_G, T_G_P, T_G_PP, T_TMAT, T_CSB2 };

/* ------------------------------------------------------------------ layout */
typedef struct {
    int ng, nn, nm, nq, n;
    int iT, iP, iN, iM, iPhi, idel, iv, idb, ivb;
} lay_t;

static lay_t layout(const orc_cfg* c) { /* perturbations.jl:114-125 */
    lay_t L;
    L.ng = c->ell_g; L.nn = c->ell_nu; L.nm = c->ell_mnu; L.nq = c->nq;
    L.iT = 0;
    L.iP = L.ng + 1;
    L.iN = 2 * (L.ng + 1);
    L.iM = L.iN + L.nn + 1;
    L.iPhi = L.iM + (L.nm + 1) * L.nq;
    L.idel = L.iPhi + 1; L.iv = L.iPhi + 2; L.idb = L.iPhi + 3; L.ivb = L.iPhi + 4;
    L.n = L.iPhi + 5;
    return L;
}

int orc_state_length(const orc_cfg* c) { return layout(c).n; }

/* the per-bin scratch arrays below are sized for ORC_MAXQ bins */
static int nq_bad(int nq) { return nq < 1 || nq > ORC_MAXQ; }
static int cfg_bad(const orc_cfg* c) { return nq_bad(c->nq) || c->ell_g < 2 || c->ell_nu < 2 || c->ell_mnu < 2; }
static void fill_nan(double* v, size_t n) { for (size_t i = 0; i < n; ++i) v[i] = NAN; }

/* ----------------------------------------------------------------- splines */
/* uniform cubic B-spline value/derivatives (util.jl:11-13) */
double orc_spline(const orc_tables* t, int tab, double x, int deriv) {
    int n = t->nx;
    double dx = (t->x1 - t->x0) / (n - 1);
    const double* c = t->coef + (size_t)tab * (n + 2);
    double tt = (x - t->x0) / dx;
    int i = (int)floor(tt);
    if (i < 0) i = 0;
    if (i > n - 2) i = n - 2;
    double s = tt - i, r = 1.0 - s;
    const double* q = c + i;
    if (deriv == 0)
        return r * r * r / 6.0 * q[0] + (2.0 / 3.0 - s * s + 0.5 * s * s * s) * q[1] +
               (2.0 / 3.0 - r * r + 0.5 * r * r * r) * q[2] + s * s * s / 6.0 * q[3];
    if (deriv == 1)
        return (-0.5 * r * r * q[0] + (-2.0 * s + 1.5 * s * s) * q[1] + (2.0 * r - 1.5 * r * r) * q[2] +
                0.5 * s * s * q[3]) / dx;
    return (r * q[0] + (-2.0 + 3.0 * s) * q[1] + (-2.0 + 3.0 * r) * q[2] + s * q[3]) / (dx * dx);
}

static double xnode(const orc_tables* t, int j) {
    return (t->x0 * (t->nx - 1 - j) + t->x1 * j) / (t->nx - 1);
}

/* ------------------------------------------------------- neutrino q-integrals */
typedef struct {
    double Tnu, lqmin, lqmax;
} qgrid_t;

static qgrid_t qgrid(const orc_params* p, const orc_tables* t) {
    qgrid_t g;
    g.Tnu = pow(p->N_nu / 3.0, 0.25) * pow(4.0 / 11.0, 1.0 / 3.0) *
            pow(15.0 / (M_PI * M_PI) * t->rho_crit * p->Omega_r, 0.25);
    g.lqmin = log10(g.Tnu / 30.0);
    g.lqmax = log10(g.Tnu * 30.0);
    return g;
}
static double xq2q(double x, double lo, double hi) { return pow(10.0, lo + (hi - lo) / 2.0 * (x + 1.0)); }
static double dxdq(double q, double lo, double hi) {
    double tu = -1.0 + 2.0 / (hi - lo) * (1.0 + lo - lo);
    return (1.0 + tu) / (q * log(10.0));
}
static double f0(double q, double Tnu) { return 2.0 / pow(2.0 * M_PI, 3) / (exp(q / Tnu) + 1.0); }
static double dlnf0dlnq(double q, double Tnu) { return -q / Tnu / (1.0 + exp(-q / Tnu)); }

void orc_rho_sigma(const orc_params* p, const orc_tables* t, int nq, double a, const double* M0,
                   const double* M2, double* rho, double* sigma) { /* perturbations.jl:127-145 */
    if (nq_bad(nq)) { *rho = *sigma = NAN; return; }
    qgrid_t g = qgrid(p, t);
    double m = p->Sum_m_nu, sr = 0.0, ss = 0.0;
    for (int i = 0; i < nq; ++i) {
        double q = xq2q(t->xq[i], g.lqmin, g.lqmax);
        double eps = sqrt(q * q + (a * m) * (a * m));
        double base = q * q * f0(q, g.Tnu) / dxdq(q, g.lqmin, g.lqmax);
        sr += base * eps * M0[i] * t->wq[i];
        ss += base * (q * q / eps) * M2[i] * t->wq[i];
    }
    *rho = 4.0 * M_PI * sr;
    *sigma = 4.0 * M_PI * ss;
}

double orc_theta(const orc_params* p, const orc_tables* t, int nq, double a, const double* M1) {
    (void)a; /* perturbations.jl:148-158 */
    if (nq_bad(nq)) return NAN;
    qgrid_t g = qgrid(p, t);
    double s = 0.0;
    for (int i = 0; i < nq; ++i) {
        double q = xq2q(t->xq[i], g.lqmin, g.lqmax);
        s += q * q * q * f0(q, g.Tnu) / dxdq(q, g.lqmin, g.lqmax) * M1[i] * t->wq[i];
    }
    return 4.0 * M_PI * s;
}

/* -------------------------------------------------------------- hierarchy! */
typedef struct {
    double a, Hc, eta, tp, tpp, csb2;
} bgx_t;

static bgx_t bgx(const orc_tables* t, double x) {
    bgx_t b;
    b.a = exp(x);
    b.Hc = orc_spline(t, T_HC, x, 0);
    b.eta = orc_spline(t, T_ETA, x, 0);
    b.tp = orc_spline(t, T_TAU_P, x, 0);
    b.tpp = orc_spline(t, T_TAU_PP, x, 0);
    b.csb2 = orc_spline(t, T_CSB2, x, 0);
    return b;
}

static int rsa_on(double k, const bgx_t* b) { /* perturbations.jl:216 */
    return (k * b->eta > 240.0) && (-b->tp * b->Hc / b->eta > 100.0);
}

static double omega_nu_rel(const orc_params* p) { /* 7*(2/3)*N_ν/8*(4/11)^(4/3)*Ω_r */
    return 7.0 * (2.0 / 3.0) * p->N_nu / 8.0 * pow(4.0 / 11.0, 4.0 / 3.0) * p->Omega_r;
}

/* Ψ and Φ′ (perturbations.jl:182-194) */
static void metric(const orc_params* p, const orc_tables* t, const lay_t* L, double k, const bgx_t* b,
                   const double* u, double* Psi, double* Phip) {
    double a = b->a, H02 = t->H0 * t->H0, Onu = omega_nu_rel(p);
    double rho, sig;
    double M0[ORC_MAXQ], M2[ORC_MAXQ];
    for (int i = 0; i < L->nq; ++i) { M0[i] = u[L->iM + i]; M2[i] = u[L->iM + 2 * L->nq + i]; }
    orc_rho_sigma(p, t, L->nq, a, M0, M2, &rho, &sig);
    double Phi = u[L->iPhi];
    double ps = -Phi - 12.0 * H02 / (k * k) / (a * a) *
                          (p->Omega_r * u[L->iT + 2] + Onu * u[L->iN + 2] + sig / t->rho_crit / 4.0);
    double Hc = b->Hc;
    double pp = ps - k * k / (3.0 * Hc * Hc) * Phi +
                H02 / (2.0 * Hc * Hc) *
                    (p->Omega_c / a * u[L->idel] + p->Omega_b / a * u[L->idb] +
                     4.0 * p->Omega_r / (a * a) * u[L->iT + 0] + 4.0 * Onu / (a * a) * u[L->iN + 0] +
                     rho / t->rho_crit / (a * a));
    *Psi = ps;
    *Phip = pp;
}

void orc_rhs(const orc_params* p, const orc_tables* t, const orc_cfg* c, double k, double x, double* u,
             double* du, int mutate) {
    lay_t L = layout(c);
    if (cfg_bad(c)) { fill_nan(du, L.n > 0 ? (size_t)L.n : 0); return; }
    bgx_t b = bgx(t, x);
    qgrid_t g = qgrid(p, t);
    double a = b.a, Hc = b.Hc, eta = b.eta, tp = b.tp, tpp = b.tpp, csb2 = b.csb2;
    double R = 4.0 * p->Omega_r / (3.0 * p->Omega_b * a);
    double Psi, Phip;
    metric(p, t, &L, k, &b, u, &Psi, &Phip);
    double Phi = u[L.iPhi], v = u[L.iv], db = u[L.idb], vb = u[L.ivb];
    const double* T = u + L.iT;
    const double* P = u + L.iP;
    const double* N = u + L.iN;
    const double* M = u + L.iM;
    double* dT = du + L.iT;
    double* dP = du + L.iP;
    double* dN = du + L.iN;
    double* dM = du + L.iM;
    int ng = L.ng, nn = L.nn, nm = L.nm, nq = L.nq;
    /* matter (perturbations.jl:197-200) */
    double ddel = k / Hc * v - 3.0 * Phip;
    double dv = -v - k / Hc * Psi;
    double ddb = k / Hc * vb - 3.0 * Phip;
    double dvb = -vb - k / Hc * (Psi + csb2 * db) + tp * R * (3.0 * T[1] + vb);
    /* massive neutrinos (203-213) */
    for (int i = 0; i < nq; ++i) {
        double q = xq2q(t->xq[i], g.lqmin, g.lqmax);
        double eps = sqrt(q * q + (a * p->Sum_m_nu) * (a * p->Sum_m_nu));
        double df0 = dlnf0dlnq(q, g.Tnu);
        dM[0 * nq + i] = -k / Hc * q / eps * M[1 * nq + i] + Phip * df0;
        dM[1 * nq + i] = k / (3.0 * Hc) * (q / eps * (M[0 * nq + i] - 2.0 * M[2 * nq + i]) - eps / q * Psi * df0);
        for (int l = 2; l <= nm - 1; ++l)
            dM[l * nq + i] = k / Hc * q / ((2 * l + 1) * eps) * (l * M[(l - 1) * nq + i] - (l + 1) * M[(l + 1) * nq + i]);
        dM[nm * nq + i] = q / eps * k / Hc * M[(nm - 1) * nq + i] - (nm + 1) / (Hc * eta) * M[nm * nq + i];
    }
    if (rsa_on(k, &b)) { /* 216-233 */
        if (mutate) {
            u[L.iT + 0] = Phi - Hc / k * tp * vb;
            u[L.iT + 1] = Hc / k * (-2.0 * Phip + tp * (Phi - csb2 * db) + Hc / k * (tpp - tp) * vb);
            u[L.iT + 2] = 0.0;
            u[L.iN + 0] = Phi;
            u[L.iN + 1] = -2.0 * Hc / k * Phip;
            u[L.iN + 2] = 0.0;
        }
        for (int l = 0; l <= nn; ++l) dN[l] = 0.0;
        for (int l = 0; l <= ng; ++l) { dT[l] = 0.0; dP[l] = 0.0; }
    } else {
        dN[0] = -k / Hc * N[1] - Phip;
        dN[1] = k / (3.0 * Hc) * N[0] - 2.0 * k / (3.0 * Hc) * N[2] + k / (3.0 * Hc) * Psi;
        for (int l = 2; l <= nn - 1; ++l) dN[l] = k / ((2 * l + 1) * Hc) * (l * N[l - 1] - (l + 1) * N[l + 1]);
        dN[nn] = k / Hc * N[nn - 1] - (nn + 1) / (Hc * eta) * N[nn];
        double Pi = T[2] + P[2] + P[0];
        dT[0] = -k / Hc * T[1] - Phip;
        dT[1] = k / (3.0 * Hc) * T[0] - 2.0 * k / (3.0 * Hc) * T[2] + k / (3.0 * Hc) * Psi + tp * (T[1] + vb / 3.0);
        for (int l = 2; l <= ng - 1; ++l)
            dT[l] = l * k / ((2 * l + 1) * Hc) * T[l - 1] - (l + 1) * k / ((2 * l + 1) * Hc) * T[l + 1] +
                    tp * (T[l] - Pi * (l == 2 ? 1.0 : 0.0) / 10.0);
        dP[0] = -k / Hc * P[1] + tp * (P[0] - Pi / 2.0);
        for (int l = 1; l <= ng - 1; ++l)
            dP[l] = l * k / ((2 * l + 1) * Hc) * P[l - 1] - (l + 1) * k / ((2 * l + 1) * Hc) * P[l + 1] +
                    tp * (P[l] - Pi * (l == 2 ? 1.0 : 0.0) / 10.0);
        dT[ng] = k / Hc * T[ng - 1] - ((ng + 1) / (Hc * eta) - tp) * T[ng];
        dP[ng] = k / Hc * P[ng - 1] - ((ng + 1) / (Hc * eta) - tp) * P[ng];
    }
    du[L.iPhi] = Phip; du[L.idel] = ddel; du[L.iv] = dv; du[L.idb] = ddb; du[L.ivb] = dvb;
}

/* Analytic ∂(hierarchy!)/∂u at (k, x): every term of the RHS is an
 * x-dependent coefficient times one state entry, so J is exact. */
typedef void (*addfn)(void* ctx, int r, int c, double v);

static void jac_entries(const orc_params* p, const orc_tables* t, const orc_cfg* cf, double k, double x,
                        addfn add, void* ctx) {
    lay_t L = layout(cf);
    bgx_t b = bgx(t, x);
    qgrid_t g = qgrid(p, t);
    double a = b.a, Hc = b.Hc, eta = b.eta, tp = b.tp, csb2 = b.csb2;
    double R = 4.0 * p->Omega_r / (3.0 * p->Omega_b * a);
    double H02 = t->H0 * t->H0, Onu = omega_nu_rel(p);
    int ng = L.ng, nn = L.nn, nm = L.nm, nq = L.nq;
    /* gradients of the Ψ and Φ′ functionals */
    int nidx = 0;
    int idx[8 + 2 * ORC_MAXQ];
    double dPsi[8 + 2 * ORC_MAXQ], dPhip[8 + 2 * ORC_MAXQ];
    double c12 = 12.0 * H02 / (k * k) / (a * a);
    double gg = H02 / (2.0 * Hc * Hc);
#define PUSH(col, vpsi, vphi) do { idx[nidx] = (col); dPsi[nidx] = (vpsi); dPhip[nidx] = (vphi); ++nidx; } while (0)
    PUSH(L.iPhi, -1.0, -1.0 - k * k / (3.0 * Hc * Hc));
    PUSH(L.iT + 2, -c12 * p->Omega_r, -c12 * p->Omega_r);
    PUSH(L.iN + 2, -c12 * Onu, -c12 * Onu);
    PUSH(L.idel, 0.0, gg * p->Omega_c / a);
    PUSH(L.idb, 0.0, gg * p->Omega_b / a);
    PUSH(L.iT + 0, 0.0, gg * 4.0 * p->Omega_r / (a * a));
    PUSH(L.iN + 0, 0.0, gg * 4.0 * Onu / (a * a));
    double qv[ORC_MAXQ], epsv[ORC_MAXQ], df0v[ORC_MAXQ];
    for (int i = 0; i < nq; ++i) {
        double q = xq2q(t->xq[i], g.lqmin, g.lqmax);
        double eps = sqrt(q * q + (a * p->Sum_m_nu) * (a * p->Sum_m_nu));
        double base = q * q * f0(q, g.Tnu) / dxdq(q, g.lqmin, g.lqmax);
        double Isig = 4.0 * M_PI * base * (q * q / eps) * t->wq[i];
        double Irho = 4.0 * M_PI * base * eps * t->wq[i];
        double s2 = -c12 * Isig / t->rho_crit / 4.0;
        PUSH(L.iM + 2 * nq + i, s2, s2);
        PUSH(L.iM + 0 * nq + i, 0.0, gg * Irho / t->rho_crit / (a * a));
        qv[i] = q; epsv[i] = eps; df0v[i] = dlnf0dlnq(q, g.Tnu);
    }
#undef PUSH
#define ROWF(row, cPsi, cPhip) do { for (int z_ = 0; z_ < nidx; ++z_) { \
        double v_ = (cPsi) * dPsi[z_] + (cPhip) * dPhip[z_]; if (v_ != 0.0) add(ctx, row, idx[z_], v_); } } while (0)
    /* Φ, matter */
    ROWF(L.iPhi, 0.0, 1.0);
    add(ctx, L.idel, L.iv, k / Hc); ROWF(L.idel, 0.0, -3.0);
    add(ctx, L.iv, L.iv, -1.0); ROWF(L.iv, -k / Hc, 0.0);
    add(ctx, L.idb, L.ivb, k / Hc); ROWF(L.idb, 0.0, -3.0);
    add(ctx, L.ivb, L.ivb, -1.0 + tp * R);
    add(ctx, L.ivb, L.idb, -k / Hc * csb2);
    add(ctx, L.ivb, L.iT + 1, 3.0 * tp * R);
    ROWF(L.ivb, -k / Hc, 0.0);
    /* massive ν */
    for (int i = 0; i < nq; ++i) {
        double qe = qv[i] / epsv[i];
        int r0 = L.iM + i, r1 = L.iM + nq + i;
        add(ctx, r0, L.iM + nq + i, -k / Hc * qe);
        ROWF(r0, 0.0, df0v[i]);
        add(ctx, r1, L.iM + i, k / (3.0 * Hc) * qe);
        add(ctx, r1, L.iM + 2 * nq + i, -2.0 * k / (3.0 * Hc) * qe);
        ROWF(r1, -k / (3.0 * Hc) * epsv[i] / qv[i] * df0v[i], 0.0);
        for (int l = 2; l <= nm - 1; ++l) {
            double cc = k / Hc * qv[i] / ((2 * l + 1) * epsv[i]);
            add(ctx, L.iM + l * nq + i, L.iM + (l - 1) * nq + i, cc * l);
            add(ctx, L.iM + l * nq + i, L.iM + (l + 1) * nq + i, -cc * (l + 1));
        }
        add(ctx, L.iM + nm * nq + i, L.iM + (nm - 1) * nq + i, qe * k / Hc);
        add(ctx, L.iM + nm * nq + i, L.iM + nm * nq + i, -(nm + 1) / (Hc * eta));
    }
    if (!rsa_on(k, &b)) {
        int iN = L.iN, iT = L.iT, iP = L.iP;
        add(ctx, iN + 0, iN + 1, -k / Hc); ROWF(iN + 0, 0.0, -1.0);
        add(ctx, iN + 1, iN + 0, k / (3.0 * Hc)); add(ctx, iN + 1, iN + 2, -2.0 * k / (3.0 * Hc));
        ROWF(iN + 1, k / (3.0 * Hc), 0.0);
        for (int l = 2; l <= nn - 1; ++l) {
            double cc = k / ((2 * l + 1) * Hc);
            add(ctx, iN + l, iN + l - 1, cc * l);
            add(ctx, iN + l, iN + l + 1, -cc * (l + 1));
        }
        add(ctx, iN + nn, iN + nn - 1, k / Hc);
        add(ctx, iN + nn, iN + nn, -(nn + 1) / (Hc * eta));
        /* photons */
        add(ctx, iT + 0, iT + 1, -k / Hc); ROWF(iT + 0, 0.0, -1.0);
        add(ctx, iT + 1, iT + 0, k / (3.0 * Hc)); add(ctx, iT + 1, iT + 2, -2.0 * k / (3.0 * Hc));
        add(ctx, iT + 1, iT + 1, tp); add(ctx, iT + 1, L.ivb, tp / 3.0);
        ROWF(iT + 1, k / (3.0 * Hc), 0.0);
        for (int l = 2; l <= ng - 1; ++l) {
            add(ctx, iT + l, iT + l - 1, l * k / ((2 * l + 1) * Hc));
            add(ctx, iT + l, iT + l + 1, -(l + 1) * k / ((2 * l + 1) * Hc));
            add(ctx, iT + l, iT + l, tp);
            if (l == 2) { add(ctx, iT + 2, iT + 2, -tp / 10.0); add(ctx, iT + 2, iP + 2, -tp / 10.0); add(ctx, iT + 2, iP + 0, -tp / 10.0); }
        }
        add(ctx, iP + 0, iP + 1, -k / Hc); add(ctx, iP + 0, iP + 0, tp);
        add(ctx, iP + 0, iT + 2, -tp / 2.0); add(ctx, iP + 0, iP + 2, -tp / 2.0); add(ctx, iP + 0, iP + 0, -tp / 2.0);
        for (int l = 1; l <= ng - 1; ++l) {
            add(ctx, iP + l, iP + l - 1, l * k / ((2 * l + 1) * Hc));
            add(ctx, iP + l, iP + l + 1, -(l + 1) * k / ((2 * l + 1) * Hc));
            add(ctx, iP + l, iP + l, tp);
            if (l == 2) { add(ctx, iP + 2, iT + 2, -tp / 10.0); add(ctx, iP + 2, iP + 2, -tp / 10.0); add(ctx, iP + 2, iP + 0, -tp / 10.0); }
        }
        add(ctx, iT + ng, iT + ng - 1, k / Hc); add(ctx, iT + ng, iT + ng, -((ng + 1) / (Hc * eta) - tp));
        add(ctx, iP + ng, iP + ng - 1, k / Hc); add(ctx, iP + ng, iP + ng, -((ng + 1) / (Hc * eta) - tp));
    }
#undef ROWF
}

typedef struct { double* J; int n; } densectx;
static void add_dense(void* c, int r, int col, double v) { densectx* d = (densectx*)c; d->J[(size_t)r * d->n + col] += v; }

void orc_jacobian(const orc_params* p, const orc_tables* t, const orc_cfg* c, double k, double x, double* J) {
    int n = orc_state_length(c);
    if (cfg_bad(c)) { fill_nan(J, n > 0 ? (size_t)n * n : 0); return; }
    memset(J, 0, sizeof(double) * (size_t)n * n);
    densectx d = {J, n};
    jac_entries(p, t, c, k, x, add_dense, &d);
}

/* ------------------------------------------------------ initial_conditions */
void orc_initial_conditions(const orc_params* p, const orc_tables* t, const orc_cfg* c, double k, double* u) {
    lay_t L = layout(c);
    if (cfg_bad(c)) { fill_nan(u, L.n > 0 ? (size_t)L.n : 0); return; }
    qgrid_t g = qgrid(p, t);
    memset(u, 0, sizeof(double) * L.n);
    double xi = t->x0;
    double Hc = orc_spline(t, T_HC, xi, 0), eta = orc_spline(t, T_ETA, xi, 0), tp = orc_spline(t, T_TAU_P, xi, 0);
    double ai = sqrt(exp(xi) * exp(xi));
    double fnu = 1.0 / (1.0 + 1.0 / (7.0 * (3.0 / 3.0) * p->N_nu / 8.0 * pow(4.0 / 11.0, 4.0 / 3.0)));
    double Rc = 1.0;
    double Phi = (4.0 * fnu + 10.0) / (4.0 * fnu + 15.0) * Rc;
    double C = -((15.0 + 4.0 * fnu) / (20.0 + 8.0 * fnu)) * Phi;
    double* T = u + L.iT; double* P = u + L.iP; double* N = u + L.iN; double* M = u + L.iM;
    T[0] = -40.0 * C / (15.0 + 4.0 * fnu) / 4.0;
    T[1] = 10.0 * C / (15.0 + 4.0 * fnu) * (k * k * eta) / (3.0 * k);
    T[2] = -8.0 * k / (15.0 * Hc * tp) * T[1];
    P[0] = (5.0 / 4.0) * T[2];
    P[1] = -k / (4.0 * Hc * tp) * T[2];
    P[2] = (1.0 / 4.0) * T[2];
    for (int l = 3; l <= L.ng; ++l) {
        T[l] = -(double)l / (2 * l + 1) * k / (Hc * tp) * T[l - 1];
        P[l] = -(double)l / (2 * l + 1) * k / (Hc * tp) * P[l - 1];
    }
    double del = 3.0 / 4.0 * (4.0 * T[0]);
    double v = -3.0 * k * T[1];
    N[0] = T[0];
    N[1] = T[1];
    N[2] = -(k * k * eta * eta) / 15.0 * 1.0 / (1.0 + 2.0 / 5.0 * fnu) * Phi / 2.0;
    for (int l = 3; l <= L.nn; ++l) N[l] = k / ((2 * l + 1) * Hc) * N[l - 1];
    for (int i = 0; i < L.nq; ++i) {
        double q = xq2q(t->xq[i], g.lqmin, g.lqmax);
        double eps = sqrt(q * q + (ai * p->Sum_m_nu) * (ai * p->Sum_m_nu));
        double df0 = dlnf0dlnq(q, g.Tnu);
        M[0 * L.nq + i] = -N[0] * df0;
        M[1 * L.nq + i] = -eps / q * N[1] * df0;
        M[2 * L.nq + i] = -N[2] * df0;
        for (int l = 3; l <= L.nm; ++l) M[l * L.nq + i] = q / eps * k / ((2 * l + 1) * Hc) * M[(l - 1) * L.nq + i];
    }
    u[L.iPhi] = Phi; u[L.idel] = del; u[L.iv] = v; u[L.idb] = del; u[L.ivb] = v;
}

/* --------------------------------------------------------- source_function */
double orc_source(const orc_params* p, const orc_tables* t, const orc_cfg* c, double k, double x,
                  const double* u, const double* du) {
    lay_t L = layout(c);
    if (cfg_bad(c)) return NAN;
    double H02 = t->H0 * t->H0;
    double Hc = orc_spline(t, T_HC, x, 0), Hcp = orc_spline(t, T_HC_P, x, 0), Hcpp = orc_spline(t, T_HC_PP, x, 0);
    double tau = orc_spline(t, T_TAU, x, 0), tp = orc_spline(t, T_TAU_P, x, 0), tpp = orc_spline(t, T_TAU_PP, x, 0);
    double gt = orc_spline(t, T_G, x, 0), gtp = orc_spline(t, T_G_P, x, 0), gtpp = orc_spline(t, T_G_PP, x, 0);
    double a = exp(x);
    double Onu = omega_nu_rel(p);
    double M0[ORC_MAXQ], M2[ORC_MAXQ], dM0[ORC_MAXQ], dM2[ORC_MAXQ];
    for (int i = 0; i < L.nq; ++i) {
        M0[i] = u[L.iM + i]; M2[i] = u[L.iM + 2 * L.nq + i];
        dM0[i] = du[L.iM + i]; dM2[i] = du[L.iM + 2 * L.nq + i];
    }
    double rho, sig, rhop, sigp;
    orc_rho_sigma(p, t, L.nq, a, M0, M2, &rho, &sig);
    orc_rho_sigma(p, t, L.nq, a, dM0, dM2, &rhop, &sigp);
    const double* T = u + L.iT; const double* P = u + L.iP; const double* N = u + L.iN;
    const double* dT = du + L.iT; const double* dP = du + L.iP; const double* dN = du + L.iN;
    double Phip = du[L.iPhi], vb = u[L.ivb], vbp = du[L.ivb];
    double Psi = -u[L.iPhi] - 12.0 * H02 / (k * k) / (a * a) * (p->Omega_r * T[2] + Onu * N[2] + sig / t->rho_crit / 4.0);
    double Psip = -Phip - 12.0 * H02 / (k * k) / (a * a) *
                              (p->Omega_r * (dT[2] - 2.0 * T[2]) + Onu * (dN[2] - 2.0 * N[2]) +
                               (sigp - 2.0 * sig) / t->rho_crit / 4.0);
    double Pi = T[2] + P[2] + P[0];
    double Pip = dT[2] + dP[2] + dP[0];
    double term1 = gt * (T[0] + Psi + Pi / 4.0) + exp(-tau) * (Psip - Phip);
    double term2 = (-1.0 / k) * (Hcp * gt * vb + Hc * gtp * vb + Hc * gt * vbp);
    double Pipp = 2.0 * k / (5.0 * Hc) * (-Hcp / Hc * T[1] + dT[1]) + (3.0 / 10.0) * (tpp * Pi + tp * Pip) -
                  3.0 * k / (5.0 * Hc) * (-Hcp / Hc * (T[3] + P[1] + P[3]) + (dT[3] + dP[1] + dP[3]));
    double term3 = (3.0 / (4.0 * k * k)) *
                   ((Hcp * Hcp + Hc * Hcpp) * gt * Pi + 3.0 * Hc * Hcp * (gtp * Pi + gt * Pip) +
                    Hc * Hc * (gtpp * Pi + 2.0 * gtp * Pip + gt * Pipp));
    return term1 + term2 + term3;
}

/* source_function_P (perturbations.jl:385-404): (3/(4y²)) g̃ Π with
 * y = k(η(x_grid[end]) − η(x)) and Π from the (RSA-mutated) u.  At the last
 * node y = 0 and the value is ±Inf (or NaN for Π = 0), as in Julia. */
double orc_source_P(const orc_params* p, const orc_tables* t, const orc_cfg* c, double k, double x,
                    const double* u) {
    (void)p;
    lay_t L = layout(c);
    double gt = orc_spline(t, T_G, x, 0);
    double y = k * (orc_spline(t, T_ETA, xnode(t, t->nx - 1), 0) - orc_spline(t, T_ETA, x, 0));
    const double* T = u + L.iT; const double* P = u + L.iP;
    double Pi = T[2] + P[2] + P[0];
    return (3.0 / (4.0 * y * y)) * gt * Pi;
}

/* ------------------------------------------------ sparse-skipping GEPP on W */
typedef struct {
    int n;
    double* W;
    int *ord, *pos, *piv;
    int *Lptr, *Lrow, *Uptr, *Ucol;
    double *Lval, *Uval, *Udiag;
    size_t cap;
    int *scratch_cols;
} splu_t;

static void splu_init(splu_t* s, const lay_t* L) {
    int n = L->n;
    s->n = n;
    s->W = (double*)calloc((size_t)n * n, sizeof(double));
    s->ord = (int*)malloc(sizeof(int) * n);
    s->pos = (int*)malloc(sizeof(int) * n);
    s->piv = (int*)malloc(sizeof(int) * n);
    s->cap = (size_t)n * n;
    s->Lptr = (int*)malloc(sizeof(int) * (n + 1));
    s->Uptr = (int*)malloc(sizeof(int) * (n + 1));
    s->Lrow = (int*)malloc(sizeof(int) * s->cap);
    s->Ucol = (int*)malloc(sizeof(int) * s->cap);
    s->Lval = (double*)malloc(sizeof(double) * s->cap);
    s->Uval = (double*)malloc(sizeof(double) * s->cap);
    s->Udiag = (double*)malloc(sizeof(double) * n);
    s->scratch_cols = (int*)malloc(sizeof(int) * n);
    /* elimination order: hierarchy tails (top ℓ first), then the rest */
    int m = 0;
    char* used = (char*)calloc(n, 1);
    for (int l = L->ng; l >= 3; --l) { s->ord[m++] = L->iT + l; used[L->iT + l] = 1; }
    for (int l = L->ng; l >= 3; --l) { s->ord[m++] = L->iP + l; used[L->iP + l] = 1; }
    for (int l = L->nn; l >= 3; --l) { s->ord[m++] = L->iN + l; used[L->iN + l] = 1; }
    for (int i = 0; i < L->nq; ++i)
        for (int l = L->nm; l >= 3; --l) { int r = L->iM + l * L->nq + i; s->ord[m++] = r; used[r] = 1; }
    for (int r = 0; r < n; ++r) if (!used[r]) s->ord[m++] = r;
    free(used);
    for (int i = 0; i < n; ++i) s->pos[s->ord[i]] = i;
}

static void splu_free(splu_t* s) {
    free(s->W); free(s->ord); free(s->pos); free(s->piv); free(s->Lptr); free(s->Uptr);
    free(s->Lrow); free(s->Ucol); free(s->Lval); free(s->Uval); free(s->Udiag); free(s->scratch_cols);
}

typedef struct { splu_t* s; double scale; } wctx;
static void add_w(void* c, int r, int col, double v) {
    wctx* w = (wctx*)c;
    w->s->W[(size_t)w->s->pos[r] * w->s->n + w->s->pos[col]] += w->scale * v;
}

/* W = I - hγ J(x), factor */
static void splu_factor(splu_t* s, const orc_params* p, const orc_tables* t, const orc_cfg* c, double k,
                        double x, double hg) {
    int n = s->n;
    double* W = s->W;
    memset(W, 0, sizeof(double) * (size_t)n * n);
    for (int i = 0; i < n; ++i) W[(size_t)i * n + i] = 1.0;
    wctx w = {s, -hg};
    jac_entries(p, t, c, k, x, add_w, &w);
    size_t nl = 0, nu = 0;
    for (int j = 0; j < n; ++j) {
        int pr = j;
        double best = fabs(W[(size_t)j * n + j]);
        for (int i = j + 1; i < n; ++i) {
            double v = fabs(W[(size_t)i * n + j]);
            if (v > best) { best = v; pr = i; }
        }
        s->piv[j] = pr;
        if (pr != j) {
            double* a = W + (size_t)j * n;
            double* b = W + (size_t)pr * n;
            for (int cc = j; cc < n; ++cc) { double tmp = a[cc]; a[cc] = b[cc]; b[cc] = tmp; }
        }
        double* prow = W + (size_t)j * n;
        double d = prow[j];
        s->Udiag[j] = d;
        s->Uptr[j] = (int)nu;
        int nc = 0;
        for (int cc = j + 1; cc < n; ++cc)
            if (prow[cc] != 0.0) { s->scratch_cols[nc++] = cc; s->Ucol[nu] = cc; s->Uval[nu] = prow[cc]; ++nu; }
        s->Lptr[j] = (int)nl;
        for (int i = j + 1; i < n; ++i) {
            double* row = W + (size_t)i * n;
            if (row[j] == 0.0) continue;
            double m = row[j] / d;
            row[j] = 0.0;
            for (int z = 0; z < nc; ++z) row[s->scratch_cols[z]] -= m * prow[s->scratch_cols[z]];
            s->Lrow[nl] = i; s->Lval[nl] = m; ++nl;
        }
    }
    s->Lptr[n] = (int)nl;
    s->Uptr[n] = (int)nu;
}

/* solve W x = b (b, x in original ordering; may alias) */
static void splu_solve(const splu_t* s, const double* b, double* x, double* tmp) {
    int n = s->n;
    for (int i = 0; i < n; ++i) tmp[i] = b[s->ord[i]];
    for (int j = 0; j < n; ++j) {
        int pr = s->piv[j];
        if (pr != j) { double t0 = tmp[j]; tmp[j] = tmp[pr]; tmp[pr] = t0; }
        double bj = tmp[j];
        for (int e = s->Lptr[j]; e < s->Lptr[j + 1]; ++e) tmp[s->Lrow[e]] -= s->Lval[e] * bj;
    }
    for (int j = n - 1; j >= 0; --j) {
        double acc = tmp[j];
        for (int e = s->Uptr[j]; e < s->Uptr[j + 1]; ++e) acc -= s->Uval[e] * tmp[s->Ucol[e]];
        tmp[j] = acc / s->Udiag[j];
    }
    for (int i = 0; i < n; ++i) x[s->ord[i]] = tmp[i];
}

/* -------------------------------------------------------------- KenCarp4 */
static const double KC_G = 0.25;
static const double KC_C[6] = {0.0, 0.5, 83.0 / 250.0, 31.0 / 50.0, 17.0 / 20.0, 1.0};
static const double KC_A[6][6] = {
    {0, 0, 0, 0, 0, 0},
    {0.25, 0, 0, 0, 0, 0},
    {8611.0 / 62500.0, -1743.0 / 31250.0, 0, 0, 0, 0},
    {5012029.0 / 34652500.0, -654441.0 / 2922500.0, 174375.0 / 388108.0, 0, 0, 0},
    {15267082809.0 / 155376265600.0, -71443401.0 / 120774400.0, 730878875.0 / 902184768.0,
     2285395.0 / 8070912.0, 0, 0},
    {82889.0 / 524892.0, 0.0, 15625.0 / 83664.0, 69875.0 / 102672.0, -2260.0 / 8211.0, 0}};
static const double KC_BH[6] = {4586570599.0 / 29645900160.0, 0.0, 178811875.0 / 945068544.0,
                                814220225.0 / 1159782912.0, -3700637.0 / 11593932.0, 61727.0 / 225920.0};

void orc_tableau(double* A, double* b, double* bh, double* c, double* gamma) {
    for (int i = 0; i < 6; ++i) {
        for (int j = 0; j < 6; ++j) A[i * 6 + j] = (i == j && i > 0) ? KC_G : KC_A[i][j];
        b[i] = (i < 5) ? KC_A[5][i] : KC_G;
        bh[i] = KC_BH[i];
        c[i] = KC_C[i];
    }
    *gamma = KC_G;
}

typedef struct obs_s {
    /* called once at the start (x0, u0) with h = 0 and after each accepted step */
    void (*fn)(struct obs_s* o, double xa, const double* ua, const double* fa, double xb, const double* ub,
               const double* fb, int n);
    void* data;
} obs_t;

static double wrms(const double* e, const double* u0, const double* u1, int n, double rt, double at) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) {
        double sc = at + rt * fmax(fabs(u0[i]), u1 ? fabs(u1[i]) : fabs(u0[i]));
        double r = e[i] / sc;
        s += r * r;
    }
    return sqrt(s / n);
}

/* KenCarp4 with Shampine's smoothed error estimate and OrdinaryDiffEq's step
 * control (perturbations.jl:25-33).  replay_xs (test infrastructure, may be
 * NULL): instead of choosing steps, take exactly the accepted steps whose end
 * points are replay_xs[0..n_replay-1] (as recorded by orc_boltsolve_steps) and
 * accept each; the discrete solution is then a smooth function of the inputs,
 * so central differences of it give the derivative of the fixed-step map. */
static int integrate(const orc_params* p, const orc_tables* t, const orc_cfg* c, double k, obs_t* obs,
                     orc_stats* st, const double* replay_xs, int n_replay) {
    lay_t L = layout(c);
    if (cfg_bad(c)) {
        memset(st, 0, sizeof(*st));
        st->status = 4;
        return 4;
    }
    int n = L.n;
    double rt = c->reltol, at = c->abstol;
    int maxsteps = c->max_steps > 0 ? c->max_steps : 100000;
    double* mem = (double*)malloc(sizeof(double) * n * 16);
    double *u = mem, *f = mem + n, *un = mem + 2 * n, *fn = mem + 3 * n, *psi = mem + 4 * n, *rhs = mem + 5 * n;
    double *err = mem + 6 * n, *tmp = mem + 7 * n, *Z[6], *ycor = mem + 14 * n, *r2 = mem + 15 * n;
    for (int i = 0; i < 6; ++i) Z[i] = mem + (8 + i) * n;
    splu_t lu;
    splu_init(&lu, &L);
    memset(st, 0, sizeof(*st));
    double x = t->x0, xend = t->x1;
    orc_initial_conditions(p, t, c, k, u);
    orc_rhs(p, t, c, k, x, u, f, 1);
    st->rhs_evals++;
    obs->fn(obs, x, u, f, x, u, f, n);
    /* Hairer's starting step (order 4) */
    double d0 = wrms(u, u, NULL, n, rt, at), d1 = wrms(f, u, NULL, n, rt, at);
    double h0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * d0 / d1;
    if (h0 > xend - x) h0 = xend - x;
    for (int i = 0; i < n; ++i) psi[i] = u[i] + h0 * f[i];
    orc_rhs(p, t, c, k, x + h0, psi, rhs, 0);
    st->rhs_evals++;
    for (int i = 0; i < n; ++i) tmp[i] = (rhs[i] - f[i]);
    double d2 = wrms(tmp, u, NULL, n, rt, at) / h0;
    double dm = fmax(d1, d2);
    double h1 = (dm <= 1e-15) ? fmax(1e-6, h0 * 1e-3) : pow(0.01 / dm, 1.0 / 5.0);
    double h = fmin(100.0 * h0, h1);
    if (h > xend - x) h = xend - x;
    int status = 0;
    long nsteps = 0;
    while (x < xend) {
        if (replay_xs) {
            if (nsteps >= n_replay) { status = 1; break; }
            h = replay_xs[nsteps] - x;
        }
        if (++nsteps > maxsteps) { status = 1; break; }
        int last = 0;
        if (h >= xend - x) { h = xend - x; last = 1; }
        for (int i = 0; i < n; ++i) Z[0][i] = h * f[i];
        for (int s = 1; s < 6; ++s) {
            for (int i = 0; i < n; ++i) {
                double acc = u[i];
                for (int j = 0; j < s; ++j) acc += KC_A[s][j] * Z[j][i];
                psi[i] = acc;
            }
            double xs = x + KC_C[s] * h;
            orc_rhs(p, t, c, k, xs, psi, rhs, 0);
            st->rhs_evals++;
            for (int i = 0; i < n; ++i) rhs[i] *= h;
            splu_factor(&lu, p, t, c, k, xs, h * KC_G);
            st->factorizations++;
            splu_solve(&lu, rhs, Z[s], tmp);
            if (c->lu_refine) {
                /* residual of z = h f(psi + γ z): r = h f(psi + γ z) - z */
                for (int i = 0; i < n; ++i) ycor[i] = psi[i] + KC_G * Z[s][i];
                orc_rhs(p, t, c, k, xs, ycor, r2, 0);
                st->rhs_evals++;
                for (int i = 0; i < n; ++i) r2[i] = h * r2[i] - Z[s][i];
                splu_solve(&lu, r2, r2, tmp);
                for (int i = 0; i < n; ++i) Z[s][i] += r2[i];
            }
        }
        for (int i = 0; i < n; ++i) {
            double acc = u[i], e = 0.0;
            for (int j = 0; j < 5; ++j) acc += KC_A[5][j] * Z[j][i];
            un[i] = acc + KC_G * Z[5][i];
            for (int j = 0; j < 6; ++j) e += ((j < 5 ? KC_A[5][j] : KC_G) - KC_BH[j]) * Z[j][i];
            err[i] = e;
        }
        /* Shampine's smoothed estimate, KenCarp4's default (smooth_est):
         * est = W⁻¹ Σ (b_j - b̂_j) z_j with W of the last stage (x + h) */
        splu_solve(&lu, err, err, tmp);
        double en = replay_xs ? 0.0 : wrms(err, u, un, n, rt, at);
        if (!isfinite(en)) {
            st->rejected++;
            h *= 0.2;
            if (h < 1e-14) { status = 3; break; }
            continue;
        }
        if (en <= 1.0) {
            double xn = replay_xs ? replay_xs[nsteps - 1] : (last ? xend : x + h);
            orc_rhs(p, t, c, k, xn, un, fn, 1); /* FSAL: RSA overwrite persists in un */
            st->rhs_evals++;
            st->accepted++;
            obs->fn(obs, x, u, f, xn, un, fn, n);
            memcpy(u, un, sizeof(double) * n);
            memcpy(f, fn, sizeof(double) * n);
            x = xn;
            double fac = (en == 0.0) ? 10.0 : 0.9 * pow(en, -0.25);
            if (fac > 10.0) fac = 10.0;
            if (fac < 0.2) fac = 0.2;
            h *= fac;
        } else {
            st->rejected++;
            double fac = 0.9 * pow(en, -0.25);
            if (fac < 0.2) fac = 0.2;
            if (fac > 1.0) fac = 1.0;
            h *= fac;
            if (h < 1e-14 * fmax(1.0, fabs(x))) { status = 2; break; }
        }
    }
    st->status = status;
    splu_free(&lu);
    free(mem);
    return status;
}

/* cubic Hermite on [xa, xb] (the interpolant used for sol(x)) */
static void hermite(double xa, const double* ua, const double* fa, double xb, const double* ub, const double* fb,
                    int n, double x, double* out) {
    double h = xb - xa;
    if (h == 0.0) { memcpy(out, ua, sizeof(double) * n); return; }
    double th = (x - xa) / h;
    if (th == 1.0) { memcpy(out, ub, sizeof(double) * n); return; }
    for (int i = 0; i < n; ++i) {
        double d = ub[i] - ua[i];
        out[i] = (1.0 - th) * ua[i] + th * ub[i] +
                 th * (th - 1.0) * ((1.0 - 2.0 * th) * d + (th - 1.0) * h * fa[i] + th * h * fb[i]);
    }
}

/* ---------------------------------------------------------------- drivers */
typedef struct {
    const double* xs;
    int ns, next;
    double* out;
    /* source mode */
    const orc_params* p;
    const orc_tables* t;
    const orc_cfg* c;
    double k;
    int mode; /* 0 save u; bit 1: temperature source into out, bit 2: polarization source into out2 */
    double* work;
    double* out2;
} save_t;

static void save_obs(obs_t* o, double xa, const double* ua, const double* fa, double xb, const double* ub,
                     const double* fb, int n) {
    save_t* s = (save_t*)o->data;
    while (s->next < s->ns) {
        double xv = s->xs[s->next];
        int take = (xb == xa) ? (xv <= xa) : (xv <= xb);
        if (!take) break;
        double* dst = (s->mode == 0) ? s->out + (size_t)s->next * n : s->work;
        hermite(xa, ua, fa, xb, ub, fb, n, xv, dst);
        if (s->mode != 0) {
            double* du = s->work + n;
            orc_rhs(s->p, s->t, s->c, s->k, xv, dst, du, 1); /* spectra.jl:16, 36 mutate u */
            if (s->mode & 1) s->out[s->next] = orc_source(s->p, s->t, s->c, s->k, xv, dst, du);
            if (s->mode & 2) s->out2[s->next] = orc_source_P(s->p, s->t, s->c, s->k, xv, dst);
        }
        s->next++;
    }
}

int orc_boltsolve(const orc_params* p, const orc_tables* t, const orc_cfg* c, double k, const double* x_save,
                  int nsave, double* out, orc_stats* st) {
    save_t s = {x_save, nsave, 0, out, p, t, c, k, 0, NULL, NULL};
    obs_t o = {save_obs, &s};
    return integrate(p, t, c, k, &o, st, NULL, 0);
}

typedef struct { double* xs; int max, cnt; } steplog_t;
static void step_obs(obs_t* o, double xa, const double* ua, const double* fa, double xb, const double* ub,
                     const double* fb, int n) {
    (void)xa; (void)ua; (void)fa; (void)ub; (void)fb; (void)n;
    steplog_t* s = (steplog_t*)o->data;
    if (s->cnt < s->max) s->xs[s->cnt] = xb;
    s->cnt++;
}

int orc_boltsolve_steps(const orc_params* p, const orc_tables* t, const orc_cfg* c, double k, double* xs,
                        int max_xs, int* n_xs, orc_stats* st) {
    steplog_t s = {xs, max_xs, 0};
    obs_t o = {step_obs, &s};
    int r = integrate(p, t, c, k, &o, st, NULL, 0);
    *n_xs = s.cnt;
    return r;
}

/* one solve, temperature (spectra.jl:6-23) and/or polarization (spectra.jl:25-42)
 * source columns at every x node; either output may be NULL */
static int source_column_run(const orc_params* p, const orc_tables* t, const orc_cfg* c, double k, double* S_T,
                             double* S_P, orc_stats* st, const double* replay_xs, int n_replay) {
    int n = orc_state_length(c);
    double* xs = (double*)malloc(sizeof(double) * t->nx);
    double* work = (double*)malloc(sizeof(double) * 2 * n);
    for (int j = 0; j < t->nx; ++j) xs[j] = xnode(t, j);
    save_t s = {xs, t->nx, 0, S_T, p, t, c, k, (S_T ? 1 : 0) | (S_P ? 2 : 0), work, S_P};
    obs_t o = {save_obs, &s};
    int r = integrate(p, t, c, k, &o, st, replay_xs, n_replay);
    free(xs);
    free(work);
    return r;
}

int orc_source_column_tp(const orc_params* p, const orc_tables* t, const orc_cfg* c, double k, double* S_T,
                         double* S_P, orc_stats* st) {
    return source_column_run(p, t, c, k, S_T, S_P, st, NULL, 0);
}

int orc_source_column_replay(const orc_params* p, const orc_tables* t, const orc_cfg* c, double k,
                             const double* step_xs, int n_steps, double* S, orc_stats* st) {
    return source_column_run(p, t, c, k, S, NULL, st, step_xs, n_steps);
}

int orc_source_column(const orc_params* p, const orc_tables* t, const orc_cfg* c, double k, double* S,
                      orc_stats* st) {
    return orc_source_column_tp(p, t, c, k, S, NULL, st);
}

int orc_source_grid(const orc_params* p, const orc_tables* t, const orc_cfg* c, const double* k, int nk,
                    double* S, int* status, int nthreads) {
    int bad = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : bad)
#endif
    for (int i = 0; i < nk; ++i) {
        orc_stats st;
        int r = orc_source_column(p, t, c, k[i], S + (size_t)i * t->nx, &st);
        if (status) status[i] = r;
        if (r) bad++;
    }
    (void)nthreads;
    return bad;
}

/* rsa_perts! (perturbations.jl:35-84) */
static void rsa_perts(const orc_params* p, const orc_tables* t, const orc_cfg* c, double k, double x, double* u) {
    lay_t L = layout(c);
    bgx_t b = bgx(t, x);
    double Psi, Phip;
    metric(p, t, &L, k, &b, u, &Psi, &Phip);
    double Hc = b.Hc, tp = b.tp, tpp = b.tpp, csb2 = b.csb2;
    double Phi = u[L.iPhi], db = u[L.idb], vb = u[L.ivb];
    double T0 = Phi - Hc / k * tp * vb;
    double T1 = Hc / k * (-2.0 * Phip + tp * (Phi - csb2 * db) + Hc / k * (tpp - tp) * vb);
    u[0] = T0; u[1] = T1; u[2] = 0.0;
    u[L.iN + 0] = Phi; u[L.iN + 1] = -2.0 * Hc / k * Phip; u[L.iN + 2] = 0.0;
    /* 1-based u[ℓ] and u[(ℓᵧ+1)+ℓ] for ℓ = 3..ℓᵧ, u[2(ℓᵧ+1)+ℓ] for ℓ = 3..ℓ_ν */
    for (int l = 3; l <= L.ng; ++l) { u[l - 1] = 0.0; u[(L.ng + 1) + l - 1] = 0.0; }
    for (int l = 3; l <= L.nn; ++l) u[2 * (L.ng + 1) + l - 1] = 0.0;
}

int orc_boltsolve_rsa(const orc_params* p, const orc_tables* t, const orc_cfg* c, double k, double* out,
                      orc_stats* st) {
    int n = orc_state_length(c), nx = t->nx;
    double* xs = (double*)malloc(sizeof(double) * nx);
    for (int j = 0; j < nx; ++j) xs[j] = xnode(t, j);
    int r = orc_boltsolve(p, t, c, k, xs, nx, out, st);
    int ihor = -1, iod = -1;
    for (int j = 0; j < nx; ++j) {
        double eta = orc_spline(t, T_ETA, xs[j], 0);
        double Hc = orc_spline(t, T_HC, xs[j], 0);
        double tp = orc_spline(t, T_TAU_P, xs[j], 0);
        if (ihor < 0 && k * eta > 240.0) ihor = j;
        if (iod < 0 && -tp * Hc / eta > 100.0) iod = j;
    }
    if (ihor < 0) ihor = nx - 1;
    if (iod < 0) iod = nx - 1;
    int isw = ihor > iod ? ihor : iod;
    double xsw = xs[isw];
    for (int j = 0; j < nx; ++j)
        if (xs[j] > xsw) rsa_perts(p, t, c, k, xs[j], out + (size_t)j * n);
    free(xs);
    return r;
}

/* Ω_ν of the massive neutrino in plin's matter sum (spectra.jl:182-186):
 * νfac = 90ζ(3)/(11π⁴) · Ω_r h²/T_γ · (N_ν/3)^(3/4), Ω_ν = Σm_ν νfac / h² */
double orc_omega_nu_planted(const orc_params* p, const orc_tables* t, int plant) {
    double Tg = pow(15.0 / (M_PI * M_PI) * t->rho_crit * p->Omega_r, 0.25);
    double zeta = 1.2020569;
    double nufac = (90.0 * zeta / (11.0 * pow(M_PI, 4))) * (p->Omega_r * p->h * p->h / Tg) * pow(p->N_nu / 3.0, 0.75);
    if (plant & 2) nufac *= 1.1;
    return p->Sum_m_nu * nufac / (p->h * p->h);
}

double orc_omega_nu(const orc_params* p, const orc_tables* t) { return orc_omega_nu_planted(p, t, 0); }

/* plin (spectra.jl:163-198): one solve to x, the massive-ν density and velocity
 * moments normalised by ρ₀ℳ(x), the comoving-gauge densities δ − 3ℋv/k, their
 * Ω-weighted sum, and P = 2π²/k³ δ_m² A (k/0.05)^(n−1) */
double orc_plin_planted(const orc_params* p, const orc_tables* t, const orc_cfg* c, double k, double x,
                        orc_stats* st, int plant) {
    lay_t L = layout(c);
    if (cfg_bad(c)) {
        if (st) { memset(st, 0, sizeof(*st)); st->status = 4; }
        return NAN;
    }
    double* u = (double*)malloc(sizeof(double) * L.n);
    orc_boltsolve(p, t, c, k, &x, 1, u, st);
    double a = exp(x);
    double rho0 = orc_spline(t, T_RHO0M, x, 0);
    double M0[ORC_MAXQ], M1[ORC_MAXQ], M2[ORC_MAXQ];
    for (int i = 0; i < L.nq; ++i) {
        M0[i] = u[L.iM + i]; M1[i] = u[L.iM + L.nq + i]; M2[i] = u[L.iM + 2 * L.nq + i];
    }
    double rho, sig;
    orc_rho_sigma(p, t, L.nq, a, M0, M2, &rho, &sig);
    double Mrho = rho / rho0;
    double Mtheta = k * orc_theta(p, t, L.nq, a, M1) / rho0 * ((plant & 8) ? 2.0 : 1.0);
    double dcN = u[L.idel], dbN = u[L.idb], vcN = u[L.iv], vbN = u[L.ivb];
    double vmN = -Mtheta / k;
    double Onu = orc_omega_nu_planted(p, t, plant);
    double Om = p->Omega_c + p->Omega_b + Onu;
    double Hc = orc_spline(t, T_HC, x, 0);
    double gs = (plant & 1) ? -3.0 : 3.0;
    double dc = dcN - gs * Hc * vcN / k;
    double db = dbN - gs * Hc * vbN / k;
    double dmn = Mrho - gs * Hc * vmN / k;
    double dm = (p->Omega_c * dc + p->Omega_b * db + Onu * dmn) / Om;
    double Pprim = p->A * pow(k / 0.05, p->n - 1.0);
    free(u);
    return (((plant & 4) ? 1.0 : 2.0) * M_PI * M_PI / (k * k * k)) * dm * dm * Pprim;
}

double orc_plin(const orc_params* p, const orc_tables* t, const orc_cfg* c, double k, double x, orc_stats* st) {
    return orc_plin_planted(p, t, c, k, x, st, 0);
}

/* ------------------------------------------------------- spherical Bessel */
void orc_sph_bessel(int lmax, double z, double* j) {
    for (int l = 0; l <= lmax; ++l) j[l] = 0.0;
    if (z == 0.0) { j[0] = 1.0; return; }
    double zm = fmax((double)lmax, z);
    int N = (int)(zm + 60.0 + 12.0 * cbrt(zm));
    /* exact low orders */
    double j0, j1;
    if (z < 1e-3) {
        double z2 = z * z;
        j0 = 1.0 - z2 / 6.0 + z2 * z2 / 120.0;
        j1 = z / 3.0 - z * z2 / 30.0 + z * z2 * z2 / 840.0;
    } else {
        double s = sin(z), co = cos(z);
        j0 = s / z;
        j1 = (s / z - co) / z;
    }
    /* Miller: f_{l-1} = (2l+1)/z f_l - f_{l+1} from f_{N+1} = 0, f_N = tiny */
    double fp = 0.0, fc = 1e-300;
    double inv = 1.0 / z;
    for (int l = N; l >= 1; --l) {
        double fm = (2 * l + 1) * inv * fc - fp;
        fp = fc;
        fc = fm;
        /* fc = f_{l-1}, fp = f_l */
        if (l - 1 <= lmax) j[l - 1] = fc;
        if (l <= lmax) j[l] = fp;
        if (fabs(fc) > 1e250) {
            fc *= 1e-250;
            fp *= 1e-250;
            int top = (N < lmax) ? N : lmax;
            for (int q = l - 1; q <= top; ++q) if (q >= 0) j[q] *= 1e-250;
        }
    }
    double scale = (fabs(j0) >= fabs(j1)) ? j0 / j[0] : j1 / j[1];
    for (int l = 0; l <= lmax; ++l) j[l] *= scale;
}

/* bessel_interpolator (spectra.jl:49-58): natural cubic spline of j_ℓ on
 * 0:Δ:zmax with Δ = zmax/5000 */
static void prefilter(const double* y, int n, double* c) {
    c[1] = y[0];
    c[n] = y[n - 1];
    int m = n - 2;
    double* cp = (double*)malloc(sizeof(double) * m);
    double* d = (double*)malloc(sizeof(double) * m);
    for (int i = 0; i < m; ++i) d[i] = 6.0 * y[i + 1];
    d[0] -= c[1];
    d[m - 1] -= c[n];
    cp[0] = 0.25;
    d[0] /= 4.0;
    for (int i = 1; i < m; ++i) {
        double den = 4.0 - cp[i - 1];
        cp[i] = 1.0 / den;
        d[i] = (d[i] - d[i - 1]) / den;
    }
    for (int i = m - 2; i >= 0; --i) d[i] -= cp[i] * d[i + 1];
    for (int i = 0; i < m; ++i) c[i + 2] = d[i];
    c[0] = 2.0 * c[1] - c[2];
    c[n + 1] = 2.0 * c[n] - c[n - 1];
    free(cp);
    free(d);
}

void orc_bessel_table(int ell, double zmax, double* coef) {
    const int nn = 5001;
    double dz = zmax / 5000.0;
    double* y = (double*)malloc(sizeof(double) * nn);
    double* jb = (double*)malloc(sizeof(double) * (ell + 1));
    for (int m = 0; m < nn; ++m) {
        orc_sph_bessel(ell, m * dz, jb);
        y[m] = jb[ell];
    }
    prefilter(y, nn, coef);
    free(y);
    free(jb);
}

double orc_bessel_table_eval(const double* c, double zmax, double z) {
    const int n = 5001;
    double dz = zmax / 5000.0;
    double tt = z / dz;
    int i = (int)floor(tt);
    if (i < 0) i = 0;
    if (i > n - 2) i = n - 2;
    double s = tt - i, r = 1.0 - s;
    const double* q = c + i;
    return r * r * r / 6.0 * q[0] + (2.0 / 3.0 - s * s + 0.5 * s * s * s) * q[1] +
           (2.0 / 3.0 - r * r + 0.5 * r * r * r) * q[2] + s * s * s / 6.0 * q[3];
}

/* ---------------------------------------------------- cltt / clte / clee */
/* C_ℓ^TT (spectra.jl:84-96), C_ℓ^TE (spectra.jl:98-114), C_ℓ^EE (116-130):
 * T_ℓ(k) = Σ_x j_ℓ(k(η₀−η(x))) S(x,k) Δx for each source (Tl, spectra.jl:70-82),
 * E-mode terms times the spin factor ℓð = √((ℓ+2)(ℓ+1)ℓ(ℓ−1)), summed over the
 * dense-k midpoints with A(k/0.05)^(n−1) dk/k and 4π.  S_P may be NULL (TT
 * only); any output may be NULL. */
int orc_cl(const orc_params* p, const orc_tables* t, const double* S, const double* S_P, const double* k_src,
           int nk_src, const double* kgrid, int nkd, const int* ells, int nell, int bessel_mode, double* cl,
           double* clte, double* clee, int nthreads, int kbeg, int kend) {
    int nx = t->nx;
    int lmax = 0;
    for (int e = 0; e < nell; ++e) if (ells[e] > lmax) lmax = ells[e];
    int i0 = -1;
    for (int j = 0; j < nx; ++j) if (xnode(t, j) > -8.0) { i0 = j; break; }
    if (i0 < 0) return -1;
    int nxi = nx - 1 - i0;
    double eta0 = t->eta0;
    double* etaj = (double*)malloc(sizeof(double) * nxi);
    double* dxj = (double*)malloc(sizeof(double) * nxi);
    for (int j = 0; j < nxi; ++j) {
        double xa = xnode(t, i0 + j), xb = xnode(t, i0 + j + 1);
        etaj[j] = orc_spline(t, T_ETA, xa, 0);
        dxj[j] = xb - xa;
    }
    double zmax = kgrid[nkd - 1] * eta0;
    double* tabs = NULL;
    if (bessel_mode == 1) {
        tabs = (double*)malloc(sizeof(double) * 5003 * (size_t)nell);
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads > 0 ? nthreads : 1)
#endif
        for (int e = 0; e < nell; ++e) orc_bessel_table(ells[e], zmax, tabs + 5003 * (size_t)e);
    }
    if (kbeg < 0) kbeg = 0;
    if (kend > nkd - 1) kend = nkd - 1;
    int nkm = kend - kbeg;
    const int nout = 3;
    double* contrib = (double*)calloc((size_t)(nkm > 0 ? nkm : 1) * nell * nout, sizeof(double));
#ifdef _OPENMP
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1)
#endif
    {
        double* jl = (double*)malloc(sizeof(double) * (lmax + 1));
        double* th = (double*)malloc(sizeof(double) * nell);
        double* te = (double*)malloc(sizeof(double) * nell);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 8)
#endif
        for (int ik = kbeg; ik < kend; ++ik) {
            double k = (kgrid[ik] + kgrid[ik + 1]) / 2.0;
            double dk = kgrid[ik + 1] - kgrid[ik];
            int ia = 0;
            while (ia < nk_src - 2 && k >= k_src[ia + 1]) ++ia;
            double tt = (k - k_src[ia]) / (k_src[ia + 1] - k_src[ia]);
            const double* Sa = S ? S + (size_t)ia * nx : NULL;
            const double* Sb = S ? S + (size_t)(ia + 1) * nx : NULL;
            const double* Pa = S_P ? S_P + (size_t)ia * nx : NULL;
            const double* Pb = S_P ? S_P + (size_t)(ia + 1) * nx : NULL;
            for (int e = 0; e < nell; ++e) th[e] = te[e] = 0.0;
            for (int j = 0; j < nxi; ++j) {
                int jx = i0 + j;
                double s = Sa ? (1.0 - tt) * Sa[jx] + tt * Sb[jx] : 0.0;
                double sp = Pa ? (1.0 - tt) * Pa[jx] + tt * Pb[jx] : 0.0;
                double z = k * (eta0 - etaj[j]);
                if (bessel_mode == 0) orc_sph_bessel(lmax, z, jl);
                for (int e = 0; e < nell; ++e) {
                    double jv = bessel_mode == 0 ? jl[ells[e]] : orc_bessel_table_eval(tabs + 5003 * (size_t)e, zmax, z);
                    th[e] += jv * s * dxj[j];
                    te[e] += jv * sp * dxj[j];
                }
            }
            double Pprim = p->A * pow(k / 0.05, p->n - 1.0);
            for (int e = 0; e < nell; ++e) {
                long long l = ells[e];
                double lf = sqrt((double)((l + 2) * (l + 1) * l * (l - 1)));
                double ep = te[e] * lf;
                double* cc = contrib + ((size_t)(ik - kbeg) * nell + e) * nout;
                cc[0] = th[e] * th[e] * Pprim * dk / k;
                cc[1] = th[e] * ep * Pprim * dk / k;
                cc[2] = ep * ep * Pprim * dk / k;
            }
        }
        free(jl);
        free(th);
        free(te);
    }
    for (int e = 0; e < nell; ++e) {
        double acc[3] = {0.0, 0.0, 0.0};
        for (int ik = 0; ik < nkm; ++ik)
            for (int o = 0; o < nout; ++o) acc[o] += contrib[((size_t)ik * nell + e) * nout + o];
        if (cl) cl[e] = 4.0 * M_PI * acc[0];
        if (clte) clte[e] = 4.0 * M_PI * acc[1];
        if (clee) clee[e] = 4.0 * M_PI * acc[2];
    }
    free(contrib);
    free(etaj);
    free(dxj);
    free(tabs);
    return 0;
}

int orc_cltt(const orc_params* p, const orc_tables* t, const double* S, const double* k_src, int nk_src,
             const double* kgrid, int nkd, const int* ells, int nell, int bessel_mode, double* cl, int nthreads,
             int kbeg, int kend) {
    return orc_cl(p, t, S, NULL, k_src, nk_src, kgrid, nkd, ells, nell, bessel_mode, cl, NULL, NULL, nthreads, kbeg,
                  kend);
}
