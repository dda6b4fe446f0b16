/*
 * Moments of spherical Bessel functions, their interpolation tables and the
 * 3rd/4th-order Filon rule -- TEST INFRASTRUCTURE ONLY (the checker of the
 * device path; see bolt_oracle.h).
 *
 * Restates Bolt.jl's src/bessel/moments.jl, interpolator.jl and integrator.jl:
 *   I_m(x) = ∫_0^x t^m j_ν(t) dt  (ν a non-negative integer, m real > −ν − 1)
 * - Lommel-function asymptotic for large x (moments.jl:5-21, 77-86,
 *   sph_j_moment_asymp): prefactor + x√(π/2)[(m+ν−1) J_{ν+½}(x) s⁽²⁾(x, m−2, ν−½)
 *   − J_{ν−½}(x) s⁽²⁾(x, m−1, ν+½)], with s⁽²⁾'s own truncation (at most 21
 *   terms, stop once a term is below 1e-16 of the sum);
 * - the Maclaurin series of ₁F₂ (moments.jl:97-102, weniger.jl:237-249) in
 *   FP64, which the tables use below their range;
 * - for the table nodes below x = 50 the reference sums ₁F₂ by Weniger's
 *   sequence transformation in Double64 (moments.jl:89-94, weniger.jl).  Here
 *   the same Maclaurin series is summed in __float128 instead, below x = 40:
 *   its terms reach about e^x/√(2πx)·x^{ν+m+1}/(2ν+1)!! against the moment,
 *   which 113 bits absorb to ~2e-16 (measured against 30-digit quadrature for
 *   ν, m ≤ 3), while the asymptotic holds to a few 1e-16 from x = 40 on for
 *   ν ≤ 3 (but the series only to ~1e-13 at x = 50).  The reference's known
 *   answers (test/testbessel.jl) pin the values.
 * - tables (interpolator.jl:85-111): nodes on LinRange(x_min, x_max, N),
 *   asymptotic at nodes ≥ 40, cubic B-spline with Line(OnGrid()) ends
 *   (natural), scaled to the nodes; below the range the FP64 Maclaurin series,
 *   above it the asymptotic (interpolator.jl:60-68);
 * - Filon (integrator.jl:5-18): a quadratic through f, f′, f″ at a, integrated
 *   against j_ν(kt) over (a, b) with the moments' differences.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "bolt_oracle.h"

#define SQRT_HALF_PI 1.2533141373155002512078826424055

/* j_0..j_n at x > 0 by upward recurrence (x ≥ n here: stable) */
static void sphj_up(int n, double x, double* j) {
    const double s = sin(x), c = cos(x);
    j[0] = s / x;
    if (n >= 1) j[1] = s / (x * x) - c / x;
    for (int l = 1; l < n; ++l) j[l + 1] = (2 * l + 1) / x * j[l] - j[l - 1];
}

/* s⁽²⁾(t, α − ½, ν) of moments.jl:5-21 */
static double s2(double t, double alpha_minus_half, double nu) {
    double s = 1.0, sk = 1.0;
    const double ti = 1.0 / t, ti2 = ti * ti, nu2 = nu * nu, am1 = alpha_minus_half - 0.5;
    for (int j = 0; j <= 20; ++j) {
        const double d = am1 - 2.0 * j;
        sk *= (nu2 - d * d) * ti2;
        s += sk;
        if (fabs(sk) < 1e-16 * fabs(s)) break;
    }
    return s * pow(t, alpha_minus_half) * sqrt(ti);
}

/* J_moment_asymp_prefactor(ν, α) · √(π/2) at ν + ½, m − ½ (moments.jl:35-36, 72-73) */
double orc_sphj_moment_asymp_prefactor(int nu, double m) {
    const double nuh = nu + 0.5, al = m - 0.5;
    return exp(log(2.0) * al + lgamma((nuh + al + 1.0) / 2.0) - lgamma((nuh - al + 1.0) / 2.0)) * SQRT_HALF_PI;
}

/* sph_j_moment_asymp (moments.jl:76-86), ν ≥ 1 */
double orc_sphj_moment_asymp(double x, int nu, double m) {
    double j[64];
    sphj_up(nu, x, j);
    /* J_{ν±½}(x) = √(2x/π) j_{ν}, j_{ν−1} */
    const double c = sqrt(x) / SQRT_HALF_PI;
    const double Jp = c * j[nu], Jm = c * j[nu - 1];
    return orc_sphj_moment_asymp_prefactor(nu, m) +
           x * SQRT_HALF_PI * ((m + nu - 1.0) * Jp * s2(x, m - 2.0, nu - 0.5) - Jm * s2(x, m - 1.0, nu + 0.5));
}

/* maclaurin_₁F₂ (weniger.jl:237-249) in FP64 */
static double maclaurin_1f2(double a, double b1, double b2, double z) {
    double S0 = 1.0, S1 = 1.0 + a * z / (b1 * b2);
    int j = 1;
    for (;;) {
        const int more = isfinite(S0) && isfinite(S1) && fabs(S0 - S1) > fmax(fabs(S0), fabs(S1)) * 10.0 * 2.220446049250313e-16;
        if (!(more || j <= 1)) break;
        double r = 1.0 / (j + 1.0);
        r *= a + j;
        r /= (b1 + j) * (b2 + j);
        const double Sn = S1 + (S1 - S0) * r * z;
        S0 = S1;
        S1 = Sn;
        ++j;
    }
    return S1;
}

/* sph_j_moment_maclaurin_₁F₂ (moments.jl:97-102) in FP64 */
double orc_sphj_moment_maclaurin(double x, int nu, double m) {
    const double nup = nu + 1.5;
    return (SQRT_HALF_PI / (m + nu + 1.0)) *
           exp((m + nu + 1.0) * log(x) - (nup - 1.0) * log(2.0) - lgamma(nup)) *
           maclaurin_1f2((1.0 + m + nu) / 2.0, (3.0 + m + nu) / 2.0, nup, -x * x / 4.0);
}

/* the same moment with the series summed in __float128:
 * I_m(x) = x^{ν+m+1}/(2ν+1)!! Σ_k (−x²/2)^k / (k! (2ν+3)(2ν+5)…(2ν+2k+1) (2k+ν+m+1)) */
double orc_sphj_moment_maclaurin_q(double x, int nu, double m) {
    if (x == 0.0) return 0.0;
    typedef __float128 Q;
    const Q z = -(Q)x * (Q)x / 2;
    Q t = 1, s = 0;
    const Q a = (Q)nu + (Q)m + 1;
    for (int k = 0; k < 4000; ++k) {
        const Q term = t / (a + 2 * k);
        s += term;
        Q at = term < 0 ? -term : term, as = s < 0 ? -s : s;
        if (k > 2 && at < as * (Q)1e-34) break;
        t = t * z / ((Q)(k + 1) * (Q)(2 * nu + 2 * k + 3));
    }
    /* x^{ν+m+1}/(2ν+1)!!, the double factorial in Q */
    Q df = 1;
    for (int i = 3; i <= 2 * nu + 1; i += 2) df *= i;
    const double pw = pow(x, nu + m + 1.0);
    return (double)((Q)pw / df * s);
}

/* ∫_0^x t^α J_ν(t) dt for half-integer ν = n + ½ (moments.jl:24-49 at ν = n + ½):
 * √(2/π) ∫_0^x t^{α+½} j_n(t) dt.  method 0: asymptotic, 1: extended-precision series */
double orc_J_moment_halfint(double x, int n, double alpha, int method) {
    const double m = alpha + 0.5;
    const double v = method == 0 ? orc_sphj_moment_asymp(x, n, m) : orc_sphj_moment_maclaurin_q(x, n, m);
    return v / SQRT_HALF_PI;
}

/* ----------------------------------------------------------------- tables */
/* Interpolations' BSpline(Cubic(Line(OnGrid()))) prefilter: c[0..n+1] */
static void prefilter_line(const double* y, int n, double* c, double* work) {
    c[1] = y[0];
    c[n] = y[n - 1];
    const int mm = n - 2;
    double* cp = work;
    double* d = work + mm;
    for (int i = 0; i < mm; ++i) d[i] = 6.0 * y[i + 1];
    d[0] -= c[1];
    d[mm - 1] -= c[n];
    cp[0] = 0.25;
    d[0] /= 4.0;
    for (int i = 1; i < mm; ++i) {
        const double den = 4.0 - cp[i - 1];
        cp[i] = 1.0 / den;
        d[i] = (d[i] - d[i - 1]) / den;
    }
    for (int i = mm - 2; i >= 0; --i) d[i] -= cp[i] * d[i + 1];
    for (int i = 0; i < mm; ++i) c[i + 2] = d[i];
    c[0] = 2.0 * c[1] - c[2];
    c[n + 1] = 2.0 * c[n] - c[n - 1];
}

/* node i of LinRange(x_min, x_max, n) */
double orc_linrange(double xmin, double xmax, int n, int i) {
    const double t = (double)i / (double)(n - 1);
    return (1.0 - t) * xmin + t * xmax;
}

int orc_moment_table_build(orc_moment_table* T, int nu, int order, double xmin, double xmax, int n, int nthreads) {
    /* the asymptotic from x = 40 on holds to double precision for ν ≤ 3 (the
     * reference builds tables for ν = 2 and 3, interpolator.jl:13-16) */
    if (nu < 1 || nu > 3 || order < 1 || order > 4 || n < 4 || !(xmax > xmin)) return 1;
    T->nu = nu;
    T->order = order;
    T->n = n;
    T->xmin = xmin;
    T->xmax = xmax;
    T->coef = (double*)malloc(sizeof(double) * (size_t)order * (n + 2));
    double* y = (double*)malloc(sizeof(double) * (size_t)order * n);
    if (!T->coef || !y) return 2;
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
    for (int i = 0; i < n; ++i) {
        const double x = orc_linrange(xmin, xmax, n, i);
        for (int m = 0; m < order; ++m)
            y[(size_t)m * n + i] = x < 40.0 ? orc_sphj_moment_maclaurin_q(x, nu, m) : orc_sphj_moment_asymp(x, nu, m);
    }
    double* work = (double*)malloc(sizeof(double) * 2 * (size_t)n);
    for (int m = 0; m < order; ++m) prefilter_line(y + (size_t)m * n, n, T->coef + (size_t)m * (n + 2), work);
    free(work);
    free(y);
    return 0;
}

void orc_moment_table_free(orc_moment_table* T) {
    free(T->coef);
    T->coef = NULL;
}

/* MomentTable(x) (interpolator.jl:60-68) */
void orc_moment_eval(const orc_moment_table* T, double x, double* out) {
    if (T->xmin <= x && x <= T->xmax) {
        const double dx = (T->xmax - T->xmin) / (T->n - 1);
        const double tt = (x - T->xmin) / dx;
        int i = (int)floor(tt);
        if (i < 0) i = 0;
        if (i > T->n - 2) i = T->n - 2;
        const double s = tt - i, r = 1.0 - s;
        const double w0 = r * r * r / 6.0, w1 = 2.0 / 3.0 - s * s + 0.5 * s * s * s,
                     w2 = 2.0 / 3.0 - r * r + 0.5 * r * r * r, w3 = s * s * s / 6.0;
        for (int m = 0; m < T->order; ++m) {
            const double* c = T->coef + (size_t)m * (T->n + 2) + i;
            out[m] = w0 * c[0] + w1 * c[1] + w2 * c[2] + w3 * c[3];
        }
    } else if (x > T->xmax) {
        for (int m = 0; m < T->order; ++m) out[m] = orc_sphj_moment_asymp(x, T->nu, m);
    } else {
        for (int m = 0; m < T->order; ++m) out[m] = orc_sphj_moment_maclaurin(x, T->nu, m);
    }
}

/* integrate_sph_bessel_filon (integrator.jl:5-18) */
double orc_filon(double f, double fp, double fpp, double k, double a, double b, const orc_moment_table* T) {
    double Ia[4], Ib[4];
    const double c2 = fpp / 2.0, afpp = a * fpp, c1 = fp - afpp, c0 = f - a * (fp - afpp / 2.0);
    const double ki = 1.0 / k, ki2 = ki * ki, ki3 = ki2 * ki;
    orc_moment_eval(T, k * b, Ib);
    orc_moment_eval(T, k * a, Ia);
    return (c0 * ki) * (Ib[0] - Ia[0]) + (c1 * ki2) * (Ib[1] - Ia[1]) + (c2 * ki3) * (Ib[2] - Ia[2]);
}

/* Σ over the intervals (t_i, t_{i+1}) of the Filon rule with f, f′, f″ at t_i
 * (the loop form, integrator.jl:20-38), for each k */
void orc_filon_sum(const orc_moment_table* T, const double* t, const double* f, const double* fp, const double* fpp,
                   int nt, const double* k, int nk, double* out, int nthreads) {
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
    for (int q = 0; q < nk; ++q) {
        double s = 0.0;
        for (int i = 0; i + 1 < nt; ++i) s += orc_filon(f[i], fp[i], fpp[i], k[q], t[i], t[i + 1], T);
        out[q] = s;
    }
}
