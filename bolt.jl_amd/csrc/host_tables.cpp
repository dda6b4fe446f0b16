// Host-side producers of the perturbation hot path's inputs, FP64 on the CPU:
// natural-unit constants, Gauss-Legendre nodes, the uniform cubic B-spline used
// for every table, Background and IonizationHistory(RECFAST).  The tables are
// what the device kernels (and the CPU oracle, in the tests) consume.
//
// Everything that depends on the cosmological parameters is a template over
// the scalar type R and is instantiated twice: R = double makes the tables,
// R = Dual makes the tables and their derivative along one parameter direction
// (forward-mode AD).  The Dual instantiation's values are bitwise the double
// instantiation's; the file is compiled without floating-point contraction so
// that fused multiply-adds cannot make the two differ.
//
// Citations are into Bolt.jl (src/...).
#include "host_tables.h"

#include <cmath>
#include <cstring>
#include <string>
#include <vector>
#include <algorithm>
#include <functional>
#include <type_traits>

namespace bolt {

// ----------------------------------------------------------------------------
// Physical constants (CODATA 2018 + IAU parsec) in Bolt's natural units:
// c = hbar = k_B = 1, lengths in Mpc (src/Bolt.jl:42-51, ionization.jl:5-28).
namespace si {
constexpr double c = 299792458.0;
constexpr double hbar = 1.054571817e-34;
constexpr double kB = 1.380649e-23;
constexpr double G = 6.67430e-11;
constexpr double eV = 1.602176634e-19;
constexpr double m_p = 1.67262192369e-27;
constexpr double sigma_T = 6.6524587321e-29;
constexpr double Mpc = 3.0856775814913673e22;  // 648000/pi au
}  // namespace si

const NaturalUnits& natural_units() {
    static const NaturalUnits u = [] {
        NaturalUnits n;
        n.km_s_Mpc_100 = 100.0e3 / si::c;                                   // Bolt.jl:49
        n.G_natural = si::G * si::hbar / (si::c * si::c * si::c) / (si::Mpc * si::Mpc);  // Bolt.jl:50
        n.mass_natural = si::eV / (si::hbar * si::c) * si::Mpc;             // Bolt.jl:51
        n.m_H = si::m_p * si::c / si::hbar * si::Mpc;                       // ionization.jl:26
        n.sigma_T = si::sigma_T / (si::Mpc * si::Mpc);                      // ionization.jl:28
        n.H0_conv = si::Mpc / si::c;                                        // recfast.jl:4 [s]
        n.K_conv = si::hbar * si::c / (si::kB * si::Mpc);                   // recfast.jl:5 [K]
        return n;
    }();
    return u;
}

// ----------------------------------------------------------------------------
// Gauss-Legendre nodes/weights on [-1,1], ascending (FastGaussQuadrature's
// gausslegendre(n), used at background.jl:105).  Newton iteration on P_n.
void gauss_legendre(int n, double* x, double* w) {
    for (int i = 0; i < n; ++i) {
        double z = std::cos(M_PI * (i + 0.75) / (n + 0.5));
        double pp = 0.0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = z;
            for (int k = 2; k <= n; ++k) {
                double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            double pn = (n == 1) ? z : p1;
            double pnm1 = (n == 1) ? 1.0 : p0;
            pp = n * (z * pn - pnm1) / (z * z - 1.0);
            double dz = pn / pp;
            z -= dz;
            if (std::fabs(dz) < 1e-16) {
                // one more evaluation for the derivative at the converged node
                p0 = 1.0; p1 = z;
                for (int k = 2; k <= n; ++k) {
                    double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                    p0 = p1; p1 = p2;
                }
                pn = (n == 1) ? z : p1; pnm1 = (n == 1) ? 1.0 : p0;
                pp = n * (z * pn - pnm1) / (z * z - 1.0);
                break;
            }
        }
        // nodes come out descending in i; store ascending
        x[n - 1 - i] = z;
        w[n - 1 - i] = 2.0 / ((1.0 - z * z) * pp * pp);
    }
}

// ----------------------------------------------------------------------------
// Uniform cubic B-spline with Line(OnGrid()) ends == natural cubic spline
// (util.jl:11-13 via Interpolations.jl BSpline(Cubic(Line(OnGrid())))).
// Coefficients c[0..n+1]; node j (0-based) sits between c[j] and c[j+2].
template <class R>
static void prefilter_t(const R* y, int n, R* c) {
    // End rows: c0 - 2c1 + c2 = 0 together with (c0 + 4c1 + c2)/6 = y0 give c1 = y0.
    c[1] = y[0];
    c[n] = y[n - 1];
    int m = n - 2;  // unknowns c[2..n-1]
    if (m > 0) {
        std::vector<double> cp(m);
        std::vector<R> d(m);
        for (int i = 0; i < m; ++i) d[i] = 6.0 * y[i + 1];
        d[0] -= c[1];
        d[m - 1] -= c[n];
        // Thomas algorithm, diagonal 4, off-diagonals 1
        double b = 4.0;
        cp[0] = 1.0 / b;
        d[0] = d[0] / b;
        for (int i = 1; i < m; ++i) {
            double den = 4.0 - cp[i - 1];
            cp[i] = 1.0 / den;
            d[i] = (d[i] - d[i - 1]) / den;
        }
        for (int i = m - 2; i >= 0; --i) d[i] -= cp[i] * d[i + 1];
        for (int i = 0; i < m; ++i) c[i + 2] = d[i];
    }
    c[0] = 2.0 * c[1] - c[2];
    c[n + 1] = 2.0 * c[n] - c[n - 1];
}

void spline_prefilter(const double* y, int n, double* c) { prefilter_t(y, n, c); }

template <class R>
static R eval_t(const R* c, int n, double x0, double dx, R x, int deriv) {
    R t = (x - x0) / dx;
    int i = (int)std::floor(val(t));
    if (i < 0) i = 0;
    if (i > n - 2) i = n - 2;
    R s = t - i, r = 1.0 - s;
    const R* p = c + i;
    if (deriv == 0) {
        R w0 = r * r * r / 6.0;
        R w1 = 2.0 / 3.0 - s * s + 0.5 * s * s * s;
        R w2 = 2.0 / 3.0 - r * r + 0.5 * r * r * r;
        R w3 = s * s * s / 6.0;
        return w0 * p[0] + w1 * p[1] + w2 * p[2] + w3 * p[3];
    } else if (deriv == 1) {
        R w0 = -0.5 * r * r;
        R w1 = -2.0 * s + 1.5 * s * s;
        R w2 = 2.0 * r - 1.5 * r * r;
        R w3 = 0.5 * s * s;
        return (w0 * p[0] + w1 * p[1] + w2 * p[2] + w3 * p[3]) / dx;
    } else {
        R w0 = r, w1 = -2.0 + 3.0 * s, w2 = -2.0 + 3.0 * r, w3 = s;
        return (w0 * p[0] + w1 * p[1] + w2 * p[2] + w3 * p[3]) / (dx * dx);
    }
}

double spline_eval(const double* c, int n, double x0, double dx, double x, int deriv) {
    return eval_t(c, n, x0, dx, x, deriv);
}

// d/dx and d2/dx2 of the spline at its own nodes (Interpolations.gradient /
// hessian at x in x_grid), used to build spline_∂ₓ / spline_∂ₓ² (util.jl:12-13).
template <class R>
static void node_derivatives(const R* c, int n, double dx, int deriv, R* out) {
    for (int j = 0; j < n; ++j) {
        const R* p = c + j;  // p[0]=c[j], p[1]=c[j+1] (node), p[2]=c[j+2]
        if (deriv == 1) out[j] = (p[2] - p[0]) / (2.0 * dx);
        else out[j] = (p[0] - 2.0 * p[1] + p[2]) / (dx * dx);
    }
}

// ----------------------------------------------------------------------------
// Cosmology helpers (background.jl:6-77)
template <class R>
CosmoT<R>::CosmoT(const ParamsT<R>& p, int nq) : par(p) {
    const NaturalUnits& u = natural_units();
    H0 = par.h * u.km_s_Mpc_100;
    rho_crit = (3.0 / (8.0 * M_PI)) * H0 * H0 / u.G_natural;
    const double zeta3 = 1.2020569;
    R Tg = pow(15.0 / (M_PI * M_PI) * rho_crit * par.Omega_r, 0.25);
    R nufac = (90.0 * zeta3 / (11.0 * std::pow(M_PI, 4))) * (par.Omega_r * par.h * par.h / Tg) *
              pow(par.N_nu / 3.0, 0.75);
    R Omega_nu = par.Sum_m_nu * nufac / (par.h * par.h);
    rad_fac = 1.0 + (2.0 / 3.0) * (7.0 * par.N_nu / 8.0) * std::pow(4.0 / 11.0, 4.0 / 3.0);
    Omega_L = 1.0 - (par.Omega_r * rad_fac + par.Omega_b + par.Omega_c + Omega_nu);
    Tnu = pow(par.N_nu / 3.0, 0.25) * std::pow(4.0 / 11.0, 1.0 / 3.0) *
          pow(15.0 / (M_PI * M_PI) * rho_crit * par.Omega_r, 0.25);
    logqmin = log10(Tnu / 30.0);
    logqmax = log10(Tnu * 30.0);
    xq.resize(nq);
    wq.resize(nq);
    gauss_legendre(nq, xq.data(), wq.data());
}

template <class R>
R CosmoT<R>::f0(R q) const {
    return 2.0 / std::pow(2.0 * M_PI, 3) / (exp(q / Tnu) + 1.0);
}

// xq2q, dxdq (util.jl:24-27)
template <class R>
static inline R from_ui(double x, R lo, R hi) { return lo + (hi - lo) / 2.0 * (x + 1.0); }
template <class R>
static inline R to_ui(R lq, R lo, R hi) { return -1.0 + 2.0 / (hi - lo) * (lq - lo); }
template <class R>
static inline R dxdq(R q, R lo, R hi) {
    return (1.0 + to_ui(R(1.0 + lo), lo, hi)) / (q * std::log(10.0));
}
template <class R>
static inline R xq2q(double x, R lo, R hi) { return pow(10.0, from_ui(x, lo, hi)); }

// ρP_0(a)[1] (background.jl:32-47)
template <class R>
R CosmoT<R>::rho0_massive(R a) const {
    R m = par.Sum_m_nu, s = 0.0;
    for (size_t i = 0; i < xq.size(); ++i) {
        R q = xq2q(xq[i], logqmin, logqmax);
        R eps = sqrt(q * q + (a * m) * (a * m));
        R I = q * q * eps * f0(q) / dxdq(q, logqmin, logqmax);
        s += I * wq[i];
    }
    return 4.0 * M_PI * pow(a, -4) * s;
}

// H_a, ℋ_a (background.jl:57-65)
template <class R>
R CosmoT<R>::H_a(R a) const {
    R rho_nu = rho0_massive(a);
    return H0 * sqrt((par.Omega_c + par.Omega_b) * pow(a, -3) + rho_nu / rho_crit +
                     par.Omega_r * pow(a, -4) * rad_fac + Omega_L);
}
template <class R>
R CosmoT<R>::Hc_a(R a) const { return a * H_a(a); }

// η(x) by GL quadrature in log10 a from -13.75 (background.jl:72-77)
template <class R>
R CosmoT<R>::eta(R x) const {
    R lo = -13.75, hi = log10(exp(x));
    R s = 0.0;
    for (size_t i = 0; i < xq.size(); ++i) {
        R ay = xq2q(xq[i], lo, hi);
        s += 1.0 / (ay * Hc_a(ay)) / dxdq(ay, lo, hi) * wq[i];
    }
    return s;
}

// ----------------------------------------------------------------------------
// Explicit Dormand-Prince 5(4) with step clipping onto requested outputs; it
// stands in for Tsit5 in RECFAST's stage solves (recfast.jl:463-471, 530-532).
// Outputs are exact to the (tight) tolerance, not Tsit5's interpolant.
//
// The independent variable and the step sizes are plain values and the error
// norm looks at the values only, so a Dual run takes the double run's steps and
// differentiates the state along them (its values are the double run's).  A
// Dual start point z0 enters by the chain rule, y(z) = S(z; z0, y0(z0)) with
// dS/dz0 = -Phi f(z0, y0): the start state's tangent becomes y0' - f(z0, y0) z0'.
// A Dual output point adds f(z_out, y) z_out'.  (Carrying tangents on the step
// sizes instead differentiates the explicit stability function at the edge of
// its stability region, which diverges on RECFAST's stiff hydrogen stage.)
// Where a stage starts on the Saha curve, the start tangent has a stiff
// transient that the explicit steps, sized for the values, resolve only
// roughly: the first few outputs of the helium stage carry tangents good to
// about 1e-4 relative; elsewhere they follow the values' accuracy.

template <class R, int N, class F>
static void dopri_outputs(F&& f, R z0_, R* y, const std::vector<R>& zout, std::vector<std::array<R, N>>& yout,
                          double rtol, double atol) {
    static const double c2 = 1.0 / 5, c3 = 3.0 / 10, c4 = 4.0 / 5, c5 = 8.0 / 9;
    static const double a21 = 1.0 / 5;
    static const double a31 = 3.0 / 40, a32 = 9.0 / 40;
    static const double a41 = 44.0 / 45, a42 = -56.0 / 15, a43 = 32.0 / 9;
    static const double a51 = 19372.0 / 6561, a52 = -25360.0 / 2187, a53 = 64448.0 / 6561,
                        a54 = -212.0 / 729;
    static const double a61 = 9017.0 / 3168, a62 = -355.0 / 33, a63 = 46732.0 / 5247,
                        a64 = 49.0 / 176, a65 = -5103.0 / 18656;
    static const double b1 = 35.0 / 384, b3 = 500.0 / 1113, b4 = 125.0 / 192,
                        b5 = -2187.0 / 6784, b6 = 11.0 / 84;
    static const double e1 = 71.0 / 57600, e3 = -71.0 / 16695, e4 = 71.0 / 1920,
                        e5 = -17253.0 / 339200, e6 = 22.0 / 525, e7 = -1.0 / 40;
    yout.resize(zout.size());
    const double z0 = val(z0_);
    double z = z0;
    double h = (zout.empty() ? 0.0 : (val(zout.back()) < z0 ? -1e-3 : 1e-3)) * std::max(1.0, std::fabs(z0));
    R k1[N], k2[N], k3[N], k4[N], k5[N], k6[N], k7[N], yt[N], yn[N];
    f(R(z), y, k1);
    if (any_tangent(z0_)) {
        for (int i = 0; i < N; ++i) add_tangent(y[i], -val(k1[i]), z0_);
        f(R(z), y, k1);
    }
    for (size_t io = 0; io < zout.size(); ++io) {
        const double zt = val(zout[io]);
        int guard = 0;
        while (z != zt) {
            if (++guard > 50000000) break;
            double dir = (zt < z) ? -1.0 : 1.0;
            if (h * dir <= 0) h = -h;
            bool last = false;
            if (std::fabs(h) >= std::fabs(zt - z)) { h = zt - z; last = true; }
            for (int i = 0; i < N; ++i) yt[i] = y[i] + h * a21 * k1[i];
            f(R(z + c2 * h), yt, k2);
            for (int i = 0; i < N; ++i) yt[i] = y[i] + h * (a31 * k1[i] + a32 * k2[i]);
            f(R(z + c3 * h), yt, k3);
            for (int i = 0; i < N; ++i) yt[i] = y[i] + h * (a41 * k1[i] + a42 * k2[i] + a43 * k3[i]);
            f(R(z + c4 * h), yt, k4);
            for (int i = 0; i < N; ++i)
                yt[i] = y[i] + h * (a51 * k1[i] + a52 * k2[i] + a53 * k3[i] + a54 * k4[i]);
            f(R(z + c5 * h), yt, k5);
            for (int i = 0; i < N; ++i)
                yt[i] = y[i] + h * (a61 * k1[i] + a62 * k2[i] + a63 * k3[i] + a64 * k4[i] + a65 * k5[i]);
            f(R(z + h), yt, k6);
            for (int i = 0; i < N; ++i)
                yn[i] = y[i] + h * (b1 * k1[i] + b3 * k3[i] + b4 * k4[i] + b5 * k5[i] + b6 * k6[i]);
            f(R(z + h), yn, k7);
            double err = 0.0;
            for (int i = 0; i < N; ++i) {
                double e = val(h * (e1 * k1[i] + e3 * k3[i] + e4 * k4[i] + e5 * k5[i] + e6 * k6[i] + e7 * k7[i]));
                double sc = atol + rtol * std::max(std::fabs(val(y[i])), std::fabs(val(yn[i])));
                err += (e / sc) * (e / sc);
            }
            err = std::sqrt(err / N);
            if (!(err == err)) err = 1e10;
            if (err <= 1.0) {
                z = last ? zt : z + h;
                for (int i = 0; i < N; ++i) { y[i] = yn[i]; k1[i] = k7[i]; }
                double fac = (err == 0.0) ? 5.0 : std::min(5.0, std::max(0.2, 0.9 * std::pow(err, -0.2)));
                if (!last) h *= fac;
            } else {
                h *= std::max(0.1, 0.9 * std::pow(err, -0.2));
            }
        }
        for (int i = 0; i < N; ++i) {
            yout[io][i] = y[i];
            add_tangent(yout[io][i], val(k1[i]), zout[io]);
        }
    }
}

// ----------------------------------------------------------------------------
// RECFAST (recfast.jl:22-121 constants, 156-310 rates, 446-536 stages)
template <class R>
struct RecfastT {
    const BackgroundT<R>* bg;
    double C = 2.99792458e8, k_B = 1.380658e-23, h_P = 6.6260755e-34, m_e = 9.1093897e-31,
           m_H = 1.673575e-27, not4 = 3.9715e0, sigma = 6.6524616e-29, a = 7.565914e-16,
           G = 6.6742e-11;
    double Lambda = 8.2245809e0, Lambda_He = 51.3e0, L_H_ion = 1.096787737e7,
           L_H_alpha = 8.225916453e6, L_He1_ion = 1.98310772e7, L_He2_ion = 4.389088863e7,
           L_He_2s = 1.66277434e7, L_He_2p = 1.71134891e7;
    double A2P_s = 1.798287e9, A2P_t = 177.58e0, L_He_2Pt = 1.690871466e7,
           L_He_2St = 1.5985597526e7, L_He2St_ion = 3.8454693845e6, sigma_He_2Ps = 1.436289e-22,
           sigma_He_2Pt = 1.484872e-22;
    double AGauss1 = -0.14e0, AGauss2 = 0.079e0, zGauss1 = 7.28e0, zGauss2 = 6.73e0,
           wGauss1 = 0.18e0, wGauss2 = 0.33e0;
    double a_PPB = 4.309, b_PPB = -0.6166, c_PPB = 0.6703, d_PPB = 0.5300;
    double a_VF = std::pow(10.0, -16.744), b_VF = 0.711, T_0 = std::pow(10.0, 0.477121),
           T_1 = std::pow(10.0, 5.114);
    double a_trip = std::pow(10.0, -16.306), b_trip = 0.761;
    double Lalpha, Lalpha_He, DeltaB, CDB, DeltaB_He, CDB_He, CB1, CB1_He1, CB1_He2, CR, CK, CK_He,
        CL, CL_He, CT, Bfact;
    double H_frac = 1e-3;
    int Hswitch = 1, Heswitch = 6;
    R Yp, OmegaB, HO, OmegaG, Tnow, mu_H, mu_T, fHe, Nnow;
    double fu, b_He = 0.86;

    RecfastT(const BackgroundT<R>* b, R Yp_, R OmegaB_, R OmegaG_, double Tnow_) : bg(b) {
        Lalpha = 1 / L_H_alpha;
        Lalpha_He = 1 / L_He_2p;
        DeltaB = h_P * C * (L_H_ion - L_H_alpha);
        CDB = DeltaB / k_B;
        DeltaB_He = h_P * C * (L_He1_ion - L_He_2s);
        CDB_He = DeltaB_He / k_B;
        CB1 = h_P * C * L_H_ion / k_B;
        CB1_He1 = h_P * C * L_He1_ion / k_B;
        CB1_He2 = h_P * C * L_He2_ion / k_B;
        CR = 2 * M_PI * (m_e / h_P) * (k_B / h_P);
        CK = std::pow(Lalpha, 3) / (8 * M_PI);
        CK_He = std::pow(Lalpha_He, 3) / (8 * M_PI);
        CL = C * h_P / (k_B * Lalpha);
        CL_He = C * h_P / (k_B / L_He_2s);
        CT = (8.0 / 3.0) * (sigma / (m_e * C)) * a;
        Bfact = h_P * C * (L_He_2p - L_He_2s) / k_B;
        Yp = Yp_;
        OmegaB = OmegaB_;
        const NaturalUnits& u = natural_units();
        HO = bg->H0 / u.H0_conv;
        OmegaG = OmegaG_;
        Tnow = (Tnow_ > 0) ? R(Tnow_) : pow(15.0 / (M_PI * M_PI) * bg->rho_crit * OmegaG, 0.25) * u.K_conv;
        mu_H = 1.0 / (1.0 - Yp);
        mu_T = not4 / (not4 - (not4 - 1) * Yp);
        fHe = Yp / (not4 * (1.0 - Yp));
        Nnow = 3.0 * HO * HO * OmegaB / (8 * M_PI * G * mu_H * m_H);
        fu = (Hswitch == 0) ? 1.14 : 1.125;
    }

    R Hc(R x) const { return bg->eval(BOLT_T_HC, x, 0); }
    R Hc_p(R x) const { return bg->eval(BOLT_T_HC_P, x, 0); }

    // ion_recfast (recfast.jl:156-310)
    void rates(const R* y, R z, R* f) const {
        R f1 = 0.0, f2 = 0.0, f3 = 0.0;
        R x_H = y[0], x_He = y[1], x = x_H + fHe * x_He, Tmat = y[2];
        R n = Nnow * pow(1.0 + z, 3);
        R n_He = fHe * Nnow * pow(1.0 + z, 3);
        R Trad = Tnow * (1.0 + z);
        R aa = 1.0 / (1.0 + z);
        R x_a = log(aa);
        R Hz = Hc(x_a) / aa / natural_units().H0_conv;
        R dHdz = (-Hc_p(x_a) + Hc(x_a)) / natural_units().H0_conv;
        R Rdown = 1e-19 * a_PPB * pow(Tmat / 1e4, b_PPB) / (1. + c_PPB * pow(Tmat / 1e4, d_PPB));
        R Rup = Rdown * pow(CR * Tmat, 1.5) * exp(-CDB / Tmat);
        R sq_0 = sqrt(Tmat / T_0), sq_1 = sqrt(Tmat / T_1);
        R Rdown_He = a_VF / (sq_0 * pow(1.0 + sq_0, 1 - b_VF));
        Rdown_He = Rdown_He / pow(1.0 + sq_1, 1 + b_VF);
        R Rup_He = Rdown_He * pow(CR * Tmat, 1.5) * exp(-CDB_He / Tmat);
        Rup_He = 4. * Rup_He;
        R He_Boltz = (Bfact / Tmat > 680.) ? R(std::exp(680.)) : exp(Bfact / Tmat);
        R K;
        if (Hswitch == 0) K = CK / Hz;
        else
            K = CK / Hz *
                (1.0 + AGauss1 * exp(-pow((log(1.0 + z) - zGauss1) / wGauss1, 2)) +
                 AGauss2 * exp(-pow((log(1.0 + z) - zGauss2) / wGauss2, 2)));
        R Rdown_trip = a_trip / (sq_0 * pow(1.0 + sq_0, 1 - b_trip));
        Rdown_trip = Rdown_trip / pow(1.0 + sq_1, 1 + b_trip);
        R Rup_trip = Rdown_trip * exp(-h_P * C * L_He2St_ion / (k_B * Tmat));
        Rup_trip = Rup_trip * pow(CR * Tmat, 1.5) * (4.0 / 3.0);
        int Heflag = ((x_He < 5.e-9) || (x_He > 0.980)) ? 0 : Heswitch;
        R K_He, CfHe_t = 0.0;
        if (Heflag == 0) {
            K_He = CK_He / Hz;
        } else {
            R tauHe_s = A2P_s * CK_He * 3 * n_He * (1.0 - x_He) / Hz;
            R pHe_s = (1.0 - exp(-tauHe_s)) / tauHe_s;
            K_He = 1.0 / (A2P_s * pHe_s * 3 * n_He * (1.0 - x_He));
            if (((Heflag == 2) || (Heflag >= 5)) && (x_H < 0.9999999)) {
                R Doppler = 2 * k_B * Tmat / (m_H * not4 * C * C);
                Doppler = C * L_He_2p * sqrt(Doppler);
                R gamma_2Ps = 3 * A2P_s * fHe * (1.0 - x_He) * C * C /
                              (std::sqrt(M_PI) * sigma_He_2Ps * 8 * M_PI * Doppler * (1.0 - x_H)) /
                              std::pow(C * L_He_2p, 2);
                double pb = 0.36, qb = b_He;
                R AHcon = A2P_s / (1.0 + pb * pow(gamma_2Ps, qb));
                K_He = 1.0 / ((A2P_s * pHe_s + AHcon) * 3 * n_He * (1.0 - x_He));
            }
            if (Heflag >= 3) {
                R tauHe_t = A2P_t * n_He * (1. - x_He) * 3;
                tauHe_t = tauHe_t / (8 * M_PI * Hz * std::pow(L_He_2Pt, 3));
                R pHe_t = (1.0 - exp(-tauHe_t)) / tauHe_t;
                double CL_PSt = h_P * C * (L_He_2Pt - L_He_2St) / k_B;
                if ((Heflag == 3) || (Heflag == 5) || (x_H > 0.99999)) {
                    CfHe_t = A2P_t * pHe_t * exp(-CL_PSt / Tmat);
                    CfHe_t = CfHe_t / (Rup_trip + CfHe_t);
                } else {
                    R Doppler = 2 * k_B * Tmat / (m_H * not4 * C * C);
                    Doppler = C * L_He_2Pt * sqrt(Doppler);
                    R gamma_2Pt = (3 * A2P_t * fHe * (1.0 - x_He) * C * C /
                                   (std::sqrt(M_PI) * sigma_He_2Pt * 8 * M_PI * Doppler * (1.0 - x_H)) /
                                   std::pow(C * L_He_2Pt, 2));
                    double pb = 0.66, qb = 0.9;
                    R AHcon = A2P_t / (1.0 + pb * pow(gamma_2Pt, qb)) / 3;
                    CfHe_t = (A2P_t * pHe_t + AHcon) * exp(-CL_PSt / Tmat);
                    CfHe_t = CfHe_t / (Rup_trip + CfHe_t);
                }
            }
        }
        R timeTh = (1.0 / (CT * pow(Trad, 4))) * (1.0 + x + fHe) / x;
        R timeH = 2.0 / (3 * HO * pow(1.0 + z, 1.5));
        if (x_H > 0.99) {
            f1 = 0.;
        } else if (x_H > 0.985) {
            f1 = (x * x_H * n * Rdown - Rup * (1.0 - x_H) * exp(-CL / Tmat)) / (Hz * (1.0 + z));
        } else {
            f1 = (((x * x_H * n * Rdown - Rup * (1.0 - x_H) * exp(-CL / Tmat)) *
                   (1.0 + K * Lambda * n * (1.0 - x_H))) /
                  (Hz * (1.0 + z) * (1.0 / fu + K * Lambda * n * (1.0 - x_H) / fu + K * Rup * n * (1.0 - x_H))));
        }
        if (x_He < 1e-15) {
            f2 = 0.;
        } else {
            f2 = (((x * x_He * n * Rdown_He - Rup_He * (1.0 - x_He) * exp(-CL_He / Tmat)) *
                   (1.0 + K_He * Lambda_He * n_He * (1.0 - x_He) * He_Boltz)) /
                  (Hz * (1.0 + z) * (1.0 + K_He * (Lambda_He + Rup_He) * n_He * (1.0 - x_He) * He_Boltz)));
            if (Heflag >= 3) {
                f2 = f2 + (x * x_He * n * Rdown_trip -
                           (1.0 - x_He) * 3 * Rup_trip * exp(-h_P * C * L_He_2St / (k_B * Tmat))) *
                              CfHe_t / (Hz * (1.0 + z));
            }
        }
        if (timeTh < H_frac * timeH) {
            R epsilon = Hz * (1.0 + x + fHe) / (CT * pow(Trad, 3) * x);
            f3 = Tnow + epsilon * ((1.0 + fHe) / (1.0 + fHe + x)) * ((f1 + fHe * f2) / x) -
                 epsilon * dHdz / Hz + 3 * epsilon / (1.0 + z);
        } else {
            f3 = CT * pow(Trad, 4) * x / (1.0 + x + fHe) * (Tmat - Trad) / (Hz * (1.0 + z)) + 2 * Tmat / (1.0 + z);
        }
        f[0] = f1; f[1] = f2; f[2] = f3;
    }

    R saha_rhs(R z, double CB) const {
        return exp(1.5 * log(CR * Tnow / (1.0 + z)) - CB / (Tnow * (1.0 + z))) / Nnow;
    }
    // He Saha total x (recfast.jl:363-366)
    R he_saha_x0(R z) const {
        R rhs = saha_rhs(z, CB1_He1) * 4;
        return 0.5 * (sqrt((rhs - 1.0) * (rhs - 1.0) + 4 * (1.0 + fHe) * rhs) - (rhs - 1.0));
    }
    R x_H0_saha(R z) const {  // recfast.jl:402-406
        R rhs = saha_rhs(z, CB1);
        return 0.5 * (sqrt(rhs * rhs + 4 * rhs) - rhs);
    }
    R end_of_saha(R z) const { return (he_saha_x0(z) - 1.0) / fHe - 0.99; }  // :373-380
    R end_He_evo(R z) const { return x_H0_saha(z) - 0.985; }               // :408

    // state of recfastsolve (recfast.jl:446-475) sampled on requested redshifts
    R z_He_evo_start = 0.0, z_H_He_evo_start = 0.0;

    // Root of g by bisection on values.  A Dual root carries the implicit
    // function's tangent -g_p / g_z (g_p the Dual part of g at the root, g_z a
    // central difference in z), which is what differentiating Bolt's bracketing
    // solve (recfast.jl:451-459) converges to.
    static R bisect(const RecfastT& r, R (RecfastT::*f)(R) const, double lo, double hi) {
        double flo = val((r.*f)(R(lo)));
        for (int it = 0; it < 200; ++it) {
            double mid = 0.5 * (lo + hi);
            double fm = val((r.*f)(R(mid)));
            if ((fm > 0) == (flo > 0)) { lo = mid; flo = fm; } else { hi = mid; }
            if (std::fabs(hi - lo) <= 1e-13 * std::max(1.0, std::fabs(mid))) break;
        }
        const double z0 = 0.5 * (lo + hi);
        if constexpr (dual_traits<R>::is_dual) {
            const double dz = 1e-6 * std::max(1.0, std::fabs(z0));
            const double gz = (val((r.*f)(R(z0 + dz))) - val((r.*f)(R(z0 - dz)))) / (2.0 * dz);
            const R g = (r.*f)(R(z0));
            R root(z0);
            for (int i = 0; i < (int)(sizeof(root.d) / sizeof(double)); ++i) root.d[i] = -g.d[i] / gz;
            return root;
        } else {
            return z0;
        }
    }

    // Evaluate X_e (pre-reionization, Xe_RECFAST) and Tmat (Tmat_RECFAST) on a
    // set of redshifts; zs need not be sorted.  With reio_dtm set, the
    // reionized matter temperature (recfast.jl:513-534) is integrated jointly
    // with the H+He+Tmat system from z_reio (which must be one of zs) down:
    // reio_dtm(z, X_e, T) is its derivative given the RECFAST X_e at z, and
    // tm_reio[i] receives it at zs[i] <= z_reio (entries above are unset).
    void solve(const std::vector<double>& zs, std::vector<R>& xe, std::vector<R>& tm,
               const std::function<R(R, R, R)>* reio_dtm = nullptr, double z_reio = 0.0,
               std::vector<R>* tm_reio = nullptr) {
        z_He_evo_start = bisect(*this, &RecfastT::end_of_saha, 3500.0, 0.0);
        z_H_He_evo_start = bisect(*this, &RecfastT::end_He_evo, val(z_He_evo_start), 0.0);
        const auto desc = [](const R& a, const R& b) { return val(a) > val(b); };
        const auto same = [](const R& a, const R& b) { return val(a) == val(b); };
        const auto above = [](const R& a, double b) { return val(a) > b; };
        // outputs per stage, descending z
        std::vector<R> zHe, zHH;
        for (double z : zs) {
            if (z <= val(z_He_evo_start) && z > val(z_H_He_evo_start)) zHe.push_back(z);
            else if (z <= val(z_H_He_evo_start)) zHH.push_back(z);
        }
        std::sort(zHe.begin(), zHe.end(), desc);
        std::sort(zHH.begin(), zHH.end(), desc);
        zHe.erase(std::unique(zHe.begin(), zHe.end(), same), zHe.end());
        zHH.erase(std::unique(zHH.begin(), zHH.end(), same), zHH.end());
        zHe.push_back(z_H_He_evo_start);
        // stage 2: He + Tmat with H in Saha (recfast.jl:461-465, 383-399)
        R y2[2];
        {
            R x0 = he_saha_x0(z_He_evo_start);
            y2[0] = (x0 - 1.0) / fHe;
            y2[1] = Tnow * (1.0 + z_He_evo_start);
        }
        auto fHe_stage = [this](R z, const R* y, R* dy) {
            R yy[3] = {x_H0_saha(z), y[0], y[1]}, f[3];
            rates(yy, z, f);
            dy[0] = f[1];
            dy[1] = f[2];
        };
        std::vector<std::array<R, 2>> oHe;
        dopri_outputs<R, 2>(fHe_stage, z_He_evo_start, y2, zHe, oHe, 1e-11, 1e-14);
        // stage 3: H + He + Tmat (recfast.jl:467-471)
        R y3[3] = {x_H0_saha(z_H_He_evo_start), oHe.back()[0], oHe.back()[1]};
        auto fHH = [this](R z, const R* y, R* dy) { rates(y, z, dy); };
        std::vector<std::array<R, 3>> oHH;
        std::vector<R> oReio;
        if (!reio_dtm) {
            dopri_outputs<R, 3>(fHH, z_H_He_evo_start, y3, zHH, oHH, 1e-11, 1e-14);
        } else {
            // down to z_reio with three components, then four
            std::vector<R> zhi, zlo;
            for (const R& z : zHH) (val(z) > z_reio ? zhi : zlo).push_back(z);
            zhi.push_back(R(z_reio));
            dopri_outputs<R, 3>(fHH, z_H_He_evo_start, y3, zhi, oHH, 1e-11, 1e-14);
            zhi.pop_back();
            oHH.pop_back();
            R y4[4] = {y3[0], y3[1], y3[2], y3[2]};
            auto f4 = [this, reio_dtm](R z, const R* y, R* dy) {
                rates(y, z, dy);
                dy[3] = (*reio_dtm)(z, y[0] + fHe * y[1], y[3]);
            };
            std::vector<std::array<R, 4>> o4;
            dopri_outputs<R, 4>(f4, R(z_reio), y4, zlo, o4, 1e-11, 1e-14);
            for (const auto& o : o4) {
                oHH.push_back({o[0], o[1], o[2]});
                oReio.push_back(o[3]);
            }
        }
        xe.resize(zs.size());
        tm.resize(zs.size());
        for (size_t i = 0; i < zs.size(); ++i) {
            double z = zs[i];
            if (z > val(z_He_evo_start)) {
                xe[i] = xe_saha_early(z);
                tm[i] = Tnow * (1.0 + z);
            } else if (z > val(z_H_He_evo_start)) {
                size_t j = std::lower_bound(zHe.begin(), zHe.end() - 1, z, above) - zHe.begin();
                xe[i] = x_H0_saha(z) + fHe * oHe[j][0];  // Xe_He_evolution (recfast.jl:313-318)
                tm[i] = oHe[j][1];
            } else {
                size_t j = std::lower_bound(zHH.begin(), zHH.end(), z, above) - zHH.begin();
                xe[i] = oHH[j][0] + fHe * oHH[j][1];
                tm[i] = oHH[j][2];
                if (tm_reio && reio_dtm && z <= z_reio) {
                    tm_reio->resize(zs.size());
                    (*tm_reio)[i] = oReio[j - (oHH.size() - oReio.size())];
                }
            }
        }
    }

    // Xe_RECFAST branches above z_He_evo_start (recfast.jl:415-427)
    R xe_saha_early(R z) const {
        if (z > 8000.) return 1.0 + 2 * fHe;
        if (z > 5000.) {
            R rhs = saha_rhs(z, CB1_He2);
            return 0.5 * (sqrt((rhs - 1.0 - fHe) * (rhs - 1.0 - fHe) + 4 * (1.0 + 2 * fHe) * rhs) - (rhs - 1.0 - fHe));
        }
        if (z > 3500.) return 1.0 + fHe;
        return he_saha_x0(z);
    }
};

// ----------------------------------------------------------------------------
template <class R>
BackgroundT<R>::BackgroundT(const ParamsT<R>& p, double x0_, double x1_, int nx_, int nq)
    : cosmo(p, nq), nx(nx_), x0(x0_), x1(x1_) {
    dx = (x1 - x0) / (nx - 1);
    H0 = cosmo.H0;
    rho_crit = cosmo.rho_crit;
    Omega_L = cosmo.Omega_L;
    coef.assign((size_t)BOLT_NTAB * (nx + 2), R(0.0));
    std::vector<double> xg(nx);
    std::vector<R> y(nx);
    for (int j = 0; j < nx; ++j) xg[j] = node(j);
    // ρ₀ℳ, ℋ, η on the grid (background.jl:106-108)
    for (int j = 0; j < nx; ++j) y[j] = cosmo.rho0_massive(R(std::exp(xg[j])));
    prefilter_t(y.data(), nx, tab(BOLT_T_RHO0M));
    for (int j = 0; j < nx; ++j) y[j] = cosmo.Hc_a(R(std::exp(xg[j])));
    prefilter_t(y.data(), nx, tab(BOLT_T_HC));
    for (int j = 0; j < nx; ++j) y[j] = cosmo.eta(R(xg[j]));
    prefilter_t(y.data(), nx, tab(BOLT_T_ETA));
    derive(BOLT_T_HC, BOLT_T_HC_P, BOLT_T_HC_PP);
    derive(BOLT_T_ETA, BOLT_T_ETA_P, BOLT_T_ETA_PP);
    eta0 = cosmo.eta(R(0.0));  // background.jl:112
}

template <class R>
double BackgroundT<R>::node(int j) const { return (x0 * (nx - 1 - j) + x1 * j) / (nx - 1); }

template <class R>
void BackgroundT<R>::derive(int src, int d1, int d2) {
    std::vector<R> d(nx);
    node_derivatives(tab(src), nx, dx, 1, d.data());
    prefilter_t(d.data(), nx, tab(d1));
    node_derivatives(tab(src), nx, dx, 2, d.data());
    prefilter_t(d.data(), nx, tab(d2));
}

template <class R>
R BackgroundT<R>::eval(int t, R x, int deriv) const {
    return eval_t(coef.data() + (size_t)t * (nx + 2), nx, x0, dx, x, deriv);
}

// IonizationHistory(𝕣::RECFAST, par, bg) (recfast.jl:674-726)
template <class R>
void BackgroundT<R>::ionization(double Tnow_override) {
    const ParamsT<R>& par = cosmo.par;
    RecfastT<R> rf(this, par.Y_p, par.Omega_b, par.Omega_r, Tnow_override);
    const NaturalUnits& u = natural_units();
    const double zinitial = 10000.0;
    const double x_init = std::log(1.0 / (zinitial + 1.0));  // z2x(zinitial)
    std::vector<double> xg(nx);
    for (int j = 0; j < nx; ++j) xg[j] = node(j);
    // redshifts needed: grid nodes after x_init, plus zre_ini = 50
    const double zre_ini = 50.0;
    std::vector<double> zs;
    for (int j = 0; j < nx; ++j)
        if (!(xg[j] < x_init)) zs.push_back(1.0 / std::exp(xg[j]) - 1.0);
    zs.push_back(zre_ini);
    // reionization (recfast.jl:478-536)
    R X_fin = 1.0 + rf.Yp / (rf.not4 * (1.0 - rf.Yp));
    const double zre = 7.6711, alpha = 1.5, dH = 0.5, zHe = 3.5, dHe = 0.5;
    R fHe_r = X_fin - 1.0;
    auto reio_xe = [&](R z, R x_orig) -> R {
        R xH = (X_fin - x_orig) / 2 *
                   (1.0 + tanh((std::pow(1 + zre, alpha) - pow(1.0 + z, alpha)) /
                               (alpha * std::pow(1 + zre, alpha - 1)) / dH)) + x_orig;
        R xHe = fHe_r / 2 * (1.0 + tanh((zHe - z) / dHe));
        return xH + xHe;
    };
    // the reionized matter temperature from z = 50 (recfast.jl:513-534), its
    // X_e taken from the RECFAST state integrated alongside
    std::function<R(R, R, R)> dtm = [&](R z, R xe_orig, R T) -> R {
        R xr = reio_xe(z, xe_orig);
        R aa = 1.0 / (1.0 + z);
        R Hz = rf.Hc(log(aa)) / aa / u.H0_conv;
        R Trad = rf.Tnow * (1.0 + z);
        return rf.CT * pow(Trad, 4) * xr / (1.0 + xr + rf.fHe) * (T - Trad) / (Hz * (1.0 + z)) + 2 * T / (1.0 + z);
    };
    std::vector<R> xe_r, tm_r, tmre_r;
    rf.solve(zs, xe_r, tm_r, &dtm, zre_ini, &tmre_r);
    std::vector<double> zlate;
    std::vector<R> tm_reio;
    for (size_t i = 0; i + 1 < zs.size(); ++i)
        if (zs[i] <= zre_ini) { zlate.push_back(zs[i]); tm_reio.push_back(tmre_r[i]); }
    {   // zlate descending, tm_reio alongside
        std::vector<size_t> o(zlate.size());
        for (size_t i = 0; i < o.size(); ++i) o[i] = i;
        std::sort(o.begin(), o.end(), [&](size_t a, size_t b) { return zlate[a] > zlate[b]; });
        std::vector<double> z2(o.size());
        std::vector<R> t2(o.size());
        for (size_t i = 0; i < o.size(); ++i) { z2[i] = zlate[o[i]]; t2[i] = tm_reio[o[i]]; }
        zlate.swap(z2);
        tm_reio.swap(t2);
    }
    // X_e(x), Tmat(x) on the grid (recfast.jl:692-696)
    std::vector<R> Xe(nx), Tm(nx);
    size_t ir = 0;
    for (int j = 0; j < nx; ++j) {
        double x = xg[j];
        R Trad = rf.Tnow * (1 + (1.0 / std::exp(x) - 1.0));
        if (x < x_init) {
            Xe[j] = 1.0 + 2 * rf.fHe;  // Xe_RECFAST(rhist, zinitial)
            Tm[j] = Trad;
        } else {
            double z = 1.0 / std::exp(x) - 1.0;
            R xo = xe_r[ir], to = tm_r[ir];
            ++ir;
            if (z > zre_ini) {
                Xe[j] = xo;
                Tm[j] = to;
            } else {
                Xe[j] = reio_xe(R(z), xo);
                size_t k = std::lower_bound(zlate.begin(), zlate.end(), z, std::greater<double>()) - zlate.begin();
                Tm[j] = tm_reio[k];
            }
        }
    }
    // τ′ on the grid and the reverse cumulative trapezoid (ionization.jl:107-129)
    std::vector<R> tp(nx), tau(nx), g(nx), y(nx);
    for (int j = 0; j < nx; ++j) {
        double a = std::exp(xg[j]);
        R nH = par.Omega_b * rho_crit / (u.m_H * a * a * a) * (1.0 - par.Y_p);
        tp[j] = -Xe[j] * nH * a * u.sigma_T / eval(BOLT_T_HC, R(xg[j]), 0);
    }
    tau[nx - 1] = 0.0;
    for (int j = nx - 2; j >= 0; --j) tau[j] = tau[j + 1] + (xg[j] - xg[j + 1]) * (tp[j] + tp[j + 1]) * 0.5;
    for (int j = 0; j < nx; ++j) g[j] = -tp[j] * exp(-tau[j]);  // g̃_function
    prefilter_t(Xe.data(), nx, tab(BOLT_T_XE));
    prefilter_t(tau.data(), nx, tab(BOLT_T_TAU));
    prefilter_t(g.data(), nx, tab(BOLT_T_G));
    prefilter_t(Tm.data(), nx, tab(BOLT_T_TMAT));
    derive(BOLT_T_TAU, BOLT_T_TAU_P, BOLT_T_TAU_PP);
    derive(BOLT_T_G, BOLT_T_G_P, BOLT_T_G_PP);
    // sound speed (recfast.jl:710-712)
    std::vector<R> dT(nx);
    node_derivatives(tab(BOLT_T_TMAT), nx, dx, 1, dT.data());
    for (int j = 0; j < nx; ++j) {
        R pre = std::pow(rf.C, -2) * rf.k_B / rf.m_H * (1.0 / rf.mu_T + (1.0 - rf.Yp) * eval(BOLT_T_XE, R(xg[j]), 0));
        y[j] = pre * (eval(BOLT_T_TMAT, R(xg[j]), 0) - 1.0 / 3.0 * dT[j]);
    }
    prefilter_t(y.data(), nx, tab(BOLT_T_CSB2));
    tau0 = tau[nx - 1];  // τ(0.)
}

template struct CosmoT<double>;
template struct CosmoT<Dual>;
template struct CosmoT<DualP>;
template struct BackgroundT<double>;
template struct BackgroundT<Dual>;
template struct BackgroundT<DualP>;

int recfast_xe(const bolt_cosmo_params& par, double Tnow, const double* z, int nz, double* out) {
    Background bg(par, -20.0, 0.0, 2001, 15);
    RecfastT<double> rf(&bg, par.Y_p, par.Omega_b, par.Omega_r, Tnow);
    std::vector<double> zs(z, z + nz), xe, tm;
    rf.solve(zs, xe, tm);
    for (int i = 0; i < nz; ++i) out[i] = xe[i];
    return 0;
}

}  // namespace bolt
