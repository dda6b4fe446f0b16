// Host-side (CPU, FP64) producers of the hot path's spline tables.
//
// Background, RECFAST and the ionization history are written once, generic
// over their scalar type R: R = double gives the tables themselves, R = Dual
// carries one tangent direction through the same arithmetic (forward-mode
// AD, what ForwardDiff.Dual does to Bolt's host code in runtests.jl:54-80).
#pragma once
#include <array>
#include <cmath>
#include <cstddef>
#include <vector>

#include "../../include/bolt_hip.h"

namespace bolt {

// ----------------------------------------------------------------------------
// Forward-mode dual number: value v and one directional derivative d.
// Every operation computes v exactly as the plain double operation does, so a
// Dual run's values are the double run's values.  Comparisons look at v only
// (branches and step-size decisions follow the primal).
struct Dual {
    double v = 0.0, d = 0.0;
    Dual() = default;
    Dual(double x) : v(x), d(0.0) {}  // constants carry no tangent
    Dual(double x, double dx) : v(x), d(dx) {}
    Dual& operator+=(const Dual& b) { d += b.d; v += b.v; return *this; }
    Dual& operator-=(const Dual& b) { d -= b.d; v -= b.v; return *this; }
    Dual& operator*=(const Dual& b) { d = d * b.v + v * b.d; v *= b.v; return *this; }
    Dual& operator/=(const Dual& b) { const double q = v / b.v; d = (d - q * b.d) / b.v; v = q; return *this; }
};
inline Dual operator-(const Dual& a) { return Dual(-a.v, -a.d); }
inline Dual operator+(Dual a, const Dual& b) { return a += b; }
inline Dual operator-(Dual a, const Dual& b) { return a -= b; }
inline Dual operator*(Dual a, const Dual& b) { return a *= b; }
inline Dual operator/(Dual a, const Dual& b) { return a /= b; }
inline bool operator<(const Dual& a, const Dual& b) { return a.v < b.v; }
inline bool operator>(const Dual& a, const Dual& b) { return a.v > b.v; }
inline bool operator<=(const Dual& a, const Dual& b) { return a.v <= b.v; }
inline bool operator>=(const Dual& a, const Dual& b) { return a.v >= b.v; }
inline bool operator==(const Dual& a, const Dual& b) { return a.v == b.v; }
inline bool operator!=(const Dual& a, const Dual& b) { return a.v != b.v; }

inline Dual exp(const Dual& a) { const double e = std::exp(a.v); return Dual(e, e * a.d); }
inline Dual log(const Dual& a) { return Dual(std::log(a.v), a.d / a.v); }
inline Dual log10(const Dual& a) { return Dual(std::log10(a.v), a.d / (a.v * std::log(10.0))); }
inline Dual sqrt(const Dual& a) { const double s = std::sqrt(a.v); return Dual(s, a.d / (2.0 * s)); }
inline Dual tanh(const Dual& a) { const double t = std::tanh(a.v); return Dual(t, a.d * (1.0 - t * t)); }
inline Dual fabs(const Dual& a) { return a.v < 0 ? -a : a; }
inline Dual pow(const Dual& a, double b) {
    const double p = std::pow(a.v, b);
    return Dual(p, a.d == 0.0 ? 0.0 : b * std::pow(a.v, b - 1.0) * a.d);
}
inline Dual pow(double a, const Dual& b) { const double p = std::pow(a, b.v); return Dual(p, p * std::log(a) * b.d); }
inline Dual pow(const Dual& a, const Dual& b) {
    if (b.d == 0.0) return pow(a, b.v);
    const double p = std::pow(a.v, b.v);
    return Dual(p, (a.d == 0.0 ? 0.0 : b.v * std::pow(a.v, b.v - 1.0) * a.d) + p * std::log(a.v) * b.d);
}
using std::exp;
using std::fabs;
using std::log;
using std::log10;
using std::pow;
using std::sqrt;
using std::tanh;

inline double val(double x) { return x; }
inline double val(const Dual& x) { return x.v; }
inline double tangent(double) { return 0.0; }
inline double tangent(const Dual& x) { return x.d; }

// ----------------------------------------------------------------------------
struct NaturalUnits {
    double km_s_Mpc_100;  // 100 km/s/Mpc in Mpc^-1
    double G_natural;     // G in Mpc^2
    double mass_natural;  // 1 eV in Mpc^-1
    double m_H;           // proton mass in Mpc^-1
    double sigma_T;       // Thomson cross section in Mpc^2
    double H0_conv;       // seconds per natural time unit (Mpc/c)
    double K_conv;        // Kelvin per natural temperature unit
};
const NaturalUnits& natural_units();

void gauss_legendre(int n, double* x, double* w);
void spline_prefilter(const double* y, int n, double* c);
double spline_eval(const double* c, int n, double x0, double dx, double x, int deriv);

// CosmoParams over R (src/Bolt.jl:56-66); built from the C struct with no tangent
template <class R>
struct ParamsT {
    R h, Omega_r, Omega_b, Omega_c, A, n, Y_p, N_nu, Sum_m_nu;
    ParamsT(const bolt_cosmo_params& p)
        : h(p.h), Omega_r(p.Omega_r), Omega_b(p.Omega_b), Omega_c(p.Omega_c), A(p.A), n(p.n), Y_p(p.Y_p),
          N_nu(p.N_nu), Sum_m_nu(p.Sum_m_nu) {}
    // value p, tangent direction dp (every field)
    ParamsT(const bolt_cosmo_params& p, const bolt_cosmo_params& dp)
        : h(p.h, dp.h), Omega_r(p.Omega_r, dp.Omega_r), Omega_b(p.Omega_b, dp.Omega_b),
          Omega_c(p.Omega_c, dp.Omega_c), A(p.A, dp.A), n(p.n, dp.n), Y_p(p.Y_p, dp.Y_p), N_nu(p.N_nu, dp.N_nu),
          Sum_m_nu(p.Sum_m_nu, dp.Sum_m_nu) {}
};

template <class R>
struct CosmoT {
    ParamsT<R> par;
    R H0, rho_crit, Omega_L, rad_fac, Tnu, logqmin, logqmax;
    std::vector<double> xq, wq;  // Gauss-Legendre nodes: no parameter dependence
    CosmoT(const ParamsT<R>& p, int nq);
    R f0(R q) const;
    R rho0_massive(R a) const;
    R H_a(R a) const;
    R Hc_a(R a) const;
    R eta(R x) const;
};

template <class R>
struct BackgroundT {
    CosmoT<R> cosmo;
    int nx;
    double x0, x1, dx;
    R H0, eta0, rho_crit, Omega_L, tau0 = R(0.0);
    std::vector<R> coef;  // [BOLT_NTAB][nx+2]
    BackgroundT(const ParamsT<R>& p, double x0, double x1, int nx, int nq);
    double node(int j) const;
    R* tab(int t) { return coef.data() + (std::size_t)t * (nx + 2); }
    const R* tab(int t) const { return coef.data() + (std::size_t)t * (nx + 2); }
    R eval(int t, R x, int deriv) const;
    void derive(int src, int d1, int d2);
    void ionization(double Tnow_override);
};

using Cosmo = CosmoT<double>;
using Background = BackgroundT<double>;
extern template struct CosmoT<double>;
extern template struct CosmoT<Dual>;
extern template struct BackgroundT<double>;
extern template struct BackgroundT<Dual>;

int recfast_xe(const bolt_cosmo_params& par, double Tnow, const double* z, int nz, double* out);

}  // namespace bolt
