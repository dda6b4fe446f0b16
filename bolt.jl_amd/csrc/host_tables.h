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
// Forward-mode dual number: value v and NP directional derivatives d[0..NP-1]
// (ForwardDiff.Dual with NP partials; NP = 1 is one direction, NP = 9 every
// CosmoParams field at once).  Every operation computes v exactly as the plain
// double operation does, and each partial exactly as the one-partial number
// would, so a run's values are the double run's and its partial i is bitwise
// the one-direction run seeded with direction i.  Comparisons look at v only
// (branches and step-size decisions follow the primal).
template <int NP>
struct DualT {
    double v = 0.0, d[NP] = {};
    DualT() = default;
    DualT(double x) : v(x) {}  // constants carry no tangent
    DualT(double x, double dx) : v(x) {
        for (int i = 0; i < NP; ++i) d[i] = dx;
    }
    DualT& operator+=(const DualT& b) {
        for (int i = 0; i < NP; ++i) d[i] += b.d[i];
        v += b.v;
        return *this;
    }
    DualT& operator-=(const DualT& b) {
        for (int i = 0; i < NP; ++i) d[i] -= b.d[i];
        v -= b.v;
        return *this;
    }
    DualT& operator*=(const DualT& b) {
        for (int i = 0; i < NP; ++i) d[i] = d[i] * b.v + v * b.d[i];
        v *= b.v;
        return *this;
    }
    DualT& operator/=(const DualT& b) {
        const double q = v / b.v;
        for (int i = 0; i < NP; ++i) d[i] = (d[i] - q * b.d[i]) / b.v;
        v = q;
        return *this;
    }
    bool any_tangent() const {
        for (int i = 0; i < NP; ++i)
            if (d[i] != 0.0) return true;
        return false;
    }
    // the value with partials d[i] = s * other.d[i] (chain rule through a scalar map)
    static DualT chain(double value, double s, const DualT& other) {
        DualT r(value);
        for (int i = 0; i < NP; ++i) r.d[i] = s * other.d[i];
        return r;
    }
};
template <int NP> inline DualT<NP> operator-(const DualT<NP>& a) { return DualT<NP>::chain(-a.v, -1.0, a); }
template <int NP> inline DualT<NP> operator+(DualT<NP> a, const DualT<NP>& b) { return a += b; }
template <int NP> inline DualT<NP> operator-(DualT<NP> a, const DualT<NP>& b) { return a -= b; }
template <int NP> inline DualT<NP> operator*(DualT<NP> a, const DualT<NP>& b) { return a *= b; }
template <int NP> inline DualT<NP> operator/(DualT<NP> a, const DualT<NP>& b) { return a /= b; }
template <int NP> inline DualT<NP> operator+(DualT<NP> a, double b) { return a += DualT<NP>(b); }
template <int NP> inline DualT<NP> operator+(double a, const DualT<NP>& b) { return DualT<NP>(a) += b; }
template <int NP> inline DualT<NP> operator-(DualT<NP> a, double b) { return a -= DualT<NP>(b); }
template <int NP> inline DualT<NP> operator-(double a, const DualT<NP>& b) { return DualT<NP>(a) -= b; }
template <int NP> inline DualT<NP> operator*(DualT<NP> a, double b) { return a *= DualT<NP>(b); }
template <int NP> inline DualT<NP> operator*(double a, const DualT<NP>& b) { return DualT<NP>(a) *= b; }
template <int NP> inline DualT<NP> operator/(DualT<NP> a, double b) { return a /= DualT<NP>(b); }
template <int NP> inline DualT<NP> operator/(double a, const DualT<NP>& b) { return DualT<NP>(a) /= b; }
#define BOLT_DUAL_CMP(op)                                                                                 \
    template <int NP> inline bool operator op(const DualT<NP>& a, const DualT<NP>& b) { return a.v op b.v; } \
    template <int NP> inline bool operator op(const DualT<NP>& a, double b) { return a.v op b; }           \
    template <int NP> inline bool operator op(double a, const DualT<NP>& b) { return a op b.v; }
BOLT_DUAL_CMP(<)
BOLT_DUAL_CMP(>)
BOLT_DUAL_CMP(<=)
BOLT_DUAL_CMP(>=)
BOLT_DUAL_CMP(==)
BOLT_DUAL_CMP(!=)
#undef BOLT_DUAL_CMP

template <int NP> inline DualT<NP> exp(const DualT<NP>& a) {
    const double e = std::exp(a.v);
    return DualT<NP>::chain(e, e, a);
}
template <int NP> inline DualT<NP> log(const DualT<NP>& a) {
    DualT<NP> r(std::log(a.v));
    for (int i = 0; i < NP; ++i) r.d[i] = a.d[i] / a.v;
    return r;
}
template <int NP> inline DualT<NP> log10(const DualT<NP>& a) {
    DualT<NP> r(std::log10(a.v));
    for (int i = 0; i < NP; ++i) r.d[i] = a.d[i] / (a.v * std::log(10.0));
    return r;
}
template <int NP> inline DualT<NP> sqrt(const DualT<NP>& a) {
    const double s = std::sqrt(a.v);
    DualT<NP> r(s);
    for (int i = 0; i < NP; ++i) r.d[i] = a.d[i] / (2.0 * s);
    return r;
}
template <int NP> inline DualT<NP> tanh(const DualT<NP>& a) {
    const double t = std::tanh(a.v);
    DualT<NP> r(t);
    for (int i = 0; i < NP; ++i) r.d[i] = a.d[i] * (1.0 - t * t);
    return r;
}
template <int NP> inline DualT<NP> fabs(const DualT<NP>& a) { return a.v < 0 ? -a : a; }
template <int NP> inline DualT<NP> pow(const DualT<NP>& a, double b) {
    DualT<NP> r(std::pow(a.v, b));
    for (int i = 0; i < NP; ++i) r.d[i] = a.d[i] == 0.0 ? 0.0 : b * std::pow(a.v, b - 1.0) * a.d[i];
    return r;
}
template <int NP> inline DualT<NP> pow(const DualT<NP>& a, int b) { return pow(a, (double)b); }
template <int NP> inline DualT<NP> pow(double a, const DualT<NP>& b) {
    const double p = std::pow(a, b.v);
    DualT<NP> r(p);
    for (int i = 0; i < NP; ++i) r.d[i] = p * std::log(a) * b.d[i];
    return r;
}
template <int NP> inline DualT<NP> pow(const DualT<NP>& a, const DualT<NP>& b) {
    if (!b.any_tangent()) return pow(a, b.v);
    const double p = std::pow(a.v, b.v);
    DualT<NP> r(p);
    for (int i = 0; i < NP; ++i)
        r.d[i] = (a.d[i] == 0.0 ? 0.0 : b.v * std::pow(a.v, b.v - 1.0) * a.d[i]) + p * std::log(a.v) * b.d[i];
    return r;
}
using Dual = DualT<1>;
constexpr int NPAR = 9;  // the fields of bolt_cosmo_params
using DualP = DualT<NPAR>;
using std::exp;
using std::fabs;
using std::log;
using std::log10;
using std::pow;
using std::sqrt;
using std::tanh;

inline double val(double x) { return x; }
template <int NP> inline double val(const DualT<NP>& x) { return x.v; }
template <class R> struct dual_traits { static constexpr bool is_dual = false; };
template <int NP> struct dual_traits<DualT<NP>> { static constexpr bool is_dual = true; };
inline bool any_tangent(double) { return false; }
template <int NP> inline bool any_tangent(const DualT<NP>& x) { return x.any_tangent(); }
// y' += slope * z' (partial by partial); nothing for plain doubles
inline void add_tangent(double&, double, double) {}
template <int NP> inline void add_tangent(DualT<NP>& y, double slope, const DualT<NP>& z) {
    for (int i = 0; i < NP; ++i) y.d[i] += slope * z.d[i];
}

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
    // value p, one tangent direction dp (every field)
    ParamsT(const bolt_cosmo_params& p, const bolt_cosmo_params& dp)
        : h(p.h, dp.h), Omega_r(p.Omega_r, dp.Omega_r), Omega_b(p.Omega_b, dp.Omega_b),
          Omega_c(p.Omega_c, dp.Omega_c), A(p.A, dp.A), n(p.n, dp.n), Y_p(p.Y_p, dp.Y_p), N_nu(p.N_nu, dp.N_nu),
          Sum_m_nu(p.Sum_m_nu, dp.Sum_m_nu) {}
    // value p, partial i = d/d(field i) for the NPAR fields in struct order
    static ParamsT seeded(const bolt_cosmo_params& p) {
        ParamsT r(p);
        R* f[NPAR] = {&r.h, &r.Omega_r, &r.Omega_b, &r.Omega_c, &r.A, &r.n, &r.Y_p, &r.N_nu, &r.Sum_m_nu};
        for (int i = 0; i < NPAR; ++i) f[i]->d[i] = 1.0;
        return r;
    }
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
extern template struct CosmoT<DualP>;
extern template struct BackgroundT<double>;
extern template struct BackgroundT<Dual>;
extern template struct BackgroundT<DualP>;

int recfast_xe(const bolt_cosmo_params& par, double Tnow, const double* z, int nz, double* out);

}  // namespace bolt
