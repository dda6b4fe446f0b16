// Host-side (CPU, FP64) producers of the hot path's spline tables.
#pragma once
#include <array>
#include <cstddef>
#include <vector>

#include "../../include/bolt_hip.h"

namespace bolt {

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

struct Cosmo {
    bolt_cosmo_params par;
    double H0, rho_crit, Omega_L, rad_fac, Tnu, logqmin, logqmax;
    std::vector<double> xq, wq;
    Cosmo(const bolt_cosmo_params& p, int nq);
    double f0(double q) const;
    double rho0_massive(double a) const;
    double H_a(double a) const;
    double Hc_a(double a) const;
    double eta(double x) const;
};

struct Background {
    Cosmo cosmo;
    int nx;
    double x0, x1, dx;
    double H0, eta0, rho_crit, Omega_L, tau0 = 0.0;
    std::vector<double> coef;  // [BOLT_NTAB][nx+2]
    Background(const bolt_cosmo_params& p, double x0, double x1, int nx, int nq);
    double node(int j) const;
    double* tab(int t) { return coef.data() + (std::size_t)t * (nx + 2); }
    const double* tab(int t) const { return coef.data() + (std::size_t)t * (nx + 2); }
    double eval(int t, double x, int deriv) const;
    void derive(int src, int d1, int d2);
    void ionization(double Tnow_override);
};

int recfast_xe(const bolt_cosmo_params& par, double Tnow, const double* z, int nz, double* out);

}  // namespace bolt
