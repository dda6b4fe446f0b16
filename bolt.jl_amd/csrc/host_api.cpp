// C ABI for the host-side producers and the shared error/version plumbing.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "host_tables.h"

namespace bolt {
thread_local std::string g_last_error;
int set_error(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}
}  // namespace bolt

using bolt::set_error;

extern "C" {

int bolt_abi_version(void) { return BOLT_ABI_VERSION; }

const char* bolt_last_error(void) { return bolt::g_last_error.c_str(); }

int bolt_state_length(const bolt_hier_cfg* c) {
    if (!c) return set_error(BOLT_E_ARG, "cfg is NULL");
    return 2 * (c->ell_g + 1) + (c->ell_nu + 1) + (c->ell_mnu + 1) * c->nq + 5;
}

int bolt_host_tables(const bolt_cosmo_params* par, double x0, double x1, int nx, int nq,
                     double* coef_out, double* scal_out, double* xq_out, double* wq_out) {
    if (!par || !coef_out || !scal_out || !xq_out || !wq_out)
        return set_error(BOLT_E_ARG, "bolt_host_tables: NULL argument");
    if (nx < 4 || nq < 1 || !(x1 > x0)) return set_error(BOLT_E_ARG, "bolt_host_tables: bad grid");
    if (!(par->Omega_r > 0) || !(par->N_nu > 0) || !(par->h > 0))
        return set_error(BOLT_E_ARG, "bolt_host_tables: Omega_r, N_nu and h must be positive");
    try {
        bolt::Background bg(*par, x0, x1, nx, nq);
        bg.ionization(-1.0);
        std::memcpy(coef_out, bg.coef.data(), sizeof(double) * bg.coef.size());
        scal_out[BOLT_S_H0] = bg.H0;
        scal_out[BOLT_S_ETA0] = bg.eta0;
        scal_out[BOLT_S_RHO_CRIT] = bg.rho_crit;
        scal_out[BOLT_S_OMEGA_L] = bg.Omega_L;
        scal_out[BOLT_S_TAU0] = bg.tau0;
        for (int i = 0; i < nq; ++i) {
            xq_out[i] = bg.cosmo.xq[i];
            wq_out[i] = bg.cosmo.wq[i];
        }
        for (double v : bg.coef)
            if (!std::isfinite(v)) return set_error(BOLT_E_NUMERIC, "bolt_host_tables: non-finite table entry");
    } catch (const std::exception& e) {
        return set_error(BOLT_E_NUMERIC, std::string("bolt_host_tables: ") + e.what());
    }
    return BOLT_OK;
}

int bolt_host_tables_batch(const bolt_cosmo_params* pars, int ncosmo, double x0, double x1, int nx, int nq,
                           int nthreads, double* coef_out, double* scal_out, double* xq_out, double* wq_out,
                           int* status) {
    if (!pars || ncosmo < 0 || !coef_out || !scal_out || !xq_out || !wq_out || !status)
        return set_error(BOLT_E_ARG, "bolt_host_tables_batch: NULL argument or negative count");
    if (nx < 4 || nq < 1 || !(x1 > x0)) return set_error(BOLT_E_ARG, "bolt_host_tables_batch: bad grid");
    if (ncosmo == 0) return BOLT_OK;
    int nt = nthreads > 0 ? nthreads : (int)std::thread::hardware_concurrency();
    nt = std::max(1, std::min({nt, ncosmo, 256}));
    const size_t tab = (size_t)BOLT_NTAB * (nx + 2);
    std::atomic<int> next(0);
    std::mutex mu;
    std::string first_error;
    auto work = [&]() {
        std::vector<double> xq(nq), wq(nq);
        for (int c = next.fetch_add(1); c < ncosmo; c = next.fetch_add(1)) {
            // each cosmology is one bolt_host_tables call: the same code path and
            // the same bits as building it alone
            const int rc = bolt_host_tables(pars + c, x0, x1, nx, nq, coef_out + tab * c,
                                            scal_out + (size_t)BOLT_NSCAL * c, xq.data(), wq.data());
            status[c] = rc;
            if (rc != BOLT_OK) {
                std::lock_guard<std::mutex> g(mu);
                if (first_error.empty())
                    first_error = "bolt_host_tables_batch: cosmology " + std::to_string(c) + ": " + bolt::g_last_error;
            } else if (c == 0) {
                std::memcpy(xq_out, xq.data(), sizeof(double) * nq);
                std::memcpy(wq_out, wq.data(), sizeof(double) * nq);
            }
        }
    };
    std::vector<std::thread> pool;
    try {
        for (int t = 1; t < nt; ++t) pool.emplace_back(work);
    } catch (const std::exception&) {
        // fewer threads than asked for: the ones started and this one finish the queue
    }
    work();
    for (auto& th : pool) th.join();
    if (status[0] != BOLT_OK) bolt::gauss_legendre(nq, xq_out, wq_out);
    if (!first_error.empty()) return set_error(BOLT_E_NUMERIC, first_error);
    return BOLT_OK;
}

int bolt_host_tables_jvp(const bolt_cosmo_params* par, const bolt_cosmo_params* dpar, double x0, double x1, int nx,
                         int nq, double* coef_out, double* scal_out, double* dcoef_out, double* dscal_out,
                         double* xq_out, double* wq_out) {
    if (!par || !dpar || !coef_out || !scal_out || !dcoef_out || !dscal_out || !xq_out || !wq_out)
        return set_error(BOLT_E_ARG, "bolt_host_tables_jvp: NULL argument");
    if (nx < 4 || nq < 1 || !(x1 > x0)) return set_error(BOLT_E_ARG, "bolt_host_tables_jvp: bad grid");
    if (!(par->Omega_r > 0) || !(par->N_nu > 0) || !(par->h > 0))
        return set_error(BOLT_E_ARG, "bolt_host_tables_jvp: Omega_r, N_nu and h must be positive");
    const double* dp = &dpar->h;
    for (int i = 0; i < 9; ++i)
        if (!std::isfinite(dp[i])) return set_error(BOLT_E_ARG, "bolt_host_tables_jvp: non-finite direction");
    // the Dual run's values are bitwise the double run's; they are taken from
    // bolt_host_tables itself so that the two entry points share one code path
    int rc = bolt_host_tables(par, x0, x1, nx, nq, coef_out, scal_out, xq_out, wq_out);
    if (rc != BOLT_OK) return rc;
    try {
        bolt::BackgroundT<bolt::Dual> bg(bolt::ParamsT<bolt::Dual>(*par, *dpar), x0, x1, nx, nq);
        bg.ionization(-1.0);
        for (size_t i = 0; i < bg.coef.size(); ++i) {
            dcoef_out[i] = bg.coef[i].d[0];
            if (!std::isfinite(dcoef_out[i]))
                return set_error(BOLT_E_NUMERIC, "bolt_host_tables_jvp: non-finite tangent");
        }
        const bolt::Dual s[BOLT_NSCAL] = {bg.H0, bg.eta0, bg.rho_crit, bg.Omega_L, bg.tau0};
        for (int i = 0; i < BOLT_NSCAL; ++i) dscal_out[i] = s[i].d[0];
    } catch (const std::exception& e) {
        return set_error(BOLT_E_NUMERIC, std::string("bolt_host_tables_jvp: ") + e.what());
    }
    return BOLT_OK;
}

int bolt_host_tables_jacobian(const bolt_cosmo_params* par, double x0, double x1, int nx, int nq, double* coef_out,
                              double* scal_out, double* dcoef_out, double* dscal_out, double* xq_out, double* wq_out) {
    if (!par || !coef_out || !scal_out || !dcoef_out || !dscal_out || !xq_out || !wq_out)
        return set_error(BOLT_E_ARG, "bolt_host_tables_jacobian: NULL argument");
    if (nx < 4 || nq < 1 || !(x1 > x0)) return set_error(BOLT_E_ARG, "bolt_host_tables_jacobian: bad grid");
    if (!(par->Omega_r > 0) || !(par->N_nu > 0) || !(par->h > 0))
        return set_error(BOLT_E_ARG, "bolt_host_tables_jacobian: Omega_r, N_nu and h must be positive");
    int rc = bolt_host_tables(par, x0, x1, nx, nq, coef_out, scal_out, xq_out, wq_out);
    if (rc != BOLT_OK) return rc;
    try {
        bolt::BackgroundT<bolt::DualP> bg(bolt::ParamsT<bolt::DualP>::seeded(*par), x0, x1, nx, nq);
        bg.ionization(-1.0);
        const size_t ntab = bg.coef.size();
        for (int p = 0; p < bolt::NPAR; ++p)
            for (size_t i = 0; i < ntab; ++i) {
                const double d = bg.coef[i].d[p];
                if (!std::isfinite(d)) return set_error(BOLT_E_NUMERIC, "bolt_host_tables_jacobian: non-finite tangent");
                dcoef_out[(size_t)p * ntab + i] = d;
            }
        const bolt::DualP s[BOLT_NSCAL] = {bg.H0, bg.eta0, bg.rho_crit, bg.Omega_L, bg.tau0};
        for (int p = 0; p < bolt::NPAR; ++p)
            for (int i = 0; i < BOLT_NSCAL; ++i) dscal_out[p * BOLT_NSCAL + i] = s[i].d[p];
    } catch (const std::exception& e) {
        return set_error(BOLT_E_NUMERIC, std::string("bolt_host_tables_jacobian: ") + e.what());
    }
    return BOLT_OK;
}

int bolt_host_recfast_xe(const bolt_cosmo_params* par, double Tnow, const double* z, int nz,
                         double* xe_out) {
    if (!par || !z || !xe_out || nz < 0) return set_error(BOLT_E_ARG, "bolt_host_recfast_xe: bad argument");
    for (int i = 0; i < nz; ++i)
        if (!(z[i] >= 0.0 && z[i] <= 10000.0))
            return set_error(BOLT_E_ARG, "bolt_host_recfast_xe: z outside [0, 10000]");
    return bolt::recfast_xe(*par, Tnow, z, nz, xe_out);
}

int bolt_host_spline_eval(const bolt_tables* t, int tab, int deriv, const double* xs, int m, double* out) {
    if (!t || !t->coef || !xs || !out || tab < 0 || tab >= BOLT_NTAB || deriv < 0 || deriv > 2)
        return set_error(BOLT_E_ARG, "bolt_host_spline_eval: bad argument");
    double dx = (t->x1 - t->x0) / (t->nx - 1);
    const double* c = t->coef + (size_t)tab * (t->nx + 2);
    for (int i = 0; i < m; ++i) out[i] = bolt::spline_eval(c, t->nx, t->x0, dx, xs[i], deriv);
    return BOLT_OK;
}

}  // extern "C"
