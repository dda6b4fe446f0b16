// AddressSanitizer / UndefinedBehaviorSanitizer run of the host-side code on the
// CPU: the host tables (Background, RECFAST, ionization history, their Dual
// instantiations, the threaded batch) and the oracle that the tests check the
// device against (hierarchy RHS, initial conditions, source columns adaptive and
// replayed, plin, C_l with exact and spline j_l), on small inputs, including the
// argument checks that reject bad sizes.  Built and run by tests/test_sanitize.py.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "bolt_hip.h"
extern "C" {
#include "bolt_oracle.h"
}

static int fails = 0;
#define CHECK(c)                                                       \
    do {                                                               \
        if (!(c)) {                                                    \
            std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            ++fails;                                                   \
        }                                                              \
    } while (0)

int main() {
    const double mnu = 0.2 * 1.602176634e-19 / (1.054571817e-34 * 299792458.0) * 3.0856775814913673e22;
    bolt_cosmo_params par = {0.67, 5.0469e-5, 0.049, 0.26, 2.1e-9, 0.965, 0.24, 3.046, mnu};
    const int nx = 2001, nq = 15;
    std::vector<double> coef(BOLT_NTAB * (nx + 2)), scal(BOLT_NSCAL), xq(nq), wq(nq);
    CHECK(bolt_host_tables(&par, -20.0, 0.0, nx, nq, coef.data(), scal.data(), xq.data(), wq.data()) == 0);
    // the tangent along a mixed direction, and the nine-partial Jacobian on a coarse grid
    bolt_cosmo_params dpar = {0.67, 0.0, 0.049, -0.26, 0.0, 0.0, 0.0, 0.3, mnu};
    std::vector<double> c2(coef.size()), s2(BOLT_NSCAL), dc(coef.size()), ds(BOLT_NSCAL);
    CHECK(bolt_host_tables_jvp(&par, &dpar, -20.0, 0.0, nx, nq, c2.data(), s2.data(), dc.data(), ds.data(),
                               xq.data(), wq.data()) == 0);
    for (size_t i = 0; i < coef.size(); ++i) CHECK(c2[i] == coef[i]);
    {
        const int nxc = 201;
        std::vector<double> cc(BOLT_NTAB * (nxc + 2)), sc(BOLT_NSCAL), dcc(9 * cc.size()), dsc(9 * BOLT_NSCAL);
        CHECK(bolt_host_tables_jacobian(&par, -20.0, 0.0, nxc, nq, cc.data(), sc.data(), dcc.data(), dsc.data(),
                                        xq.data(), wq.data()) == 0);
    }
    // three cosmologies on two host threads
    {
        bolt_cosmo_params ps[3] = {par, par, par};
        ps[1].h = 0.7;
        ps[2].Sum_m_nu = 0.0;
        const int nxb = 401;
        std::vector<double> cb(3 * BOLT_NTAB * (nxb + 2)), sb(3 * BOLT_NSCAL), xb(nq), wb(nq);
        int st[3];
        CHECK(bolt_host_tables_batch(ps, 3, -20.0, 0.0, nxb, nq, 2, cb.data(), sb.data(), xb.data(), wb.data(), st) ==
              0);
    }
    // rejected arguments
    CHECK(bolt_host_tables(&par, -20.0, 0.0, 3, nq, coef.data(), scal.data(), xq.data(), wq.data()) == BOLT_E_ARG);
    CHECK(bolt_host_tables(&par, 0.0, -20.0, nx, nq, coef.data(), scal.data(), xq.data(), wq.data()) == BOLT_E_ARG);
    {
        double zs[4] = {100.0, 800.0, 1100.0, 1500.0}, xe[4];
        CHECK(bolt_host_recfast_xe(&par, -1.0, zs, 4, xe) == 0);
    }

    // the oracle on these tables
    orc_params op = {par.h, par.Omega_r, par.Omega_b, par.Omega_c, par.A, par.n, par.Y_p, par.N_nu, par.Sum_m_nu};
    orc_tables ot = {nx, nq, -20.0, 0.0, scal[BOLT_S_H0], scal[BOLT_S_ETA0], scal[BOLT_S_RHO_CRIT], xq.data(),
                     wq.data(), coef.data()};
    orc_cfg cfg = {8, 8, 10, nq, 1e-6, 1e-6, 0, 1};
    const int n = orc_state_length(&cfg);
    std::vector<double> u(n), du(n);
    orc_initial_conditions(&op, &ot, &cfg, 0.05, u.data());
    orc_rhs(&op, &ot, &cfg, 0.05, -10.0, u.data(), du.data(), 1);
    std::vector<double> J((size_t)n * n);
    orc_jacobian(&op, &ot, &cfg, 0.05, -3.0, J.data());
    orc_stats st;
    std::vector<double> S(nx), S2(nx), SP(nx);
    const double ks[3] = {1e-3, 0.02, 0.2};
    std::vector<double> grid(3 * (size_t)nx);
    for (int i = 0; i < 3; ++i) {
        CHECK(orc_source_column_tp(&op, &ot, &cfg, ks[i], grid.data() + (size_t)i * nx, SP.data(), &st) == 0);
    }
    std::vector<double> xs(100000);
    int nxs = 0;
    CHECK(orc_boltsolve_steps(&op, &ot, &cfg, 0.02, xs.data(), (int)xs.size(), &nxs, &st) == 0);
    CHECK(orc_source_column_replay(&op, &ot, &cfg, 0.02, xs.data(), nxs, S2.data(), &st) == 0);
    for (int j = 0; j < nx; ++j) CHECK(std::fabs(S2[j] - grid[(size_t)nx + j]) <= 1e-8 * (1.0 + std::fabs(S2[j])));
    orc_cfg pc = {12, 12, 10, nq, 1e-5, 1e-6, 0, 1};
    CHECK(std::isfinite(orc_plin(&op, &ot, &pc, 0.05, 0.0, &st)));
    {
        std::vector<double> kd(60);
        for (int i = 0; i < 60; ++i) kd[i] = 1e-4 + (0.2 - 1e-4) * ((i + 1) / 60.0) * ((i + 1) / 60.0);
        int ells[4] = {2, 10, 50, 120};
        double cl[4], cl1[4];
        CHECK(orc_cltt(&op, &ot, grid.data(), ks, 3, kd.data(), 60, ells, 4, 0, cl, 2, 0, 59) == 0);
        CHECK(orc_cltt(&op, &ot, grid.data(), ks, 3, kd.data(), 60, ells, 4, 1, cl1, 2, 0, 59) == 0);
        for (int i = 0; i < 4; ++i) CHECK(std::isfinite(cl[i]) && std::isfinite(cl1[i]));
    }
    // a momentum-bin count beyond the oracle's fixed arrays is refused
    orc_cfg bad = cfg;
    bad.nq = ORC_MAXQ + 1;
    CHECK(orc_source_column(&op, &ot, &bad, 0.02, S.data(), &st) == 4);
    std::printf("sanitized host run: %s\n", fails ? "FAILED" : "ok");
    return fails ? 1 : 0;
}
