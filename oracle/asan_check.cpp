// asan_check.cpp -- runs every entry point of the CPU oracle on small lattices
// under the host address and undefined-behaviour sanitizers (make -C oracle
// asan-check).  Test infrastructure, like the oracle itself: it checks memory
// safety of the checker, plus the invariants any run must keep (finite fields,
// mass conserved on periodic and walled lattices).
#include "cpmp_oracle.h"

#include <cmath>
#include <cstdio>
#include <vector>

namespace {

int failures = 0;

void check(bool ok, const char *what, const char *detail = "") {
    if (!ok) {
        std::fprintf(stderr, "asan_check: %s: %s\n", what, detail);
        ++failures;
    }
}

orc_params params(int nx, int ny, int nz, int eos, double tr, int nwall, int scheme, int lap, int forcing,
                  double k1) {
    orc_params p{};
    p.nx = nx; p.ny = ny; p.nz = nz;
    p.nwall_lo = p.nwall_hi = nwall;
    p.eos = eos;
    p.mu_form = 0;
    p.lap_points = lap;
    p.wall_scheme = scheme;
    double Tc = 0, Rc = 0;
    orc_set_multiphase(eos, tr, &p.A, &p.B, &Tc, &Rc);
    p.T = tr * Tc;
    p.K = 0.2; p.Ka = 0.001; p.K1 = k1; p.K2 = -k1;
    const double tau = 0.8;
    for (int q = 0; q < 19; ++q) p.Ts[q] = 1.0;
    p.Ts[2] = p.Ts[4] = p.Ts[6] = p.Ts[8] = 1.1;
    p.Ts[9] = p.Ts[11] = p.Ts[13] = p.Ts[14] = p.Ts[15] = 1.0 / tau;
    p.forcing = forcing;
    return p;
}

double mass(orc_state *s) {
    double r[3];
    orc_reduce(s, r);
    return r[0];
}

void run_case(const char *name, const orc_params &p, int steps, bool adjoint) {
    orc_state *s = orc_create(&p);
    check(s != nullptr, name, "orc_create rejected the parameters");
    if (!s) return;
    double g = 0.01 / p.B, l = (p.eos == ORC_CSE ? 0.5 : 0.8) / p.B;   // guesses, replaced by the coexistence densities
    check(orc_maxwell(p.eos, p.A, p.B, p.T, &g, &l) == 0 && g < l, name, "no coexistence densities");
    const int n = p.nx * p.ny * p.nz, nxy = p.nx * p.ny;
    std::vector<double> wl(nxy), wh(nxy, 0.01);
    for (int i = 0; i < nxy; ++i) wl[i] = -0.02 + 0.04 * i / nxy;
    orc_set_wall_mu(s, wl.data(), wh.data());
    orc_init_reference(s, p.nx / 4.0, 5.0, g, l, 0.0);
    const double m0 = mass(s);
    for (int st = 0; st < 5; ++st) orc_stage(s, st);
    orc_step(s, steps);
    const double m1 = mass(s);
    char msg[128];
    std::snprintf(msg, sizeof msg, "mass %.15g -> %.15g", m0, m1);
    check(std::isfinite(m1) && std::fabs(m1 - m0) <= 1e-9 * m0, name, msg);
    std::vector<double> rho(n), f(19 * (size_t)n), v(3 * (size_t)n);
    orc_get(s, 1, rho.data());
    orc_get(s, 3, f.data());
    orc_get(s, 4, v.data());
    bool finite = true;
    for (double x : rho) finite = finite && std::isfinite(x);
    check(finite, name, "non-finite density");
    double ca[8];
    if (p.nwall_lo >= 1) orc_contact_angle(&p, rho.data(), g, l, p.ny / 2, ca);
    if (adjoint) {
        std::vector<double> target(rho), glo(nxy), ghi(nxy);
        double J = 0, gka = 0;
        orc_init_reference(s, p.nx / 4.0, 5.0, g, l, 0.0);
        const int rc = orc_adjoint(s, 4, target.data(), &J, glo.data(), ghi.data(), &gka);
        std::snprintf(msg, sizeof msg, "adjoint rc %d J %g", rc, J);
        check(rc == 0 ? std::isfinite(J) && std::isfinite(gka) : (p.forcing != 0 || p.wall_scheme == 2), name, msg);
    }
    orc_set(s, 1, rho.data());
    orc_init_from_rho(s, rho.data());
    orc_destroy(s);
}

}  // namespace

int main() {
    int ex[19], ey[19], ez[19], re[19];
    double w[19], M[361], R[361];
    orc_tables(ex, ey, ez, re, w, M, R);
    for (int i = 0; i < 1000; ++i) check(std::fabs(orc_noise(7, (uint64_t)i)) <= 1.0, "noise");
    check(orc_splitmix64(1234567) == 6457827717110365317ull, "splitmix64");
    for (int e = 0; e < 6; ++e) {
        double A, B, Tc, Rc;
        orc_set_multiphase(e, 0.8, &A, &B, &Tc, &Rc);
        const double d = 0.3 / B;   // inside every EOS's domain (B rho < 1)
        check(std::isfinite(orc_pressure(e, A, B, 0.8 * Tc, d)), "pressure");
        check(std::isfinite(orc_mu_eos(e, 0, A, B, 0.8 * Tc, d)) && std::isfinite(orc_mu_eos(e, 1, A, B, 0.8 * Tc, d)),
              "mu_eos");
    }
    run_case("PRW walls, layered", params(12, 10, 20, ORC_PRW, 0.7, 3, 1, 5, 0, 0.0), 6, true);
    run_case("PRW walls 2+2, simple, 7-point, Qp", params(10, 9, 18, ORC_PRW, 0.7, 2, 0, 7, 0, 0.003), 6, true);
    run_case("vdW walls, ChemBoundaryComplex, Guo", params(10, 8, 18, ORC_VDW, 0.8, 3, 2, 5, 1, 0.0), 6, true);
    run_case("RKE periodic", params(12, 8, 12, ORC_RKE, 0.85, 0, 1, 5, 0, 0.0), 6, true);
    run_case("CSE quasi-2-D", params(16, 1, 14, ORC_CSE, 0.85, 3, 1, 5, 0, 0.0), 6, false);
    if (failures) {
        std::fprintf(stderr, "asan_check: %d failures\n", failures);
        return 1;
    }
    std::printf("asan_check: all oracle entry points ran clean\n");
    return 0;
}
