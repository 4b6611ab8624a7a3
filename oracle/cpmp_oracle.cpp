/*
 * cpmp_oracle.cpp -- CPU FP64 (OpenMP) restatement of the reference step.
 *
 * TEST INFRASTRUCTURE ONLY (checker + CPU baseline); see cpmp_oracle.h.
 *
 * Reference files restated (zhou-yiling/3DCPMPmodel):
 *   Common.h:94-100      D3Q19 velocities, opposite map, weights
 *   Common.h:104-146     MRT matrix M and its inverse R
 *   Common.h:164-246     index macro, periodic neighbour macros, stencils
 *   FluidCuda.cu:189-193 Feq
 *   FluidCuda.cu:199-241 SetFlowField (CONTACTANGLE branch)
 *   FluidCuda.cu:245-347 LocalCollideMrt / GlobalCollide
 *   FluidCuda.cu:351-372 MacroCalculate
 *   FluidCuda.cu:376-401 ChemBoundary (COMPLEXCHEMBOUNDARY / SIMPLECHEMBOUNDARY)
 *   FluidCuda.cu:431-441 NonidealForce
 *   FluidCuda.cu:444-463 ChemPotential
 *   FluidCuda.cu:557-689 ContactAngle
 *   FluidCuda.cu:1275-1292 step order
 *   FluidCuda.cu:1314-1353 SetMultiphase constants
 *   FluidCuda.cu:1561-1573 EosPressure
 *   FluidCuda.cu:1626-1657 CalcMacroCPU
 *
 * Build decisions taken where the reference is ambiguous (the same in the HIP build):
 *   - ChemBoundary: the reference kernel reads neighbour columns that other
 *     threads write in the same launch (a race).  Here the fill runs as ordered
 *     layer passes: layer nwall-1 from the first fluid layer, then each deeper
 *     layer from the one just completed.
 *   - Non-PR EOS: the reference kernel evaluates the PR mu for every EOS.
 *     mu_form=1 reproduces that; mu_form=0 uses each EOS's own closed-form mu.
 *   - The Qp term (FluidCuda.cu:260-266) is evaluated only when K1 or K2 is
 *     non-zero; with K1=K2=0 it is a signed zero and leaves every moment unchanged.
 *   - nwall_lo = nwall_hi = 0 gives a fully periodic box (the reference always
 *     has 3+3 layers).
 *   - wall_scheme 2 runs ChemBoundaryComplex (commented out in the reference,
 *     FluidCuda.cu:403-429, 1281-1283) on level tags the reference never assigns:
 *     a solid node's level is its distance in layers to the fluid (LEVEL1..3 of
 *     Common.h:40-42 for the reference's three layers, deeper walls continue).
 *   - forcing 1 replaces the exact-difference force term by Guo's (Guo, Zheng and
 *     Shi 2002) in MRT form: equilibrium velocity (j + F/2)/rho, source
 *     (I - S/2) M Fbar with Fbar_f = w_f [3 (e_f - u) + 9 (e_f.u) e_f].F.
 *
 * Pinning (tests/test_oracle_*.py):
 *   tables      M rows orthogonal with the row norms of the d'Humieres basis, R*M = I
 *   Feq / Meq   M*Feq(rho,u) == Meq(rho,u); moment identities
 *   EOS         CoCurve golden rows (equal p, equal mu); Gibbs-Duhem by finite differences
 *   stencils    polynomial exactness (Laplacian to degree 5, gradient to degree 6)
 *   wall fill   constant and in-plane-linear fields reproduced; brute-force matrix form
 *   collide     mass/momentum identities, rest state fixed point, dense-matrix brute force
 *   reduce      exact sums on integer-valued fields
 *   contact ang wedges and caps of known angle, height and base (exact), tanh-profiled
 *               spherical caps within the method's discretisation bound
 *   init        SetFlowField against a NumPy restatement with the float32 radius
 *   noise       published SplitMix64 output sequence; uniformity of the [-1,1) map
 *   Guo         mass exact, momentum + F, agreement with the exact-difference force to O(F^2)
 *   adjoint     central finite differences of orc_step
 */
#include "cpmp_oracle.h"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>

namespace {

constexpr int Q = 19;
constexpr int FLUID = 1;
constexpr int SOLID = 2;

/* D3Q19 ordering of Common.h:96-98: 0 rest; 1-6 +-x,+-y,+-z; 7-10 xy diagonals;
   11-14 xz diagonals; 15-18 yz diagonals. */
const int EX[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
const int EY[Q] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
const int EZ[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};

struct Tables {
    int re[Q];
    double w[Q];
    double M[Q][Q];
    double R[Q][Q];
    Tables() {
        for (int i = 0; i < Q; ++i) {
            for (int j = 0; j < Q; ++j)
                if (EX[j] == -EX[i] && EY[j] == -EY[i] && EZ[j] == -EZ[i]) re[i] = j;
            int c2 = EX[i] * EX[i] + EY[i] * EY[i] + EZ[i] * EZ[i];
            w[i] = c2 == 0 ? 1. / 3 : (c2 == 1 ? 1. / 18 : 1. / 36);   /* Common.h:100 */
        }
        /* The 19 rows of the d'Humieres moment basis (Common.h:104-124), each a
           polynomial of the lattice velocity. */
        for (int i = 0; i < Q; ++i) {
            double x = EX[i], y = EY[i], z = EZ[i];
            double c2 = x * x + y * y + z * z;
            double row[Q] = {
                1.0,
                19 * c2 - 30,
                (21 * c2 * c2 - 53 * c2 + 24) / 2,
                x, (5 * c2 - 9) * x,
                y, (5 * c2 - 9) * y,
                z, (5 * c2 - 9) * z,
                3 * x * x - c2, (3 * c2 - 5) * (3 * x * x - c2),
                y * y - z * z, (3 * c2 - 5) * (y * y - z * z),
                x * y, y * z, x * z,
                (y * y - z * z) * x, (z * z - x * x) * y, (x * x - y * y) * z};
            for (int r = 0; r < Q; ++r) M[r][i] = row[r];
        }
        /* R = M^-1 = M^T diag(1/|row|^2) because the rows are orthogonal. */
        for (int r = 0; r < Q; ++r) {
            double n2 = 0;
            for (int i = 0; i < Q; ++i) n2 += M[r][i] * M[r][i];
            for (int i = 0; i < Q; ++i) R[i][r] = M[r][i] / n2;
        }
    }
};
const Tables TB;

inline int wrap(int a, int n) { a %= n; return a < 0 ? a + n : a; }

/* Feq, FluidCuda.cu:189-193 */
inline double feq(int f, double den, double vx, double vy, double vz) {
    double dot = vx * EX[f] + vy * EY[f] + vz * EZ[f];
    return den * TB.w[f] * (1.0 + 3.0 * dot + 4.5 * dot * dot - 1.5 * (vx * vx + vy * vy + vz * vz));
}

/* clamp of FluidCuda.cu:450-451 and its derivative */
inline double clamp_rho(double d) {
    const double R0 = 9.7, W0 = 1.5;
    if (d > R0) d = R0 + (d - R0) / (1.0 + (d - R0) * W0);
    return d;
}
inline double clamp_drho(double d) {
    const double R0 = 9.7, W0 = 1.5;
    if (d > R0) { double q = 1.0 + (d - R0) * W0; return 1.0 / (q * q); }
    return 1.0;
}

const double S2 = 1.4142135623730950488016887242097;   /* Common.h:153 */

/* mu of each EOS; mu_form 1 = FluidCuda.cu:453 for every EOS. den is already clamped. */
double mu_eos(int eos, int form, double A, double B, double T, double Den) {
    if (form == 1 || eos == ORC_PRW || eos == ORC_PRM) {
        /* FluidCuda.cu:453, literal */
        return T * log(Den / (1.0 - B * Den)) - A / (S2 * 2 * B) * log((S2 - 1 + B * Den) / (S2 + 1 - B * Den))
             + T / (1.0 - B * Den) - A * Den / (1.0 + B * Den * 2 - (B * Den) * (B * Den));
    }
    switch (eos) {
    case ORC_VDW:
        return T * log(Den / (1.0 - B * Den)) + T / (1.0 - B * Den) - 2.0 * A * Den;
    case ORC_RKE: case ORC_RKS:
        return T * log(Den / (1.0 - B * Den)) + T / (1.0 - B * Den)
             - A / B * log(1.0 + B * Den) - A * Den / (1.0 + B * Den);
    case ORC_CSE: {
        double e = B * Den, om = 1.0 - e;
        return T * (log(Den) + (8.0 * e - 9.0 * e * e + 3.0 * e * e * e) / (om * om * om)) - 2.0 * A * Den;
    }
    }
    return NAN;
}
/* d mu_eos / d Den (same form selection) */
double dmu_eos(int eos, int form, double A, double B, double T, double Den) {
    if (form == 1 || eos == ORC_PRW || eos == ORC_PRM) {
        double bd = B * Den, om = 1.0 - bd, den2 = 1.0 + 2 * bd - bd * bd;
        double d_log1 = 1.0 / Den + B / om;
        double d_log2 = B / (S2 - 1 + bd) + B / (S2 + 1 - bd);
        double d_rep = T * B / (om * om);
        double d_att = A * (den2 - Den * (2 * B - 2 * B * bd)) / (den2 * den2);
        return T * d_log1 - A / (S2 * 2 * B) * d_log2 + d_rep - d_att;
    }
    switch (eos) {
    case ORC_VDW: {
        double om = 1.0 - B * Den;
        return T * (1.0 / Den + B / om) + T * B / (om * om) - 2.0 * A;
    }
    case ORC_RKE: case ORC_RKS: {
        double om = 1.0 - B * Den, op = 1.0 + B * Den;
        return T * (1.0 / Den + B / om) + T * B / (om * om) - A / op - A / (op * op);
    }
    case ORC_CSE: {
        double e = B * Den, om = 1.0 - e;
        double num = 8.0 * e - 9.0 * e * e + 3.0 * e * e * e;
        double dnum = (8.0 - 18.0 * e + 9.0 * e * e) * B;
        return T * (1.0 / Den + (dnum * om + 3.0 * B * num) / (om * om * om * om)) - 2.0 * A;
    }
    }
    return NAN;
}

double pressure(int eos, double A, double B, double T, double d) {
    switch (eos) {
    case ORC_VDW: return T * d / (1.0 - B * d) - A * d * d;                       /* FluidCuda.cu:1567 */
    case ORC_PRW: case ORC_PRM:
        return T * d / (1.0 - B * d) - A * d * d / (1.0 + B * d * 2 - (B * d) * (B * d)); /* :1569 */
    case ORC_RKE: case ORC_RKS: return T * d / (1.0 - B * d) - A * d * d / (1.0 + B * d);
    case ORC_CSE: {
        double e = B * d;
        return d * T * (1 + e + e * e - e * e * e) / ((1 - e) * (1 - e) * (1 - e)) - A * d * d;
    }
    }
    return NAN;
}

}  // namespace

/* ======================================================================== */
struct orc_state {
    orc_params p;
    int64_t N;
    std::vector<int8_t> type;
    std::vector<double> rho, mu, f, ftmp, vx, vy, vz, mvx, mvy, mvz, fx, fy, fz, mu_lo, mu_hi;
    std::vector<int> sx[7], sy[7], sz[7];   /* wrapped coordinate for shift s-3 */

    int64_t I(int i, int j, int k) const { return ((int64_t)i * p.ny + j) * p.nz + k; }
    bool solid_k(int k) const { return k < p.nwall_lo || k > p.nz - 1 - p.nwall_hi; }

    /* value of a at (i,j,k) shifted by s*(dx,dy,dz), s in [-3,3] */
    double sh(const double *a, int i, int j, int k, int dx, int dy, int dz, int s) const {
        return a[I(sx[3 + s * dx][i], sy[3 + s * dy][j], sz[3 + s * dz][k])];
    }
    /* one 6th-order directional difference (the five groups of Gradx7 etc., Common.h:222-241) */
    double dd(const double *a, int i, int j, int k, int dx, int dy, int dz) const {
        return -sh(a, i, j, k, dx, dy, dz, -3) + sh(a, i, j, k, dx, dy, dz, -2) * 9
               - sh(a, i, j, k, dx, dy, dz, -1) * 45 + sh(a, i, j, k, dx, dy, dz, 1) * 45
               - sh(a, i, j, k, dx, dy, dz, 2) * 9 + sh(a, i, j, k, dx, dy, dz, 3);
    }
    double gx7(const double *a, int i, int j, int k) const {   /* Common.h:222-227 */
        return (dd(a, i, j, k, 1, 0, 0) * 2 + dd(a, i, j, k, 1, 0, 1) + dd(a, i, j, k, 1, 0, -1)
                + dd(a, i, j, k, 1, 1, 0) + dd(a, i, j, k, 1, -1, 0)) / 360;
    }
    double gy7(const double *a, int i, int j, int k) const {   /* Common.h:229-234 */
        return (dd(a, i, j, k, 0, 1, 0) * 2 + dd(a, i, j, k, 0, 1, 1) + dd(a, i, j, k, 0, 1, -1)
                + dd(a, i, j, k, 1, 1, 0) + dd(a, i, j, k, -1, 1, 0)) / 360;
    }
    double gz7(const double *a, int i, int j, int k) const {   /* Common.h:236-241 */
        return (dd(a, i, j, k, 0, 0, 1) * 2 + dd(a, i, j, k, 1, 0, 1) + dd(a, i, j, k, -1, 0, 1)
                + dd(a, i, j, k, 0, 1, 1) + dd(a, i, j, k, 0, -1, 1)) / 360;
    }
    /* GradD5, Common.h:189-192 */
    double lap5(const double *a, int i, int j, int k) const {
        auto ax = [&](int dx, int dy, int dz) {
            return -sh(a, i, j, k, dx, dy, dz, -2) + sh(a, i, j, k, dx, dy, dz, -1) * 16
                   + sh(a, i, j, k, dx, dy, dz, 1) * 16 - sh(a, i, j, k, dx, dy, dz, 2);
        };
        return (-a[I(i, j, k)] * 90 + ax(1, 0, 0) + ax(0, 1, 0) + ax(0, 0, 1)) / 12;
    }
    /* GradD7 numerator, Common.h:243-246 (the macro's /180 is applied by the caller) */
    double lap7num(const double *a, int i, int j, int k) const {
        auto ax = [&](int dx, int dy, int dz) {
            return 2 * sh(a, i, j, k, dx, dy, dz, -3) - 27 * sh(a, i, j, k, dx, dy, dz, -2)
                   + 270 * sh(a, i, j, k, dx, dy, dz, -1) + 270 * sh(a, i, j, k, dx, dy, dz, 1)
                   - 27 * sh(a, i, j, k, dx, dy, dz, 2) + 2 * sh(a, i, j, k, dx, dy, dz, 3);
        };
        return -a[I(i, j, k)] * 490 * 3 + ax(1, 0, 0) + ax(0, 1, 0) + ax(0, 0, 1);
    }
    /* -Ka * Laplacian term exactly as the reference composes it (FluidCuda.cu:457,460) */
    double ka_lap(const double *a, int i, int j, int k) const {
        if (p.lap_points == 7) return p.Ka * lap7num(a, i, j, k) / 180;
        return p.Ka * lap5(a, i, j, k);
    }
    /* in-plane diagonal average of layer kk at column (i,j), FluidCuda.cu:382 */
    double wavg(const double *a, int i, int j, int kk) const {
        int i1 = sx[2][i], i2 = sx[4][i], j1 = sy[2][j], j2 = sy[4][j];
        return (a[I(i, j, kk)] * 2 + a[I(i1, j1, kk)] + a[I(i1, j2, kk)] + a[I(i2, j1, kk)] + a[I(i2, j2, kk)]) / 6;
    }
};

namespace {

void build_shifts(orc_state *s) {
    for (int t = 0; t < 7; ++t) {
        s->sx[t].resize(s->p.nx); s->sy[t].resize(s->p.ny); s->sz[t].resize(s->p.nz);
        for (int i = 0; i < s->p.nx; ++i) s->sx[t][i] = wrap(i + t - 3, s->p.nx);
        for (int j = 0; j < s->p.ny; ++j) s->sy[t][j] = wrap(j + t - 3, s->p.ny);
        for (int k = 0; k < s->p.nz; ++k) s->sz[t][k] = wrap(k + t - 3, s->p.nz);
    }
}

/* ChemBoundary, FluidCuda.cu:376-401, as ordered layer passes. */
void stage_wall(orc_state *s) {
    const orc_params &p = s->p;
    double *rho = s->rho.data(), *mu = s->mu.data();
    const int lo = p.nwall_lo, hi = p.nwall_hi, nz = p.nz;
    if (lo == 0 && hi == 0) return;
    const int nl = std::max(lo, hi);
    if (p.wall_scheme == 2) {
        /* ChemBoundaryComplex(Type, Dens, LEVELd, LEVEL(d-1)), d = 1, 2, ...: the
           average over f = 1..18 of the neighbours tagged d-1, weights Alpha */
        auto level = [&](int k) { return k < lo ? lo - k : (k > nz - 1 - hi ? k - (nz - hi) + 1 : 0); };
        for (int d = 1; d <= nl; ++d) {
            #pragma omp parallel for collapse(2) schedule(static)
            for (int i = 0; i < p.nx; ++i)
                for (int j = 0; j < p.ny; ++j)
                    for (int k = 0; k < nz; ++k) {
                        if (level(k) != d) continue;
                        double avg_den = 0, w = 0;
                        for (int f = 1; f < Q; ++f) {
                            const int ii = wrap(i + EX[f], p.nx), jj = wrap(j + EY[f], p.ny), kk = wrap(k + EZ[f], nz);
                            if (level(kk) == d - 1) {
                                avg_den += TB.w[f] * rho[s->I(ii, jj, kk)];
                                w += TB.w[f];
                            }
                        }
                        rho[s->I(i, j, k)] = avg_den / w;
                    }
        }
    }
    for (int pass = 0; pass < (p.wall_scheme == 2 ? 0 : nl); ++pass) {
        #pragma omp parallel for collapse(2) schedule(static)
        for (int i = 0; i < p.nx; ++i)
            for (int j = 0; j < p.ny; ++j) {
                if (pass < lo) {
                    int kd = lo - 1 - pass;                       /* layer written */
                    int ks = p.wall_scheme == 1 ? kd + 1 : lo;    /* layer read */
                    rho[s->I(i, j, kd)] = s->wavg(rho, i, j, ks);
                }
                if (pass < hi) {
                    int kd = nz - hi + pass;
                    int ks = p.wall_scheme == 1 ? kd - 1 : nz - 1 - hi;
                    rho[s->I(i, j, kd)] = s->wavg(rho, i, j, ks);
                }
            }
        /* the layer pass completes before the next one reads it */
    }
    #pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < p.nx; ++i)
        for (int j = 0; j < p.ny; ++j) {
            for (int k = 0; k < lo; ++k) mu[s->I(i, j, k)] = s->mu_lo[(int64_t)i * p.ny + j];
            for (int k = nz - hi; k < nz; ++k) mu[s->I(i, j, k)] = s->mu_hi[(int64_t)i * p.ny + j];
        }
}

/* ChemPotential, FluidCuda.cu:444-463 */
void stage_mu(orc_state *s) {
    const orc_params &p = s->p;
    const double *rho = s->rho.data();
    double *mu = s->mu.data();
    #pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < p.nx; ++i)
        for (int j = 0; j < p.ny; ++j)
            for (int k = 0; k < p.nz; ++k) {
                int64_t I = s->I(i, j, k);
                if (s->type[I] != FLUID) continue;
                double Den = clamp_rho(rho[I]);
                double m = mu_eos(p.eos, p.mu_form, p.A, p.B, p.T, Den);
                mu[I] = p.K * p.K * m - s->ka_lap(rho, i, j, k);
            }
}

/* NonidealForce, FluidCuda.cu:431-441 */
void stage_force(orc_state *s) {
    const orc_params &p = s->p;
    const double *rho = s->rho.data(), *mu = s->mu.data();
    #pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < p.nx; ++i)
        for (int j = 0; j < p.ny; ++j)
            for (int k = 0; k < p.nz; ++k) {
                int64_t I = s->I(i, j, k);
                if (s->type[I] != FLUID) continue;
                s->fx[I] = -s->gx7(mu, i, j, k) * rho[I] + s->gx7(rho, i, j, k) / 3;
                s->fy[I] = -s->gy7(mu, i, j, k) * rho[I] + s->gy7(rho, i, j, k) / 3;
                s->fz[I] = -s->gz7(mu, i, j, k) * rho[I] + s->gz7(rho, i, j, k) / 3;
            }
}

/* GlobalCollide + LocalCollideMrt, FluidCuda.cu:245-347, then the swap of :1290 */
void stage_collide(orc_state *s) {
    const orc_params &p = s->p;
    const int64_t N = s->N;
    const double *rho = s->rho.data();
    double *dist = s->f.data(), *temp = s->ftmp.data();
    const bool qp_on = (p.K1 != 0.0 || p.K2 != 0.0);
    const bool guo = p.forcing == 1;
    #pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < p.nx; ++i)
        for (int j = 0; j < p.ny; ++j)
            for (int k = 0; k < p.nz; ++k) {
                int64_t I = s->I(i, j, k);
                if (s->type[I] != FLUID) continue;
                const double Vx0 = s->vx[I], Vy0 = s->vy[I], Vz0 = s->vz[I];
                s->mvx[I] = Vx0 + s->fx[I] / rho[I] * 0.5;
                s->mvy[I] = Vy0 + s->fy[I] / rho[I] * 0.5;
                s->mvz[I] = Vz0 + s->fz[I] / rho[I] * 0.5;
                /* Guo: the equilibrium velocity carries half the force */
                const double Vx = guo ? s->mvx[I] : Vx0, Vy = guo ? s->mvy[I] : Vy0, Vz = guo ? s->mvz[I] : Vz0;
                double EVx = Vx + s->fx[I] / rho[I];
                double EVy = Vy + s->fy[I] / rho[I];
                double EVz = Vz + s->fz[I] / rho[I];

                /* LocalCollideMrt */
                double Den = rho[I];
                double Meq[Q], Mf[Q], Qp[Q] = {0};
                if (qp_on) {
                    double gx = s->gx7(rho, i, j, k), gy = s->gy7(rho, i, j, k), gz = s->gz7(rho, i, j, k);
                    Qp[1] = -19 * (p.K1 + 3 * p.K2) * (gx * gx + gy * gy + gz * gz);
                    Qp[9] = -p.K1 * (2 * gx * gx - gy * gy - gz * gz);
                    Qp[11] = -p.K1 * (gy * gy - gz * gz);
                    Qp[13] = -p.K1 * gx * gy;
                    Qp[14] = -p.K1 * gz * gy;
                    Qp[15] = -p.K1 * gx * gz;
                }
                double usq = Vx * Vx + Vy * Vy + Vz * Vz;
                Meq[0] = Den;
                Meq[1] = Den * (-11.0 + 19.0 * usq);
                Meq[2] = Den * (3.0 - 11.0 / 2 * usq);
                Meq[3] = Den * Vx;
                Meq[4] = Den * Vx * -2 / 3;
                Meq[5] = Den * Vy;
                Meq[6] = Den * Vy * -2 / 3;
                Meq[7] = Den * Vz;
                Meq[8] = Den * Vz * -2 / 3;
                Meq[9] = Den * (Vx * Vx * 2 - Vy * Vy - Vz * Vz);
                Meq[10] = Den * (Vx * Vx * 2 - Vy * Vy - Vz * Vz) / -2;
                Meq[11] = Den * (Vy * Vy - Vz * Vz);
                Meq[12] = Den * (Vy * Vy - Vz * Vz) / -2;
                Meq[13] = Den * Vx * Vy;
                Meq[14] = Den * Vy * Vz;
                Meq[15] = Den * Vx * Vz;
                Meq[16] = Meq[17] = Meq[18] = 0;
                for (int f = 0; f < Q; ++f) {
                    Mf[f] = 0;
                    for (int q = 0; q < Q; ++q) Mf[f] += TB.M[f][q] * dist[q * N + I];
                }
                for (int f = 0; f < Q; ++f)
                    Mf[f] = Mf[f] - p.Ts[f] * (Mf[f] - Meq[f]) + Qp[f] * p.Ts[f];
                if (guo) {
                    /* (I - S/2) M Fbar */
                    const double Fx = s->fx[I], Fy = s->fy[I], Fz = s->fz[I];
                    double Fb[Q];
                    for (int q = 0; q < Q; ++q) {
                        const double eu = EX[q] * Vx + EY[q] * Vy + EZ[q] * Vz;
                        Fb[q] = TB.w[q] * (3 * ((EX[q] - Vx) * Fx + (EY[q] - Vy) * Fy + (EZ[q] - Vz) * Fz)
                                           + 9 * eu * (EX[q] * Fx + EY[q] * Fy + EZ[q] * Fz));
                    }
                    for (int f = 0; f < Q; ++f) {
                        double G = 0;
                        for (int q = 0; q < Q; ++q) G += TB.M[f][q] * Fb[q];
                        Mf[f] += (1 - p.Ts[f] / 2) * G;
                    }
                }
                double post[Q];
                for (int f = 0; f < Q; ++f) {
                    post[f] = 0;
                    for (int q = 0; q < Q; ++q) post[f] += TB.R[f][q] * Mf[q];
                }
                /* EDM force + push streaming with half-way bounce-back */
                for (int f = 0; f < Q; ++f) {
                    int ii = s->sx[3 + EX[f]][i], jj = s->sy[3 + EY[f]][j], kk = s->sz[3 + EZ[f]][k];
                    double Df = guo ? post[f] : post[f] + feq(f, Den, EVx, EVy, EVz) - feq(f, Den, Vx, Vy, Vz);
                    int64_t T = s->I(ii, jj, kk);
                    if (s->type[T] == FLUID) temp[f * N + T] = Df;
                    else temp[(int64_t)TB.re[f] * N + I] = Df;
                }
            }
    std::swap(s->f, s->ftmp);
}

/* MacroCalculate, FluidCuda.cu:351-372 */
void stage_macro(orc_state *s) {
    const int64_t N = s->N;
    const double *d = s->f.data();
    #pragma omp parallel for schedule(static)
    for (int64_t I = 0; I < N; ++I) {
        if (s->type[I] != FLUID) continue;
        double den = 0;
        for (int f = 0; f < Q; ++f) den += d[f * N + I];
        s->rho[I] = den;
        auto DF = [&](int q) { return d[(int64_t)q * N + I]; };
        s->vx[I] = (DF(1) + DF(7) + DF(9) + DF(11) + DF(13) - DF(2) - DF(8) - DF(10) - DF(12) - DF(14)) / den;
        s->vy[I] = (DF(3) + DF(7) + DF(8) + DF(15) + DF(17) - DF(4) - DF(9) - DF(10) - DF(16) - DF(18)) / den;
        s->vz[I] = (DF(5) + DF(11) + DF(12) + DF(15) + DF(16) - DF(6) - DF(13) - DF(14) - DF(17) - DF(18)) / den;
    }
}

uint64_t splitmix64(uint64_t x) {
    /* SplitMix64 (Steele, Lea, Flood 2014) */
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

}  // namespace

/* ======================================================================== */
extern "C" {

int orc_set_multiphase(int eos, double Tr, double *A, double *B, double *Tc, double *Rc) {
    /* FluidCuda.cu:1318-1344 */
    switch (eos) {
    case ORC_VDW: *A = 9.0 / 49; *B = 2.0 / 21; *Tc = 4.0 / 7; *Rc = 7.0 / 2; break;
    case ORC_CSE: *A = 1.0; *B = 1.0; *Tc = 0.09432870; *Rc = 0.13044388; break;
    case ORC_RKE:
        *A = 2.0 / 49; *B = 2.0 / 21; *Tc = 0.196133; *Rc = 2.729171;
        *A *= 1.0 / sqrt(*Tc * Tr);
        break;
    case ORC_RKS: {
        *A = 2.0 / 49; *B = 2.0 / 21; *Tc = 0.086861; *Rc = 2.729171; double w = 0.344;
        double m = 1.0 + (0.480 + 1.574 * w - 0.176 * w * w) * (1.0 - sqrt(Tr));
        *A *= m * m;
    } break;
    case ORC_PRW: case ORC_PRM: {
        *A = 2.0 / 49; *B = 2.0 / 21; *Tc = 0.072919; *Rc = 2.65730416;
        double w = eos == ORC_PRW ? 0.344 : 0.011;
        double m = 1.0 + (0.37464 + 1.54226 * w - 0.26992 * w * w) * (1.0 - sqrt(Tr));
        *A *= m * m;
    } break;
    default: return -1;
    }
    return 0;
}

double orc_pressure(int eos, double A, double B, double T, double rho) { return pressure(eos, A, B, T, rho); }
double orc_mu_eos(int eos, int form, double A, double B, double T, double rho) {
    return mu_eos(eos, form, A, B, T, rho);
}

/* Maxwell construction by Newton on (rho_g, rho_l): p(g)=p(l), mu(g)=mu(l).
   The caller supplies starting values in *rho_g, *rho_l. */
int orc_maxwell(int eos, double A, double B, double T, double *rho_g, double *rho_l) {
    double g = *rho_g, l = *rho_l;
    for (int it = 0; it < 200; ++it) {
        double pg = pressure(eos, A, B, T, g), pl = pressure(eos, A, B, T, l);
        double mg = mu_eos(eos, 0, A, B, T, g), ml = mu_eos(eos, 0, A, B, T, l);
        double dmg = dmu_eos(eos, 0, A, B, T, g), dml = dmu_eos(eos, 0, A, B, T, l);
        /* dp/drho = rho * dmu/drho (Gibbs-Duhem) */
        double a11 = g * dmg, a12 = -l * dml, a21 = dmg, a22 = -dml;
        double r1 = pg - pl, r2 = mg - ml;
        double det = a11 * a22 - a12 * a21;
        if (det == 0 || !std::isfinite(det)) return -1;
        double dg = (r1 * a22 - a12 * r2) / det, dl = (a11 * r2 - a21 * r1) / det;
        double sc = 1.0;
        while ((g - sc * dg <= 0 || l - sc * dl <= 0 || B * (l - sc * dl) >= 1.0) && sc > 1e-6) sc *= 0.5;
        g -= sc * dg; l -= sc * dl;
        if (fabs(dg) < 1e-15 * g && fabs(dl) < 1e-15 * l) break;
    }
    *rho_g = g; *rho_l = l;
    return (std::isfinite(g) && std::isfinite(l)) ? 0 : -1;
}

void orc_tables(int *ex, int *ey, int *ez, int *re, double *w, double *M, double *R) {
    for (int i = 0; i < Q; ++i) {
        ex[i] = EX[i]; ey[i] = EY[i]; ez[i] = EZ[i]; re[i] = TB.re[i]; w[i] = TB.w[i];
        for (int j = 0; j < Q; ++j) { M[i * Q + j] = TB.M[i][j]; R[i * Q + j] = TB.R[i][j]; }
    }
}

orc_state *orc_create(const orc_params *p) {
    if (!p || p->nx <= 0 || p->ny <= 0 || p->nz <= 0) return nullptr;
    if (p->nwall_lo < 0 || p->nwall_hi < 0 || p->nwall_lo + p->nwall_hi >= p->nz) return nullptr;
    orc_state *s = new orc_state();
    s->p = *p;
    s->N = (int64_t)p->nx * p->ny * p->nz;
    const int64_t N = s->N;
    s->type.assign(N, FLUID);
    for (auto *v : {&s->rho, &s->mu, &s->vx, &s->vy, &s->vz, &s->mvx, &s->mvy, &s->mvz, &s->fx, &s->fy, &s->fz})
        v->assign(N, 0.0);
    s->f.assign(Q * N, 0.0);
    s->ftmp.assign(Q * N, 0.0);
    s->mu_lo.assign((int64_t)p->nx * p->ny, 0.0);
    s->mu_hi.assign((int64_t)p->nx * p->ny, 0.0);
    build_shifts(s);
    for (int i = 0; i < p->nx; ++i)
        for (int j = 0; j < p->ny; ++j)
            for (int k = 0; k < p->nz; ++k)
                s->type[s->I(i, j, k)] = s->solid_k(k) ? SOLID : FLUID;
    return s;
}

void orc_destroy(orc_state *s) { delete s; }

int orc_get(orc_state *s, int field, double *out) {
    const int64_t N = s->N;
    switch (field) {
    case ORC_TYPE: for (int64_t I = 0; I < N; ++I) out[I] = s->type[I]; break;
    case ORC_RHO: std::copy(s->rho.begin(), s->rho.end(), out); break;
    case ORC_MU: std::copy(s->mu.begin(), s->mu.end(), out); break;
    case ORC_F: std::copy(s->f.begin(), s->f.end(), out); break;
    case ORC_VEL:
        std::copy(s->vx.begin(), s->vx.end(), out);
        std::copy(s->vy.begin(), s->vy.end(), out + N);
        std::copy(s->vz.begin(), s->vz.end(), out + 2 * N); break;
    case ORC_MV:
        std::copy(s->mvx.begin(), s->mvx.end(), out);
        std::copy(s->mvy.begin(), s->mvy.end(), out + N);
        std::copy(s->mvz.begin(), s->mvz.end(), out + 2 * N); break;
    case ORC_FORCE:
        std::copy(s->fx.begin(), s->fx.end(), out);
        std::copy(s->fy.begin(), s->fy.end(), out + N);
        std::copy(s->fz.begin(), s->fz.end(), out + 2 * N); break;
    default: return -1;
    }
    return 0;
}

int orc_set(orc_state *s, int field, const double *in) {
    const int64_t N = s->N;
    switch (field) {
    case ORC_RHO: std::copy(in, in + N, s->rho.begin()); break;
    case ORC_MU: std::copy(in, in + N, s->mu.begin()); break;
    case ORC_F: std::copy(in, in + Q * N, s->f.begin()); break;
    case ORC_VEL:
        std::copy(in, in + N, s->vx.begin());
        std::copy(in + N, in + 2 * N, s->vy.begin());
        std::copy(in + 2 * N, in + 3 * N, s->vz.begin()); break;
    case ORC_FORCE:
        std::copy(in, in + N, s->fx.begin());
        std::copy(in + N, in + 2 * N, s->fy.begin());
        std::copy(in + 2 * N, in + 3 * N, s->fz.begin()); break;
    default: return -1;
    }
    return 0;
}

int orc_set_wall_mu(orc_state *s, const double *lo, const double *hi) {
    const int64_t n = (int64_t)s->p.nx * s->p.ny;
    std::copy(lo, lo + n, s->mu_lo.begin());
    std::copy(hi ? hi : lo, (hi ? hi : lo) + n, s->mu_hi.begin());
    return 0;
}

int orc_init_reference(orc_state *s, double radius, double width, double den_g, double den_l, double basept_init) {
    const orc_params &p = s->p;
    const int64_t N = s->N;
    const int DX = p.nx, DY = p.ny;
    for (int i = 0; i < p.nx; ++i)
        for (int j = 0; j < p.ny; ++j)
            for (int k = 0; k < p.nz; ++k) {
                int64_t I = s->I(i, j, k);
                if (s->solid_k(k)) {
                    s->rho[I] = s->mu[I] = basept_init;
                } else {
                    s->mu[I] = 0;
                    double sq = (double(i) - DX / 2) * (double(i) - DX / 2) + (double(j) - DY / 2) * (double(j) - DY / 2)
                                + (double(k) - (radius + 2)) * (double(k) - (radius + 2));
                    double r = (double)sqrtf((float)sq);   /* FluidCuda.cu:215 applies sqrtf */
                    s->rho[I] = (den_l + den_g) / 2 - (den_l - den_g) / 2 * tanh(double(r - radius) * 2 / width);
                }
                for (int f = 0; f < Q; ++f) s->f[f * N + I] = feq(f, s->rho[I], 0, 0, 0);
                s->vx[I] = s->vy[I] = s->vz[I] = 0;
                s->fx[I] = s->fy[I] = s->fz[I] = 0;
            }
    return 0;
}

int orc_init_from_rho(orc_state *s, const double *rho) {
    const int64_t N = s->N;
    for (int64_t I = 0; I < N; ++I) {
        s->rho[I] = rho[I];
        for (int f = 0; f < Q; ++f) s->f[f * N + I] = feq(f, rho[I], 0, 0, 0);
        s->vx[I] = s->vy[I] = s->vz[I] = 0;
        s->fx[I] = s->fy[I] = s->fz[I] = 0;
        s->mvx[I] = s->mvy[I] = s->mvz[I] = 0;
    }
    return 0;
}

int orc_stage(orc_state *s, int stage) {
    switch (stage) {
    case 0: stage_wall(s); break;
    case 1: stage_mu(s); break;
    case 2: stage_force(s); break;
    case 3: stage_collide(s); break;
    case 4: stage_macro(s); break;
    default: return -1;
    }
    return 0;
}

int orc_step(orc_state *s, int nsteps) {
    for (int n = 0; n < nsteps; ++n) {
        stage_wall(s);     /* FluidCuda.cu:1279 */
        stage_mu(s);       /* :1285 */
        stage_force(s);    /* :1288 */
        stage_collide(s);  /* :1289-1290 */
        stage_macro(s);    /* :1291 */
    }
    return 0;
}

int orc_reduce(orc_state *s, double *out) {
    /* CalcMacroCPU, FluidCuda.cu:1626-1657: sequential FP64 sum in (i,j,k) order */
    double mass = 0, bad = 0, ms = 0;
    for (int64_t I = 0; I < s->N; ++I) {
        if (s->type[I] != FLUID) continue;
        double d = s->rho[I];
        if (d != d || d < 0) bad += 1;
        mass += d;
    }
    for (int64_t I = 0; I < s->N; ++I) {
        if (s->type[I] != FLUID) continue;
        double m = s->mvx[I] * s->mvx[I] + s->mvy[I] * s->mvy[I] + s->mvz[I] * s->mvz[I];
        if (m > ms) ms = m;
    }
    out[0] = mass; out[1] = bad; out[2] = sqrt(ms);
    return 0;
}

int orc_contact_angle(const orc_params *p, const double *rho, double den_g, double den_l, int tier, double *out) {
    /* ContactAngle(side=0, tier), FluidCuda.cu:557-689, on the XZ plane y=tier.  The
       reference scans rows z = 2..5 and extrapolates to z = 2, the last of its three
       solid layers; here the rows are offset by the bottom wall depth (begin = lo - 1),
       which is the reference for lo = 3. */
    const int DX = p->nx, DY = p->ny, DZ = p->nz;
    if (p->nwall_lo < 1 || p->nwall_lo + 4 >= DZ) return -1;
    const double PI2 = 3.14159265358979323846264338327950288;
    const double Mid = (den_g + den_l) / 2;
    auto T = [&](int i, int k) { return rho[((int64_t)i * DY + tier) * DZ + k]; };
    double Px[8] = {0};
    const int begin = p->nwall_lo - 1;
    for (int j = begin; j <= begin + 3; ++j) {
        int n = (j - begin) * 2;
        for (int i = 0; i < DX / 2 + 2; ++i) {
            double t1 = T(i + 1, j), t2 = T(i + 2, j);
            if (t1 < Mid && t2 >= Mid) { Px[n] = double(i + 1) + (t1 - Mid) / (t1 - t2); break; }
        }
        for (int i = DX - 2; i > DX / 2 - 2; --i) {
            double t1 = T(i - 1, j), t2 = T(i - 2, j);
            if (t1 < Mid && t2 >= Mid) { Px[n + 1] = double(i - 1) - (t1 - Mid) / (t1 - t2); break; }
        }
    }
    double y0 = begin, y1 = begin + 1, y2 = begin + 2;
    double CaLeft, CaRight, CaCap = 0, XLeft, XRight, H = NAN, L = 0;
    int valid = 0;
    if (Px[0] != 0.0 && Px[1] != 0.0) {
        valid = 1;
        XLeft = Px[2] * ((y0 - y2) / (y1 - y2)) + Px[4] * ((y0 - y1) / (y2 - y1));
        XRight = Px[3] * ((y0 - y2) / (y1 - y2)) + Px[5] * ((y0 - y1) / (y2 - y1));
        double Ca = atan(double(y1 - y0) / (Px[2] - XLeft));
        if (Ca < 0) Ca += PI2;
        CaLeft = Ca / PI2 * 180;
        Ca = atan(-double(y1 - y0) / (Px[3] - XRight));
        if (Ca < 0) Ca += PI2;
        CaRight = Ca / PI2 * 180;
        L = XRight - XLeft;
        for (int i = int((Px[4] + Px[5]) / 2 + 0.5), j = begin; j < DZ - 2; ++j) {
            double t = T(i, j), t1 = T(i, j + 1);
            if (t >= Mid && t1 < Mid) H = double(j) + (t - Mid) / (t - t1) - y0;
        }
        if (L <= 0) { L = Px[3] - Px[2]; H = H - 0.5; }
        if (XLeft > DY / 2 && XRight < DY / 2) { XLeft = DY / 2; XRight = DY / 2; }
        double R = (H * H * 4 + L * L) / H / 8;
        CaCap = atan(L / (R - H) / 2);
        if (CaCap < 0) CaCap += PI2;
        CaCap = CaCap / PI2 * 180;
    } else {
        XLeft = DY / 2; XRight = DY / 2; CaLeft = 0; CaRight = 0;
    }
    out[0] = CaLeft; out[1] = CaRight; out[2] = CaCap; out[3] = XLeft; out[4] = XRight; out[5] = H; out[6] = L;
    out[7] = valid;
    return 0;
}

uint64_t orc_splitmix64(uint64_t x) { return splitmix64(x); }

double orc_noise(uint64_t seed, uint64_t index) {
    uint64_t r = splitmix64(seed ^ splitmix64(index));
    return (double)(r >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
}

}  // extern "C"

/* ======================================================================== */
/* Discrete adjoint (reverse mode of orc_step).  Derivation:
 *   state between steps: (f, rho_fluid, V_fluid); solid rho is rebuilt by the
 *   wall fill each step, solid mu is the wall field.
 *   M  : rho' = sum f', V' = j(f')/rho'        -> f'_bar += rho'_bar + V'_bar.(e-V')/rho'
 *   S  : push/bounce permutation                -> pull of f'_bar
 *   X  : Df = R[(I-S)Mf + S Meq(rho,V) + S Qp(G rho)] + Feq(rho,V+F/rho) - Feq(rho,V)
 *   F  : F = -rho G(mu) + G(rho)/3              -> G^T = -G (antisymmetric stencils)
 *   C  : mu = K^2 mu_eos(clamp rho) - Ka Lap(rho) (fluid); mu = mu_w (solid)
 *   W  : solid rho = layered in-plane averages of the fluid layer (symmetric avg)
 */
namespace {

struct Tape { std::vector<double> rho, vx, vy, vz; };

/* forward one step, recording rho (after the wall fill) and V at the start */
void fwd_record(orc_state *s, Tape &t) {
    stage_wall(s);
    t.rho = s->rho; t.vx = s->vx; t.vy = s->vy; t.vz = s->vz;
    stage_mu(s); stage_force(s); stage_collide(s); stage_macro(s);
}

}  // namespace

extern "C" int orc_adjoint(orc_state *s, int nsteps, const double *target, double *J,
                           double *g_mu_lo, double *g_mu_hi, double *g_ka) {
    const orc_params &p = s->p;
    if (p.forcing != 0 || p.wall_scheme == 2) return -1;
    const int64_t N = s->N;
    const int64_t NXY = (int64_t)p.nx * p.ny;
    std::vector<Tape> tape(nsteps + 1);
    for (int n = 0; n < nsteps; ++n) fwd_record(s, tape[n]);
    /* final state (rho^N, V^N) */
    tape[nsteps].rho = s->rho; tape[nsteps].vx = s->vx; tape[nsteps].vy = s->vy; tape[nsteps].vz = s->vz;

    double j = 0;
    std::vector<double> rb(N, 0.0), vxb(N, 0.0), vyb(N, 0.0), vzb(N, 0.0), fb(Q * N, 0.0);
    for (int64_t I = 0; I < N; ++I)
        if (s->type[I] == FLUID) { double d = s->rho[I] - target[I]; j += 0.5 * d * d; rb[I] = d; }
    *J = j;
    std::fill(g_mu_lo, g_mu_lo + NXY, 0.0);
    std::fill(g_mu_hi, g_mu_hi + NXY, 0.0);
    double gka = 0;

    /* scratch */
    std::vector<double> mu(N), Fx(N), Fy(N), Fz(N), fpb(Q * N), mub(N), Wx(N), Wy(N), Wz(N), Px(N), Py(N), Pz(N);
    std::vector<double> rbn(N), rho_save(N);
    const bool qp_on = (p.K1 != 0.0 || p.K2 != 0.0);

    for (int n = nsteps - 1; n >= 0; --n) {
        const Tape &cur = tape[n], &nxt = tape[n + 1];
        /* ---- reverse M of step n: f^{n+1}_bar from rho^{n+1}_bar, V^{n+1}_bar ---- */
        #pragma omp parallel for schedule(static)
        for (int64_t I = 0; I < N; ++I) {
            if (s->type[I] != FLUID) { for (int q = 0; q < Q; ++q) fpb[q * N + I] = 0; continue; }
            double r = nxt.rho[I];
            for (int q = 0; q < Q; ++q)
                fpb[q * N + I] = fb[q * N + I] + rb[I]
                    + (vxb[I] * (EX[q] - nxt.vx[I]) + vyb[I] * (EY[q] - nxt.vy[I]) + vzb[I] * (EZ[q] - nxt.vz[I])) / r;
        }
        /* ---- recompute mu, F of step n from the tape ---- */
        std::copy(cur.rho.begin(), cur.rho.end(), s->rho.begin());
        stage_mu(s);   /* fluid mu; solid mu from the wall fields */
        for (int i = 0; i < p.nx; ++i)
            for (int jj = 0; jj < p.ny; ++jj)
                for (int k = 0; k < p.nz; ++k) {
                    int64_t I = s->I(i, jj, k);
                    if (k < p.nwall_lo) s->mu[I] = s->mu_lo[(int64_t)i * p.ny + jj];
                    else if (k > p.nz - 1 - p.nwall_hi) s->mu[I] = s->mu_hi[(int64_t)i * p.ny + jj];
                }
        stage_force(s);
        /* ---- reverse S and X per fluid node ---- */
        std::fill(rbn.begin(), rbn.end(), 0.0);
        std::fill(Px.begin(), Px.end(), 0.0); std::fill(Py.begin(), Py.end(), 0.0); std::fill(Pz.begin(), Pz.end(), 0.0);
        std::fill(Wx.begin(), Wx.end(), 0.0); std::fill(Wy.begin(), Wy.end(), 0.0); std::fill(Wz.begin(), Wz.end(), 0.0);
        std::vector<double> nvxb(N, 0.0), nvyb(N, 0.0), nvzb(N, 0.0), nfb(Q * N, 0.0);
        #pragma omp parallel for collapse(2) schedule(static)
        for (int i = 0; i < p.nx; ++i)
            for (int jj = 0; jj < p.ny; ++jj)
                for (int k = 0; k < p.nz; ++k) {
                    int64_t I = s->I(i, jj, k);
                    if (s->type[I] != FLUID) continue;
                    double Db[Q];
                    for (int f = 0; f < Q; ++f) {
                        int ii = s->sx[3 + EX[f]][i], j2 = s->sy[3 + EY[f]][jj], kk = s->sz[3 + EZ[f]][k];
                        int64_t T = s->I(ii, j2, kk);
                        Db[f] = (s->type[T] == FLUID) ? fpb[f * N + T] : fpb[(int64_t)TB.re[f] * N + I];
                    }
                    double rho = cur.rho[I], ux = cur.vx[I], uy = cur.vy[I], uz = cur.vz[I];
                    double Fxa = s->fx[I], Fya = s->fy[I], Fza = s->fz[I];
                    double ex = ux + Fxa / rho, ey = uy + Fya / rho, ez = uz + Fza / rho;
                    /* m*_bar = R^T Db */
                    double mb[Q];
                    for (int a = 0; a < Q; ++a) { mb[a] = 0; for (int f = 0; f < Q; ++f) mb[a] += TB.R[f][a] * Db[f]; }
                    /* f_bar = M^T (1-S) mb */
                    for (int q = 0; q < Q; ++q) {
                        double acc = 0;
                        for (int a = 0; a < Q; ++a) acc += TB.M[a][q] * (1.0 - p.Ts[a]) * mb[a];
                        nfb[q * N + I] = acc;
                    }
                    double qv[Q];
                    for (int a = 0; a < Q; ++a) qv[a] = p.Ts[a] * mb[a];
                    /* Meq(rho,u) partials */
                    double usq = ux * ux + uy * uy + uz * uz;
                    double g9 = 2 * ux * ux - uy * uy - uz * uz, g11 = uy * uy - uz * uz;
                    double meq_over_rho[Q] = {1, -11 + 19 * usq, 3 - 5.5 * usq, ux, -2.0 / 3 * ux, uy, -2.0 / 3 * uy,
                                              uz, -2.0 / 3 * uz, g9, -0.5 * g9, g11, -0.5 * g11, ux * uy, uy * uz, ux * uz, 0, 0, 0};
                    double rbar = 0, uxb = 0, uyb = 0, uzb = 0;
                    for (int a = 0; a < Q; ++a) rbar += meq_over_rho[a] * qv[a];
                    uxb += rho * (38 * ux * qv[1] - 11 * ux * qv[2] + qv[3] - 2.0 / 3 * qv[4] + 4 * ux * qv[9] - 2 * ux * qv[10]
                                  + uy * qv[13] + uz * qv[15]);
                    uyb += rho * (38 * uy * qv[1] - 11 * uy * qv[2] + qv[5] - 2.0 / 3 * qv[6] - 2 * uy * qv[9] + uy * qv[10]
                                  + 2 * uy * qv[11] - uy * qv[12] + ux * qv[13] + uz * qv[14]);
                    uzb += rho * (38 * uz * qv[1] - 11 * uz * qv[2] + qv[7] - 2.0 / 3 * qv[8] - 2 * uz * qv[9] + uz * qv[10]
                                  - 2 * uz * qv[11] + uz * qv[12] + uy * qv[14] + ux * qv[15]);
                    /* EDM: Feq(rho,EV) - Feq(rho,u) */
                    double evb_x = 0, evb_y = 0, evb_z = 0;
                    for (int f = 0; f < Q; ++f) {
                        double w = TB.w[f];
                        double de = ex * EX[f] + ey * EY[f] + ez * EZ[f];
                        double du = ux * EX[f] + uy * EY[f] + uz * EZ[f];
                        double esq = ex * ex + ey * ey + ez * ez;
                        double pe = w * (1 + 3 * de + 4.5 * de * de - 1.5 * esq);
                        double pu = w * (1 + 3 * du + 4.5 * du * du - 1.5 * usq);
                        rbar += Db[f] * (pe - pu);
                        double ce = rho * w * (3 + 9 * de), cu = rho * w * (3 + 9 * du);
                        evb_x += Db[f] * (ce * EX[f] - 3 * rho * w * ex);
                        evb_y += Db[f] * (ce * EY[f] - 3 * rho * w * ey);
                        evb_z += Db[f] * (ce * EZ[f] - 3 * rho * w * ez);
                        uxb -= Db[f] * (cu * EX[f] - 3 * rho * w * ux);
                        uyb -= Db[f] * (cu * EY[f] - 3 * rho * w * uy);
                        uzb -= Db[f] * (cu * EZ[f] - 3 * rho * w * uz);
                    }
                    /* EV = u + F/rho */
                    uxb += evb_x; uyb += evb_y; uzb += evb_z;
                    double Fxb = evb_x / rho, Fyb = evb_y / rho, Fzb = evb_z / rho;
                    rbar -= (evb_x * Fxa + evb_y * Fya + evb_z * Fza) / (rho * rho);
                    nvxb[I] = uxb; nvyb[I] = uyb; nvzb[I] = uzb;
                    /* F = -rho G(mu) + G(rho)/3: rho_bar from the -rho factor */
                    double gmx = s->gx7(s->mu.data(), i, jj, k), gmy = s->gy7(s->mu.data(), i, jj, k), gmz = s->gz7(s->mu.data(), i, jj, k);
                    rbar -= gmx * Fxb + gmy * Fyb + gmz * Fzb;
                    /* fields whose stencil transposes are taken below */
                    Px[I] = rho * Fxb; Py[I] = rho * Fyb; Pz[I] = rho * Fzb;   /* mu_bar += sum G(P) */
                    double wx = Fxb / 3, wy = Fyb / 3, wz = Fzb / 3;          /* rho_bar -= sum G(W) */
                    if (qp_on) {
                        const double *rr = cur.rho.data();
                        double gx = s->gx7(rr, i, jj, k), gy = s->gy7(rr, i, jj, k), gz = s->gz7(rr, i, jj, k);
                        double c1 = -19 * (p.K1 + 3 * p.K2);
                        wx += qv[1] * 2 * c1 * gx + qv[9] * (-4 * p.K1 * gx) + qv[13] * (-p.K1 * gy) + qv[15] * (-p.K1 * gz);
                        wy += qv[1] * 2 * c1 * gy + qv[9] * (2 * p.K1 * gy) + qv[11] * (-2 * p.K1 * gy) + qv[13] * (-p.K1 * gx)
                              + qv[14] * (-p.K1 * gz);
                        wz += qv[1] * 2 * c1 * gz + qv[9] * (2 * p.K1 * gz) + qv[11] * (2 * p.K1 * gz) + qv[14] * (-p.K1 * gy)
                              + qv[15] * (-p.K1 * gx);
                    }
                    Wx[I] = wx; Wy[I] = wy; Wz[I] = wz;
                    rbn[I] += rbar;
                }
        /* ---- stencil transposes (gather form over all nodes, fields zero off-fluid) ---- */
        #pragma omp parallel for collapse(2) schedule(static)
        for (int i = 0; i < p.nx; ++i)
            for (int jj = 0; jj < p.ny; ++jj)
                for (int k = 0; k < p.nz; ++k) {
                    int64_t I = s->I(i, jj, k);
                    mub[I] = s->gx7(Px.data(), i, jj, k) + s->gy7(Py.data(), i, jj, k) + s->gz7(Pz.data(), i, jj, k);
                }
        /* rho_bar from the force stencil on rho (all nodes, incl. solid) */
        std::vector<double> rball(rbn);
        #pragma omp parallel for collapse(2) schedule(static)
        for (int i = 0; i < p.nx; ++i)
            for (int jj = 0; jj < p.ny; ++jj)
                for (int k = 0; k < p.nz; ++k) {
                    int64_t I = s->I(i, jj, k);
                    rball[I] -= s->gx7(Wx.data(), i, jj, k) + s->gy7(Wy.data(), i, jj, k) + s->gz7(Wz.data(), i, jj, k);
                }
        /* ---- reverse C: fluid mu = K^2 mu_eos(clamp rho) - Ka Lap(rho) ---- */
        std::vector<double> mubf(N, 0.0);
        for (int64_t I = 0; I < N; ++I) if (s->type[I] == FLUID) mubf[I] = mub[I];
        double kas = 0;
        #pragma omp parallel for collapse(2) schedule(static) reduction(+:kas)
        for (int i = 0; i < p.nx; ++i)
            for (int jj = 0; jj < p.ny; ++jj)
                for (int k = 0; k < p.nz; ++k) {
                    int64_t I = s->I(i, jj, k);
                    double lapT = (p.lap_points == 7) ? s->lap7num(mubf.data(), i, jj, k) / 180 : s->lap5(mubf.data(), i, jj, k);
                    rball[I] -= p.Ka * lapT;
                    if (s->type[I] == FLUID) {
                        double r = cur.rho[I];
                        double c = clamp_rho(r);
                        rball[I] += p.K * p.K * dmu_eos(p.eos, p.mu_form, p.A, p.B, p.T, c) * clamp_drho(r) * mubf[I];
                        double lap = (p.lap_points == 7) ? s->lap7num(cur.rho.data(), i, jj, k) / 180 : s->lap5(cur.rho.data(), i, jj, k);
                        kas += -mubf[I] * lap;
                    }
                }
        gka += kas;
        /* solid mu = wall field */
        for (int i = 0; i < p.nx; ++i)
            for (int jj = 0; jj < p.ny; ++jj) {
                for (int k = 0; k < p.nwall_lo; ++k) g_mu_lo[(int64_t)i * p.ny + jj] += mub[s->I(i, jj, k)];
                for (int k = p.nz - p.nwall_hi; k < p.nz; ++k) g_mu_hi[(int64_t)i * p.ny + jj] += mub[s->I(i, jj, k)];
            }
        /* ---- reverse W: solid layers, deepest first ---- */
        const int lo = p.nwall_lo, hi = p.nwall_hi, nl = std::max(lo, hi);
        for (int pass = nl - 1; pass >= 0; --pass) {
            std::vector<double> add(N, 0.0);
            for (int i = 0; i < p.nx; ++i)
                for (int jj = 0; jj < p.ny; ++jj) {
                    if (pass < lo) {
                        int kd = lo - 1 - pass, ks = p.wall_scheme == 1 ? kd + 1 : lo;
                        add[s->I(i, jj, ks)] += s->wavg(rball.data(), i, jj, kd);
                    }
                    if (pass < hi) {
                        int kd = p.nz - hi + pass, ks = p.wall_scheme == 1 ? kd - 1 : p.nz - 1 - hi;
                        add[s->I(i, jj, ks)] += s->wavg(rball.data(), i, jj, kd);
                    }
                }
            for (int i = 0; i < p.nx; ++i)
                for (int jj = 0; jj < p.ny; ++jj) {
                    if (pass < lo) rball[s->I(i, jj, lo - 1 - pass)] = 0;
                    if (pass < hi) rball[s->I(i, jj, p.nz - hi + pass)] = 0;
                }
            for (int64_t I = 0; I < N; ++I) rball[I] += add[I];
        }
        for (int64_t I = 0; I < N; ++I) if (s->type[I] != FLUID) rball[I] = 0;
        rb.swap(rball);
        vxb.swap(nvxb); vyb.swap(nvyb); vzb.swap(nvzb);
        fb.swap(nfb);
    }
    *g_ka = gka;
    /* leave the state at the end of the forward run */
    std::copy(tape[nsteps].rho.begin(), tape[nsteps].rho.end(), s->rho.begin());
    return 0;
}
