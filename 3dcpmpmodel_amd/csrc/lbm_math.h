// lbm_math.h -- per-node arithmetic of the chemical-potential multiphase D3Q19
// MRT step: lattice constants, moment transforms, equilibrium moments, the
// MRT + exact-difference-forcing collision, the EOS chemical potentials.
//
// Everything here is __host__ __device__ so that the library can also run it
// on the CPU (cpmp_selftest_* entry points) for parity checks without a GPU.
//
// Reference: Common.h:94-146 (D3Q19 velocities, weights, M, R=M^-1),
// FluidCuda.cu:189-193 (Feq), :245-306 (LocalCollideMrt), :311-347 (GlobalCollide),
// :444-463 (ChemPotential).
#pragma once

#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>

#define CPMP_HD __host__ __device__ __forceinline__

namespace cpmp {

constexpr int Q = 19;

// D3Q19 velocity ordering of Common.h:96-98: 0 rest; 1-6 +-x,+-y,+-z;
// 7-10 xy diagonals; 11-14 xz diagonals; 15-18 yz diagonals.
CPMP_HD constexpr int cx(int q) {
    return (q == 1 || q == 7 || q == 9 || q == 11 || q == 13) ? 1
         : (q == 2 || q == 8 || q == 10 || q == 12 || q == 14) ? -1 : 0;
}
CPMP_HD constexpr int cy(int q) {
    return (q == 3 || q == 7 || q == 8 || q == 15 || q == 17) ? 1
         : (q == 4 || q == 9 || q == 10 || q == 16 || q == 18) ? -1 : 0;
}
CPMP_HD constexpr int cz(int q) {
    return (q == 5 || q == 11 || q == 12 || q == 15 || q == 16) ? 1
         : (q == 6 || q == 13 || q == 14 || q == 17 || q == 18) ? -1 : 0;
}
// opposite direction (Common.h:99)
CPMP_HD constexpr int opp(int q) {
    return q == 0 ? 0 : (q <= 6 ? ((q & 1) ? q + 1 : q - 1)
                                : (q == 7 ? 10 : q == 8 ? 9 : q == 9 ? 8 : q == 10 ? 7
                                 : q == 11 ? 14 : q == 12 ? 13 : q == 13 ? 12 : q == 14 ? 11
                                 : q == 15 ? 18 : q == 16 ? 17 : q == 17 ? 16 : 15));
}
// lattice weights (Common.h:100)
CPMP_HD constexpr double wq(int q) { return q == 0 ? 1.0 / 3 : (q <= 6 ? 1.0 / 18 : 1.0 / 36); }
// squared norms of the 19 moment rows (rows of M are orthogonal, R = M^T diag(1/n2))
CPMP_HD constexpr double n2(int r) {
    return r == 0 ? 19 : r == 1 ? 2394 : r == 2 ? 252 : (r >= 3 && r <= 8) ? ((r & 1) ? 10 : 40)
         : r == 9 ? 36 : r == 10 ? 72 : r == 11 ? 12 : r == 12 ? 24 : (r <= 15) ? 4 : 8;
}

// m = M f.  Opposite-pair sums and differences first, then the moment rows.
template <class T>
CPMP_HD void moments(const T f[Q], T m[Q]) {
    const T s1 = f[1] + f[2], d1 = f[1] - f[2];
    const T s3 = f[3] + f[4], d3 = f[3] - f[4];
    const T s5 = f[5] + f[6], d5 = f[5] - f[6];
    // xy diagonals 7(1,1) 8(-1,1) 9(1,-1) 10(-1,-1)
    const T u1 = f[7] + f[10], u2 = f[7] - f[10], v1 = f[8] + f[9], v2 = f[8] - f[9];
    const T Dxy = u1 + v1, cxy = u1 - v1, axy = u2 - v2, bxy = u2 + v2;
    // xz diagonals 11(1,1) 12(-1,1) 13(1,-1) 14(-1,-1)
    const T p1 = f[11] + f[14], p2 = f[11] - f[14], r1 = f[12] + f[13], r2 = f[12] - f[13];
    const T Dxz = p1 + r1, cxz = p1 - r1, axz = p2 - r2, bxz = p2 + r2;
    // yz diagonals 15(1,1) 16(-1,1) 17(1,-1) 18(-1,-1)
    const T g1 = f[15] + f[18], g2 = f[15] - f[18], h1 = f[16] + f[17], h2 = f[16] - f[17];
    const T Dyz = g1 + h1, cyz = g1 - h1, ayz = g2 - h2, byz = g2 + h2;
    const T Ax = s1 + s3 + s5, Dg = Dxy + Dxz + Dyz;
    m[0] = f[0] + Ax + Dg;
    m[1] = T(-30) * f[0] - T(11) * Ax + T(8) * Dg;
    m[2] = T(12) * f[0] - T(4) * Ax + Dg;
    const T Xd = axy + axz, Yd = bxy + ayz, Zd = bxz + byz;
    m[3] = d1 + Xd;
    m[4] = T(-4) * d1 + Xd;
    m[5] = d3 + Yd;
    m[6] = T(-4) * d3 + Yd;
    m[7] = d5 + Zd;
    m[8] = T(-4) * d5 + Zd;
    const T e9 = Dxy + Dxz - T(2) * Dyz;
    m[9] = T(2) * s1 - s3 - s5 + e9;
    m[10] = T(-4) * s1 + T(2) * (s3 + s5) + e9;
    const T e11 = Dxy - Dxz;
    m[11] = s3 - s5 + e11;
    m[12] = T(-2) * (s3 - s5) + e11;
    m[13] = cxy;
    m[14] = cyz;
    m[15] = cxz;
    m[16] = axy - axz;
    m[17] = ayz - bxy;
    m[18] = bxz - byz;
}

// f = R m = M^T diag(1/n2) m, given a[r] = m[r] / n2[r].  Z16: a[16..18] are zero
// and their terms are left out.
template <class T, bool Z16 = false>
CPMP_HD void inverse_scaled(const T a[Q], T f[Q]) {
    f[0] = a[0] - T(30) * a[1] + T(12) * a[2];
    const T Ca = a[0] - T(11) * a[1] - T(4) * a[2];
    const T Cd = a[0] + T(8) * a[1] + a[2];
    const T qx = a[3] - T(4) * a[4], qy = a[5] - T(4) * a[6], qz = a[7] - T(4) * a[8];
    const T ax1 = Ca + T(2) * a[9] - T(4) * a[10];
    const T ay1 = Ca - a[9] + T(2) * a[10] + a[11] - T(2) * a[12];
    const T az1 = Ca - a[9] + T(2) * a[10] - a[11] + T(2) * a[12];
    f[1] = ax1 + qx; f[2] = ax1 - qx;
    f[3] = ay1 + qy; f[4] = ay1 - qy;
    f[5] = az1 + qz; f[6] = az1 - qz;
    const T X = a[3] + a[4], Y = a[5] + a[6], Z = a[7] + a[8];
    const T Exy = Cd + a[9] + a[10] + a[11] + a[12];
    const T Exz = Cd + a[9] + a[10] - a[11] - a[12];
    const T Eyz = Cd - T(2) * (a[9] + a[10]);
    if constexpr (Z16) {
        f[7] = Exy + X + Y + a[13];
        f[8] = Exy - X + Y - a[13];
        f[9] = Exy + X - Y - a[13];
        f[10] = Exy - X - Y + a[13];
        f[11] = Exz + X + Z + a[15];
        f[12] = Exz - X + Z - a[15];
        f[13] = Exz + X - Z - a[15];
        f[14] = Exz - X - Z + a[15];
        f[15] = Eyz + Y + Z + a[14];
        f[16] = Eyz - Y + Z - a[14];
        f[17] = Eyz + Y - Z - a[14];
        f[18] = Eyz - Y - Z + a[14];
        return;
    }
    // xy: f7 (+,+) f8 (-,+) f9 (+,-) f10 (-,-)
    f[7] = Exy + X + Y + a[13] + a[16] - a[17];
    f[8] = Exy - X + Y - a[13] - a[16] - a[17];
    f[9] = Exy + X - Y - a[13] + a[16] + a[17];
    f[10] = Exy - X - Y + a[13] - a[16] + a[17];
    // xz: f11 (+,+) f12 (-,+) f13 (+,-) f14 (-,-)
    f[11] = Exz + X + Z + a[15] - a[16] + a[18];
    f[12] = Exz - X + Z - a[15] + a[16] + a[18];
    f[13] = Exz + X - Z - a[15] - a[16] - a[18];
    f[14] = Exz - X - Z + a[15] + a[16] - a[18];
    // yz: f15 (+,+) f16 (-,+) f17 (+,-) f18 (-,-)
    f[15] = Eyz + Y + Z + a[14] + a[17] - a[18];
    f[16] = Eyz - Y + Z - a[14] - a[17] - a[18];
    f[17] = Eyz + Y - Z - a[14] + a[17] + a[18];
    f[18] = Eyz - Y - Z + a[14] - a[17] + a[18];
}

// Rows whose relaxation rate is 1 in the reference set (FluidCuda.cu:154-157):
// 0, 1, 3, 5, 7, 10, 12, 16, 17, 18.  With rate 1 a row's post-collision moment is
// its forced equilibrium alone (and 0 for rows 16-18).
CPMP_HD constexpr bool unit_rate_row(int r) {
    return r == 0 || r == 1 || r == 3 || r == 5 || r == 7 || r == 10 || r == 12 || r >= 16;
}

// Equilibrium moments Meq(rho,u) of FluidCuda.cu:268-286 (rows 16-18 are 0).
template <class T>
CPMP_HD void meq(T rho, T ux, T uy, T uz, T m[16]) {
    const T usq = ux * ux + uy * uy + uz * uz;
    const T g9 = T(2) * ux * ux - uy * uy - uz * uz, g11 = uy * uy - uz * uz;
    m[0] = rho;
    m[1] = rho * (T(-11) + T(19) * usq);
    m[2] = rho * (T(3) - T(5.5) * usq);
    m[3] = rho * ux;
    m[4] = rho * ux * T(-2.0 / 3.0);
    m[5] = rho * uy;
    m[6] = rho * uy * T(-2.0 / 3.0);
    m[7] = rho * uz;
    m[8] = rho * uz * T(-2.0 / 3.0);
    m[9] = rho * g9;
    m[10] = rho * g9 * T(-0.5);
    m[11] = rho * g11;
    m[12] = rho * g11 * T(-0.5);
    m[13] = rho * ux * uy;
    m[14] = rho * uy * uz;
    m[15] = rho * ux * uz;
}

// Momentum j = sum e f (the x/y/z sums of MacroCalculate, FluidCuda.cu:364-366),
// pairwise so that a rest state gives exactly zero.
template <class T>
CPMP_HD void momentum(const T f[Q], T &jx, T &jy, T &jz) {
    jx = (f[1] - f[2]) + ((f[7] - f[10]) - (f[8] - f[9])) + ((f[11] - f[14]) - (f[12] - f[13]));
    jy = (f[3] - f[4]) + ((f[7] - f[10]) + (f[8] - f[9])) + ((f[15] - f[18]) - (f[16] - f[17]));
    jz = (f[5] - f[6]) + ((f[11] - f[14]) + (f[12] - f[13])) + ((f[15] - f[18]) + (f[16] - f[17]));
}

// Collision of GlobalCollide/LocalCollideMrt (FluidCuda.cu:245-347) in moment space:
//   m* = m - S(m - Meq(rho,u)) + S Qp + [Meq(rho,u+F/rho) - Meq(rho,u)]
//      = (I - S)(m - Meq(rho,u)) + Meq(rho,u+F/rho) + S Qp
//   out = R m*
// The bracket is M (Feq(rho,EV) - Feq(rho,u)), the exact-difference force term of
// FluidCuda.cu:341, since M Feq(rho,v) = Meq(rho,v) for this basis.  u = j/rho with
// j = (m3, m5, m7) and rho the stored density; it is returned in u[3].  Each row
// is finished as soon as its equilibrium values are formed, which keeps the
// number of live registers low.
// 1/x: in FP32 device code the hardware reciprocal (v_rcp_f32, 1 ulp), the
// correctly rounded quotient elsewhere (FP64, host)
template <class T>
CPMP_HD T recip(T x) {
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (sizeof(T) == 4) return __builtin_amdgcn_rcpf(x);
#endif
    return T(1) / x;
}

// UR: the rates of the unit_rate_row rows are 1 (the caller has checked); those rows
// skip (I - S)(m - Meq), which is exactly 0, and their moments are not formed.
//
// GUO: Guo forcing (Guo, Zheng and Shi, Phys. Rev. E 65, 046308 (2002)) in its MRT
// form instead of the exact-difference term:
//   g  = (j + F/2) / rho
//   m* = m - S(m - Meq(rho,g)) + (I - S/2) G + S Qp,   out = R m*
// with G = M Fbar, Fbar_q = w_q [3 (e_q - g) + 9 (e_q.g) e_q].F.  Since Fbar is the
// derivative of Feq(rho, g) along F/rho and M Feq = Meq, G is the derivative of the
// Meq rows along F/rho: G0 = 0, G1 = 38 g.F, G2 = -11 g.F, G3 = Fx, G4 = -2/3 Fx
// (y, z alike), G9 = 2 (2 gx Fx - gy Fy - gz Fz), G10 = -G9/2, G11 = 2 (gy Fy - gz Fz),
// G12 = -G11/2, G13 = gx Fy + gy Fx, G14 = gy Fz + gz Fy, G15 = gx Fz + gz Fx,
// G16..18 = 0.  Mass is conserved and the momentum gains F exactly.  u returns j/rho
// in both forms.
template <class T, bool UR = false, bool GUO = false>
CPMP_HD void collide(const T f[Q], T rho, T Fx, T Fy, T Fz, const T *ts, const T *qp, T out[Q], T u[3]) {
    T m[Q];
    moments(f, m);
    const T inv = recip(rho);
    const T ux = m[3] * inv, uy = m[5] * inv, uz = m[7] * inv;
    u[0] = ux; u[1] = uy; u[2] = uz;
    T a[Q];
    auto row = [&](int r, T mequ, T meqe) {
        T v = (UR && unit_rate_row(r)) ? meqe : (T(1) - ts[r]) * (m[r] - mequ) + meqe;
        if (qp) v += ts[r] * qp[r];
        a[r] = v * T(1.0 / n2(r));
    };
    const T c23 = T(-2.0 / 3.0);
    if constexpr (GUO) {
        const T h = T(0.5) * inv;
        const T gx = ux + Fx * h, gy = uy + Fy * h, gz = uz + Fz * h;
        const T gg = gx * gx + gy * gy + gz * gz, gF = gx * Fx + gy * Fy + gz * Fz;
        const T g9 = T(2) * gx * gx - gy * gy - gz * gz, g11 = gy * gy - gz * gz;
        const T G9 = T(2) * (T(2) * gx * Fx - gy * Fy - gz * Fz), G11 = T(2) * (gy * Fy - gz * Fz);
        // row r: Meq_r(g) and Meq_r(g) + (1 - s_r/2) G_r
        auto grow = [&](int r, T mq, T G) { row(r, mq, mq + (T(1) - T(0.5) * ts[r]) * G); };
        row(0, rho, rho);
        grow(1, rho * (T(-11) + T(19) * gg), T(38) * gF);
        grow(2, rho * (T(3) - T(5.5) * gg), T(-11) * gF);
        grow(3, rho * gx, Fx);
        grow(4, rho * gx * c23, Fx * c23);
        grow(5, rho * gy, Fy);
        grow(6, rho * gy * c23, Fy * c23);
        grow(7, rho * gz, Fz);
        grow(8, rho * gz * c23, Fz * c23);
        grow(9, rho * g9, G9);
        grow(10, rho * g9 * T(-0.5), G9 * T(-0.5));
        grow(11, rho * g11, G11);
        grow(12, rho * g11 * T(-0.5), G11 * T(-0.5));
        grow(13, rho * gx * gy, gx * Fy + gy * Fx);
        grow(14, rho * gy * gz, gy * Fz + gz * Fy);
        grow(15, rho * gx * gz, gx * Fz + gz * Fx);
#pragma unroll
        for (int r = 16; r < Q; ++r) a[r] = UR ? T(0) : (T(1) - ts[r]) * m[r] * T(1.0 / n2(r));
        inverse_scaled<T, UR>(a, out);
        return;
    }
    const T ex = ux + Fx * inv, ey = uy + Fy * inv, ez = uz + Fz * inv;
    const T uu = ux * ux + uy * uy + uz * uz, ee = ex * ex + ey * ey + ez * ez;
    const T u9 = T(2) * ux * ux - uy * uy - uz * uz, e9 = T(2) * ex * ex - ey * ey - ez * ez;
    const T u11 = uy * uy - uz * uz, e11 = ey * ey - ez * ez;
    row(0, rho, rho);
    row(1, rho * (T(-11) + T(19) * uu), rho * (T(-11) + T(19) * ee));
    row(2, rho * (T(3) - T(5.5) * uu), rho * (T(3) - T(5.5) * ee));
    row(3, rho * ux, rho * ex);
    row(4, rho * ux * c23, rho * ex * c23);
    row(5, rho * uy, rho * ey);
    row(6, rho * uy * c23, rho * ey * c23);
    row(7, rho * uz, rho * ez);
    row(8, rho * uz * c23, rho * ez * c23);
    row(9, rho * u9, rho * e9);
    row(10, rho * u9 * T(-0.5), rho * e9 * T(-0.5));
    row(11, rho * u11, rho * e11);
    row(12, rho * u11 * T(-0.5), rho * e11 * T(-0.5));
    row(13, rho * ux * uy, rho * ex * ey);
    row(14, rho * uy * uz, rho * ey * ez);
    row(15, rho * ux * uz, rho * ex * ez);
#pragma unroll
    for (int r = 16; r < Q; ++r) a[r] = UR ? T(0) : (T(1) - ts[r]) * m[r] * T(1.0 / n2(r));
    inverse_scaled<T, UR>(a, out);
}

// Qp of FluidCuda.cu:260-266 from the density gradient (K1, K2 tuning).
template <class T>
CPMP_HD void qp_terms(T gx, T gy, T gz, T K1, T K2, T qp[Q]) {
#pragma unroll
    for (int r = 0; r < Q; ++r) qp[r] = T(0);
    qp[1] = T(-19) * (K1 + T(3) * K2) * (gx * gx + gy * gy + gz * gz);
    qp[9] = -K1 * (T(2) * gx * gx - gy * gy - gz * gz);
    qp[11] = -K1 * (gy * gy - gz * gz);
    qp[13] = -K1 * gx * gy;
    qp[14] = -K1 * gz * gy;
    qp[15] = -K1 * gx * gz;
}

// ---- EOS ----
// ids as in cpmp.h
constexpr int EOS_VDW = 0, EOS_CSE = 1, EOS_RKE = 2, EOS_RKS = 3, EOS_PRW = 4, EOS_PRM = 5;

// density clamp of FluidCuda.cu:450-451
CPMP_HD double clamp_rho(double d) {
    return d > 9.7 ? 9.7 + (d - 9.7) / (1.0 + (d - 9.7) * 1.5) : d;
}
CPMP_HD double clamp_drho(double d) {
    if (d > 9.7) { const double q = 1.0 + (d - 9.7) * 1.5; return 1.0 / (q * q); }
    return 1.0;
}

constexpr double kSqrt2 = 1.4142135623730950488016887242097;

// Bulk chemical potential of each EOS (without K^2).  PR: FluidCuda.cu:453.
// The others are the closed forms of integral p'(rho)/rho drho.  R is the
// arithmetic type (double everywhere except the bulk of the FP32 forms).
template <class R>
CPMP_HD R mu_eos_t(int eos, int form, R A, R B, R T, R d) {
    const R one = R(1), s2 = R(kSqrt2);
    if (form == 1 || eos == EOS_PRW || eos == EOS_PRM) {
        const R bd = B * d;
        return T * log(d / (one - bd)) - A / (s2 * R(2) * B) * log((s2 - one + bd) / (s2 + one - bd))
             + T / (one - bd) - A * d / (one + bd * R(2) - bd * bd);
    }
    if (eos == EOS_VDW) {
        const R bd = B * d;
        return T * log(d / (one - bd)) + T / (one - bd) - R(2) * A * d;
    }
    if (eos == EOS_RKE || eos == EOS_RKS) {
        const R bd = B * d;
        return T * log(d / (one - bd)) + T / (one - bd) - A / B * log(one + bd) - A * d / (one + bd);
    }
    // CSE: Carnahan-Starling hard spheres with B = b/4
    const R e = B * d, om = one - e;
    return T * (log(d) + (R(8) * e - R(9) * e * e + R(3) * e * e * e) / (om * om * om)) - R(2) * A * d;
}
CPMP_HD double mu_eos(int eos, int form, double A, double B, double T, double d) {
    return mu_eos_t<double>(eos, form, A, B, T, d);
}

// d mu_eos / d rho
CPMP_HD double dmu_eos(int eos, int form, double A, double B, double T, double d) {
    if (form == 1 || eos == EOS_PRW || eos == EOS_PRM) {
        const double bd = B * d, om = 1.0 - bd, q = 1.0 + 2 * bd - bd * bd;
        return T * (1.0 / d + B / om) - A / (kSqrt2 * 2 * B) * (B / (kSqrt2 - 1 + bd) + B / (kSqrt2 + 1 - bd))
             + T * B / (om * om) - A * (q - d * (2 * B - 2 * B * bd)) / (q * q);
    }
    if (eos == EOS_VDW) {
        const double om = 1.0 - B * d;
        return T * (1.0 / d + B / om) + T * B / (om * om) - 2.0 * A;
    }
    if (eos == EOS_RKE || eos == EOS_RKS) {
        const double om = 1.0 - B * d, op = 1.0 + B * d;
        return T * (1.0 / d + B / om) + T * B / (om * om) - A / op - A / (op * op);
    }
    const double e = B * d, om = 1.0 - e;
    const double num = 8.0 * e - 9.0 * e * e + 3.0 * e * e * e;
    const double dnum = (8.0 - 18.0 * e + 9.0 * e * e) * B;
    return T * (1.0 / d + (dnum * om + 3.0 * B * num) / (om * om * om * om)) - 2.0 * A;
}

CPMP_HD double pressure_eos(int eos, double A, double B, double T, double d) {
    const double bd = B * d;
    if (eos == EOS_VDW) return T * d / (1.0 - bd) - A * d * d;
    if (eos == EOS_PRW || eos == EOS_PRM) return T * d / (1.0 - bd) - A * d * d / (1.0 + bd * 2 - bd * bd);
    if (eos == EOS_RKE || eos == EOS_RKS) return T * d / (1.0 - bd) - A * d * d / (1.0 + bd);
    const double e = bd;
    return d * T * (1 + e + e * e - e * e * e) / ((1 - e) * (1 - e) * (1 - e)) - A * d * d;
}

// SplitMix64 (Steele, Lea, Flood 2014); noise U in [-1,1) keyed by (seed, index)
CPMP_HD uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
CPMP_HD double noise_u(uint64_t seed, uint64_t index) {
    const uint64_t r = splitmix64(seed ^ splitmix64(index));
    return (double)(r >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
}

}  // namespace cpmp
