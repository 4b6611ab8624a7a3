// adjoint.hip -- discrete adjoint of the step (no reference counterpart).
//
// Objective J = 1/2 sum_fluid (rho^N - target)^2 after N steps; gradients with
// respect to the wall chemical-potential fields (bottom, top) and kappa (Ka).
//
// The state carried between steps is (f, rho) with u = j(f)/rho.  Reverse of one
// step n, for adjoint state (fbar, rbar, ubar) of step n+1:
//   moments : rho' = sum f', u' = j(f')/rho'   -> fhat = fbar + rbar + ubar.(e - u')/rho'
//   stream  : push / half-way bounce-back       -> pull of fhat
//   collide : m* = m - S(m - Meq(rho,u)) + S Qp + Meq(rho,u+F/rho) - Meq(rho,u), out = R m*
//             -> fbar^n = M^T (I-S) R^T D, plus rho, u, F, grad(rho) cotangents
//   force   : F = -rho G(mu) + G(rho)/3         -> the gradient stencils are antisymmetric,
//             so their transposes are -G, applied in gather form
//   mu      : mu = K^2 mu_eos(clamp rho) - Ka Lap(rho) (fluid), mu = mu_w (solid)
//   wall    : solid rho = layered diagonal averages of the first fluid layer (symmetric)
// The forward trajectory is recomputed segment by segment from population
// checkpoints; the tape holds rho (after the wall fill), u, grad mu and grad rho of
// every step of the segment, in the storage scalar type of the form (the values the
// forward collide used).  Trailing steps that fit the tape budget are taped during
// the first forward run and not recomputed; segments whose tape does not fit the
// memory limit are split into sub-segments with their own checkpoints.  The
// reverse sweep runs in FP64 in every form.
#include "internal.h"
#include "lbm_math.h"
#include "device_common.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <type_traits>
#include <vector>

namespace cpmp {
namespace {

// sum_r w_r dMeq_r/drho and sum_r w_r dMeq_r/dv (rows 0..15 of FluidCuda.cu:268-286)
__device__ __forceinline__ void meq_vjp(double rho, double vx, double vy, double vz, const double *w, double &drho,
                                        double &dvx, double &dvy, double &dvz) {
    const double usq = vx * vx + vy * vy + vz * vz, g9 = 2 * vx * vx - vy * vy - vz * vz, g11 = vy * vy - vz * vz;
    const double t = 2.0 / 3.0;
    drho = w[0] + w[1] * (-11 + 19 * usq) + w[2] * (3 - 5.5 * usq) + (w[3] - t * w[4]) * vx + (w[5] - t * w[6]) * vy
           + (w[7] - t * w[8]) * vz + (w[9] - 0.5 * w[10]) * g9 + (w[11] - 0.5 * w[12]) * g11 + w[13] * vx * vy
           + w[14] * vy * vz + w[15] * vx * vz;
    dvx = rho * (38 * vx * w[1] - 11 * vx * w[2] + w[3] - t * w[4] + 4 * vx * w[9] - 2 * vx * w[10] + vy * w[13] + vz * w[15]);
    dvy = rho * (38 * vy * w[1] - 11 * vy * w[2] + w[5] - t * w[6] - 2 * vy * w[9] + vy * w[10] + 2 * vy * w[11]
                 - vy * w[12] + vx * w[13] + vz * w[14]);
    dvz = rho * (38 * vz * w[1] - 11 * vz * w[2] + w[7] - t * w[8] - 2 * vz * w[9] + vz * w[10] - 2 * vz * w[11]
                 + vz * w[12] + vy * w[14] + vx * w[15]);
}

// J and its seed: rb = rho - target on fluid nodes
template <class TT>
__global__ void __launch_bounds__(TPB) k_adj_J(const DevParams P, const TT *__restrict__ rho,
                                               const double *__restrict__ target, double *__restrict__ rb,
                                               double *__restrict__ part) {
    const long long I = (long long)blockIdx.x * TPB + threadIdx.x;
    double v = 0;
    if (I < P.N) {
        if (is_solid(P, (int)((unsigned long long)I % (unsigned)P.nz))) {
            rb[I] = 0;
        } else {
            const double d = (double)rho[I] - target[I];
            rb[I] = d;
            v = 0.5 * d * d;
        }
    }
    const double s = block_sum(v);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// Reverse of moments + streaming + collision at each fluid node of step n.  The
// cotangent arriving along q from the node y that pulls it is
//   fhat_q(y) = fbar_q(y) + rbar(y) + ubar(y).(c_q - u(y))/rho(y)
// with (fbar, ubar) of step n+1 at y and rho, u of the state y pulled; the part
// without rbar is local to y and was written by the previous reverse step as fh_in
// (fbar_q + ubar.(c_q - u)/rho), so each pulled value costs two loads.  This kernel
// writes the same folded form for step n: fh_out_q = fbar_q + ubar.(c_q - u^n)/rho^n.
// TT: the tape's type; TH: the storage type of the folded population cotangents
// (FP64; FP32 for the adjoint of an FP32 run), arithmetic in FP64.
template <class TT, class TH>
__global__ void __launch_bounds__(TPB) k_adj_collide(const DevParams P, const TH *__restrict__ fh_in,
                                                     const double *__restrict__ rb_in,
                                                     const TT *__restrict__ rho0, const TT *__restrict__ u0,
                                                     const TT *__restrict__ grads, TH *__restrict__ fh_out,
                                                     double *__restrict__ rb_loc,
                                                     double *__restrict__ Pf, double *__restrict__ Wf) {
    const NodeIdx nd = node_from_grid(P);
    if (!nd.ok) return;
    const long long I = nd.I;
    const long long N = P.N;
    const Coord c{nd.x, nd.y, nd.z};
    if (is_solid(P, c.z)) {
        rb_loc[I] = 0;
        for (int a = 0; a < 3; ++a) { Pf[a * N + I] = 0; Wf[a * N + I] = 0; }
#pragma unroll
        for (int q = 0; q < Q; ++q) fh_out[(long long)q * N + I] = TH(0);
        return;
    }
    int xs[7], ys[7], zs[7];
    wrap7(P, c, xs, ys, zs);
    // pulled cotangents of the populations written by this node
    double D[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int tz = zs[3 + cz(q)];
        const bool bounce = is_solid(P, tz);
        const long long y = bounce ? I : ((long long)xs[3 + cx(q)] * P.ny + ys[3 + cy(q)]) * P.nz + tz;
        const int qq = bounce ? opp(q) : q;   // bounced back: pulled by this node as opp(q), c_opp(q) = -c_q
        D[q] = (double)fh_in[(long long)qq * N + y] + rb_in[y];
    }
    double m[Q], mb[Q], b[Q];
    moments(D, m);
#pragma unroll
    for (int r = 0; r < Q; ++r) {
        mb[r] = m[r] * (1.0 / n2(r));
        b[r] = (1.0 - P.ts[r]) * mb[r];
    }
    double fbar[Q];
    inverse_scaled(b, fbar);

    const double rho = (double)rho0[I], ux = (double)u0[I], uy = (double)u0[N + I], uz = (double)u0[2 * N + I];
    // grad mu^n and grad rho^n as the forward collide evaluated them (taped)
    const double gmx = (double)grads[I], gmy = (double)grads[N + I], gmz = (double)grads[2 * N + I];
    const double grx = (double)grads[3 * N + I], gry = (double)grads[4 * N + I], grz = (double)grads[5 * N + I];
    const double Fx = -gmx * rho + grx / 3, Fy = -gmy * rho + gry / 3, Fz = -gmz * rho + grz / 3;
    const double ex = ux + Fx / rho, ey = uy + Fy / rho, ez = uz + Fz / rho;
    double wu[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) wu[r] = (P.ts[r] - 1.0) * mb[r];
    double dru, dux, duy, duz, dre, dex, dey, dez;
    meq_vjp(rho, ux, uy, uz, wu, dru, dux, duy, duz);
    meq_vjp(rho, ex, ey, ez, mb, dre, dex, dey, dez);
    const double Fbx = dex / rho, Fby = dey / rho, Fbz = dez / rho;
    double rbar = dru + dre - (dex * Fx + dey * Fy + dez * Fz) / (rho * rho);
    rbar -= gmx * Fbx + gmy * Fby + gmz * Fbz;
    {
        const double ubx = (dux + dex) / rho, uby = (duy + dey) / rho, ubz = (duz + dez) / rho;
        const double ubu = ubx * ux + uby * uy + ubz * uz;
#pragma unroll
        for (int q = 0; q < Q; ++q)
            fh_out[(long long)q * N + I] = (TH)(fbar[q] + (cx(q) * ubx + cy(q) * uby + cz(q) * ubz - ubu));
    }
    double wx = Fbx / 3, wy = Fby / 3, wz = Fbz / 3;
    if (P.qp_on) {
        const double K1 = P.K1, c1 = -19 * (P.K1 + 3 * P.K2);
        const double q1 = P.ts[1] * mb[1], q9 = P.ts[9] * mb[9], q11 = P.ts[11] * mb[11];
        const double q13 = P.ts[13] * mb[13], q14 = P.ts[14] * mb[14], q15 = P.ts[15] * mb[15];
        wx += q1 * 2 * c1 * grx - q9 * 4 * K1 * grx - q13 * K1 * gry - q15 * K1 * grz;
        wy += q1 * 2 * c1 * gry + q9 * 2 * K1 * gry - q11 * 2 * K1 * gry - q13 * K1 * grx - q14 * K1 * grz;
        wz += q1 * 2 * c1 * grz + q9 * 2 * K1 * grz + q11 * 2 * K1 * grz - q14 * K1 * gry - q15 * K1 * grx;
    }
    rb_loc[I] = rbar;
    Pf[I] = rho * Fbx; Pf[N + I] = rho * Fby; Pf[2 * N + I] = rho * Fbz;
    Wf[I] = wx; Wf[N + I] = wy; Wf[2 * N + I] = wz;
}

// sum_a G_a(A_a) for the isotropic gradient G (axis and face-diagonal 6-point
// differences dd, weights 2:1:1:1:1, /360).  Each face diagonal appears in two
// components; since dd(a, -u) = -dd(a, u) the pair is one difference of a sum or
// difference of two components, e.g. dd(Ax, (1,-1,0)) + dd(Ay, (-1,1,0)) =
// dd(Ax - Ay, (1,-1,0)): nine differences instead of fifteen.
__device__ __forceinline__ double grad_divergence(const Stencil<double, double> &st, const double *ax,
                                                  const double *ay, const double *az) {
    auto dd2 = [&](const double *a, const double *b, double sg, int ux, int uy, int uz) {
        auto v = [&](int s) { return st.at(a, s * ux, s * uy, s * uz) + sg * st.at(b, s * ux, s * uy, s * uz); };
        return -v(-3) + 9 * v(-2) - 45 * v(-1) + 45 * v(1) - 9 * v(2) + v(3);
    };
    const double axis = st.dd(ax, 1, 0, 0) + st.dd(ay, 0, 1, 0) + st.dd(az, 0, 0, 1);
    const double diag = dd2(ax, ay, 1, 1, 1, 0) + dd2(ax, ay, -1, 1, -1, 0) + dd2(ax, az, 1, 1, 0, 1)
                        + dd2(ax, az, -1, 1, 0, -1) + dd2(ay, az, 1, 0, 1, 1) + dd2(ay, az, -1, 0, 1, -1);
    return (2 * axis + diag) * (1.0 / 360.0);
}

// transposed force stencils, all nodes: mub = sum_a G_a(P_a), rb_new = rb_loc - sum_a G_a(W_a)
__global__ void __launch_bounds__(TPB) k_adj_stencil(const DevParams P, const double *__restrict__ Pf,
                                                     const double *__restrict__ Wf, const double *__restrict__ rb_loc,
                                                     double *__restrict__ mub, double *__restrict__ rb_new) {
    const NodeIdx nd = node_from_grid(P);
    if (!nd.ok) return;
    const long long I = nd.I;
    const long long N = P.N;
    const Coord c{nd.x, nd.y, nd.z};
    int xs[7], ys[7], zs[7];
    wrap7(P, c, xs, ys, zs);
    const Stencil<double, double> st{P, xs, ys, zs};
    mub[I] = grad_divergence(st, Pf, Pf + N, Pf + 2 * N);
    rb_new[I] = rb_loc[I] - grad_divergence(st, Wf, Wf + N, Wf + 2 * N);
}

// The same transposed stencils as a march along x, one field triple per launch
// (Pf -> mub, Wf -> rb_new).  A workgroup owns a tile of kAdjTY x kAdjTZ columns
// and a chunk [xa, xb) of planes; it streams the planes xa-3 .. xb+2 of the
// triple (with a halo of 3 in y and z) through LDS one at a time, and each thread
// adds the contributions of the plane in LDS to the seven output planes that
// reach it (x' - 3 .. x' + 3), kept as running sums in registers; output plane
// x' - 3 is then complete and written.  Every value is read from memory once per
// chunk instead of once per stencil point.  With the coefficients c_s of the
// 6-point difference (c_1 = 45, c_2 = -9, c_3 = 1, c_-s = -c_s) the
// contribution of source plane x' = x + s to output x is
//   c_s [2 Ax(y,z) + (Ax+Ay)(y+s,z) + (Ax-Ay)(y-s,z) + (Ax+Az)(y,z+s) + (Ax-Az)(y,z-s)]
// and the in-plane directions add to output x' itself
//   sum_s c_s [2 Ay(y+s,z) + 2 Az(y,z+s) + (Ay+Az)(y+s,z+s) + (Ay-Az)(y+s,z-s)],
// the sum of grad_divergence above (summed in another order).
// SUB: out = base - sum / 360 (rb_new from rb_loc); otherwise out = sum / 360.
constexpr int kAdjTY = 8, kAdjTZ = 32;   // tile of the reverse march kernels (one thread per column)
constexpr int kAdjHalo = 3;
constexpr int kAdjEY = kAdjTY + 2 * kAdjHalo, kAdjEZ = kAdjTZ + 2 * kAdjHalo, kAdjEA = kAdjEY * kAdjEZ;
constexpr int kAdjSlots = (kAdjEA + TPB - 1) / TPB;
static_assert(kAdjTY * kAdjTZ == TPB, "one thread per tile column");

template <bool SUB>
__device__ __forceinline__ void adj_stencil_march(const DevParams &P, const double *__restrict__ A,
                                                  const double *__restrict__ base, double *__restrict__ out, int ntz,
                                                  int nty, int xc, int b, double (*sf)[kAdjEA]) {
    const long long N = P.N, plane = (long long)P.ny * P.nz;
    const int tid = threadIdx.x, tz = tid % kAdjTZ, ty = tid / kAdjTZ;
    const int zt = b % ntz;
    b /= ntz;
    const int yt = b % nty, ch = b / nty;
    const int z0 = zt * kAdjTZ, y0 = yt * kAdjTY, xa = ch * xc, xb = min(xa + xc, P.nx);
    const int y = y0 + ty, z = z0 + tz;
    const bool own = y < P.ny && z < P.nz;
    // LDS slots this thread fills: position i = tid + TPB r of the haloed tile
    long long off[kAdjSlots];
    int pos[kAdjSlots];
#pragma unroll
    for (int r = 0; r < kAdjSlots; ++r) {
        const int i = tid + TPB * r;
        pos[r] = i < kAdjEA ? i : -1;
        const int ii = i < kAdjEA ? i : 0;
        off[r] = (long long)wrapc(y0 - kAdjHalo + ii / kAdjEZ, P.ny) * P.nz + wrapc(z0 - kAdjHalo + ii % kAdjEZ, P.nz);
    }
    double pre[kAdjSlots][3];
    auto fetch = [&](int xp) {   // wraps past the end (loads unused then)
        const double *src = A + (long long)wrapc(xp, P.nx) * plane;
#pragma unroll
        for (int r = 0; r < kAdjSlots; ++r)
#pragma unroll
            for (int a = 0; a < 3; ++a) pre[r][a] = src[a * N + off[r]];
    };
    double acc[7];
#pragma unroll
    for (int k = 0; k < 7; ++k) acc[k] = 0;
    const int c0 = (ty + kAdjHalo) * kAdjEZ + tz + kAdjHalo;
    // the planes alternate between two LDS buffers: a wave stores plane xp + 1 while
    // slower waves still read plane xp, so one barrier per plane suffices
    int buf = 0;
    fetch(xa - kAdjHalo);
    for (int xp = xa - kAdjHalo; xp <= xb + kAdjHalo - 1; ++xp) {
        double(*sb)[kAdjEA] = sf + 3 * buf;
#pragma unroll
        for (int r = 0; r < kAdjSlots; ++r)
            if (pos[r] >= 0)
#pragma unroll
                for (int f = 0; f < 3; ++f) sb[f][pos[r]] = pre[r][f];
        __syncthreads();
        fetch(xp + 1);     // next plane in flight during the arithmetic
        const double *ax = sb[0], *ay = sb[1], *az = sb[2];
        const double x0 = 2 * ax[c0];
        double inpl = 0;
#pragma unroll
        for (int s = 1; s <= 3; ++s) {
            const double cs = s == 1 ? 45.0 : (s == 2 ? -9.0 : 1.0);
            const int dy = s * kAdjEZ;
            const double axp = ax[c0 + dy], axm = ax[c0 - dy], ayp = ay[c0 + dy], aym = ay[c0 - dy];
            const double axr = ax[c0 + s], axl = ax[c0 - s], azr = az[c0 + s], azl = az[c0 - s];
            // source plane x' = x + s (output x = x' - s, slot 3 - s) and x' = x - s (slot 3 + s)
            const double vp = x0 + (axp + ayp) + (axm - aym) + (axr + azr) + (axl - azl);
            const double vm = x0 + (axm + aym) + (axp - ayp) + (axl + azl) + (axr - azr);
            acc[3 - s] += cs * vp;
            acc[3 + s] -= cs * vm;
            // in-plane directions, output x'
            const double dpp = ay[c0 + dy + s] + az[c0 + dy + s], dmm = ay[c0 - dy - s] + az[c0 - dy - s];
            const double dpm = ay[c0 + dy - s] - az[c0 + dy - s], dmp = ay[c0 - dy + s] - az[c0 - dy + s];
            inpl += cs * (2 * (ayp - aym) + 2 * (azr - azl) + (dpp - dmm) + (dpm - dmp));
        }
        acc[3] += inpl;
        if (xp >= xa + kAdjHalo && own) {   // output plane xp - 3 is complete
            const long long I = (long long)(xp - kAdjHalo) * plane + (long long)y * P.nz + z;
            const double v = acc[0] * (1.0 / 360.0);
            out[I] = SUB ? base[I] - v : v;
        }
#pragma unroll
        for (int k = 0; k < 6; ++k) acc[k] = acc[k + 1];
        acc[6] = 0;
        buf ^= 1;
    }
}

// Both transposed stencils of one reverse step in one launch: workgroups
// 0 .. nb-1 take P_f -> mu cotangent, nb .. 2 nb-1 take W_f -> rho cotangent
// (independent outputs; one launch fills the device twice as well on small lattices).
__global__ void __launch_bounds__(TPB) k_adj_stencil_march(const DevParams P, const double *__restrict__ Pf,
                                                           const double *__restrict__ Wf,
                                                           const double *__restrict__ rb_loc, double *__restrict__ mub,
                                                           double *__restrict__ rb_new, int ntz, int nty, int xc,
                                                           int nb) {
    __shared__ double sf[2 * 3][kAdjEA];   // x, y, z components of two planes
    const int b = blockIdx.x;
    if (b < nb) adj_stencil_march<false>(P, Pf, nullptr, mub, ntz, nty, xc, b, sf);
    else adj_stencil_march<true>(P, Wf, rb_loc, rb_new, ntz, nty, xc, b - nb, sf);
}

// transposed chemical potential: Laplacian of the fluid mu cotangent, EOS slope,
// and the kappa gradient (per-block partial sums)
template <class TT>
__global__ void __launch_bounds__(TPB) k_adj_mu(const DevParams P, const TT *__restrict__ rho0,
                                                const double *__restrict__ mub, double *__restrict__ rb_new,
                                                double *__restrict__ kapart) {
    const NodeIdx nd = node_from_grid(P);
    const long long I = nd.I;
    double kap = 0;
    if (nd.ok) {
        const Coord c{nd.x, nd.y, nd.z};
        int xs[7], ys[7], zs[7];
        wrap7(P, c, xs, ys, zs);
        auto idx = [&](int dx, int dy, int dz) { return ((long long)xs[dx + 3] * P.ny + ys[dy + 3]) * P.nz + zs[dz + 3]; };
        auto mm = [&](int dx, int dy, int dz) { return is_solid(P, zs[dz + 3]) ? 0.0 : mub[idx(dx, dy, dz)]; };
        auto rr = [&](int dx, int dy, int dz) { return (double)rho0[idx(dx, dy, dz)]; };
        auto lap = [&](auto &&v) {
            if (P.lap_points == 7) {
                auto ax = [&](int ux, int uy, int uz) {
                    return 2 * v(-3 * ux, -3 * uy, -3 * uz) - 27 * v(-2 * ux, -2 * uy, -2 * uz) + 270 * v(-ux, -uy, -uz)
                           + 270 * v(ux, uy, uz) - 27 * v(2 * ux, 2 * uy, 2 * uz) + 2 * v(3 * ux, 3 * uy, 3 * uz);
                };
                return (-v(0, 0, 0) * 490 * 3 + ax(1, 0, 0) + ax(0, 1, 0) + ax(0, 0, 1)) / 180;
            }
            auto ax = [&](int ux, int uy, int uz) {
                return -v(-2 * ux, -2 * uy, -2 * uz) + v(-ux, -uy, -uz) * 16 + v(ux, uy, uz) * 16 - v(2 * ux, 2 * uy, 2 * uz);
            };
            return (-v(0, 0, 0) * 90 + ax(1, 0, 0) + ax(0, 1, 0) + ax(0, 0, 1)) / 12;
        };
        double acc = -P.Ka * lap(mm);
        if (!is_solid(P, c.z)) {
            const double r = (double)rho0[I], mu_b = mub[I];
            acc += P.K2s * dmu_eos(P.eos, P.mu_form, P.A, P.B, P.T, clamp_rho(r)) * clamp_drho(r) * mu_b;
            kap = -mu_b * lap(rr);
        }
        rb_new[I] += acc;
    }
    // each block adds its partial to its own slot, step after step (one writer per
    // slot and launch: a fixed order); the slots are summed once at the end
    const double s = block_sum(kap);
    if (threadIdx.x == 0) kapart[blockIdx.x + (long long)gridDim.x * (blockIdx.y + (long long)gridDim.y * blockIdx.z)] += s;
}

// The same transposed chemical potential as a march along x over the tiles of
// k_adj_stencil_march: the fluid-masked mu cotangent and rho (tape) of planes
// x-H..x+H with an in-plane halo of H (H = 2 for the 5-point Laplacian, 3 for the
// 7-point one) sit in an LDS ring.  Same arithmetic per node; the kappa partial
// sums are accumulated over the chunk and added once per workgroup.
template <class TT, int H>
__global__ void __launch_bounds__(TPB) k_adj_mu_march(const DevParams P, const TT *__restrict__ rho0,
                                                      const double *__restrict__ mub, double *__restrict__ rb_new,
                                                      double *__restrict__ kapart, int ntz, int nty, int xc) {
    constexpr int R = 2 * H + 1, EY = kAdjTY + 2 * H, EZ = kAdjTZ + 2 * H, EA = EY * EZ;
    constexpr int SL = (EA + TPB - 1) / TPB;
    __shared__ double sm[R][EA], sr[R][EA];
    const long long plane = (long long)P.ny * P.nz;
    const int tid = threadIdx.x, tz = tid % kAdjTZ, ty = tid / kAdjTZ;
    int b = blockIdx.x;
    const int zt = b % ntz;
    b /= ntz;
    const int yt = b % nty, ch = b / nty;
    const int z0 = zt * kAdjTZ, y0 = yt * kAdjTY, xa = ch * xc, xb = min(xa + xc, P.nx);
    const int y = y0 + ty, z = z0 + tz;
    const bool own = y < P.ny && z < P.nz;
    long long off[SL];
    int pos[SL];
    bool sol[SL];
#pragma unroll
    for (int r = 0; r < SL; ++r) {
        const int i = tid + TPB * r;
        pos[r] = i < EA ? i : -1;
        const int ii = i < EA ? i : 0;
        const int gz = wrapc(z0 - H + ii % EZ, P.nz);
        off[r] = (long long)wrapc(y0 - H + ii / EZ, P.ny) * P.nz + gz;
        sol[r] = is_solid(P, gz);
    }
    double pm[SL], pr[SL];
    auto fetch = [&](int xp) {   // wraps past the chunk (loads unused then)
        const long long base = (long long)wrapc(xp, P.nx) * plane;
#pragma unroll
        for (int r = 0; r < SL; ++r) {
            pm[r] = mub[base + off[r]];
            pr[r] = (double)rho0[base + off[r]];
        }
    };
    auto store = [&](int slot) {
#pragma unroll
        for (int r = 0; r < SL; ++r)
            if (pos[r] >= 0) {
                sm[slot][pos[r]] = sol[r] ? 0.0 : pm[r];   // the Laplacian sees the fluid cotangent only
                sr[slot][pos[r]] = pr[r];
            }
    };
    for (int k = -H; k < H; ++k) { fetch(xa + k); store(k + H); }
    fetch(xa + H);
    const int c0 = (ty + H) * EZ + tz + H;
    double kap = 0;
    for (int x = xa; x < xb; ++x) {
        __syncthreads();                // plane x-H-1 is no longer read
        store((x - xa + 2 * H) % R);    // plane x+H
        __syncthreads();
        fetch(x + H + 1);
        if (own) {
            int slot[R];
#pragma unroll
            for (int d = 0; d < R; ++d) slot[d] = (x - xa + d) % R;
            auto mm = [&](int dx, int dy, int dz) { return sm[slot[dx + H]][c0 + dy * EZ + dz]; };
            auto rr = [&](int dx, int dy, int dz) { return sr[slot[dx + H]][c0 + dy * EZ + dz]; };
            auto lap = [&](auto &&v) {
                if constexpr (H == 3) {
                    auto ax = [&](int ux, int uy, int uz) {
                        return 2 * v(-3 * ux, -3 * uy, -3 * uz) - 27 * v(-2 * ux, -2 * uy, -2 * uz) + 270 * v(-ux, -uy, -uz)
                               + 270 * v(ux, uy, uz) - 27 * v(2 * ux, 2 * uy, 2 * uz) + 2 * v(3 * ux, 3 * uy, 3 * uz);
                    };
                    return (-v(0, 0, 0) * 490 * 3 + ax(1, 0, 0) + ax(0, 1, 0) + ax(0, 0, 1)) / 180;
                } else {
                    auto ax = [&](int ux, int uy, int uz) {
                        return -v(-2 * ux, -2 * uy, -2 * uz) + v(-ux, -uy, -uz) * 16 + v(ux, uy, uz) * 16
                               - v(2 * ux, 2 * uy, 2 * uz);
                    };
                    return (-v(0, 0, 0) * 90 + ax(1, 0, 0) + ax(0, 1, 0) + ax(0, 0, 1)) / 12;
                }
            };
            const long long I = (long long)x * plane + (long long)y * P.nz + z;
            double acc = -P.Ka * lap(mm);
            if (!is_solid(P, z)) {
                const double r = rr(0, 0, 0), mu_b = mm(0, 0, 0);
                acc += P.K2s * dmu_eos(P.eos, P.mu_form, P.A, P.B, P.T, clamp_rho(r)) * clamp_drho(r) * mu_b;
                kap += -mu_b * lap(rr);
            }
            rb_new[I] += acc;
        }
    }
    const double sk = block_sum(kap);
    if (threadIdx.x == 0) kapart[blockIdx.x] += sk;
}

// transposed wall fill, one layer pass (run deepest pass first).  The deepest
// pass also adds the wall-field gradients (solid mu = mu_w of the column) when
// mub is given.
__global__ void __launch_bounds__(TPB) k_adj_wall_pass(const DevParams P, double *__restrict__ rb, int pass,
                                                       const double *__restrict__ mub, double *__restrict__ glo,
                                                       double *__restrict__ ghi) {
    const long long col = (long long)blockIdx.x * TPB + threadIdx.x;
    if (col >= (long long)P.nx * P.ny) return;
    if (mub) {
        const double *m = mub + col * P.nz;
        double a = 0, b = 0;
        for (int k = 0; k < P.lo; ++k) a += m[k];
        for (int k = P.nz - P.hi; k < P.nz; ++k) b += m[k];
        glo[col] += a;
        ghi[col] += b;
    }
    const int x = (int)(col / P.ny), y = (int)(col % P.ny);
    const int x1 = wrapc(x - 1, P.nx), x2 = wrapc(x + 1, P.nx), y1 = wrapc(y - 1, P.ny), y2 = wrapc(y + 1, P.ny);
    auto at = [&](int xx, int yy, int k) { return ((long long)xx * P.ny + yy) * P.nz + k; };
    auto avg = [&](int k) {
        return (rb[at(x, y, k)] * 2 + rb[at(x1, y1, k)] + rb[at(x1, y2, k)] + rb[at(x2, y1, k)] + rb[at(x2, y2, k)]) / 6;
    };
    if (pass < P.lo) {
        const int kd = P.lo - 1 - pass, ks = P.wall_scheme ? kd + 1 : P.lo;
        rb[at(x, y, ks)] += avg(kd);
    }
    if (pass < P.hi) {
        const int kd = P.nz - P.hi + pass, ks = P.wall_scheme ? kd - 1 : P.nz - 1 - P.hi;
        rb[at(x, y, ks)] += avg(kd);
    }
}

// The transposed fill in one launch for walls of at most R (<= 3) layers, over
// 16 x 16 tiles of columns (blockIdx.z: 0 bottom wall, 1 top wall).  Layered
// scheme: the first fluid layer gains A(rb(1) + A(rb(2) + A rb(3))), rb(d) the
// cotangent of the solid layer at depth d and A the symmetric diagonal average;
// layer d is staged in LDS on the tile plus a halo of d columns and the sums are
// formed in the order of the pass kernel above (deepest first), so the results
// are identical.  Simple scheme: the fluid node gains the sum of A rb(d), deepest
// first.  Only the fluid layer is written, and only at the tile's own columns.
template <int R>
__global__ void __launch_bounds__(kWallTile * kWallTile) k_adj_wall_tile(const DevParams P, double *__restrict__ rb,
                                                                          const double *__restrict__ mub,
                                                                          double *__restrict__ glo,
                                                                          double *__restrict__ ghi) {
    constexpr int T = kWallTile, E = T + 2 * R, NT = T * T;
    __shared__ double L[R + 1][E * E];
    const int top = blockIdx.z;
    const int depth = top ? P.hi : P.lo;
    if (depth == 0) return;   // uniform over the workgroup
    const int kf = top ? P.nz - 1 - P.hi : P.lo, step = top ? -1 : 1;
    const int x0 = blockIdx.x * T, y0 = blockIdx.y * T;
    const int x = x0 + (int)threadIdx.x % T, y = y0 + (int)threadIdx.x / T;
    const bool own = x < P.nx && y < P.ny;
    auto at = [&](int xx, int yy, int k) { return ((long long)wrapc(xx, P.nx) * P.ny + wrapc(yy, P.ny)) * P.nz + k; };
    const long long col = (long long)x * P.ny + y;
    if (own) {   // wall-field gradient: solid mu = mu_w of the column
        const double *m = mub + col * P.nz + (top ? P.nz - P.hi : 0);
        double g = 0;
#pragma unroll
        for (int k = 0; k < R; ++k) {   // loads issued together (layers past depth read layer 0, add 0)
            const double t = m[k < depth ? k : 0];
            g += k < depth ? t : 0.0;
        }
        (top ? ghi : glo)[col] += g;
    }
    if (!P.wall_scheme) {
        if (!own) return;
        double acc = rb[at(x, y, kf)];
        for (int d = depth; d >= 1; --d) {
            const int k = kf - d * step;
            acc += (rb[at(x, y, k)] * 2 + rb[at(x - 1, y - 1, k)] + rb[at(x - 1, y + 1, k)] + rb[at(x + 1, y - 1, k)]
                    + rb[at(x + 1, y + 1, k)]) / 6;
        }
        rb[at(x, y, kf)] = acc;
        return;
    }
    // layer d on the patch [R - d, E - R + d)^2 (0 elsewhere); every load of the
    // thread in flight at once
    constexpr int LA = (R + 1) * E * E, PER = (LA + NT - 1) / NT;
    double v[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const int i = (int)threadIdx.x + j * NT;
        const int d = i / (E * E), e = i - d * (E * E), a = e % E, b = e / E;
        const bool use = i < LA && d <= depth && a >= R - d && a < E - R + d && b >= R - d && b < E - R + d;
        v[j] = use ? rb[at(x0 - R + a, y0 - R + b, kf - d * step)] : 0.0;
    }
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const int i = (int)threadIdx.x + j * NT;
        if (i < LA) (&L[0][0])[i] = v[j];
    }
    __syncthreads();
    for (int d = depth - 1; d >= 0; --d) {
        const double *c = L[d + 1];
        for (int i = threadIdx.x; i < E * E; i += NT) {
            const int a = i % E, b = i / E;
            if (a < R - d || a >= E - R + d || b < R - d || b >= E - R + d) continue;
            L[d][i] += (c[i] * 2 + c[i - 1 - E] + c[i - 1 + E] + c[i + 1 - E] + c[i + 1 + E]) / 6;
        }
        __syncthreads();
    }
    if (own) rb[at(x, y, kf)] = L[0][(y - y0 + R) * E + (x - x0 + R)];
}

// acc[0] += sum(part[0..nb)) in a fixed order
__global__ void __launch_bounds__(TPB) k_sum_into(const double *__restrict__ part, long long nb, double *__restrict__ acc) {
    double s = 0;
    for (long long b = threadIdx.x; b < nb; b += TPB) s += part[b];
    s = block_sum(s);
    if (threadIdx.x == 0) acc[0] += s;
}

}  // namespace
}  // namespace cpmp

using namespace cpmp;

#define AHIP(expr)                                                                                \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess) { rc = set_error(CPMP_EHIP, "%s: %s", #expr, hipGetErrorString(e_)); goto done; } \
    } while (0)

namespace {

// Memory plan of one adjoint call.  The trailing ntaped steps are taped during the
// first forward run; the n0 steps before them are covered by checkpoints every C
// steps and recomputed segment by segment with a working tape of S entries.  When
// S < C a segment is recomputed twice: a first pass stores sub-checkpoints every
// S steps (and tapes the last sub-segment), then every earlier sub-segment is
// restored and taped in turn.  With whole segments taped (S = C) and room for a
// second tape, the reverse sweep overlaps the recompute of the next segment (pipe).
struct AdjPlan {
    int C = 0, ntaped = 0, n0 = 0, nseg = 0, S = 0, nsubck = 0, tape_len = 1;
    bool pipe = false;
    size_t per_ck = 0;       // bytes of one checkpoint: populations + rho in the storage form
    size_t per_tape = 0;     // bytes of one tape entry: 10 values of the storage scalar type per node
    size_t nbuf = 0;         // doubles of the cotangent and work buffers
    long long nkap = 0;      // kappa partial-sum slots
    size_t fixed = 0;        // bytes besides checkpoints and tape
    size_t total = 0;        // device bytes allocated by the call
    long long recomputed = 0;
};

// Tiles and chunks of the x-march reverse kernels (k_adj_stencil_march,
// k_adj_mu_march): the longest chunk of 32, 16 or 8 planes that gives at least
// 1024 workgroups, else 8 (shorter chunks re-read too many halo planes: measured
// slower than the one-thread-per-node kernels at 96x96x64, which chunks of 8 beat);
// CPMP_ADJ_XC sets a chunk length.
void adj_march_geometry(const DevParams &P, int &ntz, int &nty, int &xc, unsigned &blocks) {
    ntz = (P.nz + kAdjTZ - 1) / kAdjTZ;
    nty = (P.ny + kAdjTY - 1) / kAdjTY;
    xc = 8;
    for (int cand = 32; cand >= 8; cand /= 2)
        if ((long long)((P.nx + cand - 1) / cand) * ntz * nty >= 1024) { xc = cand; break; }
    if (const char *xe = std::getenv("CPMP_ADJ_XC")) xc = std::max(1, std::min(64, std::atoi(xe)));
    blocks = (unsigned)((P.nx + xc - 1) / xc * ntz * nty);
}

int sub_count(int L, int S) { return (L + S - 1) / S; }

size_t plan_total(const AdjPlan &q, int S, int *nsubck, int *tape_len) {
    const int nsub = q.n0 > 0 ? sub_count(std::min(q.C, q.n0), S) : 0;
    *nsubck = std::max(nsub - 2, 0);
    *tape_len = std::max(std::max(q.ntaped, q.n0 > 0 ? S : 0), 1);
    return (size_t)std::max(q.nseg, 1) * q.per_ck + (size_t)*nsubck * q.per_ck + (size_t)*tape_len * q.per_tape + q.fixed;
}

int make_plan(cpmp_ctx *c, const cpmp_adj_desc *d, AdjPlan &q) {
    if (c->form != CPMP_FORM_FP64 && c->form != CPMP_FORM_FP32)
        return set_error(CPMP_EINVAL, "the adjoint runs in the FP64 and FP32 forms");
    if (d->nsteps < 1 || d->checkpoint_every < 1) return set_error(CPMP_EINVAL, "nsteps and checkpoint_every must be >= 1");
    if (c->p.forcing != 0) return set_error(CPMP_EINVAL, "the adjoint is derived for the exact-difference force term");
    if (c->p.wall_scheme == 2) return set_error(CPMP_EINVAL, "the adjoint covers the layered and simple wall schemes");
    const long long N = c->N, NXY = c->NXY;
    const dim3 ng = node_grid(c->dp);
    const long long nb = blocks_for(N), nbg = (long long)ng.x * ng.y * ng.z;
    q.C = std::min(d->checkpoint_every, d->nsteps);
    q.per_ck = (size_t)N * (19 * c->pop_bytes + c->scal_bytes);
    q.per_tape = 10 * (size_t)N * c->scal_bytes;
    int ntz_a, nty_a, xc_a;
    unsigned mblocks;
    adj_march_geometry(c->dp, ntz_a, nty_a, xc_a, mblocks);
    q.nkap = std::max<long long>(nbg, mblocks);   // kappa partial sums: one slot per workgroup
    // folded population cotangents in the storage precision of the form
    const size_t hb = c->form == CPMP_FORM_FP64 ? sizeof(double) : sizeof(float);
    q.nbuf = 2 * ((19 * (size_t)N * hb + 7) / 8) + (size_t)N * (3 + 6 + 1) + std::max(nb, nbg) + q.nkap + 2 * NXY + 4;
    // buffers, the final state in the storage form, the target
    q.fixed = q.nbuf * sizeof(double) + q.per_ck + (size_t)N * sizeof(double);
    size_t fr = 0, tot = 0;
    const bool need_free = d->tape_budget_bytes < 0 || d->memory_limit_bytes == 0;
    if (need_free) {
        const hipError_t e = hipMemGetInfo(&fr, &tot);
        if (e != hipSuccess) return set_error(CPMP_EHIP, "hipMemGetInfo: %s", hipGetErrorString(e));
    }
    long long budget = d->tape_budget_bytes < 0 ? (long long)(fr / 2) : d->tape_budget_bytes;
    q.ntaped = (int)std::min<long long>(budget / (long long)q.per_tape, d->nsteps);
    q.n0 = d->nsteps - q.ntaped;
    q.nseg = q.n0 > 0 ? (q.n0 + q.C - 1) / q.C : 0;
    // memory limit: given, or the free device memory less 2 GiB
    const size_t margin = (size_t)2 << 30;
    const size_t limit = d->memory_limit_bytes > 0 ? (size_t)d->memory_limit_bytes : (fr > margin ? fr - margin : 0);
    // forward steps run during the reverse sweep with sub-segments of S steps: a
    // segment of L steps costs L + (ceil(L / S) - 1) S (the last sub-segment is
    // taped in the first pass)
    const int Lend = q.nseg > 0 ? q.n0 - (q.nseg - 1) * q.C : 0;
    auto recompute = [&](int S) {
        auto seg = [&](long long L) { return L + (long long)(sub_count((int)L, S) - 1) * S; };
        return q.nseg > 0 ? (long long)(q.nseg - 1) * seg(q.C) + seg(Lend) : 0LL;
    };
    // the schedule with the fewest recomputed steps that fits, then the least memory
    int best = -1;
    long long best_rc = 0;
    size_t best_total = 0, least = SIZE_MAX;
    for (int S = q.n0 > 0 ? std::min(q.C, q.n0) : 0; S >= (q.n0 > 0 ? 1 : 0); --S) {
        int nsc, tl;
        const size_t total = plan_total(q, S, &nsc, &tl);
        least = std::min(least, total);
        if (total > limit) continue;
        const long long rcs = recompute(S);
        if (best < 0 || rcs < best_rc || (rcs == best_rc && total < best_total)) {
            best = S;
            best_rc = rcs;
            best_total = total;
        }
        if (q.n0 == 0) break;
    }
    if (best < 0)
        return set_error(CPMP_ENOMEM, "adjoint needs at least %zu bytes of device memory, limit %zu "
                         "(fewer checkpoints or a smaller tape budget)", least, limit);
    q.S = best;
    q.total = plan_total(q, best, &q.nsubck, &q.tape_len);
    q.recomputed = best_rc;
    // the last segment is taped during the first forward run when the working tape
    // holds it beside the taped steps: no memory beyond the schedule's, one segment
    // fewer to recompute
    if (q.nseg > 0) {
        const int Lend = q.n0 - (q.nseg - 1) * q.C;
        if (q.ntaped + Lend <= q.tape_len) {
            q.ntaped += Lend;
            q.n0 -= Lend;
            q.nseg -= 1;
            q.recomputed -= Lend;
            q.total = plan_total(q, best, &q.nsubck, &q.tape_len);
        }
    }
    // overlap: at least two blocks to reverse (the taped steps, the segments), whole
    // segments taped, and a second tape within the limit; CPMP_ADJ_PIPE=0 turns it off
    const char *pe = std::getenv("CPMP_ADJ_PIPE");
    const int blocks = q.nseg + (q.ntaped > 0 ? 1 : 0);
    const size_t piped = q.total + (size_t)q.tape_len * q.per_tape;
    if (!(pe && std::strcmp(pe, "0") == 0) && q.nseg > 0 && blocks >= 2 && q.S >= std::min(q.C, q.n0) &&
        piped <= limit) {
        q.pipe = true;
        q.total = piped;
    }
    return CPMP_OK;
}

}  // namespace

extern "C" int cpmp_adjoint_plan(cpmp_ctx *c, const cpmp_adj_desc *d, cpmp_adj_plan *out) {
    if (!c || !c->stream || !d || !out) return set_error(CPMP_EINVAL, "bad arguments");
    if (hipSetDevice(c->p.device) != hipSuccess) return set_error(CPMP_EHIP, "hipSetDevice");
    AdjPlan q;
    const int rc = make_plan(c, d, q);
    if (rc != CPMP_OK) return rc;
    out->device_bytes = (int64_t)q.total;
    out->taped_steps = q.ntaped;
    out->segments = q.nseg;
    out->sub_length = q.S;
    out->sub_checkpoints = q.nsubck;
    out->recomputed_steps = q.recomputed;
    out->overlapped = q.pipe ? 1 : 0;
    return CPMP_OK;
}

extern "C" int cpmp_adjoint(cpmp_ctx *c, const cpmp_adj_desc *d, double *J, double *g_lo, double *g_hi, double *g_ka) {
    if (!c || !c->stream || !d || !d->target || !J) return set_error(CPMP_EINVAL, "bad arguments");
    if (hipSetDevice(c->p.device) != hipSuccess) return set_error(CPMP_EHIP, "hipSetDevice");
    AdjPlan q;
    int rc = make_plan(c, d, q);
    if (rc != CPMP_OK) return rc;
    const long long N = c->N, NXY = c->NXY;
    const unsigned nb = blocks_for(N), nbc = blocks_for(NXY);
    const dim3 ng = node_grid(c->dp);
    const long long nbg = (long long)ng.x * ng.y * ng.z;   // blocks of the node grid
    const hipStream_t s = c->stream;
    const DevParams &P = c->dp;
    const dim3 gwt((P.nx + kWallTile - 1) / kWallTile, (P.ny + kWallTile - 1) / kWallTile, 2), bwt(kWallTile * kWallTile);
    const int passes = std::max(P.lo, P.hi);
    int ntz_a, nty_a, xc_a;
    unsigned march_blocks;
    adj_march_geometry(P, ntz_a, nty_a, xc_a, march_blocks);
    const char *sv = std::getenv("CPMP_ADJ_STENCIL");
    const bool march = !(sv && std::strcmp(sv, "v1") == 0);
    // the chemical-potential march pays only with long chunks (measured: 5 % less
    // reverse time at 512x512x256, chunks of 32; 2 % more at 96x96x64, chunks of 8)
    const bool march_mu = march && (xc_a >= 16 || (sv && std::strcmp(sv, "march") == 0));
    const int64_t step0 = c->step;
    const size_t fbytes = 19 * (size_t)N * c->pop_bytes, rbytes = (size_t)N * c->scal_bytes;
    const size_t hb = c->form == CPMP_FORM_FP64 ? sizeof(double) : sizeof(float);   // population cotangents
    // tape entry of step n, in the storage scalar type (FP64 or FP32): rho^n after
    // the wall fill (N), then u^n (3N), grad mu^n (3N) and grad rho^n (3N) as the
    // forward collide evaluated them
    const size_t sb = c->scal_bytes;
    char *ckpt = nullptr, *subck = nullptr, *fin = nullptr, *tape = nullptr, *tape2 = nullptr;
    double *buf = nullptr, *tgt = nullptr;
    // overlapped schedule: the recompute runs on the context's state on stream s2
    hipStream_t s2 = nullptr;
    hipEvent_t ev_first = nullptr, ev_tape[2] = {nullptr, nullptr}, ev_rev[2] = {nullptr, nullptr};
    void *fbA, *fbB;
    double *rb, *rbn, *rbl, *Pf, *Wf, *mub, *part, *kpart, *gl, *gh, *acc;
    double acc_h[2] = {0, 0};
    {
        AHIP(hipMalloc((void **)&ckpt, (size_t)std::max(q.nseg, 1) * q.per_ck));
        if (q.nsubck > 0) AHIP(hipMalloc((void **)&subck, (size_t)q.nsubck * q.per_ck));
        AHIP(hipMalloc((void **)&fin, q.per_ck));
        AHIP(hipMalloc((void **)&tape, (size_t)q.tape_len * q.per_tape));
        if (q.pipe) {
            AHIP(hipMalloc((void **)&tape2, (size_t)q.tape_len * q.per_tape));
            AHIP(hipStreamCreateWithFlags(&s2, hipStreamNonBlocking));
            AHIP(hipEventCreateWithFlags(&ev_first, hipEventDisableTiming));
            for (int i = 0; i < 2; ++i) {
                AHIP(hipEventCreateWithFlags(&ev_tape[i], hipEventDisableTiming));
                AHIP(hipEventCreateWithFlags(&ev_rev[i], hipEventDisableTiming));
            }
        }
        AHIP(hipMalloc((void **)&buf, q.nbuf * sizeof(double)));
        AHIP(hipMalloc((void **)&tgt, N * sizeof(double)));
        {
            double *p = buf;
            // folded population cotangents: two sets of 19 N values of hb bytes
            fbA = p; p += (19 * N * hb + 7) / 8; fbB = p; p += (19 * N * hb + 7) / 8;
            rb = p; p += N; rbn = p; p += N; rbl = p; p += N;
            Pf = p; p += 3 * N; Wf = p; p += 3 * N; mub = p; p += N;
            part = p; p += std::max<long long>(nb, nbg); kpart = p; p += q.nkap;
            gl = p; p += NXY; gh = p; p += NXY; acc = p; p += 4;
        }
        AHIP(hipMemcpyAsync(tgt, d->target, N * sizeof(double), hipMemcpyHostToDevice, s));
        AHIP(hipMemsetAsync(buf, 0, q.nbuf * sizeof(double), s));

        // state = populations of the current set + rho, in the storage form.  The
        // forward work (save, restore, steps) runs on the context's stream: s, or s2
        // while an overlapped recompute is issued.
        auto save = [&](char *dst) -> hipError_t {
            hipError_t e = hipMemcpyAsync(dst, c->f[c->cur], fbytes, hipMemcpyDeviceToDevice, c->stream);
            return e != hipSuccess ? e : hipMemcpyAsync(dst + fbytes, c->rho, rbytes, hipMemcpyDeviceToDevice, c->stream);
        };
        auto restore = [&](const char *src) -> hipError_t {
            hipError_t e = hipMemcpyAsync(c->f[c->cur], src, fbytes, hipMemcpyDeviceToDevice, c->stream);
            return e != hipSuccess ? e : hipMemcpyAsync(c->rho, src + fbytes, rbytes, hipMemcpyDeviceToDevice, c->stream);
        };
        // one forward step that records its tape entry tr
        auto taped_step = [&](char *tr) -> hipError_t {
            hipError_t e;
            if ((e = launch_wall(c)) != hipSuccess) return e;
            if ((e = launch_mu(c, tr)) != hipSuccess) return e;   // also copies rho (after the wall fill) to tr
            if ((e = launch_collide(c, false, tr + N * sb)) != hipSuccess) return e;
            return launch_macro(c);
        };
        auto tape_steps = [&](char *tp, int L) -> hipError_t {
            for (int k = 0; k < L; ++k) {
                const hipError_t e = taped_step(tp + k * q.per_tape);
                if (e != hipSuccess) return e;
            }
            return hipSuccess;
        };
        // reverse sweep over L taped steps (entries 0..L-1 of tape tp), on stream s
        auto reverse = [&](const char *tp, int L) -> hipError_t {
            for (int k = L - 1; k >= 0; --k) {
                const char *t0 = tp + k * q.per_tape;
                auto tape_kernels = [&](auto *tt) {
                    using TT = std::remove_const_t<std::remove_pointer_t<decltype(tt)>>;
                    if (hb == sizeof(double))
                        hipLaunchKernelGGL((k_adj_collide<TT, double>), node_grid(P), dim3(TPB), 0, s, P,
                                           (const double *)fbA, (const double *)rb, tt, tt + N, tt + 4 * N, (double *)fbB,
                                           rbl, Pf, Wf);
                    else
                        hipLaunchKernelGGL((k_adj_collide<TT, float>), node_grid(P), dim3(TPB), 0, s, P,
                                           (const float *)fbA, (const double *)rb, tt, tt + N, tt + 4 * N, (float *)fbB,
                                           rbl, Pf, Wf);
                    if (march)
                        hipLaunchKernelGGL(k_adj_stencil_march, dim3(2 * march_blocks), dim3(TPB), 0, s, P,
                                           (const double *)Pf, (const double *)Wf, (const double *)rbl, mub, rbn, ntz_a,
                                           nty_a, xc_a, (int)march_blocks);
                    else
                        hipLaunchKernelGGL(k_adj_stencil, node_grid(P), dim3(TPB), 0, s, P, (const double *)Pf,
                                           (const double *)Wf, (const double *)rbl, mub, rbn);
                    if (!march_mu)
                        hipLaunchKernelGGL(k_adj_mu<TT>, node_grid(P), dim3(TPB), 0, s, P, tt, (const double *)mub, rbn, kpart);
                    else if (P.lap_points == 7)
                        hipLaunchKernelGGL((k_adj_mu_march<TT, 3>), dim3(march_blocks), dim3(TPB), 0, s, P, tt,
                                           (const double *)mub, rbn, kpart, ntz_a, nty_a, xc_a);
                    else
                        hipLaunchKernelGGL((k_adj_mu_march<TT, 2>), dim3(march_blocks), dim3(TPB), 0, s, P, tt,
                                           (const double *)mub, rbn, kpart, ntz_a, nty_a, xc_a);
                };
                if (sb == sizeof(double)) tape_kernels((const double *)t0);
                else tape_kernels((const float *)t0);
                if (passes == 1)
                    hipLaunchKernelGGL(k_adj_wall_tile<1>, gwt, bwt, 0, s, P, rbn, (const double *)mub, gl, gh);
                else if (passes == 2)
                    hipLaunchKernelGGL(k_adj_wall_tile<2>, gwt, bwt, 0, s, P, rbn, (const double *)mub, gl, gh);
                else if (passes == 3)
                    hipLaunchKernelGGL(k_adj_wall_tile<3>, gwt, bwt, 0, s, P, rbn, (const double *)mub, gl, gh);
                else
                    for (int pass = passes - 1; pass >= 0; --pass)
                        hipLaunchKernelGGL(k_adj_wall_pass, dim3(nbc), dim3(TPB), 0, s, P, rbn, pass,
                                           pass == passes - 1 ? (const double *)mub : nullptr, gl, gh);
                const hipError_t e = hipGetLastError();
                if (e != hipSuccess) return e;
                std::swap(fbA, fbB);
                std::swap(rb, rbn);
            }
            return hipSuccess;
        };

        // ---- forward: checkpoints over the first n0 steps, then taped steps ----
        for (int sg = 0; sg < q.nseg; ++sg) {
            AHIP(save(ckpt + sg * q.per_ck));
            const int L = std::min(q.C, q.n0 - sg * q.C);
            for (int k = 0; k < L; ++k) AHIP(launch_step(c, false));
        }
        AHIP(tape_steps(tape, q.ntaped));
        AHIP(save(fin));
        if (sb == sizeof(double))
            hipLaunchKernelGGL(k_adj_J<double>, dim3(nb), dim3(TPB), 0, s, P, (const double *)c->rho, (const double *)tgt, rb, part);
        else
            hipLaunchKernelGGL(k_adj_J<float>, dim3(nb), dim3(TPB), 0, s, P, (const float *)c->rho, (const double *)tgt, rb, part);
        AHIP(hipGetLastError());
        hipLaunchKernelGGL(k_sum_into, dim3(1), dim3(TPB), 0, s, (const double *)part, (long long)nb, acc);
        AHIP(hipGetLastError());

        // ---- reverse: the taped steps, then the checkpoint segments ----
        if (q.pipe) {
            // Block b (the taped steps first if any, then segments nseg-1 .. 0) uses
            // tape b % 2.  Stream s2 recomputes block b+1 from its checkpoint while s
            // sweeps block b backwards; s2 refills a tape once the sweep of the block
            // two before is done, and s starts a block once its tape is recorded.
            const int first_seg = q.ntaped > 0 ? 1 : 0, nblk = q.nseg + first_seg;
            char *tb[2] = {tape, tape2};
            auto seg_of = [&](int b) { return q.nseg - 1 - (b - first_seg); };
            auto len_of = [&](int b) { return b < first_seg ? q.ntaped : std::min(q.C, q.n0 - seg_of(b) * q.C); };
            AHIP(hipEventRecord(ev_first, s));     // forward run, final state and J issued
            AHIP(hipStreamWaitEvent(s2, ev_first, 0));
            auto recompute_block = [&](int b) -> hipError_t {
                const int i = b & 1;
                hipError_t e = hipSuccess;
                if (b >= 2) e = hipStreamWaitEvent(s2, ev_rev[i], 0);
                c->stream = s2;
                if (e == hipSuccess) e = restore(ckpt + seg_of(b) * q.per_ck);
                if (e == hipSuccess) e = tape_steps(tb[i], len_of(b));
                c->stream = s;
                if (e == hipSuccess) e = hipEventRecord(ev_tape[i], s2);
                return e;
            };
            if (first_seg == 0) AHIP(recompute_block(0));
            for (int b = 0; b < nblk; ++b) {
                if (b + 1 < nblk) AHIP(recompute_block(b + 1));
                const int i = b & 1;
                if (b >= first_seg) AHIP(hipStreamWaitEvent(s, ev_tape[i], 0));
                AHIP(reverse(tb[i], len_of(b)));
                AHIP(hipEventRecord(ev_rev[i], s));
            }
        } else {
            if (q.ntaped > 0) AHIP(reverse(tape, q.ntaped));
            for (int sg = q.nseg - 1; sg >= 0; --sg) {
                const char *ck = ckpt + sg * q.per_ck;
                const int L = std::min(q.C, q.n0 - sg * q.C), S = q.S, nsub = sub_count(L, S);
                AHIP(restore(ck));
                if (nsub == 1) {
                    AHIP(tape_steps(tape, L));
                    AHIP(reverse(tape, L));
                    continue;
                }
                // first pass: sub-checkpoints at S, 2S, .. (nsub-2)S, last sub-segment taped
                for (int j = 1; j < nsub; ++j) {
                    for (int k = 0; k < S; ++k) AHIP(launch_step(c, false));
                    if (j <= nsub - 2) AHIP(save(subck + (j - 1) * q.per_ck));
                }
                const int Llast = L - (nsub - 1) * S;
                AHIP(tape_steps(tape, Llast));
                AHIP(reverse(tape, Llast));
                for (int j = nsub - 2; j >= 0; --j) {
                    AHIP(restore(j == 0 ? ck : subck + (j - 1) * q.per_ck));
                    AHIP(tape_steps(tape, S));
                    AHIP(reverse(tape, S));
                }
            }
        }
    }
    hipLaunchKernelGGL(k_sum_into, dim3(1), dim3(TPB), 0, s, (const double *)kpart, q.nkap, acc + 1);
    AHIP(hipGetLastError());
    AHIP(hipMemcpyAsync(acc_h, acc, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    if (g_lo) AHIP(hipMemcpyAsync(g_lo, gl, NXY * sizeof(double), hipMemcpyDeviceToHost, s));
    if (g_hi) AHIP(hipMemcpyAsync(g_hi, gh, NXY * sizeof(double), hipMemcpyDeviceToHost, s));
    // leave the context at the end of the forward run
    {
        const hipError_t e = hipMemcpyAsync(c->f[c->cur], fin, fbytes, hipMemcpyDeviceToDevice, s);
        AHIP(e);
    }
    AHIP(hipMemcpyAsync(c->rho, fin + fbytes, rbytes, hipMemcpyDeviceToDevice, s));
    AHIP(hipStreamSynchronize(s));
    *J = acc_h[0];
    if (g_ka) *g_ka = acc_h[1];
    c->step = step0 + d->nsteps;
    c->mv_valid = false;
done:
    c->stream = s;
    if (s2) {
        (void)hipStreamSynchronize(s2);
        (void)hipStreamDestroy(s2);
    }
    for (hipEvent_t ev : {ev_first, ev_tape[0], ev_tape[1], ev_rev[0], ev_rev[1]})
        if (ev) (void)hipEventDestroy(ev);
    if (tape2) (void)hipFree(tape2);
    if (ckpt) (void)hipFree(ckpt);
    if (subck) (void)hipFree(subck);
    if (fin) (void)hipFree(fin);
    if (tape) (void)hipFree(tape);
    if (buf) (void)hipFree(buf);
    if (tgt) (void)hipFree(tgt);
    return rc;
}
