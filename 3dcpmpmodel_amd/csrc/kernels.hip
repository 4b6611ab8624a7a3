// kernels.hip -- gfx950 kernels of the chemical-potential multiphase D3Q19 MRT
// step and their launchers.
//
// Layout: reference order I(x,y,z) = (x*ny + y)*nz + z (Common.h:164), z fastest,
// populations SoA f[q*N + I].  Node types are implicit: solid iff z lies in the
// wall layers (the geometry of SetFlowField, FluidCuda.cu:204), so no type array
// is read.  Storage forms: FP64 (double f, rho, mu), FP32 (float), FP16
// (f stored as f - w_q*rho_ref in _Float16, rho_ref the density that the node
// pulling the value had in the step that wrote it; rho, mu float), FP32A (FP64
// storage as FP64, FP32 arithmetic in the collide).
#include "internal.h"
#include "lbm_math.h"
#include "device_common.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <type_traits>

namespace cpmp {

namespace {

template <class TF_, class TS_, class TC_>
struct Form { using TF = TF_; using TS = TS_; using TC = TC_; };
using FormF64 = Form<double, double, double>;
using FormF32 = Form<float, float, float>;
using FormF16 = Form<_Float16, float, float>;
using FormF32A = Form<double, double, float>;   // FP32 arithmetic on FP64 storage

template <class TF> constexpr bool kShifted = std::is_same<TF, _Float16>::value;

template <class Fn>
hipError_t dispatch(int form, Fn &&fn) {
    switch (form) {
    case CPMP_FORM_FP64: return fn(FormF64{});
    case CPMP_FORM_FP32: return fn(FormF32{});
    case CPMP_FORM_FP16: return fn(FormF16{});
    case CPMP_FORM_FP32A: return fn(FormF32A{});
    }
    return hipErrorInvalidValue;
}

// relaxation rates in the arithmetic precision, read from the kernel argument
template <class TC> __device__ __forceinline__ const TC *rates(const DevParams &P);
template <> __device__ __forceinline__ const double *rates<double>(const DevParams &P) { return P.ts; }
template <> __device__ __forceinline__ const float *rates<float>(const DevParams &P) { return P.tsf; }

// FP16 storage holds f_q - w_q rho_ref with rho_ref the density of the node that
// pulls the value, in the step that wrote it (the density the collide kernel read
// there); the stored numbers are then the small deviations from rest, not w_q rho.
// Solid nodes' own slots are stored unshifted (rho_ref = 0).  The other forms
// store f_q and ignore rho_ref.
template <class TF, class TC, class R>
__device__ __forceinline__ TC pop_load(const TF *p, long long i, int q, R rho_ref) {
    if constexpr (kShifted<TF>) return (TC)p[i] + (TC)wq(q) * (TC)rho_ref;
    else return (TC)p[i];
}
template <class TF, class TC, class R>
__device__ __forceinline__ TF pop_val(TC v, int q, R rho_ref) {
    if constexpr (kShifted<TF>) return (TF)(v - (TC)wq(q) * (TC)rho_ref);
    else return (TF)v;
}
// rho_ref of a pulling node from the previous-density buffer (FP16 form only)
template <class TF, class TS>
__device__ __forceinline__ TS ref_at(const TS *rref, long long I) {
    if constexpr (kShifted<TF>) return rref[I];
    else return TS(0);
}


// ---------------------------------------------------------------------------
// Wall fill, one pass (ChemBoundary, FluidCuda.cu:376-401, as ordered layers).
// Pass p writes bottom layer lo-1-p and top layer nz-hi+p from the layer next to
// it (layered scheme) or from the first fluid layer (simple scheme).  Pass 0 also
// writes the wall chemical potential into every solid node of the column.
//
// Scheme 2 is ChemBoundaryComplex (FluidCuda.cu:403-429, launched per level at
// :1281-1283): a solid node at distance d from the fluid (tag LEVELd, Common.h:40-42;
// fluid is level 0) gets sum_f Alpha_f rho(x + e_f) / sum_f Alpha_f over the
// directions f = 1..18, in that order, whose neighbour (periodic in x, y, z) has level
// d - 1.  Pass p fills level p + 1 of both walls; it reads only level p nodes, which
// no thread of the pass writes.
__device__ __forceinline__ int wall_level(const DevParams &P, int z) {
    return z < P.lo ? P.lo - z : (z >= P.nz - P.hi ? z - (P.nz - P.hi) + 1 : 0);
}
template <class TS>
__device__ __forceinline__ double level_average(const DevParams &P, const TS *rho, int x, int y, int z, int level) {
    double s = 0, w = 0;
#pragma unroll
    for (int f = 1; f < Q; ++f) {
        if (cz(f) == 0) continue;   // same z: the same level, never level - 1
        const int zz = wrapc(z + cz(f), P.nz);
        if (wall_level(P, zz) != level - 1) continue;
        const int xx = wrapc(x + cx(f), P.nx), yy = wrapc(y + cy(f), P.ny);
        s += wq(f) * (double)rho[((long long)xx * P.ny + yy) * P.nz + zz];
        w += wq(f);
    }
    return s / w;
}

template <class TS>
__global__ void __launch_bounds__(TPB) k_wall(const DevParams P, TS *__restrict__ rho, TS *__restrict__ mu,
                                              const double *__restrict__ muw, int pass) {
    const long long col = (long long)blockIdx.x * TPB + threadIdx.x;
    const long long nxy = (long long)P.nx * P.ny;
    if (col >= nxy) return;
    const int x = (int)(col / P.ny), y = (int)(col % P.ny);
    const int x1 = wrapc(x - 1, P.nx), x2 = wrapc(x + 1, P.nx), y1 = wrapc(y - 1, P.ny), y2 = wrapc(y + 1, P.ny);
    auto at = [&](int xx, int yy, int k) { return ((long long)xx * P.ny + yy) * P.nz + k; };
    auto avg = [&](int k) {
        return ((double)rho[at(x, y, k)] * 2 + (double)rho[at(x1, y1, k)] + (double)rho[at(x1, y2, k)]
                + (double)rho[at(x2, y1, k)] + (double)rho[at(x2, y2, k)]) / 6;
    };
    if (P.wall_scheme == 2) {
        if (pass < P.lo) rho[at(x, y, P.lo - 1 - pass)] = (TS)level_average(P, rho, x, y, P.lo - 1 - pass, pass + 1);
        if (pass < P.hi) rho[at(x, y, P.nz - P.hi + pass)] = (TS)level_average(P, rho, x, y, P.nz - P.hi + pass, pass + 1);
    } else {
        if (pass < P.lo) {
            const int kd = P.lo - 1 - pass, ks = P.wall_scheme ? kd + 1 : P.lo;
            rho[at(x, y, kd)] = (TS)avg(ks);
        }
        if (pass < P.hi) {
            const int kd = P.nz - P.hi + pass, ks = P.wall_scheme ? kd - 1 : P.nz - 1 - P.hi;
            rho[at(x, y, kd)] = (TS)avg(ks);
        }
    }
    if (pass == 0) {
        const double ml = muw[col], mh = muw[nxy + col];
        for (int k = 0; k < P.lo; ++k) mu[at(x, y, k)] = (TS)ml;
        for (int k = P.nz - P.hi; k < P.nz; ++k) mu[at(x, y, k)] = (TS)mh;
    }
}

// The same fill in one launch for walls of at most R (<= 3) layers.  With the
// layered scheme layer lo-d is A^d applied to the first fluid layer, A the diagonal
// average above, so a 16 x 16 tile of columns needs the fluid layer on the tile
// plus a halo of R columns.  The workgroup stages that patch in LDS and recomputes
// the layers on shrinking patches with the same arithmetic (including the rounding
// to the storage type) as the pass kernels, so the results are identical.
// blockIdx.z selects the wall (0: bottom, 1: top).

template <int R, class TS>
__global__ void __launch_bounds__(kWallTile * kWallTile) k_wall_tile(const DevParams P, TS *__restrict__ rho,
                                                                      TS *__restrict__ mu, const double *__restrict__ muw) {
    constexpr int T = kWallTile, E = T + 2 * R, NT = T * T;
    __shared__ double buf[2][E * E];
    const int top = blockIdx.z;
    const int depth = top ? P.hi : P.lo;
    if (depth == 0) return;   // uniform over the workgroup
    const int kf = top ? P.nz - 1 - P.hi : P.lo, step = top ? -1 : 1;
    const int x0 = blockIdx.x * T, y0 = blockIdx.y * T;
    const int x = x0 + (int)threadIdx.x % T, y = y0 + (int)threadIdx.x / T;
    const bool own = x < P.nx && y < P.ny;
    auto at = [&](int xx, int yy, int k) { return ((long long)wrapc(xx, P.nx) * P.ny + wrapc(yy, P.ny)) * P.nz + k; };
    if (own) {
        const double m = muw[(long long)top * P.nx * P.ny + (long long)x * P.ny + y];
        for (int d = 0; d < depth; ++d) mu[at(x, y, kf - (d + 1) * step)] = (TS)m;
    }
    if (!P.wall_scheme) {   // simple scheme: every layer from the first fluid layer
        if (!own) return;
        const double v = ((double)rho[at(x, y, kf)] * 2 + (double)rho[at(x - 1, y - 1, kf)] + (double)rho[at(x - 1, y + 1, kf)]
                          + (double)rho[at(x + 1, y - 1, kf)] + (double)rho[at(x + 1, y + 1, kf)]) / 6;
        for (int d = 1; d <= depth; ++d) rho[at(x, y, kf - d * step)] = (TS)v;
        return;
    }
    for (int i = threadIdx.x; i < E * E; i += NT) buf[0][i] = (double)rho[at(x0 - R + i % E, y0 - R + i / E, kf)];
    __syncthreads();
    for (int d = 1; d <= depth; ++d) {
        const double *cur = buf[(d - 1) & 1];
        double *nxt = buf[d & 1];
        // layer d is needed on the patch [d, E - d)^2
        for (int i = threadIdx.x; i < E * E; i += NT) {
            const int a = i % E, b = i / E;
            if (a < d || a >= E - d || b < d || b >= E - d) continue;
            nxt[i] = (double)(TS)((cur[i] * 2 + cur[i - 1 - E] + cur[i - 1 + E] + cur[i + 1 - E] + cur[i + 1 + E]) / 6);
        }
        __syncthreads();
        if (own) rho[at(x, y, kf - d * step)] = (TS)nxt[(y - y0 + R) * E + (x - x0 + R)];
    }
}

// ---------------------------------------------------------------------------
// Chemical potential (ChemPotential, FluidCuda.cu:444-463):
//   mu = K^2 mu_eos(clamp(rho)) - Ka * Laplacian(rho).
// FP64 form: all in FP64, the Laplacian as the reference writes it.
// FP32 storage forms (mu is stored in FP32): the Laplacian is summed over the
// differences rho_k - rho_0 (its coefficients sum to zero, so this is the same
// operator, and the differences carry no cancellation), and mu_eos is evaluated
// in FP32 with the hardware log2 and reciprocal (relative error of a few ulp,
// below the FP32 rounding of the stored mu).

// mu_eos in FP32 with v_log_f32 / v_rcp_f32; same closed forms as mu_eos_t
__device__ __forceinline__ float ln_hw(float x) { return __builtin_amdgcn_logf(x) * 0.693147180559945f; }
__device__ __forceinline__ float rcp_hw(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float mu_eos_fast(int eos, int form, float A, float B, float T, float d) {
    const float s2 = 1.41421356f;
    const float bd = B * d, rom = rcp_hw(1.0f - bd);
    if (form == 1 || eos == EOS_PRW || eos == EOS_PRM) {
        return T * ln_hw(d * rom) - A * rcp_hw(s2 * 2.0f * B) * ln_hw((s2 - 1.0f + bd) * rcp_hw(s2 + 1.0f - bd))
               + T * rom - A * d * rcp_hw(1.0f + bd * 2.0f - bd * bd);
    }
    if (eos == EOS_VDW) return T * ln_hw(d * rom) + T * rom - 2.0f * A * d;
    if (eos == EOS_RKE || eos == EOS_RKS) {
        const float op = 1.0f + bd;
        return T * ln_hw(d * rom) + T * rom - A * rcp_hw(B) * ln_hw(op) - A * d * rcp_hw(op);
    }
    return T * (ln_hw(d) + (8.0f * bd - 9.0f * bd * bd + 3.0f * bd * bd * bd) * rom * rom * rom) - 2.0f * A * d;
}

// FP32 mu of an interface node from its FP32 Laplacian term Ka * lap: the EOS part
// in FP64, rounded once
__device__ __forceinline__ float mu_band(const DevParams &P, float d, float kal) {
    return (float)(P.K2s * mu_eos_t<double>(P.eos, P.mu_form, P.A, P.B, P.T, clamp_rho((double)d)) - (double)kal);
}
__device__ __forceinline__ bool in_band(const DevParams &P, float d) { return d > P.band_lo && d < P.band_hi; }

// mu at one fluid node; get(dx, dy, dz) returns rho at the offset node.  FP32:
// interface nodes (in_band) get the EOS part in FP64 if EOS64, else the FP32 value;
// kal_out (if given) receives the FP32 Laplacian term.
template <class TS, bool EOS64 = true, class Get>
__device__ __forceinline__ TS mu_node(const DevParams &P, Get get, float *kal_out = nullptr) {
    if constexpr (sizeof(TS) == 8) {
        const double d = get(0, 0, 0);
        double kal;
        if (P.lap_points == 7) {   // GradD7, Common.h:243-246
            auto ax = [&](int ux, int uy, int uz) {
                return 2 * get(-3 * ux, -3 * uy, -3 * uz) - 27 * get(-2 * ux, -2 * uy, -2 * uz)
                       + 270 * get(-ux, -uy, -uz) + 270 * get(ux, uy, uz) - 27 * get(2 * ux, 2 * uy, 2 * uz)
                       + 2 * get(3 * ux, 3 * uy, 3 * uz);
            };
            kal = P.Ka * (-d * 490 * 3 + ax(1, 0, 0) + ax(0, 1, 0) + ax(0, 0, 1)) / 180;
        } else {                   // GradD5, Common.h:189-192
            auto ax = [&](int ux, int uy, int uz) {
                return -get(-2 * ux, -2 * uy, -2 * uz) + get(-ux, -uy, -uz) * 16 + get(ux, uy, uz) * 16
                       - get(2 * ux, 2 * uy, 2 * uz);
            };
            kal = P.Ka * ((-d * 90 + ax(1, 0, 0) + ax(0, 1, 0) + ax(0, 0, 1)) / 12);
        }
        return P.K2s * mu_eos_t<double>(P.eos, P.mu_form, P.A, P.B, P.T, clamp_rho(d)) - kal;
    } else {
        const float d = get(0, 0, 0);
        auto df = [&](int dx, int dy, int dz) { return (float)get(dx, dy, dz) - d; };
        float lap;
        if (P.lap_points == 7) {
            auto ax = [&](int ux, int uy, int uz) {
                return 2.0f * (df(-3 * ux, -3 * uy, -3 * uz) + df(3 * ux, 3 * uy, 3 * uz))
                       - 27.0f * (df(-2 * ux, -2 * uy, -2 * uz) + df(2 * ux, 2 * uy, 2 * uz))
                       + 270.0f * (df(-ux, -uy, -uz) + df(ux, uy, uz));
            };
            lap = (ax(1, 0, 0) + ax(0, 1, 0) + ax(0, 0, 1)) * (1.0f / 180.0f);
        } else {
            auto ax = [&](int ux, int uy, int uz) {
                return 16.0f * (df(-ux, -uy, -uz) + df(ux, uy, uz)) - (df(-2 * ux, -2 * uy, -2 * uz) + df(2 * ux, 2 * uy, 2 * uz));
            };
            lap = (ax(1, 0, 0) + ax(0, 1, 0) + ax(0, 0, 1)) * (1.0f / 12.0f);
        }
        const float kal = (float)P.Ka * lap;
        if (kal_out) *kal_out = kal;
        if (EOS64 && in_band(P, d)) return mu_band(P, d, kal);
        const float dc = d > 9.7f ? 9.7f + (d - 9.7f) * rcp_hw(1.0f + (d - 9.7f) * 1.5f) : d;   // clamp_rho
        const float m = mu_eos_fast(P.eos, P.mu_form, (float)P.A, (float)P.B, (float)P.T, dc);
        return (float)P.K2s * m - kal;
    }
}

// The chemical-potential march finishes its interface nodes at the end: each wave
// appends (node, density, Laplacian term) to its run as it goes (lanes in lane
// order; count is the wave's running total, the same in every lane), then sets
// their mu with the FP64 EOS part, 64 at a time.
__device__ __forceinline__ void band_append(bool in, unsigned idx, float d, float kal, uint4 *run, unsigned &count) {
    const unsigned long long m = __ballot(in);
    const int lane = (int)__lane_id();
    if (in) run[count + (unsigned)__popcll(m & ((1ull << lane) - 1ull))] =
        make_uint4(idx, __float_as_uint(d), __float_as_uint(kal), 0u);
    count += (unsigned)__popcll(m);
}
// A wave finishes the first kBandWave entries of its run; the rest (the interface
// crosses some waves' 64 columns for most of their planes: up to ~1800 entries at 256^3
// against a mean of ~70 per wave) goes to k_band_overflow in descriptors of kBandDesc
// entries, so that no wave holds the end of the launch for tens of rounds of the FP64
// EOS.  run_index: the run's first slot in the band array; ovf_cap: the descriptor
// slots (never exceeded while each launch's counter starts at zero; past it a wave
// marks its reserved slots empty and finishes all its entries itself).
__device__ __forceinline__ void band_finish(const DevParams &P, const uint4 *run, unsigned count, float *mu,
                                            unsigned run_index, uint2 *ovf, unsigned *ovf_cnt, unsigned ovf_cap) {
    if (count == 0) return;   // uniform over the wave
    __threadfence_block();    // entries written by other lanes are visible
    if (count > kBandWave) {  // uniform
        const unsigned extra = count - kBandWave, nd = (extra + kBandDesc - 1) / kBandDesc;
        unsigned d0 = 0;
        if (__lane_id() == 0) d0 = atomicAdd(ovf_cnt, nd);
        d0 = (unsigned)__shfl((int)d0, 0);
        const bool fits = d0 <= ovf_cap && nd <= ovf_cap - d0;
        for (unsigned j = __lane_id(); j < nd && d0 + j < ovf_cap; j += 64)
            ovf[d0 + j] = fits ? make_uint2(run_index + kBandWave + kBandDesc * j, min(kBandDesc, extra - kBandDesc * j))
                               : make_uint2(0u, 0u);
        if (fits) count = kBandWave;
    }
    for (unsigned i = __lane_id(); i < count; i += 64) {
        const uint4 e = run[i];
        mu[e.x] = mu_band(P, __uint_as_float(e.y), __uint_as_float(e.z));
    }
}

// The interface entries beyond the first kBandWave of the runs of one mu launch: one
// workgroup per descriptor of at most kBandDesc entries.  Also clears the descriptor
// counter that the next launch posts to (cnt_next), so that no memset is needed.
__global__ void __launch_bounds__(kBandDesc) k_band_overflow(const DevParams P, const uint4 *__restrict__ band,
                                                             const uint2 *__restrict__ ovf,
                                                             const unsigned *__restrict__ cnt,
                                                             unsigned *__restrict__ cnt_next, float *__restrict__ mu,
                                                             unsigned cap) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *cnt_next = 0u;
    const unsigned n = min(*cnt, cap);
    for (unsigned d = blockIdx.x; d < n; d += gridDim.x) {
        const uint2 ds = ovf[d];
        if (threadIdx.x < ds.y) {
            const uint4 e = band[ds.x + threadIdx.x];
            mu[e.x] = mu_band(P, __uint_as_float(e.y), __uint_as_float(e.z));
        }
    }
}

// one thread per node, stencil values from the caches
// rho_solid (when given) is the buffer the next collide sums the density into;
// its solid nodes receive this step's wall values.
template <class TS>
__global__ void __launch_bounds__(TPB) k_mu(const DevParams P, const TS *__restrict__ rho, TS *__restrict__ mu,
                                            TS *__restrict__ rho_solid) {
    const NodeIdx nd = node_from_grid(P);
    if (!nd.ok) return;
    const int I = (int)nd.I, pl = P.ny * P.nz;
    if (is_solid(P, nd.z)) {
        if (rho_solid) rho_solid[I] = rho[I];
        return;
    }
    // 32-bit offsets of coordinate + s, s = -3..3 (periodic wrap folded in)
    int ox[7], oy[7], oz[7];
#pragma unroll
    for (int s = 0; s < 7; ++s) {
        ox[s] = (wrap_off(nd.x, s - 3, P.nx) - nd.x) * pl;
        oy[s] = (wrap_off(nd.y, s - 3, P.ny) - nd.y) * P.nz;
        oz[s] = wrap_off(nd.z, s - 3, P.nz) - nd.z;
    }
    mu[I] = mu_node<TS>(P, [&](int dx, int dy, int dz) { return rho[I + ox[dx + 3] + oy[dy + 3] + oz[dz + 3]]; });
}


// ---------------------------------------------------------------------------
// Fused non-ideal force + MRT collision + exact-difference forcing + push
// streaming with half-way bounce-back (NonidealForce + GlobalCollide,
// FluidCuda.cu:311-347, 431-441).  u = j(f)/rho with rho the stored density.
// GUO: the Guo force term instead of the exact-difference one (collide()).
template <class TF, class TS, class TC, bool MV, bool GUO = false>
__global__ void __launch_bounds__(TPB) k_collide(const DevParams P, const TF *__restrict__ fin, TF *__restrict__ fout,
                                                 const TS *__restrict__ rho, const TS *__restrict__ mu,
                                                 double *__restrict__ mv, TS *__restrict__ uout,
                                                 const TS *__restrict__ rref) {
    const NodeIdx nd = node_from_grid(P);
    if (!nd.ok) return;
    const long long I = nd.I;
    const Coord c{nd.x, nd.y, nd.z};
    // a solid node's slots are never pulled: left as they are (storing zeros into them,
    // as the march collide does for its whole rows, measured 6 % slower here)
    if (is_solid(P, c.z)) return;
    const long long N = P.N;
    int xs[7], ys[7], zs[7];
#pragma unroll
    for (int s = 0; s < 7; ++s) {
        xs[s] = wrapc(c.x + s - 3, P.nx);
        ys[s] = wrapc(c.y + s - 3, P.ny);
        zs[s] = wrapc(c.z + s - 3, P.nz);
    }
    const Stencil<TS, TC> st{P, xs, ys, zs};
    const unsigned Nu = (unsigned)N;
    const Nbr nb = make_nbr(P, c.x, c.y, c.z);
    const unsigned Iu = nb.site;

    TC f[Q];
    const TS ref = ref_at<TF>(rref, I);
#pragma unroll
    for (int q = 0; q < Q; ++q) f[q] = pop_load<TF, TC>(fin, pull_index(nb, Nu, q), q, ref);

    const TC r = (TC)rho[I];
    TC gmx, gmy, gmz, grx, gry, grz;
    st.grad(mu, gmx, gmy, gmz);
    st.grad(rho, grx, gry, grz);
    const TC third = TC(1.0 / 3.0);
    const TC Fx = -gmx * r + grx * third, Fy = -gmy * r + gry * third, Fz = -gmz * r + grz * third;
    TC out[Q], u[3];
    if (P.qp_on) {
        TC qp[Q];
        qp_terms(grx, gry, grz, (TC)P.K1, (TC)P.K2, qp);
        collide<TC, false, GUO>(f, r, Fx, Fy, Fz, rates<TC>(P), qp, out, u);
    } else {
        collide<TC, false, GUO>(f, r, Fx, Fy, Fz, rates<TC>(P), (const TC *)nullptr, out, u);
    }
    if (uout) {   // adjoint tape: u, grad mu, grad rho (9 N values)
        const TC t9[9] = {u[0], u[1], u[2], gmx, gmy, gmz, grx, gry, grz};
#pragma unroll
        for (int a = 0; a < 9; ++a) uout[a * N + I] = (TS)t9[a];
    }
    if (MV) {
        const TC h = TC(0.5) / r;
        mv[I] = (double)(u[0] + Fx * h);
        mv[N + I] = (double)(u[1] + Fy * h);
        mv[2 * N + I] = (double)(u[2] + Fz * h);
    }
    // post-collision values stay at this node; the next step pulls them (FP16:
    // shifted by the density the puller has now, or this node's when bounced back)
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        TS tref = TS(0);
        if constexpr (kShifted<TF>) {
            const int tz = zs[3 + cz(q)];
            tref = is_solid(P, tz) ? rho[I] : rho[((long long)xs[3 + cx(q)] * P.ny + ys[3 + cy(q)]) * P.nz + tz];
        }
        fout[(unsigned)q * Nu + Iu] = pop_val<TF, TC>(out[q], q, tref);
    }
}

// ---------------------------------------------------------------------------
// The same fused step as k_collide, organised for gfx950: a workgroup owns an
// in-plane tile of MTY x MTZ nodes (z contiguous, 32 lanes per row) and marches
// along x over P.xc planes (march_chunk).  rho and mu of planes x-3..x+3 with an in-plane halo
// of 3 sit in an LDS ring, so each value is fetched from L2/HBM about once per
// tile instead of 54 times per stencil; the populations of plane x are loaded
// (coalesced) before the ring update so their latency overlaps the barrier.
constexpr int MTY = TILE_Y, MTZ = TILE_Z, MH = 3, MXC = 32;   // MXC: longest chunk (P.xc <= MXC)
constexpr int MT = MTY * MTZ;   // threads of a march workgroup: one per tile column
constexpr int MLY = MTY + 2 * MH, MLZ = MTZ + 2 * MH, MLP = MLY * MLZ;

template <class TS, class TC>
struct LdsStencil {
    const TS (*a)[MLP];
    const int *slot;   // ring slot of plane x + dx, dx = -3..3
    int base;          // (ty + MH) * MLZ + tz + MH
    __device__ __forceinline__ TC at(int dx, int dy, int dz) const { return (TC)a[slot[dx + 3]][base + dy * MLZ + dz]; }
    __device__ __forceinline__ TC dd(int ux, int uy, int uz) const {
        return -at(-3 * ux, -3 * uy, -3 * uz) + TC(9) * at(-2 * ux, -2 * uy, -2 * uz) - TC(45) * at(-ux, -uy, -uz)
               + TC(45) * at(ux, uy, uz) - TC(9) * at(2 * ux, 2 * uy, 2 * uz) + at(3 * ux, 3 * uy, 3 * uz);
    }
    // nine directional differences: each face diagonal serves two components
    // (dd(-u) = -dd(u) for the antisymmetric 6-point difference)
    __device__ __forceinline__ void grad(TC &gx, TC &gy, TC &gz) const {
        const TC s = TC(1.0 / 360.0);
        const TC xy = dd(1, 1, 0), xmy = dd(1, -1, 0), xz = dd(1, 0, 1), xmz = dd(1, 0, -1);
        const TC yz = dd(0, 1, 1), ymz = dd(0, 1, -1);
        gx = (TC(2) * dd(1, 0, 0) + xz + xmz + xy + xmy) * s;
        gy = (TC(2) * dd(0, 1, 0) + yz + ymz + xy - xmy) * s;
        gz = (TC(2) * dd(0, 0, 1) + xz - xmz + yz - ymz) * s;
    }
};

// rho and mu of one node side by side in LDS: one 8-byte (16-byte in FP64) read
// serves both stencils
template <class TS>
struct __align__(2 * sizeof(TS)) RhoMu {
    TS r, m;
};

// The isotropic gradients of rho and mu (Stencil above) from the RhoMu ring.
template <class TS, class TC>
struct LdsStencil2 {
    const RhoMu<TS> (*a)[MLP];
    const int *slot;   // ring slot of plane x + dx, dx = -3..3
    int base;          // (ty + MH) * MLZ + tz + MH
    __device__ __forceinline__ RhoMu<TS> at(int dx, int dy, int dz) const { return a[slot[dx + 3]][base + dy * MLZ + dz]; }
    // 6-point differences of rho (dr) and mu (dm) along u
    __device__ __forceinline__ void dd(int ux, int uy, int uz, TC &dr, TC &dm) const {
        const RhoMu<TS> m3 = at(-3 * ux, -3 * uy, -3 * uz), m2 = at(-2 * ux, -2 * uy, -2 * uz), m1 = at(-ux, -uy, -uz);
        const RhoMu<TS> p1 = at(ux, uy, uz), p2 = at(2 * ux, 2 * uy, 2 * uz), p3 = at(3 * ux, 3 * uy, 3 * uz);
        dr = -(TC)m3.r + TC(9) * (TC)m2.r - TC(45) * (TC)m1.r + TC(45) * (TC)p1.r - TC(9) * (TC)p2.r + (TC)p3.r;
        dm = -(TC)m3.m + TC(9) * (TC)m2.m - TC(45) * (TC)m1.m + TC(45) * (TC)p1.m - TC(9) * (TC)p2.m + (TC)p3.m;
    }
    __device__ __forceinline__ void grad(TC &grx, TC &gry, TC &grz, TC &gmx, TC &gmy, TC &gmz) const {
        const TC s = TC(1.0 / 360.0);
        TC rx, mx, ry, my, rz, mz, rxy, mxy, rxmy, mxmy, rxz, mxz, rxmz, mxmz, ryz, myz, rymz, mymz;
        dd(1, 0, 0, rx, mx);
        dd(0, 1, 0, ry, my);
        dd(0, 0, 1, rz, mz);
        dd(1, 1, 0, rxy, mxy);
        dd(1, -1, 0, rxmy, mxmy);
        dd(1, 0, 1, rxz, mxz);
        dd(1, 0, -1, rxmz, mxmz);
        dd(0, 1, 1, ryz, myz);
        dd(0, 1, -1, rymz, mymz);
        grx = (TC(2) * rx + rxz + rxmz + rxy + rxmy) * s;
        gry = (TC(2) * ry + ryz + rymz + rxy - rxmy) * s;
        grz = (TC(2) * rz + rxz - rxmz + ryz - rymz) * s;
        gmx = (TC(2) * mx + mxz + mxmz + mxy + mxmy) * s;
        gmy = (TC(2) * my + myz + mymz + mxy - mxmy) * s;
        gmz = (TC(2) * mz + mxz - mxmz + myz - mymz) * s;
    }
};

constexpr int AY = MTY + 2, AZ = MTZ + 2, AP = AY * AZ, ARING = AP - MTY * MTZ;

// accumulator position (ly, lz) of ring entry r: rows ly = 0 and ly = AY-1, then
// the columns lz = 0 and lz = AZ-1 of rows 1..MTY
__device__ __forceinline__ int ring_pos(int r) {
    if (r < AZ) return r;
    if (r < 2 * AZ) return (AY - 1) * AZ + (r - AZ);
    return (1 + ((r - 2 * AZ) >> 1)) * AZ + (((r - 2 * AZ) & 1) ? AZ - 1 : 0);
}
__device__ __forceinline__ int ring_index(int ly, int lz) {
    if (ly == 0) return lz;
    if (ly == AY - 1) return AZ + lz;
    return 2 * AZ + 2 * (ly - 1) + (lz == AZ - 1 ? 1 : 0);
}

// Border exchange of the summed density, by receiving tile.  Each tile owns ARING
// slots per plane for what its neighbours' collides add to its border nodes:
// row 0 from the tile below (y-1), row MTY-1 from the tile above, column 0 from the
// z-1 tile, column MTZ-1 from the z+1 tile, and the four corners from the diagonal
// tiles.  A workgroup thus reads its own border contributions as one contiguous run.
__device__ __forceinline__ int slot_y0(int tz) { return tz; }
__device__ __forceinline__ int slot_y1(int tz) { return MTZ + tz; }
__device__ __forceinline__ int slot_z0(int ty) { return 2 * MTZ + ty; }
__device__ __forceinline__ int slot_z1(int ty) { return 2 * MTZ + MTY + ty; }
// corner of the receiving tile fed by the diagonal neighbour at (dy, dz)
__device__ __forceinline__ int slot_c(int dy, int dz) { return 2 * MTZ + 2 * MTY + (dy > 0 ? 2 : 0) + (dz > 0 ? 1 : 0); }
// receiving tile (offset dy, dz) and slot of the emitting tile's accumulator border
// position (ly, lz)
__device__ __forceinline__ void ring_dest(int ly, int lz, int &dy, int &dz, int &slot) {
    if (ly == 0) {
        dy = -1;
        if (lz == 0) { dz = -1; slot = slot_c(1, 1); }
        else if (lz == AZ - 1) { dz = 1; slot = slot_c(1, -1); }
        else { dz = 0; slot = slot_y1(lz - 1); }
    } else if (ly == AY - 1) {
        dy = 1;
        if (lz == 0) { dz = -1; slot = slot_c(-1, 1); }
        else if (lz == AZ - 1) { dz = 1; slot = slot_c(-1, -1); }
        else { dz = 0; slot = slot_y0(lz - 1); }
    } else {
        dy = 0;
        if (lz == 0) { dz = -1; slot = slot_z1(ly - 1); }
        else { dz = 1; slot = slot_z0(ly - 1); }
    }
}

// Chemical potential as a march over the same tiles: rho of planes x-3..x+3 with an
// in-plane halo of 3 sits in an LDS ring, so the 13 (or 19) stencil values of every
// node come from LDS and each rho value is fetched from memory about once per tile.
// Arithmetic of mu_node, as k_mu.
// FIX (periodic FP32 lattices with the density summed by the collide kernel): rho
// holds the previous collide's tile sums, which this kernel completes on the fly
// with the terms k_density_fix adds (the ring entry and, on a chunk's first / last
// plane, the partial-plane entry of the y-, z- and diagonal neighbour tile holding
// the node, then the tile's own partial-plane entry: the same order, with +0 for
// absent terms, so the same values) as each plane enters the ring; the completed
// density of the workgroup's own tile and chunk is written to rho_out.  This
// replaces the separate completion pass.
template <class TS, bool FIX, int H>
__global__ void __launch_bounds__(MT) k_mu_march(const DevParams P, const TS *__restrict__ rho, TS *__restrict__ mu,
                                                  int ntz, int nty, TS *__restrict__ rho_solid, int xc,
                                                  const TS *__restrict__ rbuf, const TS *__restrict__ cbuf,
                                                  TS *__restrict__ rho_out, int nch, uint4 *__restrict__ band,
                                                  uint2 *__restrict__ band_ovf, unsigned *__restrict__ band_cnt,
                                                  unsigned band_cap) {
    // ring of 2H+2 planes with an in-plane halo of H (2 for the 5-point Laplacian, 3 for
    // the 7-point one): the stencil reads 2H+1 of them, and the extra slot lets the
    // plane entering the ring be stored while slower waves still read the plane that
    // left it, so one barrier per plane suffices
    constexpr int R = 2 * H + 2, LZ = MTZ + 2 * H, LP = (MTY + 2 * H) * LZ;
    __shared__ TS s_rho[R][LP];
    // FP32: this wave's run of interface nodes and their count
    const unsigned run_index = (unsigned)(((size_t)blockIdx.x * (MT / 64) + threadIdx.x / 64) * 64 * xc);
    uint4 *band_run = band ? band + run_index : nullptr;
    unsigned band_count = 0;
    int t = blockIdx.x;
    if ((gridDim.x & 7) == 0) t = (t & 7) * (gridDim.x >> 3) + (t >> 3);
    const int tzi = t % ntz;
    t /= ntz;
    const int tyi = t % nty;
    const int ch = t / nty;
    const int y0 = tyi * MTY, z0 = tzi * MTZ;
    const int xa = ch * xc, xb = min(xa + xc, P.nx);
    const int ty = threadIdx.x / MTZ, tz = threadIdx.x % MTZ;
    const int y = y0 + ty, z = z0 + tz;
    const bool inr = (y < P.ny) && (z < P.nz);
    const bool active = inr && !is_solid(P, z);
    constexpr int HE = (LP + MT - 1) / MT;
    int hoff[HE];
    // FIX: in-plane part of the ring index of the (up to three) neighbouring tiles
    // holding each halo node (y-, z-, diagonal neighbour), -1 where absent
    int ro[FIX ? HE : 1][3];
    // and of their partial-plane entries (slot 3: the node's own tile), without the
    // chunk and side, which are uniform per plane
    int co[FIX ? HE : 1][4];
    const long long plane_ring = (long long)nty * ntz * ARING;
#pragma unroll
    for (int h = 0; h < HE; ++h) {
        const int e = threadIdx.x + h * MT;
        const int ly = e / LZ, lz = e - ly * LZ;
        const int hy = e < LP ? wrapc(y0 - H + ly, P.ny) : 0;
        const int hz = e < LP ? wrapc(z0 - H + lz, P.nz) : 0;
        hoff[h] = hy * P.nz + hz;
        if constexpr (FIX) {
            const int gyi = hy / MTY, gzi = hz / MTZ, gy = hy % MTY, gz = hz % MTZ;
            const bool ey0 = gy == 0, ey1 = gy == MTY - 1, ez0 = gz == 0, ez1 = gz == MTZ - 1;
            const int own_ring = (gyi * ntz + gzi) * ARING;
            auto set = [&](int i, bool on, int dy, int dz, int ly2, int lz2, int slot) {
                if (!on) return;
                const int yy = (gyi + dy + nty) % nty, zz = (gzi + dz + ntz) % ntz;
                ro[h][i] = own_ring + slot;
                co[h][i] = (yy * ntz + zz) * 2 * AP + ly2 * AZ + lz2;
            };
            ro[h][0] = ro[h][1] = ro[h][2] = -1;
            co[h][0] = co[h][1] = co[h][2] = 0;
            set(0, ey0, -1, 0, AY - 1, gz + 1, slot_y0(gz));
            set(0, ey1, 1, 0, 0, gz + 1, slot_y1(gz));
            set(1, ez0, 0, -1, gy + 1, AZ - 1, slot_z0(gy));
            set(1, ez1, 0, 1, gy + 1, 0, slot_z1(gy));
            set(2, ey0 && ez0, -1, -1, AY - 1, AZ - 1, slot_c(-1, -1));
            set(2, ey0 && ez1, -1, 1, AY - 1, 0, slot_c(-1, 1));
            set(2, ey1 && ez0, 1, -1, 0, AZ - 1, slot_c(1, -1));
            set(2, ey1 && ez1, 1, 1, 0, 0, slot_c(1, 1));
            co[h][3] = (gyi * ntz + gzi) * 2 * AP + (gy + 1) * AZ + gz + 1;
        }
    }
    // planes xg of the march, xa - H .. xb + H + 1, wrapped into [0, nx) without a
    // division once the lattice is at least 8 planes long
    const bool long_x = P.nx >= 8;
    auto wrap_x = [&](int xg) { return long_x ? (xg < 0 ? xg + P.nx : (xg >= P.nx ? xg - P.nx : xg)) : wrapc(xg, P.nx); };
    // two planes in flight: the halo of planes x+H+1 and x+H+2 while plane x
    // computes (register sets A and B alternate; the loop is unrolled by two so
    // that they keep fixed registers)
    TS hA[HE], hB[HE];
    TS fA[FIX ? HE : 1][3], fB[FIX ? HE : 1][3];
    auto fetch_plane = [&](int xg, TS *hr, TS (*fv)[3]) {   // wraps past the chunk (loads unused then)
        xg = wrap_x(xg);
        const TS *rp = rho + (long long)xg * P.ny * P.nz;
        const TS *fp = FIX ? rbuf + (long long)xg * plane_ring : nullptr;   // ring entries of plane xg
#pragma unroll
        for (int h = 0; h < HE; ++h)
            if ((h + 1) * MT <= LP || threadIdx.x + h * MT < LP) {
                hr[h] = rp[hoff[h]];
                if constexpr (FIX) {
#pragma unroll
                    for (int i = 0; i < 3; ++i) fv[h][i] = ro[h][i] >= 0 ? fp[ro[h][i]] : TS(0);
                }
            }
    };
    // the planes enter the ring in order xa - H, xa - H + 1, ...: the wrapped plane,
    // its position in its chunk and its chunk advance by one per stored plane
    int st_xw = wrap_x(xa - H), st_k = 0, st_ch = 0;
    if constexpr (FIX) {
        st_k = st_xw % xc;
        st_ch = st_xw / xc;
    }
    auto store_plane = [&](int slot, const TS *hr, TS (*fv)[3]) {
        const int k = st_k;
        const bool chunk_end = k == 0 || k == xc - 1;   // partial-plane terms too (uniform)
#pragma unroll
        for (int h = 0; h < HE; ++h)
            if ((h + 1) * MT <= LP || threadIdx.x + h * MT < LP) {
                TS v = hr[h];
                if constexpr (FIX) {
                    // k_density_fix's order: ring and partial-plane entry of each
                    // neighbour, then the own partial-plane entry (+0 off the chunk ends)
                    TS a = TS(0);
                    if (chunk_end) {
                        const int cs = k == 0 ? (st_ch == 0 ? nch - 1 : st_ch - 1) : (st_ch == nch - 1 ? 0 : st_ch + 1);
                        const long long ub = ((long long)cs * nty * ntz * 2 + (k == 0 ? 1 : 0)) * AP;
#pragma unroll
                        for (int i = 0; i < 3; ++i) {
                            const TS c = cbuf[ub + co[h][i]];
                            a += fv[h][i];
                            a += ro[h][i] >= 0 ? c : TS(0);
                        }
                        a += cbuf[ub + co[h][3]];
                    } else {
                        a += fv[h][0];
                        a += fv[h][1];
                        a += fv[h][2];
                    }
                    v = v + a;
                }
                s_rho[slot][threadIdx.x + h * MT] = v;
            }
        // next plane
        st_xw = st_xw + 1 == P.nx ? 0 : st_xw + 1;
        if constexpr (FIX) {
            if (++st_k == xc || st_xw == 0) {
                st_k = 0;
                st_ch = st_xw == 0 ? 0 : st_ch + 1;
            }
        }
    };
    for (int k = -H; k < H; ++k) { fetch_plane(xa + k, hA, fA); store_plane(k + H, hA, fA); }
    fetch_plane(xa + H, hA, fA);
    fetch_plane(xa + H + 1, hB, fB);
    const int base = (ty + H) * LZ + tz + H;
    const int ym = y < P.ny ? y : 0, zm = z < P.nz ? z : 0;
    // ring slot of plane x - H is rs0 (plane p sits in slot (p - xa + H) % R)
    int rs0 = 0;
    auto plane_step = [&](int x, TS *hr, TS (*fv)[3]) {
        int slot[R];
#pragma unroll
        for (int dd = 0; dd < R; ++dd) slot[dd] = rs0 + dd >= R ? rs0 + dd - R : rs0 + dd;
        store_plane(slot[2 * H], hr, fv);                    // plane x+H
        __syncthreads();
        fetch_plane(x + 2 + H, hr, fv);                      // plane x+H+2
        const long long I = ((long long)x * P.ny + ym) * P.nz + zm;
        float kal = 0.0f;
        if (active) {
            mu[I] = mu_node<TS, false>(P, [&](int dx, int dy, int dz) { return s_rho[slot[dx + H]][base + dy * LZ + dz]; },
                                       &kal);
        } else if (rho_solid && inr) {
            rho_solid[I] = s_rho[slot[H]][base];
        }
        if constexpr (sizeof(TS) == 4)
            if (band) {
                const float dcen = s_rho[slot[H]][base];
                band_append(active && in_band(P, dcen), (unsigned)I, dcen, kal, band_run, band_count);
            }
        // rho_out (if given) receives the density of the node: with FIX the completed
        // one, without a copy of rho (the adjoint's tape)
        if (rho_out && inr) rho_out[I] = s_rho[slot[H]][base];
        rs0 = rs0 + 1 == R ? 0 : rs0 + 1;
    };
    int x = xa;
    for (; x + 1 < xb; x += 2) {
        plane_step(x, hA, fA);
        plane_step(x + 1, hB, fB);
    }
    if (x < xb) plane_step(x, hA, fA);
    if constexpr (sizeof(TS) == 4)
        if (band) band_finish(P, band_run, band_count, (float *)mu, run_index, band_ovf, band_cnt, band_cap);
}

// With RS the kernel also sums the next step's density: every post-collision
// value (as stored) is added to the node that will pull it (MacroCalculate,
// FluidCuda.cu:351-372, summed on the emitting side).  Populations along x stay
// in the thread's column (registers); in-plane shifts go to an LDS accumulator
// ring of planes x-1, x, x+1 with a one-node border.  A plane is complete once
// the plane after it has collided: its tile nodes are stored to rho_next (the
// sums over this tile and chunk), its border ring to rbuf, and the two partial
// planes outside the chunk (xa-1, xb) to cbuf.  k_density_fix then adds the
// rings of the neighbouring tiles and the partial planes of the neighbouring
// chunks, in a fixed order, with no atomics to global memory.  Requires
// ny % MTY == nz % MTZ == nx % P.xc == 0 (fuse_geometry).
template <class TF, class TS, class TC, bool MV, bool RS, bool UR, bool NW, bool QPR, bool GUO>
__global__ void __launch_bounds__(MT) k_collide_march(const DevParams P, const TF *__restrict__ fin,
                                                       TF *__restrict__ fout, const TS *__restrict__ rho,
                                                       const TS *__restrict__ mu, double *__restrict__ mv,
                                                       TS *__restrict__ uout, int ntz, int nty,
                                                       TS *__restrict__ rho_next, TS *__restrict__ rbuf,
                                                       TS *__restrict__ cbuf, int remap, const TS *__restrict__ rref) {
    // ring of the planes x-3 .. x+4 (plane p in slot (p - xa + 3) & 7): the stencil
    // reads seven, the eighth receives the next plane while slower waves still read
    // the plane that left it
    __shared__ RhoMu<TS> s_rm[8][MLP];
    // in-plane populations 3..18 of plane x, double-buffered by the parity of x - xa
    // (a wave posting plane x+1 does not wait for slower waves collecting plane x),
    // so that one barrier per plane suffices
    __shared__ TC s_mail[RS ? 2 : 1][RS ? Q - 3 : 1][RS ? MT : 1];
    const long long N = P.N;
    // tile order: plain, or (remap) each XCD a contiguous range of tiles --
    // workgroups are dispatched round-robin over the 8 XCDs
    int t = blockIdx.x;
    if (remap == 1 && (gridDim.x & 7) == 0) t = (t & 7) * (gridDim.x >> 3) + (t >> 3);
    if (remap == 2) {
        // y-row order: workgroup b runs on XCD b % 8, which takes the tile rows
        // tyi = b % 8 (mod 8), the z tiles of a row consecutively, so the tiles that
        // share the partial lines of the z-shifted pulls sit in one L2 (the host
        // selects this only when nty % 8 == 0)
        const int k = t & 7, m = t >> 3, rows = nty >> 3;
        const int tz_ = m % ntz, j = (m / ntz) % rows, ch_ = m / (ntz * rows);
        t = (ch_ * nty + k + 8 * j) * ntz + tz_;
    }
    const int tile = t;
    const int tzi = t % ntz;
    t /= ntz;
    const int tyi = t % nty;
    const int ch = t / nty;
    const int y0 = tyi * MTY, z0 = tzi * MTZ;
    const int xa = ch * P.xc, xb = min(xa + P.xc, P.nx);
    const int ty = threadIdx.x / MTZ, tz = threadIdx.x % MTZ;
    const int y = y0 + ty, z = z0 + tz;
    const bool inr = (y < P.ny) && (z < P.nz);
    const bool active = inr && (NW || !is_solid(P, z));
    // The lanes of solid nodes store zeros into their own slots (which are never
    // pulled) in the same instructions as the fluid lanes of their row, so that a
    // population row of the tile is written whole: a partial line costs the memory
    // a read-modify-write (measured 9-10 % of the collide on walled lattices).

    // Halo entries of one plane owned by this thread (e = tid + h*MT) and their
    // in-plane offsets; the plane index is added per x.
    constexpr int HE = (MLP + MT - 1) / MT;
    int hoff[HE];
#pragma unroll
    for (int h = 0; h < HE; ++h) {
        const int e = threadIdx.x + h * MT;
        const int ly = e / MLZ, lz = e - ly * MLZ;
        hoff[h] = e < MLP ? wrapc(y0 - MH + ly, P.ny) * P.nz + wrapc(z0 - MH + lz, P.nz) : 0;
    }
    TS hr[HE], hm[HE];
    auto fetch_plane = [&](int xg) {        // global -> registers
        const long long pb = (long long)wrapc(xg, P.nx) * P.ny * P.nz;
#pragma unroll
        for (int h = 0; h < HE; ++h)
            if ((h + 1) * MT <= MLP || threadIdx.x + h * MT < MLP) { hr[h] = rho[pb + hoff[h]]; hm[h] = mu[pb + hoff[h]]; }
    };
    auto store_plane = [&](int slot) {      // registers -> LDS ring slot
#pragma unroll
        for (int h = 0; h < HE; ++h)
            if ((h + 1) * MT <= MLP || threadIdx.x + h * MT < MLP) s_rm[slot][threadIdx.x + h * MT] = RhoMu<TS>{hr[h], hm[h]};
    };
    for (int k = -MH; k <= MH; ++k) { fetch_plane(xa + k); store_plane(k + MH); }

    const int ym = y < P.ny ? y : 0, zm = z < P.nz ? z : 0;
    const int base = (ty + MH) * MLZ + tz + MH;
    const unsigned plane = (unsigned)(P.ny * P.nz), own_site = pop_site(P, 0, ym, zm);
    // populations arriving at (xg, ym, zm), gathered from their emitting nodes;
    // the in-plane neighbour offsets and wall flags do not change along x
    Nbr nb = make_nbr(P, xa, ym, zm);
    if (NW) nb.sb[0] = nb.sb[2] = false;   // no solid layers: nothing bounces back
    // FP16 without solid layers: each pull is a per-plane base (uniform) plus a byte
    // offset fixed along x, and each store one base per plane plus a 32-bit offset,
    // which takes 64-bit address arithmetic and scalar-register spills off the VALU
    // the FP16 collide is bound by (pulls 1.8 %, stores 3.5 % of its step; 19 N values
    // of 2 bytes fit 32-bit offsets since the context requires 19 N < 2^31).
    // The memory-bound FP32 and FP64 collides measured no gain.
    constexpr bool VO = NW && kShifted<TF>;
    unsigned vo[VO ? Q : 1];
    if constexpr (VO) {
#pragma unroll
        for (int q = 0; q < Q; ++q)
            vo[q] = ((unsigned)q * (unsigned)N + (unsigned)((int)own_site + nb.oyz[(1 - cy(q)) * 3 + 1 - cz(q)])) *
                    (unsigned)sizeof(TF);
    }
    auto gather = [&](int xg, TC *dst) {
        xg = wrapc(xg, P.nx);
        if constexpr (VO) {
            const char *b0 = (const char *)fin + (size_t)xg * plane * sizeof(TF);
            const char *bm = (const char *)fin + (size_t)wrap_off(xg, -1, P.nx) * plane * sizeof(TF);
            const char *bp = (const char *)fin + (size_t)wrap_off(xg, 1, P.nx) * plane * sizeof(TF);
            const TS ref = ref_at<TF>(rref, ((long long)xg * P.ny + ym) * P.nz + zm);
            // The empty asm makes the offsets values of this plane: otherwise the
            // compiler hoists their zero extension out of the x loop, keeps 19
            // 64-bit register pairs and adds each to its base in 64 bits instead of
            // using the scalar-base form of the load (FP16 step 0.8 % faster).
            static_assert(Q == 19, "one asm operand per population");
            unsigned o[Q];
#pragma unroll
            for (int q = 0; q < Q; ++q) o[q] = vo[q];
            asm volatile("" : "+v"(o[0]), "+v"(o[1]), "+v"(o[2]), "+v"(o[3]), "+v"(o[4]), "+v"(o[5]), "+v"(o[6]),
                         "+v"(o[7]), "+v"(o[8]), "+v"(o[9]), "+v"(o[10]), "+v"(o[11]), "+v"(o[12]), "+v"(o[13]),
                         "+v"(o[14]), "+v"(o[15]), "+v"(o[16]), "+v"(o[17]), "+v"(o[18]));
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const char *b = cx(q) > 0 ? bm : cx(q) < 0 ? bp : b0;   // pulled from x - c_x
                const TF v = __builtin_nontemporal_load((const TF *)(b + o[q]));
                dst[q] = pop_load<TF, TC>(&v, 0, q, ref);
            }
            return;
        }
        nb.site = (unsigned)xg * plane + own_site;
        nb.ox[0] = (wrap_off(xg, -1, P.nx) - xg) * (int)plane;
        nb.ox[2] = (wrap_off(xg, 1, P.nx) - xg) * (int)plane;
        const TS ref = ref_at<TF>(rref, ((long long)xg * P.ny + ym) * P.nz + zm);
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            // Non-temporal loads: a line is read by one workgroup only, except the
            // partial lines of the z-shifted pulls, which the neighbouring z tile
            // reads too and the L2 serves.  FP32 keeps ordinary loads for those (all
            // 19 non-temporal: 2 % slower; the 9 pulls without a z shift: 1 % faster
            // per step at 256^3); FP16 and FP64 storage take the hint on every pull
            // (2.3 % and 1-4 % faster steps)
            const unsigned i = pull_index(nb, (unsigned)N, q);
            if (sizeof(TF) == 4 && cz(q) != 0) {
                dst[q] = pop_load<TF, TC>(fin, i, q, ref);
            } else {
                const TF v = __builtin_nontemporal_load(&fin[i]);
                dst[q] = pop_load<TF, TC>(&v, 0, q, ref);
            }
        }
    };
    // software pipeline: the populations of plane x+1 and the halo of plane x+4
    // are in flight while plane x computes
    TC fnext[Q];
    // the loads below are issued by every thread on every path (inactive threads
    // read valid slots of their clamped coordinates and discard them): the
    // compiler then needs to wait only for the oldest loads at the loop head
    // instead of draining the population stores of the previous plane
    fetch_plane(xa + MH + 1);
    gather(xa, fnext);

    // density accumulation (RS).  Populations 0..2 stay in the node's column
    // (registers acc_*, planes x-1, x, x+1); populations 3..18 leave the column:
    // each node posts them in the mailbox s_mail (0 where bounced back or solid)
    // and every node, and every ring entry of the tile border (threads < ARING,
    // registers rng_*), collects what arrives from this tile in a fixed order.
    const int own = (ty + 1) * AZ + tz + 1;
    const int rpos = ring_pos(threadIdx.x < ARING ? threadIdx.x : 0);
    const int rly = rpos / AZ, rlz = rpos - rly * AZ;
    const bool bb_m = !NW && is_solid(P, wrap_off(zm, -1, P.nz)), bb_p = !NW && is_solid(P, wrap_off(zm, 1, P.nz));
    TC acc_prev = TC(0), acc_cur = TC(0), acc_next = TC(0);
    TC rng_prev = TC(0), rng_cur = TC(0), rng_next = TC(0);
    if constexpr (RS) {
        if (!active) {
#pragma unroll
            for (int q = 3; q < Q; ++q) s_mail[0][q - 3][threadIdx.x] = s_mail[1][q - 3][threadIdx.x] = TC(0);
        }
    }
    int mb = 0;   // mailbox buffer of plane x
    // arrivals from the mailbox of plane x at this thread's own node: sources
    // outside the tile read a valid slot and contribute 0 (no branches)
    const bool sy_lo = ty >= 1, sy_hi = ty <= MTY - 2, sz_lo = tz >= 1, sz_hi = tz <= MTZ - 2;
    auto collect_own = [&](TC &pm, TC &pc, TC &pn) {
#pragma unroll
        for (int q = 3; q < Q; ++q) {
            const bool ok = (cy(q) > 0 ? sy_lo : cy(q) < 0 ? sy_hi : true) && (cz(q) > 0 ? sz_lo : cz(q) < 0 ? sz_hi : true);
            const int src = ok ? (int)threadIdx.x - cy(q) * MTZ - cz(q) : (int)threadIdx.x;
            const TC v = ok ? s_mail[mb][q - 3][src] : TC(0);
            if (cx(q) == 0) pc += v;
            else if (cx(q) > 0) pn += v;
            else pm += v;
        }
    };
    // arrivals at a ring position (ly, lz) of the tile border (accumulator frame)
    // (branch-free: sources outside the tile read slot 0 and contribute 0)
    auto collect = [&](int ly, int lz, TC &pm, TC &pc, TC &pn) {
#pragma unroll
        for (int q = 3; q < Q; ++q) {
            const int sy = ly - 1 - cy(q), sz = lz - 1 - cz(q);
            const bool in = sy >= 0 && sy < MTY && sz >= 0 && sz < MTZ;
            const TC v = s_mail[mb][q - 3][in ? sy * MTZ + sz : 0];
            const TC w = in ? v : TC(0);
            if (cx(q) == 0) pc += w;
            else if (cx(q) > 0) pn += w;
            else pm += w;
        }
    };
    // plane xp (xa-1 <= xp <= xb) is complete over this tile and chunk: write it out
    int rdst = 0;   // receiving tile's slot (in-plane part) of this thread's ring entry
    if (threadIdx.x < ARING) {
        int ddy, ddz, sl;
        ring_dest(rly, rlz, ddy, ddz, sl);
        rdst = (((tyi + ddy + nty) % nty) * ntz + (tzi + ddz + ntz) % ntz) * ARING + sl;
    }
    auto flush = [&](int xp, TC own_v, TC ring_v) {
        if (xp >= xa && xp < xb) {
            if (active) rho_next[((long long)xp * P.ny + y) * P.nz + z] = (TS)own_v;
            if (threadIdx.x < ARING) rbuf[(long long)xp * nty * ntz * ARING + rdst] = (TS)ring_v;
        } else {
            TS *cb = cbuf + ((long long)tile * 2 + (xp < xa ? 0 : 1)) * AP;
            cb[own] = (TS)own_v;
            if (threadIdx.x < ARING) cb[rpos] = (TS)ring_v;
        }
    };

    __syncthreads();   // planes xa-3 .. xa+3 in the ring
    for (int x = xa; x < xb; ++x) {
        const long long I = ((long long)x * P.ny + ym) * P.nz + zm;
        TC f[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) f[q] = fnext[q];
        // plane x+4 into the slot of plane x-4, which every wave finished reading
        // (plane x-1) before the barrier that ended the previous iteration; it is
        // read from the next iteration on, after this iteration's barrier
        store_plane((x - xa + 2 * MH + 1) & 7);
        fetch_plane(x + 2 + MH);   // past the chunk on the last passes (wrapped, unused)
        gather(x + 1, fnext);
        // NW: fluid lanes only; otherwise the solid lanes too, which store zeros
        // (fluid: constant true under NW, the zero path compiles away)
        if (NW ? active : inr) {
            const bool fluid = NW || active;
            int slot[7];
#pragma unroll
            for (int d = 0; d < 7; ++d) slot[d] = (x - xa + d) & 7;
            TC out[Q], r = TC(0);
            if (fluid) {
                const LdsStencil2<TS, TC> st{s_rm, slot, base};
                r = (TC)st.at(0, 0, 0).r;
                TC gmx, gmy, gmz, grx, gry, grz;
                st.grad(grx, gry, grz, gmx, gmy, gmz);
                const TC third = TC(1.0 / 3.0);
                const TC Fx = -gmx * r + grx * third, Fy = -gmy * r + gry * third, Fz = -gmz * r + grz * third;
                TC u[3];
                if (QPR && P.qp_on) {   // QPR = false: the Qp term is off (K1 = K2 = 0), compiled out
                    TC qp[Q];
                    qp_terms(grx, gry, grz, (TC)P.K1, (TC)P.K2, qp);
                    collide<TC, UR, GUO>(f, r, Fx, Fy, Fz, rates<TC>(P), qp, out, u);
                } else {
                    collide<TC, UR, GUO>(f, r, Fx, Fy, Fz, rates<TC>(P), (const TC *)nullptr, out, u);
                }
                if (uout) {   // adjoint tape: u, grad mu, grad rho (9 N values)
                    const TC t9[9] = {u[0], u[1], u[2], gmx, gmy, gmz, grx, gry, grz};
#pragma unroll
                    for (int a = 0; a < 9; ++a) uout[a * N + I] = (TS)t9[a];
                }
                if (MV) {
                    const TC h = TC(0.5) / r;
                    mv[I] = (double)(u[0] + Fx * h);
                    mv[N + I] = (double)(u[1] + Fy * h);
                    mv[2 * N + I] = (double)(u[2] + Fz * h);
                }
            } else {
#pragma unroll
                for (int q = 0; q < Q; ++q) out[q] = TC(0);
            }
            // aligned stores: the next step pulls these post-collision values
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                // FP16: shifted by the density its puller has in this step (this
                // node's when bounced back), read from the LDS ring; a solid node's
                // zeros unshifted
                const bool bounced = (cz(q) < 0 && bb_m) || (cz(q) > 0 && bb_p);
                TC tref = TC(0);
                if constexpr (kShifted<TF>)
                    tref = !fluid ? TC(0) : bounced ? r : (TC)s_rm[slot[3 + cx(q)]][base + cy(q) * MLZ + cz(q)].r;
                const TF sv = pop_val<TF, TC>(out[q], q, tref);
                // streaming store: the populations are read again only next step,
                // after far more data than the caches hold has passed through (a
                // wave's store covers whole lines in every form: FP16 2 % faster per
                // step than with ordinary stores)
                if constexpr (NW && kShifted<TF>) {
                    char *sb = (char *)fout + (size_t)x * plane * sizeof(TF);   // uniform per plane
                    TF *dp = (TF *)(sb + ((unsigned)q * (unsigned)N + own_site) * (unsigned)sizeof(TF));
                    __builtin_nontemporal_store(sv, dp);
                } else {
                    const unsigned at = (unsigned)q * (unsigned)N + (unsigned)x * plane + own_site;
                    __builtin_nontemporal_store(sv, &fout[at]);
                }
                if (RS) {
                    // the value the next step pulls: as stored (and decoded); a solid
                    // node's slots are not pulled and add nothing
                    TC w;
                    w = pop_load<TF, TC>(&sv, 0, q, tref);
                    if (!fluid) w = TC(0);
                    const int qx = cx(q), qy = cy(q), qz = cz(q);
                    if (qy == 0 && qz == 0) {           // resolved per q at compile time
                        if (qx == 0) acc_cur += w;
                        else if (qx > 0) acc_next += w;
                        else acc_prev += w;
                    } else {
                        // bounced back: pulled by this node as opp(q) (selects, no branch)
                        const bool bnc = (qz < 0 && bb_m) || (qz > 0 && bb_p);
                        acc_cur += bnc ? w : TC(0);
                        s_mail[mb][q - 3][threadIdx.x] = bnc ? TC(0) : w;
                    }
                }
            }
        }
        __syncthreads();                       // ring plane x+4 stored, mailbox of plane x posted
        if (RS) {
            collect_own(acc_prev, acc_cur, acc_next);
            if (threadIdx.x < ARING) collect(rly, rlz, rng_prev, rng_cur, rng_next);
            flush(x - 1, acc_prev, rng_prev);  // plane x-1 is complete
            acc_prev = acc_cur;
            acc_cur = acc_next;
            acc_next = TC(0);
            rng_prev = rng_cur;
            rng_cur = rng_next;
            rng_next = TC(0);
            mb ^= 1;
        }
    }
    if (RS) {
        flush(xb - 1, acc_prev, rng_prev);
        flush(xb, acc_cur, rng_cur);
    }
}

// Completes the density summed by k_collide_march<..., RS>: a fluid node on the
// border of its tile receives the ring entries of the (up to three) neighbouring
// tiles that hold it, and a node of the first / last plane of a chunk receives
// the partial plane xb / xa-1 of the previous / next chunk (its own tile's
// entry and the neighbouring tiles' ring entries).  Tiles and chunks wrap
// periodically, so a lattice one tile wide reads its own ring.
// Border items of one tile: for each of its P.xc planes the ETILE border positions
// (rows ty = 0 and MTY-1, then the z ends of the rows between), then on the first
// and last plane the interior positions.
constexpr int ETILE = 2 * MTZ + 2 * (MTY - 2), ITILE = (MTY - 2) * (MTZ - 2);
constexpr int FIX_ITEMS_MAX = MXC * ETILE + 2 * ITILE;

template <class TS>
__global__ void __launch_bounds__(MT) k_density_fix(const DevParams P, TS *__restrict__ rho, const TS *__restrict__ rbuf,
                                                     const TS *__restrict__ cbuf, int ntz, int nty, int nch) {
    int t = blockIdx.x;
    const int tzi = t % ntz;
    t /= ntz;
    const int tyi = t % nty;
    const int ch = t / nty;
    // neighbouring tiles (dy, dz in -1..1, index (dy+1)*3 + dz+1): base of their
    // ring entries in this chunk, and of their partial planes in the previous
    // (xb side) and next (xa-1 side) chunk
    const int cp = (ch + nch - 1) % nch, cn = (ch + 1) % nch;
    long long rbase[9], pbase[9], nbase[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const int yy = (tyi + i / 3 - 1 + nty) % nty, zz = (tzi + i % 3 - 1 + ntz) % ntz;
        rbase[i] = ((long long)yy * ntz + zz) * ARING;   // in-plane part of the ring index
        pbase[i] = ((((long long)cp * nty + yy) * ntz + zz) * 2 + 1) * AP;
        nbase[i] = ((((long long)cn * nty + yy) * ntz + zz) * 2 + 0) * AP;
    }
    // every item of this thread at once (all loads in flight, with 4-byte densities
    // the densities the items complete among them: 0.7 % off the FP16 step; in FP64
    // they would cost 20 VGPRs and 2 of 8 waves per SIMD), then the updates
    constexpr int PER = (FIX_ITEMS_MAX + MT - 1) / MT;
    const int items = P.xc * ETILE + 2 * ITILE;
    const long long plane_ring = (long long)nty * ntz * ARING;
    TS add[PER], old[PER];
    long long at[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const int it = threadIdx.x + j * MT;
        add[j] = TS(0);
        at[j] = -1;
        if (it >= items) continue;
        int k, ty, tz;
        if (it < P.xc * ETILE) {
            k = it / ETILE;
            const int e = it - k * ETILE;
            if (e < MTZ) { ty = 0; tz = e; }
            else if (e < 2 * MTZ) { ty = MTY - 1; tz = e - MTZ; }
            else { ty = 1 + ((e - 2 * MTZ) >> 1); tz = ((e - 2 * MTZ) & 1) ? MTZ - 1 : 0; }
        } else {
            const int jj = it - P.xc * ETILE;
            k = jj < ITILE ? 0 : P.xc - 1;
            const int e = jj < ITILE ? jj : jj - ITILE;
            ty = 1 + e / (MTZ - 2);
            tz = 1 + e % (MTZ - 2);
        }
        const int z = tzi * MTZ + tz;
        if (is_solid(P, z)) continue;
        const bool ey0 = ty == 0, ey1 = ty == MTY - 1, ez0 = tz == 0, ez1 = tz == MTZ - 1;
        const bool lo = k == 0, hi = k == P.xc - 1;
        TS a = TS(0);
        // neighbouring tiles whose border ring holds this node (ring position ly,
        // lz): their ring entry of this plane, and on the chunk's first / last
        // plane their entry of the previous / next chunk's partial plane
        auto from = [&](bool on, int dy, int dz, int ly, int lz, int slot) {
            if (!on) return;
            const int i = (dy + 1) * 3 + dz + 1;
            a += rbuf[(long long)(ch * P.xc + k) * plane_ring + rbase[4] + slot];
            if (lo) a += cbuf[pbase[i] + ly * AZ + lz];
            if (hi) a += cbuf[nbase[i] + ly * AZ + lz];
        };
        from(ey0, -1, 0, AY - 1, tz + 1, slot_y0(tz));
        from(ey1, 1, 0, 0, tz + 1, slot_y1(tz));
        from(ez0, 0, -1, ty + 1, AZ - 1, slot_z0(ty));
        from(ez1, 0, 1, ty + 1, 0, slot_z1(ty));
        from(ey0 && ez0, -1, -1, AY - 1, AZ - 1, slot_c(-1, -1));
        from(ey0 && ez1, -1, 1, AY - 1, 0, slot_c(-1, 1));
        from(ey1 && ez0, 1, -1, 0, AZ - 1, slot_c(1, -1));
        from(ey1 && ez1, 1, 1, 0, 0, slot_c(1, 1));
        // this tile's own entry of the neighbouring chunks' partial planes
        if (lo) a += cbuf[pbase[4] + (ty + 1) * AZ + tz + 1];
        if (hi) a += cbuf[nbase[4] + (ty + 1) * AZ + tz + 1];
        add[j] = a;
        at[j] = ((long long)(ch * P.xc + k) * P.ny + tyi * MTY + ty) * P.nz + z;
        if constexpr (sizeof(TS) == 4) old[j] = rho[at[j]];
    }
#pragma unroll
    for (int j = 0; j < PER; ++j)
        if (at[j] >= 0) rho[at[j]] = (sizeof(TS) == 4 ? old[j] : rho[at[j]]) + add[j];
}

// ---------------------------------------------------------------------------
// Density (MacroCalculate, FluidCuda.cu:351-372): rho = sum_q f_q on fluid nodes.
// rho_in: the density of the step that wrote f (FP16 decoding); rho_out may be rho_in
template <class TF, class TS, class TC>
__global__ void __launch_bounds__(TPB) k_macro_into(const DevParams P, const TF *__restrict__ f, const TS *rho_in,
                                                    TS *rho_out) {
    const NodeIdx nd = node_from_grid(P);
    if (!nd.ok) return;
    const long long I = nd.I;
    if (is_solid(P, nd.z)) return;
    const Nbr nb = make_nbr(P, nd.x, nd.y, nd.z);
    const TS ref = ref_at<TF>(rho_in, I);
    TC s = TC(0);
#pragma unroll
    for (int q = 0; q < Q; ++q) s += pop_load<TF, TC>(f, pull_index(nb, (unsigned)P.N, q), q, ref);
    rho_out[I] = (TS)s;
}

// velocity u = j/rho, rho = sum f (fluid nodes; 0 on solid)
template <class TF, class TS, class TC>
__global__ void __launch_bounds__(TPB) k_velocity(const DevParams P, const TF *__restrict__ fp, double *__restrict__ out,
                                                  const TS *__restrict__ rref) {
    const NodeIdx nd = node_from_grid(P);
    if (!nd.ok) return;
    const long long I = nd.I;
    if (is_solid(P, nd.z)) { out[I] = out[P.N + I] = out[2 * P.N + I] = 0; return; }
    const Nbr nb = make_nbr(P, nd.x, nd.y, nd.z);
    TC f[Q], s = TC(0);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        f[q] = pop_load<TF, TC>(fp, pull_index(nb, (unsigned)P.N, q), q, ref_at<TF>(rref, I));
        s += f[q];
    }
    // the same arithmetic as the velocity of collide(): moment rows 3, 5, 7 times
    // the reciprocal density (the adjoint's tape entries then agree bit for bit
    // whichever kernel recorded them)
    TC m[Q];
    moments(f, m);
    const TC inv = TC(1) / s;
    out[I] = (double)(m[3] * inv);
    out[P.N + I] = (double)(m[5] * inv);
    out[2 * P.N + I] = (double)(m[7] * inv);
}

// force field F = -rho grad(mu) + grad(rho)/3 (NonidealForce) into 3N doubles
template <class TS, class TC>
__global__ void __launch_bounds__(TPB) k_force(const DevParams P, const TS *__restrict__ rho, const TS *__restrict__ mu,
                                               double *__restrict__ out) {
    const NodeIdx nd = node_from_grid(P);
    if (!nd.ok) return;
    const long long I = nd.I;
    const Coord c{nd.x, nd.y, nd.z};
    if (is_solid(P, c.z)) { out[I] = out[P.N + I] = out[2 * P.N + I] = 0; return; }
    int xs[7], ys[7], zs[7];
#pragma unroll
    for (int s = 0; s < 7; ++s) {
        xs[s] = wrapc(c.x + s - 3, P.nx); ys[s] = wrapc(c.y + s - 3, P.ny); zs[s] = wrapc(c.z + s - 3, P.nz);
    }
    const Stencil<TS, TC> st{P, xs, ys, zs};
    TC gmx, gmy, gmz, grx, gry, grz;
    st.grad(mu, gmx, gmy, gmz);
    st.grad(rho, grx, gry, grz);
    const TC r = (TC)rho[I], third = TC(1.0 / 3.0);
    out[I] = (double)(-gmx * r + grx * third);
    out[P.N + I] = (double)(-gmy * r + gry * third);
    out[2 * P.N + I] = (double)(-gmz * r + grz * third);
}

// ---------------------------------------------------------------------------
// Initial conditions (SetFlowField, FluidCuda.cu:199-241)

// populations of a node at rest, Feq(d, 0) = w_q d (FluidCuda.cu:238), written in
// the pull storage convention
template <class TF>
__device__ __forceinline__ void store_feq(const DevParams &P, TF *f, const NodeIdx &nd, double d) {
    const unsigned Nu = (unsigned)P.N;
    const Nbr nb = make_nbr(P, nd.x, nd.y, nd.z);
    if (is_solid(P, nd.z)) {   // own slots, unshifted
#pragma unroll
        for (int q = 0; q < Q; ++q) f[(unsigned)q * Nu + nb.site] = pop_val<TF, double>(d * wq(q), q, 0.0);
        return;
    }
    // FP16: shifted by this node's density (the previous-density buffer is set to
    // the initial density by the caller), so the stored deviations are 0
#pragma unroll
    for (int q = 0; q < Q; ++q) f[pull_index(nb, Nu, q)] = pop_val<TF, double>(d * wq(q), q, d);
}

template <class TF, class TS>
__global__ void __launch_bounds__(TPB) k_init_reference(const DevParams P, TF *__restrict__ f, TS *__restrict__ rho,
                                                        TS *__restrict__ mu, double radius, double width, double rg,
                                                        double rl, double solid) {
    const NodeIdx nd = node_from_grid(P);
    if (!nd.ok) return;
    const long long I = nd.I;
    const Coord c{nd.x, nd.y, nd.z};
    double d;
    if (is_solid(P, c.z)) {
        d = solid;
        mu[I] = (TS)solid;
    } else {
        mu[I] = (TS)0;
        const double ddx = double(c.x) - P.nx / 2, ddy = double(c.y) - P.ny / 2, ddz = double(c.z) - (radius + 2);
        const double r = (double)sqrtf((float)(ddx * ddx + ddy * ddy + ddz * ddz));   // FluidCuda.cu:215
        d = (rl + rg) / 2 - (rl - rg) / 2 * tanh(double(r - radius) * 2 / width);
    }
    rho[I] = (TS)d;
    store_feq(P, f, nd, d);
}

// DROPLETSPLASH branch of SetFlowField (FluidCuda.cu:219-233): a droplet centred
// in the box above a liquid film below z = film; solid rho = mu = 0
template <class TF, class TS>
__global__ void __launch_bounds__(TPB) k_init_splash(const DevParams P, TF *__restrict__ f, TS *__restrict__ rho,
                                                     TS *__restrict__ mu, double radius, double width, double film,
                                                     double rg, double rl) {
    const NodeIdx nd = node_from_grid(P);
    if (!nd.ok) return;
    const long long I = nd.I;
    const Coord c{nd.x, nd.y, nd.z};
    double d = 0;
    mu[I] = (TS)0;
    if (!is_solid(P, c.z)) {
        const double ddx = double(c.x) - P.nx / 2, ddy = double(c.y) - P.ny / 2, ddz = double(c.z) - P.nz / 2;
        const double r = (double)sqrtf((float)(ddx * ddx + ddy * ddy + ddz * ddz));   // FluidCuda.cu:230
        d = rl + (rl - rg) / 2 * (tanh((radius - r) * 2 / width) + tanh((film - double(c.z)) * 2 / width));
    }
    rho[I] = (TS)d;
    store_feq(P, f, nd, d);
}

template <class TF, class TS>
__global__ void __launch_bounds__(TPB) k_init_spheres(const DevParams P, TF *__restrict__ f, TS *__restrict__ rho,
                                                      TS *__restrict__ mu, const cpmp_init_desc D) {
    const NodeIdx nd = node_from_grid(P);
    if (!nd.ok) return;
    const long long I = nd.I;
    const Coord c{nd.x, nd.y, nd.z};
    double d;
    if (is_solid(P, c.z)) {
        d = D.solid_rho;
        mu[I] = (TS)D.solid_rho;
    } else {
        double phi = 0;
        for (int s = 0; s < D.nspheres; ++s) {
            const double dx = c.x - D.cx[s], dy = c.y - D.cy[s], dz = c.z - D.cz[s];
            const double r = sqrt(dx * dx + dy * dy + dz * dz);
            phi = fmax(phi, 0.5 * (1.0 - tanh(2.0 * (r - D.radius[s]) / D.width)));
        }
        if (D.invert) phi = 1.0 - phi;
        d = D.rho_g + (D.rho_l - D.rho_g) * phi;
        if (D.noise_amp != 0.0) d += D.noise_amp * noise_u(D.seed, (uint64_t)I);
        mu[I] = (TS)0;
    }
    rho[I] = (TS)d;
    store_feq(P, f, nd, d);
}

template <class TF, class TS>
__global__ void __launch_bounds__(TPB) k_feq_from_rho(const DevParams P, TF *__restrict__ f, const TS *__restrict__ rho) {
    const NodeIdx nd = node_from_grid(P);
    if (!nd.ok) return;
    store_feq(P, f, nd, (double)rho[nd.I]);
}

// ---------------------------------------------------------------------------
// conversions between the storage forms and FP64 host values.  Host values are
// post-stream populations (the reference's Dist after the swap, FluidCuda.cu:1290);
// storage is the pull convention, so fluid nodes un-stream / gather through
// pull_index and solid nodes keep their own slots.
template <class TF, class TS>
__global__ void __launch_bounds__(TPB) k_pop_from_double(const DevParams P, int q, const double *__restrict__ src,
                                                         TF *__restrict__ dst, const TS *__restrict__ rref) {
    const NodeIdx nd = node_from_grid(P);
    if (!nd.ok) return;
    const unsigned Nu = (unsigned)P.N;
    const Nbr nb = make_nbr(P, nd.x, nd.y, nd.z);
    const bool solid = is_solid(P, nd.z);
    const unsigned slot = solid ? (unsigned)q * Nu + nb.site : pull_index(nb, Nu, q);
    dst[slot] = pop_val<TF, double>(src[nd.I], q, solid ? 0.0 : (double)ref_at<TF>(rref, nd.I));
}
template <class TF, class TS>
__global__ void __launch_bounds__(TPB) k_pop_to_double(const DevParams P, int q, const TF *__restrict__ src,
                                                       double *__restrict__ dst, const TS *__restrict__ rref) {
    const NodeIdx nd = node_from_grid(P);
    if (!nd.ok) return;
    const unsigned Nu = (unsigned)P.N;
    const Nbr nb = make_nbr(P, nd.x, nd.y, nd.z);
    const bool solid = is_solid(P, nd.z);
    const unsigned slot = solid ? (unsigned)q * Nu + nb.site : pull_index(nb, Nu, q);
    dst[nd.I] = pop_load<TF, double>(src, slot, q, solid ? 0.0 : (double)ref_at<TF>(rref, nd.I));
}
// ---------------------------------------------------------------------------
// Observables on the device: the XZ plane y = tier (the contact-angle scan reads
// only that plane), and per-column sums for the 3-D cap fit.
template <class TS>
__global__ void __launch_bounds__(TPB) k_plane_y(const DevParams P, const TS *__restrict__ rho, int tier,
                                                 double *__restrict__ out) {
    const long long e = (long long)blockIdx.x * TPB + threadIdx.x;
    if (e >= (long long)P.nx * P.nz) return;
    const long long i = e / P.nz, k = e - i * P.nz;
    out[e] = (double)rho[(i * P.ny + tier) * P.nz + k];
}

// per column (x, y): 1 when the first fluid layer is at or above the interface
// density, and the highest interface crossing above the wall plane z = lo - 1/2;
// block partials (sum, max) in part[2 b], part[2 b + 1]
template <class TS>
__global__ void __launch_bounds__(TPB) k_cap_columns(const DevParams P, const TS *__restrict__ rho, double mid,
                                                     double *__restrict__ part) {
    const long long col = (long long)blockIdx.x * TPB + threadIdx.x;
    double area = 0, h = 0;
    if (col < (long long)P.nx * P.ny) {
        const TS *c = rho + col * P.nz;
        const double zw = P.lo - 0.5;
        if ((double)c[P.lo] >= mid) area = 1;
        for (int k = P.lo; k < P.nz - P.hi - 1; ++k) {
            const double a = (double)c[k], b = (double)c[k + 1];
            if (a >= mid && b < mid) {
                const double z = k + (a - mid) / (a - b);
                if (z - zw > h) h = z - zw;
            }
        }
    }
    const double sa = block_sum(area);
    __shared__ double smax[TPB / 64];
    double m = h;
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) smax[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double mm = 0;
        for (int w = 0; w < TPB / 64; ++w) mm = fmax(mm, smax[w]);
        part[2 * blockIdx.x] = sa;
        part[2 * blockIdx.x + 1] = mm;
    }
}

// Liquid fraction clamp((rho - rho_g) / (rho_l - rho_g), 0, 1) summed over the fluid
// nodes of plane x (one contiguous ny*nz block of rho); block partials in part[b]
template <class TS>
__global__ void __launch_bounds__(TPB) k_plane_fraction(const DevParams P, const TS *__restrict__ rho, int x, double rg,
                                                        double rl, double *__restrict__ part) {
    const long long j = (long long)blockIdx.x * TPB + threadIdx.x;
    double v = 0;
    if (j < (long long)P.ny * P.nz && !is_solid(P, (int)(j % P.nz))) {
        const double phi = ((double)rho[(long long)x * P.ny * P.nz + j] - rg) / (rl - rg);
        v = phi < 0 ? 0.0 : (phi > 1 ? 1.0 : phi);
    }
    const double s = block_sum(v);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// raw copy of every population slot between the device layout (pop_site) and the
// reference node order
template <class TF>
__global__ void __launch_bounds__(TPB) k_pop_relayout(const DevParams P, const TF *__restrict__ src, TF *__restrict__ dst,
                                                      int to_site) {
    const NodeIdx nd = node_from_grid(P);
    if (!nd.ok) return;
    const unsigned Nu = (unsigned)P.N, site = pop_site(P, nd.x, nd.y, nd.z), I = (unsigned)nd.I;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        if (to_site) dst[(unsigned)q * Nu + site] = src[(unsigned)q * Nu + I];
        else dst[(unsigned)q * Nu + I] = src[(unsigned)q * Nu + site];
    }
}
template <class A, class B>
__global__ void __launch_bounds__(TPB) k_convert(long long n, const A *__restrict__ src, B *__restrict__ dst) {
    const long long I = (long long)blockIdx.x * TPB + threadIdx.x;
    if (I < n) dst[I] = (B)src[I];
}
__global__ void __launch_bounds__(TPB) k_type(const DevParams P, double *__restrict__ out) {
    const NodeIdx nd = node_from_grid(P);
    if (nd.ok) out[nd.I] = is_solid(P, nd.z) ? 2.0 : 1.0;
}

// ---------------------------------------------------------------------------
// Reductions (CalcMacroCPU, FluidCuda.cu:1626-1657) in FP64: fixed grid, fixed
// per-thread order, wavefront shuffles + LDS, then one block over the partials.
constexpr int RED_BLOCKS = 512;
constexpr int RED_VALS = 12;   // mass, bad, first_bad, rmin, rmax, maxspeed2, mom x/y/z, kin, count, -

// kinds: 0 sum, 1 min, 2 max
__device__ __forceinline__ void block_reduce(double (&v)[RED_VALS], const int (&kind)[RED_VALS], double *out) {
    __shared__ double sh[TPB / 64][RED_VALS];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < RED_VALS; ++k) {
        double x = kind[k] == 0 ? wave_sum(v[k]) : kind[k] == 1 ? wave_min(v[k]) : wave_max(v[k]);
        if (lane == 0) sh[wv][k] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < RED_VALS; ++k) {
            double x = sh[0][k];
            for (int w = 1; w < TPB / 64; ++w)
                x = kind[k] == 0 ? x + sh[w][k] : kind[k] == 1 ? fmin(x, sh[w][k]) : fmax(x, sh[w][k]);
            out[k] = x;
        }
    }
}

template <class TS>
__global__ void __launch_bounds__(TPB) k_reduce(const DevParams P, const TS *__restrict__ rho,
                                                const double *__restrict__ mv, int has_mv, double *__restrict__ part) {
    const int kind[RED_VALS] = {0, 0, 1, 1, 2, 2, 0, 0, 0, 0, 0, 0};
    double v[RED_VALS] = {0, 0, 1e300, 1e300, -1e300, 0, 0, 0, 0, 0, 0, 0};
    const long long stride = (long long)gridDim.x * TPB;
    for (long long I = (long long)blockIdx.x * TPB + threadIdx.x; I < P.N; I += stride) {
        if (is_solid(P, (int)((unsigned long long)I % (unsigned)P.nz))) continue;
        const double d = (double)rho[I];
        v[0] += d;
        if (!(d >= 0)) { v[1] += 1; v[2] = fmin(v[2], (double)I); }
        v[3] = fmin(v[3], d);
        v[4] = fmax(v[4], d);
        v[10] += 1;
        if (has_mv) {
            const double a = mv[I], b = mv[P.N + I], e = mv[2 * P.N + I];
            const double m2 = a * a + b * b + e * e;
            v[5] = fmax(v[5], m2);
            v[6] += d * a; v[7] += d * b; v[8] += d * e;
            v[9] += 0.5 * d * m2;
        }
    }
    block_reduce(v, kind, part + (long long)blockIdx.x * RED_VALS);
}
__global__ void __launch_bounds__(TPB) k_reduce_final(const double *__restrict__ part, int nb, double *__restrict__ out) {
    const int kind[RED_VALS] = {0, 0, 1, 1, 2, 2, 0, 0, 0, 0, 0, 0};
    double v[RED_VALS] = {0, 0, 1e300, 1e300, -1e300, 0, 0, 0, 0, 0, 0, 0};
    for (int b = threadIdx.x; b < nb; b += TPB) {
#pragma unroll
        for (int k = 0; k < RED_VALS; ++k) {
            const double x = part[(long long)b * RED_VALS + k];
            v[k] = kind[k] == 0 ? v[k] + x : kind[k] == 1 ? fmin(v[k], x) : fmax(v[k], x);
        }
    }
    block_reduce(v, kind, out);
}


}  // namespace

// ===========================================================================
// launchers
hipError_t launch_wall(cpmp_ctx *c) {
    const DevParams &P = c->dp;
    if (P.lo == 0 && P.hi == 0) return hipSuccess;
    const int passes = std::max(P.lo, P.hi);
    return dispatch(c->form, [&](auto fm) {
        using TS = typename decltype(fm)::TS;
        const dim3 g(blocks_for(c->NXY));
        const dim3 gt((P.nx + kWallTile - 1) / kWallTile, (P.ny + kWallTile - 1) / kWallTile, 2), bt(kWallTile * kWallTile);
        TS *rho = (TS *)c->rho, *mu = (TS *)c->mu;
        if (P.wall_scheme == 2)   // ChemBoundaryComplex: one launch per level, as the reference
            for (int p = 0; p < passes; ++p)
                hipLaunchKernelGGL(k_wall<TS>, g, dim3(TPB), 0, c->stream, P, rho, mu, c->mu_w, p);
        else if (passes == 1) hipLaunchKernelGGL((k_wall_tile<1, TS>), gt, bt, 0, c->stream, P, rho, mu, c->mu_w);
        else if (passes == 2) hipLaunchKernelGGL((k_wall_tile<2, TS>), gt, bt, 0, c->stream, P, rho, mu, c->mu_w);
        else if (passes == 3) hipLaunchKernelGGL((k_wall_tile<3, TS>), gt, bt, 0, c->stream, P, rho, mu, c->mu_w);
        else   // deeper walls: one launch per layer
            for (int p = 0; p < passes; ++p)
                hipLaunchKernelGGL(k_wall<TS>, g, dim3(TPB), 0, c->stream, P, rho, mu, c->mu_w, p);
        return hipGetLastError();
    });
}

// CPMP_COLLIDE=v1 / CPMP_MU=v1 select the one-thread-per-node kernels (stencils
// from cache) instead of the LDS marches
static int variant(const char *var) {
    const char *e = std::getenv(var);
    return (e && std::strcmp(e, "v1") == 0) ? 1 : 2;
}
static int mu_variant() {
    static const int v = variant("CPMP_MU");
    return v;
}

// FP32 storage forms with the march collide sum the density inside the collide
// kernel when the tiles cover the lattice exactly (a context created with
// CPMP_FUSE=0 in the environment keeps the separate density pass)
bool fuse_geometry(int nx, int ny, int nz, int xc, size_t *ring_elems, size_t *chunk_elems) {
    if (xc < 2 || xc > MXC || nx % xc || ny % MTY || nz % MTZ) return false;
    const size_t tiles_yz = (size_t)(ny / MTY) * (nz / MTZ);
    *ring_elems = (size_t)nx * tiles_yz * ARING;
    *chunk_elems = (size_t)(nx / xc) * tiles_yz * 2 * AP;
    return true;
}

// Planes per workgroup of the march kernels: the longest chunk (32, 16, ..., 2)
// that divides nx and still gives 1024 workgroups.  A longer chunk loads a
// smaller share of halo planes and chunk-end partial planes; below about 1024
// workgroups the launches no longer spread evenly over the CUs.  Measured per
// step against the longest chunk giving 2048 workgroups (FP32 / FP16 / FP64):
// 128^3 -12 / -8 / -8 %, 192^3 -3 / -8 / -4 %, 32x256x256 -12 / -7 / -12 %,
// 256x256x32 -15 / -11 / -13 %, 128x128x96 with walls (FP64) -13 %; 256^3 and
// 512x512x256 keep 32.  FP32 storage forms on lattices with fewer than 2048
// workgroups of 2 planes take single planes, which leaves the fused density path
// (it needs chunks of 2 or more) for the separate density pass: 6 % faster per
// step than chunks of 2 at 64^3, 16 % at 96x96x64; 128x128x96 (3072 such
// workgroups) takes chunks of 4, 8 % faster than single planes in FP32 and FP16
// (FP64 lattices below 2048 workgroups of 2 planes take the per-node collide
// instead, node_collide).  CPMP_XC overrides.
int march_chunk(int nx, int ny, int nz, int form) {
    if (const char *e = std::getenv("CPMP_XC")) {
        const int v = std::atoi(e);
        if (v >= 1 && v <= MXC) return v;
    }
    const long long tiles = (long long)((ny + MTY - 1) / MTY) * ((nz + MTZ - 1) / MTZ);
    if (!fp64_storage(form) && (long long)((nx + 1) / 2) * tiles < 2048) return 1;   // small lattice
    for (int c = MXC; c >= 2; c /= 2)
        if (nx % c == 0 && (nx / c) * tiles >= 1024) return c;
    for (int c = 2; c <= MXC; c *= 2)
        if (nx % c == 0) return c;
    for (int c = MXC; c >= 2; c /= 2)
        if ((nx + c - 1) / c * tiles >= 1024) return c;
    return 2;
}
bool density_fused(const cpmp_ctx *c) { return c->rbuf && !c->collide_v1; }

// FP64 lattices with fewer than 2048 march workgroups of 2 planes: the per-node
// collide (2 waves per SIMD, no LDS ring) measured 8 % faster per step at 64^3 and
// 96x96x64, the march faster from 128x128x96 on
bool node_collide(int nx, int ny, int nz, int form) {
    if (const char *e = std::getenv("CPMP_COLLIDE")) {
        if (std::strcmp(e, "v1") == 0) return true;
        if (std::strcmp(e, "march") == 0) return false;
    }
    const long long tiles = (long long)((ny + MTY - 1) / MTY) * ((nz + MTZ - 1) / MTZ);
    return fp64_storage(form) && (long long)((nx + 1) / 2) * tiles < 2048;
}

// previous-density buffer the FP16 form shifts its populations by (null otherwise)
template <class TS> static const TS *ref_buffer(const cpmp_ctx *c) {
    return c->form == CPMP_FORM_FP16 ? (const TS *)c->rho_next : nullptr;
}

// after a mu march with interface runs: the entries its waves posted as overflow, then
// the other descriptor counter for the next launch
static hipError_t launch_band_overflow(cpmp_ctx *c) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int par = c->band_par;
    // a grid of 2048 workgroups (8 per CU) strides over the descriptors: ~2100 at 256^3
    hipLaunchKernelGGL(k_band_overflow, dim3(2048), dim3(kBandDesc), 0, c->stream, c->dp, (const uint4 *)c->band,
                       (const uint2 *)c->band_ovf, (const unsigned *)(c->band_cnt + par), c->band_cnt + (par ^ 1),
                       (float *)c->mu, c->band_cap);
    c->band_par = par ^ 1;
    return hipGetLastError();
}

hipError_t launch_mu(cpmp_ctx *c, void *rho_copy) {
    return dispatch(c->form, [&](auto fm) {
        using TS = typename decltype(fm)::TS;
        TS *solid = (TS *)c->rho_next;   // double-buffered density: solid values follow the step
        if (mu_variant() == 1) {
            if (rho_copy) {
                const hipError_t e = hipMemcpyAsync(rho_copy, c->rho, (size_t)c->N * sizeof(TS), hipMemcpyDeviceToDevice,
                                                    c->stream);
                if (e != hipSuccess) return e;
            }
            hipLaunchKernelGGL(k_mu<TS>, node_grid(c->dp), dim3(TPB), 0, c->stream, c->dp, (const TS *)c->rho,
                               (TS *)c->mu, solid);
        } else {
            const int xc = c->dp.xc;   // shorter chunks measured slower (0.077 vs 0.071 ms at 256^3)
            const int ntz = (c->p.nz + MTZ - 1) / MTZ, nty = (c->p.ny + MTY - 1) / MTY, nch = (c->p.nx + xc - 1) / xc;
            auto go = [&](auto h_t) {
                hipLaunchKernelGGL((k_mu_march<TS, false, decltype(h_t)::value>), dim3((unsigned)(ntz * nty * nch)),
                                   dim3(MT), 0, c->stream, c->dp, (const TS *)c->rho, (TS *)c->mu, ntz, nty, solid, xc,
                                   (const TS *)nullptr, (const TS *)nullptr, (TS *)rho_copy, nch, c->band, c->band_ovf,
                                   c->band_cnt ? c->band_cnt + c->band_par : nullptr, c->band_cap);
            };
            if (c->dp.lap_points == 7) go(std::integral_constant<int, 3>{});
            else go(std::integral_constant<int, 2>{});
            if (c->band) return launch_band_overflow(c);
        }
        return hipGetLastError();
    });
}

hipError_t launch_collide(cpmp_ctx *c, bool want_mv, void *uout_v, bool keep_sums) {
    const bool rs = density_fused(c);
    const hipError_t e = dispatch(c->form, [&](auto fm) {
        using TF = typename decltype(fm)::TF;
        using TS = typename decltype(fm)::TS;
        using TC = typename decltype(fm)::TC;
        const TF *fin = (const TF *)c->f[c->cur];
        TF *fout = (TF *)c->f[c->cur ^ 1];
        const TS *rho = (const TS *)c->rho, *mu = (const TS *)c->mu;
        TS *rn = (TS *)c->rho_next, *rb = (TS *)c->rbuf, *cb = (TS *)c->cbuf;
        TS *uout = (TS *)uout_v;
        // Tiles go to the XCDs by tile rows: XCD k takes the rows tyi = k (mod 8)
        // with their z tiles consecutively, so the tiles sharing the partial lines
        // of the z-shifted pulls share one L2 (measured 3 % faster than plain
        // round-robin at 256^3, which CPMP_REMAP=0 selects; CPMP_REMAP=1 gives each
        // XCD a contiguous range of tiles).  Lattices whose tile-row count is not a
        // multiple of 8 keep round-robin.
        static const int remap_env = [] {
            const char *e = std::getenv("CPMP_REMAP");
            return e ? std::atoi(e) : 2;
        }();
        const int remap = remap_env == 2 ? ((((c->p.ny + MTY - 1) / MTY) % 8 == 0) ? 2 : 0) : (remap_env == 1 ? 1 : 0);
        const bool guo = c->dp.forcing == 1;
        if (c->collide_v1) {
            auto v1 = [&](auto mv_t, auto guo_t) {
                hipLaunchKernelGGL((k_collide<TF, TS, TC, decltype(mv_t)::value, decltype(guo_t)::value>), node_grid(c->dp),
                                   dim3(TPB), 0, c->stream, c->dp, fin, fout, rho, mu, c->mv, uout, ref_buffer<TS>(c));
            };
            if (guo) { if (want_mv) v1(std::true_type{}, std::true_type{}); else v1(std::false_type{}, std::true_type{}); }
            else { if (want_mv) v1(std::true_type{}, std::false_type{}); else v1(std::false_type{}, std::false_type{}); }
        } else {
            const int ntz = (c->p.nz + MTZ - 1) / MTZ, nty = (c->p.ny + MTY - 1) / MTY, nch = (c->p.nx + c->dp.xc - 1) / c->dp.xc;
            const dim3 grid((unsigned)(ntz * nty * nch));
            constexpr bool can_rs = true;
            // MV, RS and the unit-rate specialisation as template arguments
            // NW: no solid layers (lo = hi = 0), the bounce-back selects compile away
            const bool nw = c->dp.lo == 0 && c->dp.hi == 0;
            // QPR: the Qp term may be on (always in the MV launches, which are rare, and in
            // FP64, whose collide measured 4 % slower without the Qp path: 234 VGPRs give it
            // 2 waves per SIMD instead of 1, and the second wave costs more than it hides)
            const bool qp_on = c->dp.qp_on != 0;
            auto go = [&](auto mv_t, auto rs_t, auto ur_t) {
                auto launch = [&](auto nw_t, auto qp_t, auto guo_t) {
                    hipLaunchKernelGGL((k_collide_march<TF, TS, TC, decltype(mv_t)::value, decltype(rs_t)::value,
                                                        decltype(ur_t)::value, decltype(nw_t)::value,
                                                        decltype(qp_t)::value, decltype(guo_t)::value>),
                                       grid, dim3(MT), 0, c->stream, c->dp, fin, fout, rho, mu, c->mv, uout, ntz, nty, rn,
                                       rb, cb, remap, ref_buffer<TS>(c));
                };
                // Guo forcing: one specialisation per (MV, RS, UR, NW), with the Qp path
                auto pick = [&](auto nw_t) {
                    if (guo) launch(nw_t, std::true_type{}, std::true_type{});
                    else if constexpr (decltype(mv_t)::value || std::is_same<TF, double>::value)
                        launch(nw_t, std::true_type{}, std::false_type{});
                    else if (qp_on) launch(nw_t, std::true_type{}, std::false_type{});
                    else launch(nw_t, std::false_type{}, std::false_type{});
                };
                if (nw) pick(std::true_type{});
                else pick(std::false_type{});
            };
            using T1 = std::true_type;
            using F0 = std::false_type;
            using RSB = std::integral_constant<bool, can_rs>;
            const bool ur = c->dp.unit_rates != 0;
            if (can_rs && rs) {
                if (want_mv) { if (ur) go(T1{}, RSB{}, T1{}); else go(T1{}, RSB{}, F0{}); }
                else { if (ur) go(F0{}, RSB{}, T1{}); else go(F0{}, RSB{}, F0{}); }
            } else {
                if (want_mv) { if (ur) go(T1{}, F0{}, T1{}); else go(T1{}, F0{}, F0{}); }
                else { if (ur) go(F0{}, F0{}, T1{}); else go(F0{}, F0{}, F0{}); }
            }
        }
        return hipGetLastError();
    });
    c->cur ^= 1;
    // the summed density is now current (keep_sums: it stays in rho_next for the
    // next mu kernel to complete)
    if (rs && !keep_sums) std::swap(c->rho, c->rho_next);
    return e;
}

hipError_t launch_macro(cpmp_ctx *c) {
    if (density_fused(c)) {   // summed by the collide kernel; add the tile and chunk borders
        return dispatch(c->form, [&](auto fm) {
            using TS = typename decltype(fm)::TS;
            const int ntz = c->p.nz / MTZ, nty = c->p.ny / MTY, nch = c->p.nx / c->dp.xc;
            hipLaunchKernelGGL(k_density_fix<TS>, dim3((unsigned)(ntz * nty * nch)), dim3(MT), 0, c->stream, c->dp,
                               (TS *)c->rho, (const TS *)c->rbuf, (const TS *)c->cbuf, ntz, nty, nch);
            return hipGetLastError();
        });
    }
    // separate density pass: in place, or (FP16, double-buffered) into the other
    // buffer, which then becomes current while this one keeps the reference density
    const hipError_t e = dispatch(c->form, [&](auto fm) {
        using TF = typename decltype(fm)::TF;
        using TS = typename decltype(fm)::TS;
        using TC = typename decltype(fm)::TC;
        TS *out = c->rho_next ? (TS *)c->rho_next : (TS *)c->rho;
        hipLaunchKernelGGL((k_macro_into<TF, TS, TC>), node_grid(c->dp), dim3(TPB), 0, c->stream, c->dp,
                           (const TF *)c->f[c->cur], (const TS *)c->rho, out);
        return hipGetLastError();
    });
    if (c->rho_next) std::swap(c->rho, c->rho_next);
    return e;
}

hipError_t launch_step(cpmp_ctx *c, bool want_mv) {
    hipError_t e;
    if ((e = launch_wall(c)) != hipSuccess) return e;
    if ((e = launch_mu(c)) != hipSuccess) return e;
    if ((e = launch_collide(c, want_mv)) != hipSuccess) return e;
    return launch_macro(c);
}

// Periodic FP32 lattices with the density summed by the collide kernel: within a
// run of steps the completion of each step's density (k_density_fix) is folded
// into the next step's mu kernel (k_mu_march<float, true>); the last step of the
// run completes its density as usual, so the state between runs is the same.
// CPMP_DEFER=0 turns this off.
static bool fix_deferrable(const cpmp_ctx *c) {
    const char *env = std::getenv("CPMP_DEFER");
    const bool off = env && std::strcmp(env, "0") == 0;
    return !off && density_fused(c) && c->form == CPMP_FORM_FP32 && c->dp.lo == 0 && c->dp.hi == 0 &&
           mu_variant() != 1 && c->dp.xc >= 2;
}

// the previous collide's tile sums (rho_next) completed into rho, and mu from them
static hipError_t launch_mu_fix(cpmp_ctx *c) {
    const int xc = c->dp.xc, ntz = c->p.nz / MTZ, nty = c->p.ny / MTY, nch = c->p.nx / xc;
    auto go = [&](auto h_t) {
        hipLaunchKernelGGL((k_mu_march<float, true, decltype(h_t)::value>), dim3((unsigned)(ntz * nty * nch)), dim3(MT), 0,
                           c->stream, c->dp, (const float *)c->rho_next, (float *)c->mu, ntz, nty, (float *)nullptr, xc,
                           (const float *)c->rbuf, (const float *)c->cbuf, (float *)c->rho, nch, c->band, c->band_ovf,
                           c->band_cnt ? c->band_cnt + c->band_par : nullptr, c->band_cap);
    };
    if (c->dp.lap_points == 7) go(std::integral_constant<int, 3>{});
    else go(std::integral_constant<int, 2>{});
    if (c->band) return launch_band_overflow(c);
    return hipGetLastError();
}

hipError_t launch_steps(cpmp_ctx *c, int nsteps) {
    const bool defer = fix_deferrable(c);
    hipError_t e;
    for (int n = 0; n < nsteps; ++n) {
        const bool last = n == nsteps - 1;
        if (defer && n > 0) {
            if ((e = launch_mu_fix(c)) != hipSuccess) return e;
        } else {
            if ((e = launch_wall(c)) != hipSuccess) return e;
            if ((e = launch_mu(c)) != hipSuccess) return e;
        }
        if (defer && !last) {
            if ((e = launch_collide(c, false, nullptr, true)) != hipSuccess) return e;
        } else {
            if ((e = launch_collide(c, last)) != hipSuccess) return e;
            if ((e = launch_macro(c)) != hipSuccess) return e;
        }
        ++c->step;
    }
    return hipSuccess;
}

hipError_t launch_init_reference(cpmp_ctx *c, double radius, double width, double rg, double rl, double solid) {
    return dispatch(c->form, [&](auto fm) {
        using TF = typename decltype(fm)::TF;
        using TS = typename decltype(fm)::TS;
        hipLaunchKernelGGL((k_init_reference<TF, TS>), node_grid(c->dp), dim3(TPB), 0, c->stream, c->dp,
                           (TF *)c->f[c->cur], (TS *)c->rho, (TS *)c->mu, radius, width, rg, rl, solid);
        return hipGetLastError();
    });
}

hipError_t launch_init_splash(cpmp_ctx *c, double radius, double width, double film, double rg, double rl) {
    return dispatch(c->form, [&](auto fm) {
        using TF = typename decltype(fm)::TF;
        using TS = typename decltype(fm)::TS;
        hipLaunchKernelGGL((k_init_splash<TF, TS>), node_grid(c->dp), dim3(TPB), 0, c->stream, c->dp,
                           (TF *)c->f[c->cur], (TS *)c->rho, (TS *)c->mu, radius, width, film, rg, rl);
        return hipGetLastError();
    });
}

hipError_t launch_init_spheres(cpmp_ctx *c, const cpmp_init_desc *d) {
    return dispatch(c->form, [&](auto fm) {
        using TF = typename decltype(fm)::TF;
        using TS = typename decltype(fm)::TS;
        hipLaunchKernelGGL((k_init_spheres<TF, TS>), node_grid(c->dp), dim3(TPB), 0, c->stream, c->dp,
                           (TF *)c->f[c->cur], (TS *)c->rho, (TS *)c->mu, *d);
        return hipGetLastError();
    });
}

hipError_t launch_feq_from_rho(cpmp_ctx *c) {
    return dispatch(c->form, [&](auto fm) {
        using TF = typename decltype(fm)::TF;
        using TS = typename decltype(fm)::TS;
        hipLaunchKernelGGL((k_feq_from_rho<TF, TS>), node_grid(c->dp), dim3(TPB), 0, c->stream, c->dp,
                           (TF *)c->f[c->cur], (const TS *)c->rho);
        return hipGetLastError();
    });
}

hipError_t launch_pop_from_double(cpmp_ctx *c, int q, const double *src, void *dst) {
    return dispatch(c->form, [&](auto fm) {
        using TF = typename decltype(fm)::TF;
        using TS = typename decltype(fm)::TS;
        hipLaunchKernelGGL((k_pop_from_double<TF, TS>), node_grid(c->dp), dim3(TPB), 0, c->stream, c->dp, q, src,
                           (TF *)dst, ref_buffer<TS>(c));
        return hipGetLastError();
    });
}

hipError_t launch_pop_to_double(cpmp_ctx *c, int q, const void *src, double *dst) {
    return dispatch(c->form, [&](auto fm) {
        using TF = typename decltype(fm)::TF;
        using TS = typename decltype(fm)::TS;
        hipLaunchKernelGGL((k_pop_to_double<TF, TS>), node_grid(c->dp), dim3(TPB), 0, c->stream, c->dp, q,
                           (const TF *)src, dst, ref_buffer<TS>(c));
        return hipGetLastError();
    });
}

hipError_t launch_pop_relayout(cpmp_ctx *c, const void *src, void *dst, bool to_device_layout) {
    return dispatch(c->form, [&](auto fm) {
        using TF = typename decltype(fm)::TF;
        hipLaunchKernelGGL(k_pop_relayout<TF>, node_grid(c->dp), dim3(TPB), 0, c->stream, c->dp, (const TF *)src,
                           (TF *)dst, to_device_layout ? 1 : 0);
        return hipGetLastError();
    });
}

hipError_t launch_scalar_from_double(cpmp_ctx *c, const double *src, void *dst) {
    return dispatch(c->form, [&](auto fm) {
        using TS = typename decltype(fm)::TS;
        hipLaunchKernelGGL((k_convert<double, TS>), dim3(blocks_for(c->N)), dim3(TPB), 0, c->stream, (long long)c->N, src,
                           (TS *)dst);
        return hipGetLastError();
    });
}

hipError_t launch_scalar_to_double(cpmp_ctx *c, const void *src, double *dst) {
    return dispatch(c->form, [&](auto fm) {
        using TS = typename decltype(fm)::TS;
        hipLaunchKernelGGL((k_convert<TS, double>), dim3(blocks_for(c->N)), dim3(TPB), 0, c->stream, (long long)c->N,
                           (const TS *)src, dst);
        return hipGetLastError();
    });
}

hipError_t launch_plane_y(cpmp_ctx *c, int tier, double *out) {
    return dispatch(c->form, [&](auto fm) {
        using TS = typename decltype(fm)::TS;
        hipLaunchKernelGGL(k_plane_y<TS>, dim3(blocks_for((long long)c->p.nx * c->p.nz)), dim3(TPB), 0, c->stream,
                           c->dp, (const TS *)c->rho, tier, out);
        return hipGetLastError();
    });
}

hipError_t launch_cap_columns(cpmp_ctx *c, double mid, double *part, unsigned *nblocks) {
    *nblocks = blocks_for(c->NXY);
    return dispatch(c->form, [&](auto fm) {
        using TS = typename decltype(fm)::TS;
        hipLaunchKernelGGL(k_cap_columns<TS>, dim3(*nblocks), dim3(TPB), 0, c->stream, c->dp, (const TS *)c->rho, mid,
                           part);
        return hipGetLastError();
    });
}

hipError_t launch_plane_fraction(cpmp_ctx *c, int x, double rg, double rl, double *part, unsigned *nblocks) {
    *nblocks = blocks_for((long long)c->p.ny * c->p.nz);
    return dispatch(c->form, [&](auto fm) {
        using TS = typename decltype(fm)::TS;
        hipLaunchKernelGGL(k_plane_fraction<TS>, dim3(*nblocks), dim3(TPB), 0, c->stream, c->dp, (const TS *)c->rho, x, rg,
                           rl, part);
        return hipGetLastError();
    });
}

hipError_t launch_velocity(cpmp_ctx *c, double *out3) {
    return dispatch(c->form, [&](auto fm) {
        using TF = typename decltype(fm)::TF;
        using TS = typename decltype(fm)::TS;
        using TC = typename decltype(fm)::TC;
        hipLaunchKernelGGL((k_velocity<TF, TS, TC>), node_grid(c->dp), dim3(TPB), 0, c->stream, c->dp,
                           (const TF *)c->f[c->cur], out3, ref_buffer<TS>(c));
        return hipGetLastError();
    });
}

hipError_t launch_force(cpmp_ctx *c, double *out3) {
    return dispatch(c->form, [&](auto fm) {
        using TS = typename decltype(fm)::TS;
        using TC = typename decltype(fm)::TC;
        hipLaunchKernelGGL((k_force<TS, TC>), node_grid(c->dp), dim3(TPB), 0, c->stream, c->dp, (const TS *)c->rho,
                           (const TS *)c->mu, out3);
        return hipGetLastError();
    });
}

hipError_t launch_type(cpmp_ctx *c, double *out) {
    hipLaunchKernelGGL(k_type, node_grid(c->dp), dim3(TPB), 0, c->stream, c->dp, out);
    return hipGetLastError();
}

hipError_t launch_reduce_into(cpmp_ctx *c, double *part, double *host_out) {
    hipError_t e = dispatch(c->form, [&](auto fm) {
        using TS = typename decltype(fm)::TS;
        hipLaunchKernelGGL(k_reduce<TS>, dim3(RED_BLOCKS), dim3(TPB), 0, c->stream, c->dp, (const TS *)c->rho, c->mv,
                           c->mv_valid ? 1 : 0, part);
        return hipGetLastError();
    });
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(TPB), 0, c->stream, part, RED_BLOCKS, part + RED_BLOCKS * RED_VALS);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(host_out, part + RED_BLOCKS * RED_VALS, RED_VALS * sizeof(double), hipMemcpyDeviceToHost,
                            c->stream)) != hipSuccess)
        return e;
    return hipStreamSynchronize(c->stream);
}

hipError_t launch_reduce(cpmp_ctx *c, double *host_out) { return launch_reduce_into(c, c->red, host_out); }

size_t reduce_scratch_doubles() { return (size_t)(RED_BLOCKS + 1) * RED_VALS; }

// host evaluation of the collision (same code as the device path, FP64)
void host_collide_fp64(const double f[19], double rho, const double F[3], const double ts[19], const double *qp,
                       double out[19], double u[3], int forcing) {
    if (forcing == 1) collide<double, false, true>(f, rho, F[0], F[1], F[2], ts, qp, out, u);
    else collide(f, rho, F[0], F[1], F[2], ts, qp, out, u);
}

}  // namespace cpmp
