// internal.h -- context and launcher declarations shared by the library's
// translation units.  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>

#include "../../include/cpmp.h"

namespace cpmp {

// Parameters as the kernels see them (one kernel argument; replaces the
// __constant__ symbols of FluidCuda.cu:67 and Common.h:96-153).
struct DevParams {
    int nx, ny, nz;
    int lo, hi;                 // solid layers at the z ends
    long long N;
    int eos, mu_form, lap_points, wall_scheme;
    int qp_on;                  // K1 or K2 non-zero
    double A, B, T, K2s, Ka;    // K2s = K*K
    double K1, K2;
    double ts[19];
    float tsf[19];              // ts in FP32 for the FP32 arithmetic forms
    double rho0;                // FP16 shift density
    int tiled, nty, ntz;        // tile-contiguous population sites (pop_site), tile counts in y, z
    int xc;                     // x planes per workgroup of the march kernels
    int unit_rates;             // the rates of rows 0, 1, 3, 5, 7, 10, 12, 16-18 are all 1 (reference set)
    int forcing;                // 0 exact-difference force term, 1 Guo
    // FP32 storage forms: fluid nodes with a density strictly inside (band_lo, band_hi)
    // -- the interface: between the Maxwell gas and liquid densities, less 2 % of their
    // gap at each end -- get the EOS chemical potential in FP64; empty (lo > hi) without
    // a two-phase region
    float band_lo, band_hi;
};

enum Kern { KW = CPMP_K_WALL, KM = CPMP_K_MU, KC = CPMP_K_COLLIDE, KR = CPMP_K_MACRO };

struct Timers {
    bool on = false;
    hipEvent_t ev[2 * CPMP_K_COUNT * 8] = {};
    double ms[CPMP_K_COUNT] = {};
    long long count[CPMP_K_COUNT] = {};
};

}  // namespace cpmp

struct cpmp_ctx {
    cpmp_params p;
    cpmp::DevParams dp;
    int64_t N = 0, NXY = 0;
    int form = 0;
    size_t pop_bytes = 0;       // bytes of one population value
    size_t scal_bytes = 0;      // bytes of one rho/mu value
    hipStream_t stream = nullptr;
    void *f[2] = {nullptr, nullptr};
    int cur = 0;
    void *rho = nullptr, *mu = nullptr;
    // FP32 storage forms on exactly tiled lattices: density summed by the collide
    // kernel into rho_next, tile rings and chunk partial planes for the fix-up pass
    void *rho_next = nullptr, *rbuf = nullptr, *cbuf = nullptr;
    // FP32 storage forms: per wave of the chemical-potential march, a run of 64 * xc
    // slots (node, density, Laplacian term) for the interface nodes it finishes in FP64
    uint4 *band = nullptr;
    // the entries of a run beyond the first kBandWave: descriptors (first slot, count of
    // at most kBandDesc) that k_band_overflow finishes after the march, and two counters
    // of posted descriptors used by alternate launches (band_par), each cleared by the
    // overflow kernel of the launch before its use
    uint2 *band_ovf = nullptr;
    unsigned *band_cnt = nullptr;
    unsigned band_cap = 0;      // descriptor slots
    int band_par = 0;
    bool collide_v1 = false;    // one-thread-per-node collide (node_collide)
    size_t fuse_bytes = 0;
    double *mv = nullptr;       // 3*N, MV of the last step
    double *mu_w = nullptr;     // 2*NXY: bottom then top
    double *red = nullptr;      // reduction partials
    double *staging = nullptr;  // N doubles
    int64_t step = 0;
    bool mv_valid = false;
    cpmp::Timers *timers = nullptr;
};

namespace cpmp {

int set_error(int code, const char *fmt, ...);

// launchers (kernels.hip); all enqueue on ctx->stream
hipError_t launch_wall(cpmp_ctx *c);
// rho_copy (optional): receives a copy of rho as the kernel reads it (the adjoint's tape)
hipError_t launch_mu(cpmp_ctx *c, void *rho_copy = nullptr);
bool density_fused(const cpmp_ctx *c);
bool fuse_geometry(int nx, int ny, int nz, int xc, size_t *ring_elems, size_t *chunk_elems);
// forms whose populations, rho and mu are stored in FP64 (FP64, FP32A)
inline bool fp64_storage(int form) { return form == CPMP_FORM_FP64 || form == CPMP_FORM_FP32A; }
constexpr unsigned kBandWave = 128;   // interface nodes a wave of the mu march finishes itself
constexpr unsigned kBandDesc = 256;   // interface nodes per overflow descriptor (one workgroup)
// x planes per march workgroup for a lattice and storage form (CPMP_XC overrides)
int march_chunk(int nx, int ny, int nz, int form);
// the one-thread-per-node collide instead of the x march (CPMP_COLLIDE=v1 / march
// override): FP64 lattices too small for 2048 march workgroups of 2 planes
bool node_collide(int nx, int ny, int nz, int form);
constexpr int kTileY = 8, kTileZ = 32;   // in-plane tile of the march kernels and of tiled population storage
constexpr int kWallTile = 16;             // columns per edge of the wall-fill tiles (16 x 16 workgroups)
// copy one population set between the device layout and the reference node order
hipError_t launch_pop_relayout(cpmp_ctx *c, const void *src, void *dst, bool to_device_layout);
// uout (when given): u, grad mu, grad rho of every node, 9 N values of the
// storage scalar type (adjoint tape)
// keep_sums (fused density): leave the tile sums in rho_next for k_mu_march<…, FIX>
hipError_t launch_collide(cpmp_ctx *c, bool want_mv, void *uout = nullptr, bool keep_sums = false);
hipError_t launch_macro(cpmp_ctx *c);
hipError_t launch_step(cpmp_ctx *c, bool want_mv);
// nsteps whole steps (MV on the last), advancing c->step
hipError_t launch_steps(cpmp_ctx *c, int nsteps);
hipError_t launch_init_reference(cpmp_ctx *c, double radius, double width, double rg, double rl, double solid);
hipError_t launch_init_splash(cpmp_ctx *c, double radius, double width, double film, double rg, double rl);
hipError_t launch_init_spheres(cpmp_ctx *c, const cpmp_init_desc *d);
hipError_t launch_feq_from_rho(cpmp_ctx *c);
hipError_t launch_pop_from_double(cpmp_ctx *c, int q, const double *src, void *dst_pop);
hipError_t launch_pop_to_double(cpmp_ctx *c, int q, const void *src_pop, double *dst);
hipError_t launch_scalar_from_double(cpmp_ctx *c, const double *src, void *dst);
hipError_t launch_scalar_to_double(cpmp_ctx *c, const void *src, double *dst);
hipError_t launch_velocity(cpmp_ctx *c, double *out3);
// XZ plane y = tier of rho into out (nx*nz doubles, out[i*nz + k])
hipError_t launch_plane_y(cpmp_ctx *c, int tier, double *out);
// per-block (area, max height) partials of the cap fit into part (2 per block)
hipError_t launch_cap_columns(cpmp_ctx *c, double mid, double *part, unsigned *nblocks);
// liquid fraction of the fluid nodes of plane x, block partials (nblocks values)
hipError_t launch_plane_fraction(cpmp_ctx *c, int x, double rg, double rl, double *part, unsigned *nblocks);
hipError_t launch_force(cpmp_ctx *c, double *out3);
hipError_t launch_reduce(cpmp_ctx *c, double *host_out /* 12 */);
// the same reduction with its partials (reduce_scratch_doubles() values) in `part`
hipError_t launch_reduce_into(cpmp_ctx *c, double *part, double *host_out /* 12 */);
// cpmp_stats from the 12 reduction values; CPMP_ENAN (naming the first bad node) when
// a density is NaN or negative
int stats_from_reduction(cpmp_ctx *c, const double *v, cpmp_stats *out);
hipError_t launch_type(cpmp_ctx *c, double *out);
size_t reduce_scratch_doubles();

// host-side CPU evaluation of the per-node collision (self-test hook)
void host_collide_fp64(const double f[19], double rho, const double F[3], const double ts[19],
                       const double *qp, double out[19], double u[3], int forcing);

}  // namespace cpmp
