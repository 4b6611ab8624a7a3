// device_common.h -- device helpers shared by the forward and adjoint kernels:
// node geometry, periodic wrap, the isotropic 6th-order gradient stencil and
// wavefront reductions.
#pragma once

#include "internal.h"
#include "lbm_math.h"

namespace cpmp {

constexpr int TPB = 256;

__device__ __forceinline__ bool is_solid(const DevParams &P, int z) { return z < P.lo || z >= P.nz - P.hi; }
__device__ __forceinline__ int wrapc(int a, int n) { a %= n; return a < 0 ? a + n : a; }
// a + d wrapped into [0, n) for |d| <= 3, without a division when n >= 3
__device__ __forceinline__ int wrap_off(int a, int d, int n) {
    int b = a + d;
    if (n >= 3) {
        b += (b < 0) ? n : 0;
        b -= (b >= n) ? n : 0;
        return b;
    }
    return wrapc(b, n);
}

struct Coord { int x, y, z; };
__device__ __forceinline__ Coord decompose(const DevParams &P, long long I) {
    Coord c;
    c.z = (int)(I % P.nz);
    const long long r = I / P.nz;
    c.y = (int)(r % P.ny);
    c.x = (int)(r / P.ny);
    return c;
}

// Per-node kernels launch on a 3-D grid (ceil(nz/64), ceil(ny/4), nx) of 64 x 4
// thread blocks, so coordinates come without integer division.
struct NodeIdx { int x, y, z; long long I; bool ok; };
// One block = 64 z x 4 y of one x plane; blocks are numbered z-block fastest,
// then y-block, then x.  Hardware hands consecutive block ids to the 8 XCDs in
// turn; when the grid divides by 8 the id is remapped so that each XCD works
// through one contiguous run of blocks, keeping z- and y-neighbour blocks (which
// read the same partial cache lines through the shifted population reads) in
// one L2.
__device__ __forceinline__ NodeIdx node_from_grid(const DevParams &P) {
    NodeIdx n;
    unsigned t = blockIdx.x;
    if ((gridDim.x & 7) == 0) t = (t & 7) * (gridDim.x >> 3) + (t >> 3);
    const unsigned nzb = ((unsigned)P.nz + 63) >> 6, nyb = ((unsigned)P.ny + 3) >> 2;
    const unsigned zb = t % nzb;
    t /= nzb;
    const unsigned yb = t % nyb;
    n.z = (int)(zb * 64 + (threadIdx.x & 63));
    n.y = (int)(yb * 4 + (threadIdx.x >> 6));
    n.x = (int)(t / nyb);
    n.ok = n.z < P.nz && n.y < P.ny;
    n.I = ((long long)n.x * P.ny + n.y) * P.nz + n.z;
    return n;
}
inline dim3 node_grid(const DevParams &P) {
    return dim3((unsigned)(((P.nz + 63) / 64) * ((P.ny + 3) / 4) * P.nx));
}

// Isotropic 6th-order gradient of Common.h:222-241 (Gradx7/y7/z7): per component
// two axis differences and four face-diagonal differences, each the 6-point
// central difference (-1, 9, -45, 0, 45, -9, 1)/60, weighted 2:1:1:1:1 and /360.
// xs/ys/zs[s+3] hold the wrapped coordinate + s.
template <class TS, class TC>
struct Stencil {
    const DevParams &P;
    const int *xs, *ys, *zs;
    __device__ __forceinline__ TC at(const TS *a, int dx, int dy, int dz) const {
        return (TC)a[((long long)xs[dx + 3] * P.ny + ys[dy + 3]) * P.nz + zs[dz + 3]];
    }
    __device__ __forceinline__ TC dd(const TS *a, int ux, int uy, int uz) const {
        return -at(a, -3 * ux, -3 * uy, -3 * uz) + TC(9) * at(a, -2 * ux, -2 * uy, -2 * uz)
               - TC(45) * at(a, -ux, -uy, -uz) + TC(45) * at(a, ux, uy, uz)
               - TC(9) * at(a, 2 * ux, 2 * uy, 2 * uz) + at(a, 3 * ux, 3 * uy, 3 * uz);
    }
    __device__ __forceinline__ TC gx(const TS *a) const {
        return (TC(2) * dd(a, 1, 0, 0) + dd(a, 1, 0, 1) + dd(a, 1, 0, -1) + dd(a, 1, 1, 0) + dd(a, 1, -1, 0)) * TC(1.0 / 360.0);
    }
    __device__ __forceinline__ TC gy(const TS *a) const {
        return (TC(2) * dd(a, 0, 1, 0) + dd(a, 0, 1, 1) + dd(a, 0, 1, -1) + dd(a, 1, 1, 0) + dd(a, -1, 1, 0)) * TC(1.0 / 360.0);
    }
    __device__ __forceinline__ TC gz(const TS *a) const {
        return (TC(2) * dd(a, 0, 0, 1) + dd(a, 1, 0, 1) + dd(a, -1, 0, 1) + dd(a, 0, 1, 1) + dd(a, 0, -1, 1)) * TC(1.0 / 360.0);
    }
    __device__ __forceinline__ void grad(const TS *a, TC &x, TC &y, TC &z) const { x = gx(a); y = gy(a); z = gz(a); }
};

__device__ __forceinline__ void wrap7(const DevParams &P, const Coord &c, int *xs, int *ys, int *zs) {
#pragma unroll
    for (int s = 0; s < 7; ++s) {
        xs[s] = wrap_off(c.x, s - 3, P.nx);
        ys[s] = wrap_off(c.y, s - 3, P.ny);
        zs[s] = wrap_off(c.z, s - 3, P.nz);
    }
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// deterministic sum over a TPB-thread block; thread 0 receives the result
__device__ __forceinline__ double block_sum(double v) {
    __shared__ double sh[TPB / 64];
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0;
    if (threadIdx.x == 0)
        for (int w = 0; w < TPB / 64; ++w) s += sh[w];
    return s;
}

inline unsigned blocks_for(long long n) { return (unsigned)((n + TPB - 1) / TPB); }

// Population storage site of node (x, y, z).  Reference order (x*ny + y)*nz + z,
// or, on lattices the 8 x 32 (y, z) tiles cover exactly in the FP32 storage
// forms, tile-contiguous within each x plane: the 256 nodes of a tile (y
// slowest, then z) are one contiguous run, tiles ordered y-tile slowest.  A
// workgroup of the march collide then streams 1 KB runs per population and
// plane.  rho, mu and all host-facing arrays keep the reference order.
constexpr int TILE_Y = kTileY, TILE_Z = kTileZ;
__device__ __forceinline__ unsigned pop_site(const DevParams &P, int x, int y, int z) {
    if (!P.tiled) return ((unsigned)x * P.ny + y) * P.nz + z;
    return (unsigned)x * (unsigned)(P.ny * P.nz) +
           (((unsigned)(y / TILE_Y) * P.ntz + (unsigned)(z / TILE_Z)) * (TILE_Y * TILE_Z) + (y % TILE_Y) * TILE_Z + z % TILE_Z);
}

// Storage site of one node, offsets of its neighbour sites (periodic wrap folded
// in; [1+dx] along x, [(1+dy)*3 + 1+dz] in the plane) and the solid flags of its
// z-1 / z+1 layers.  Indices are 32-bit: the host guarantees 19 N < 2^31.
struct Nbr {
    unsigned site;
    int ox[3], oyz[9];
    bool sb[3];
};
__device__ __forceinline__ Nbr make_nbr(const DevParams &P, int x, int y, int z) {
    Nbr n;
    const int pl = P.ny * P.nz;
    n.site = pop_site(P, x, y, z);
    n.ox[0] = (wrap_off(x, -1, P.nx) - x) * pl; n.ox[1] = 0; n.ox[2] = (wrap_off(x, 1, P.nx) - x) * pl;
    const int ys[3] = {wrap_off(y, -1, P.ny), y, wrap_off(y, 1, P.ny)};
    const int zs[3] = {wrap_off(z, -1, P.nz), z, wrap_off(z, 1, P.nz)};
    const int own = (int)pop_site(P, 0, y, z);
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) n.oyz[a * 3 + b] = (int)pop_site(P, 0, ys[a], zs[b]) - own;
    n.sb[0] = is_solid(P, zs[0]); n.sb[1] = false; n.sb[2] = is_solid(P, zs[2]);
    return n;
}
// Storage convention ("pull"): population q is stored post-collision at the node
// that emitted it.  The value arriving at fluid node I along q comes from
// I - c_q, or, when I - c_q is solid, is opp(q) bounced back at I (half-way
// bounce-back).
__device__ __forceinline__ unsigned pull_index(const Nbr &n, unsigned Nu, int q) {
    return n.sb[1 - cz(q)] ? (unsigned)opp(q) * Nu + n.site
                           : (unsigned)q * Nu + (unsigned)((int)n.site + n.ox[1 - cx(q)] + n.oyz[(1 - cy(q)) * 3 + 1 - cz(q)]);
}
// Over fluid nodes and all q, pull_index is one-to-one and never lands on a
// solid node's slot, so writing a post-stream value f_q(I) to pull_index(I, q)
// ("un-streaming") is conflict-free; solid nodes keep their own slots
// q*N + site.

}  // namespace cpmp
