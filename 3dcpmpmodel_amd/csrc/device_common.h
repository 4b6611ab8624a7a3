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

}  // namespace cpmp
