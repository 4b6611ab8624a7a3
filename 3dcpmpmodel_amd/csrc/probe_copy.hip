// probe_copy.hip -- reference kernels of known HBM traffic on a 256^3 lattice of
// 19 FP32 population arrays, used to calibrate rocprofv3's FETCH_SIZE /
// WRITE_SIZE counters and to measure the attainable bandwidth of the access
// patterns of the step:
//   copy  : out[q][I] = in[q][I]                 (aligned, 2*19*4 B per node)
//   pull  : out[q][I] = in[q][I - c_q (wrapped)] (shifted reads, aligned writes)
//   push  : out[q][I + c_q (wrapped)] = in[q][I] (aligned reads, shifted writes)
//   sum19 : out[I] = sum_q in[q][I]              (read-only, 20*4 B per node)
//   pullx : pull on a 1-D grid whose block ids are remapped so that each XCD
//           takes one contiguous run of z-fastest blocks (shared partial lines
//           of the shifted reads stay in one L2)
//   sumpx : sum over the 19 pulled values (the density pass), remapped grid
//   flat  : copy on a 1-D grid, one population per loop pass (contiguous 1 KB per
//           workgroup and population); flatnt: the same with non-temporal stores
//   marchr: the access pattern of the march collide without its arithmetic: 8 x 32
//           (y, z) tiles marching 32 x planes, pull reads, non-temporal stores,
//           reference node order; marcht: the same on tile-contiguous sites
// Prints one line per pattern: ms per launch and GB/s of the known traffic.
//   probe_copy [reps]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

namespace {

constexpr int NX = 256, NY = 256, NZ = 256;
constexpr unsigned N = (unsigned)NX * NY * NZ;

__device__ __forceinline__ int cxq(int q) { return (q == 1 || q == 7 || q == 9 || q == 11 || q == 13) ? 1 : (q == 2 || q == 8 || q == 10 || q == 12 || q == 14) ? -1 : 0; }
__device__ __forceinline__ int cyq(int q) { return (q == 3 || q == 7 || q == 8 || q == 15 || q == 17) ? 1 : (q == 4 || q == 9 || q == 10 || q == 16 || q == 18) ? -1 : 0; }
__device__ __forceinline__ int czq(int q) { return (q == 5 || q == 11 || q == 12 || q == 15 || q == 16) ? 1 : (q == 6 || q == 13 || q == 14 || q == 17 || q == 18) ? -1 : 0; }

__device__ __forceinline__ unsigned shifted(int x, int y, int z, int q, int s) {
    const int xx = (x + s * cxq(q) + NX) % NX, yy = (y + s * cyq(q) + NY) % NY, zz = (z + s * czq(q) + NZ) % NZ;
    return ((unsigned)xx * NY + yy) * NZ + zz;
}

// MODE 0 copy, 1 push, 2 pull
template <int MODE>
__global__ void __launch_bounds__(256) k_pattern(const float *__restrict__ in, float *__restrict__ out) {
    const int z = blockIdx.x * 64 + (threadIdx.x & 63), y = blockIdx.y * 4 + (threadIdx.x >> 6), x = blockIdx.z;
    const unsigned I = ((unsigned)x * NY + y) * NZ + z;
    float v[19];
#pragma unroll
    for (int q = 0; q < 19; ++q) v[q] = MODE == 2 ? in[(unsigned)q * N + shifted(x, y, z, q, -1)] : in[(unsigned)q * N + I];
#pragma unroll
    for (int q = 0; q < 19; ++q) {
        if (MODE == 1) out[(unsigned)q * N + shifted(x, y, z, q, 1)] = v[q];
        else out[(unsigned)q * N + I] = v[q];
    }
}

// MODE 0 pull copy, 1 pulled sum; grid of NZ/64 * NY/4 * NX blocks
template <int MODE>
__global__ void __launch_bounds__(256) k_pull_xcd(const float *__restrict__ in, float *__restrict__ out) {
    unsigned t = blockIdx.x;
    t = (t & 7) * (gridDim.x >> 3) + (t >> 3);
    const int z = (t % (NZ / 64)) * 64 + (threadIdx.x & 63);
    t /= NZ / 64;
    const int y = (t % (NY / 4)) * 4 + (threadIdx.x >> 6), x = t / (NY / 4);
    const unsigned I = ((unsigned)x * NY + y) * NZ + z;
    float v[19];
#pragma unroll
    for (int q = 0; q < 19; ++q) v[q] = in[(unsigned)q * N + shifted(x, y, z, q, -1)];
    if (MODE == 1) {
        float s = 0;
#pragma unroll
        for (int q = 0; q < 19; ++q) s += v[q];
        out[I] = s;
    } else {
#pragma unroll
        for (int q = 0; q < 19; ++q) out[(unsigned)q * N + I] = v[q];
    }
}

template <bool NT>
__global__ void __launch_bounds__(256) k_flat(const float *__restrict__ in, float *__restrict__ out) {
    const unsigned i = blockIdx.x * 256u + threadIdx.x;
#pragma unroll
    for (int q = 0; q < 19; ++q) {
        const float v = in[(unsigned)q * N + i];
        if (NT) __builtin_nontemporal_store(v, &out[(unsigned)q * N + i]);
        else out[(unsigned)q * N + i] = v;
    }
}

// node site: reference order, or tile-contiguous 8 x 32 tiles within each x plane
template <bool TILED>
__device__ __forceinline__ unsigned site(int x, int y, int z) {
    if (!TILED) return ((unsigned)x * NY + y) * NZ + z;
    return (unsigned)x * NY * NZ + (unsigned)(((y / 8) * (NZ / 32) + z / 32) * 256 + (y % 8) * 32 + z % 32);
}
template <bool TILED>
__global__ void __launch_bounds__(256) k_march(const float *__restrict__ in, float *__restrict__ out) {
    constexpr int ntz = NZ / 32, nty = NY / 8, XC = 32;
    int t = blockIdx.x;
    const int tzi = t % ntz;
    t /= ntz;
    const int tyi = t % nty, ch = t / nty;
    const int y = tyi * 8 + threadIdx.x / 32, z = tzi * 32 + threadIdx.x % 32;
    const int ym = (y + NY - 1) % NY, yp = (y + 1) % NY, zm = (z + NZ - 1) % NZ, zp = (z + 1) % NZ;
    for (int x = ch * XC; x < ch * XC + XC; ++x) {
        const int xm = (x + NX - 1) % NX, xp = (x + 1) % NX;
        float v[19];
#pragma unroll
        for (int q = 0; q < 19; ++q) {
            const int sx = cxq(q) > 0 ? xm : cxq(q) < 0 ? xp : x;
            const int sy = cyq(q) > 0 ? ym : cyq(q) < 0 ? yp : y;
            const int sz = czq(q) > 0 ? zm : czq(q) < 0 ? zp : z;
            v[q] = in[(unsigned)q * N + site<TILED>(sx, sy, sz)];
        }
        const unsigned I = site<TILED>(x, y, z);
#pragma unroll
        for (int q = 0; q < 19; ++q) __builtin_nontemporal_store(v[q], &out[(unsigned)q * N + I]);
    }
}

__global__ void __launch_bounds__(256) k_sum19(const float *__restrict__ in, float *__restrict__ out) {
    const int z = blockIdx.x * 64 + (threadIdx.x & 63), y = blockIdx.y * 4 + (threadIdx.x >> 6), x = blockIdx.z;
    const unsigned I = ((unsigned)x * NY + y) * NZ + z;
    float s = 0;
#pragma unroll
    for (int q = 0; q < 19; ++q) s += in[(unsigned)q * N + I];
    out[I] = s;
}

}  // namespace

int main(int argc, char **argv) {
    const int reps = argc > 1 ? std::atoi(argv[1]) : 20;
    float *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, 19ull * N * 4) != hipSuccess || hipMalloc(&b, 19ull * N * 4) != hipSuccess) {
        std::printf("allocation failed\n");
        return 1;
    }
    hipMemset(a, 0, 19ull * N * 4);
    hipMemset(b, 0, 19ull * N * 4);
    const dim3 g(NZ / 64, NY / 4, NX), t(256);
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    const char *names[10] = {"copy", "push", "pull", "sum19", "pullx", "sumpx", "flat", "flatnt", "marchr", "marcht"};
    const dim3 g1((NZ / 64) * (NY / 4) * NX), gf(N / 256), gm((NX / 32) * (NY / 8) * (NZ / 32));
    for (int mode = 0; mode < 10; ++mode) {
        float ms = 0;
        for (int pass = 0; pass < 2; ++pass) {   // first pass warms up
            hipEventRecord(e0);
            for (int i = 0; i < reps; ++i) {
                if (mode == 0) hipLaunchKernelGGL(k_pattern<0>, g, t, 0, 0, a, b);
                if (mode == 1) hipLaunchKernelGGL(k_pattern<1>, g, t, 0, 0, a, b);
                if (mode == 2) hipLaunchKernelGGL(k_pattern<2>, g, t, 0, 0, a, b);
                if (mode == 3) hipLaunchKernelGGL(k_sum19, g, t, 0, 0, a, b);
                if (mode == 4) hipLaunchKernelGGL(k_pull_xcd<0>, g1, t, 0, 0, a, b);
                if (mode == 5) hipLaunchKernelGGL(k_pull_xcd<1>, g1, t, 0, 0, a, b);
                if (mode == 6) hipLaunchKernelGGL(k_flat<false>, gf, t, 0, 0, a, b);
                if (mode == 7) hipLaunchKernelGGL(k_flat<true>, gf, t, 0, 0, a, b);
                if (mode == 8) hipLaunchKernelGGL(k_march<false>, gm, t, 0, 0, a, b);
                if (mode == 9) hipLaunchKernelGGL(k_march<true>, gm, t, 0, 0, a, b);
            }
            hipEventRecord(e1);
            hipEventSynchronize(e1);
            hipEventElapsedTime(&ms, e0, e1);
        }
        const double bytes = (mode == 3 || mode == 5) ? 20.0 * 4 * N : 2.0 * 19 * 4 * N;
        std::printf("%-6s %9.4f ms/launch %8.1f GB/s  known bytes/launch %.0f\n", names[mode], ms / reps,
                    bytes / (ms / reps * 1e-3) / 1e9, bytes);
    }
    const hipError_t e = hipGetLastError();
    std::printf("status %s\n", hipGetErrorString(e));
    hipFree(a);
    hipFree(b);
    return e == hipSuccess ? 0 : 1;
}
