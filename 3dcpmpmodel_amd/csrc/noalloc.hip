// noalloc.hip -- the reductions of cpmp_reduce with their partial sums in a
// file-scope device array instead of context memory, for callers that cannot
// allocate device memory around the call.
//
// The array (`scratch`) holds (RED_BLOCKS + 1) * RED_VALS doubles: the per-block
// partials of the fixed 512-block reduction grid and the final values.  Every
// call on one GPU uses that GPU's copy and overwrites all of it (each block writes
// its own partials, the final kernel writes the totals), so nothing has to start
// at zero; calls that overlap on one GPU (two host threads, or two streams in
// flight) race on it.
#include "internal.h"

#include <mutex>

namespace {

constexpr size_t kScratchDoubles = 513 * 12;   // (reduction blocks + 1) x values per block
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kScratchDoubles];

// device address of `scratch` on each device, looked up once per device with that
// device current (the host-side name is only a stand-in for the symbol)
double *g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;

int scratch_for_device(int dev, double **out) {
    if (dev < 0 || dev >= kMaxDevices) return cpmp::set_error(CPMP_EINVAL, "device %d outside 0..%d", dev, kMaxDevices - 1);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return cpmp::set_error(CPMP_EHIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        g_scratch_ptr[dev] = (double *)p;
    }
    *out = g_scratch_ptr[dev];
    return CPMP_OK;
}

}  // namespace

extern "C" int cpmp_reduce_noalloc(cpmp_ctx *c, cpmp_stats *out) {
    if (!c || !c->stream || !out) return cpmp::set_error(CPMP_EINVAL, "bad context or output");
    const size_t need = cpmp::reduce_scratch_doubles();
    if (need > kScratchDoubles)
        return cpmp::set_error(CPMP_EINVAL, "the reduction needs %zu doubles of scratch, the static array holds %zu", need,
                               kScratchDoubles);
    hipError_t e = hipSetDevice(c->p.device);
    if (e != hipSuccess) return cpmp::set_error(CPMP_EHIP, "hipSetDevice: %s", hipGetErrorString(e));
    double *part = nullptr;
    if (const int rc = scratch_for_device(c->p.device, &part)) return rc;
    double v[12];
    if ((e = cpmp::launch_reduce_into(c, part, v)) != hipSuccess)
        return cpmp::set_error(CPMP_EHIP, "reduction: %s", hipGetErrorString(e));
    return cpmp::stats_from_reduction(c, v, out);
}
