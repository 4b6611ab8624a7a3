// plt.cpp -- Tecplot binary (#!TDV112) writer for the density field.
//
// File layout produced by the reference's SaveTask (FluidCuda.cu:901-927) through
// ordered_tec (ordered_tec.cpp:674-717 header, :990-1013 zone header,
// :1015-1147 zone data): one ordered zone, variables i, j, k as int16 and
// Density as float64, I index = z (fastest), J = y, K = x.
#include "internal.h"

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

using namespace cpmp;

namespace {

struct Out {
    FILE *fp;
    bool ok = true;
    void raw(const void *p, size_t n) { ok = ok && std::fwrite(p, 1, n, fp) == n; }
    void i32(int32_t v) { raw(&v, 4); }
    void f32(float v) { raw(&v, 4); }
    void f64(double v) { raw(&v, 8); }
    void str(const std::string &s) {   // one int32 per character, then 0
        for (unsigned char ch : s) i32((int32_t)(signed char)ch);
        i32(0);
    }
};

}  // namespace

extern "C" int cpmp_write_plt(cpmp_ctx *c, const char *path, const char *title, int decimate, int decimate_linear) {
    if (!c || !path) return set_error(CPMP_EINVAL, "bad context or path");
    if (decimate < 1) return set_error(CPMP_EINVAL, "decimate %d", decimate);
    const int nx = c->p.nx, ny = c->p.ny, nz = c->p.nz;
    const int lx = nx / decimate, ly = ny / decimate, lz = nz / decimate;
    if (lx < 1 || ly < 1 || lz < 1) return set_error(CPMP_EINVAL, "decimated zone is empty");
    if (lx > 32767 || ly > 32767 || lz > 32767) return set_error(CPMP_EINVAL, "index exceeds int16");
    std::vector<double> rho(c->N);
    int rc = cpmp_get_field(c, CPMP_FIELD_RHO, rho.data());
    if (rc) return rc;
    const int64_t n = (int64_t)lx * ly * lz;
    std::vector<int16_t> vi(n), vj(n), vk(n);
    std::vector<double> d(n);
    for (int i = 0; i < lx; ++i)
        for (int j = 0; j < ly; ++j)
            for (int k = 0; k < lz; ++k) {
                const int64_t m = ((int64_t)i * ly + j) * lz + k;
                vi[m] = (int16_t)i; vj[m] = (int16_t)j; vk[m] = (int16_t)k;
                int64_t src;
                if (decimate == 1) src = m;
                else if (decimate_linear) src = (((int64_t)i * ny + j) * nz + k) * decimate;   // FluidCuda.cu:470
                else src = (((int64_t)i * decimate * ny + (int64_t)j * decimate) * nz + (int64_t)k * decimate);
                d[m] = rho[src];
            }
    std::string name(path);
    const size_t slash = name.find_last_of('/');
    std::string zone = slash == std::string::npos ? name : name.substr(slash + 1);
    if (zone.size() > 4 && zone.compare(zone.size() - 4, 4, ".plt") == 0) zone.resize(zone.size() - 4);

    FILE *fp = std::fopen(path, "wb");
    if (!fp) return set_error(CPMP_EIO, "cannot write %s", path);
    Out o{fp};
    // header
    o.raw("#!TDV112", 8);
    o.i32(1);                         // byte order
    o.i32(0);                         // file type FULL
    o.str(title ? title : "Model3D");
    o.i32(4);
    o.str("i"); o.str("j"); o.str("k"); o.str("Density");
    // zone header
    o.f32(299.0f);
    o.str(zone);
    o.i32(-1);                        // parent zone
    o.i32(-1);                        // strand id
    o.f64(0.0);                       // solution time
    o.i32(-1);                        // unused
    o.i32(0);                         // ordered zone
    o.i32(0);                         // data at nodes
    o.i32(0);                         // no face neighbours
    o.i32(0);                         // no misc connections
    o.i32(lz); o.i32(ly); o.i32(lx);  // IMax JMax KMax
    o.i32(0);                         // no auxiliary data
    o.f32(357.0f);                    // end of header
    // data section
    o.f32(299.0f);
    o.i32(4); o.i32(4); o.i32(4); o.i32(2);   // int16, int16, int16, float64
    o.i32(0); o.i32(0); o.i32(-1);
    auto mm16 = [&](const std::vector<int16_t> &v) {
        int16_t a = v[0], b = v[0];
        for (int16_t x : v) { a = x < a ? x : a; b = x > b ? x : b; }
        o.f64(a); o.f64(b);
    };
    mm16(vi); mm16(vj); mm16(vk);
    {
        double a = d[0], b = d[0];
        for (double x : d) { a = x < a ? x : a; b = x > b ? x : b; }
        o.f64(a); o.f64(b);
    }
    o.raw(vi.data(), n * 2); o.raw(vj.data(), n * 2); o.raw(vk.data(), n * 2);
    o.raw(d.data(), n * 8);
    const bool ok = o.ok;
    if (std::fclose(fp) != 0 || !ok) return set_error(CPMP_EIO, "short write to %s", path);
    return CPMP_OK;
}
