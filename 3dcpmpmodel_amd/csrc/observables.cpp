// observables.cpp -- droplet observables: the reference's contact-angle
// measurement (on the one XZ plane it scans, extracted on the device), a 3-D
// spherical-cap fit (column reduction on the device), single-node probes and the
// Laplace pressure jump.
#include "internal.h"
#include "lbm_math.h"

#include <algorithm>
#include <cmath>
#include <vector>

using namespace cpmp;

namespace {

#define OBS_HIP(expr)                                                                         \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess) return set_error(CPMP_EHIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

}  // namespace

extern "C" {

// ContactAngle(side=0, tier), FluidCuda.cu:557-689.  Interface x on rows z=2..5 by
// linear interpolation at (rhoG+rhoL)/2 (:595-623), linear extrapolation of rows
// z=3,4 to z=2 (:634-637), angles from the extrapolated foot (:641-649), and the
// spherical-cap angle from height H and base L (:656-680).  The reference rows fit
// its three bottom wall layers (z = 2 is the last solid layer); the rows here are
// offset by the bottom wall depth (begin = nwall_lo - 1), and a lattice without a
// bottom wall is rejected.  When no interface crosses the last row scanned for H,
// H is NaN (the reference leaves it unset).
int cpmp_contact_angle(cpmp_ctx *c, int tier, double rho_g, double rho_l, double *out) {
    if (!c || !out) return set_error(CPMP_EINVAL, "bad context or output");
    const int DX = c->p.nx, DY = c->p.ny, DZ = c->p.nz;
    if (tier < 0 || tier >= DY) return set_error(CPMP_EINVAL, "tier %d outside [0,%d)", tier, DY);
    if (DX < 8 || DZ < 8) return set_error(CPMP_EINVAL, "domain too small for the contact-angle scan");
    const int begin = c->p.nwall_lo - 1;   // last solid layer: the row the foot is extrapolated to
    if (begin < 0 || begin + 5 >= DZ) return set_error(CPMP_EINVAL, "the contact-angle scan needs a bottom wall (nwall_lo=%d)", c->p.nwall_lo);
    OBS_HIP(hipSetDevice(c->p.device));
    std::vector<double> plane((size_t)DX * DZ);
    OBS_HIP(launch_plane_y(c, tier, c->staging));
    OBS_HIP(hipMemcpyAsync(plane.data(), c->staging, plane.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    OBS_HIP(hipStreamSynchronize(c->stream));
    const double PI = 3.14159265358979323846264338327950288;
    const double Mid = (rho_g + rho_l) / 2;
    auto T = [&](int i, int k) { return plane[(size_t)i * DZ + k]; };
    double Px[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int j = begin; j <= begin + 3; ++j) {
        const int n = (j - begin) * 2;
        for (int i = 0; i < DX / 2 + 2; ++i) {
            const double t1 = T(i + 1, j), t2 = T(i + 2, j);
            if (t1 < Mid && t2 >= Mid) { Px[n] = double(i + 1) + (t1 - Mid) / (t1 - t2); break; }
        }
        for (int i = DX - 2; i > DX / 2 - 2; --i) {
            const double t1 = T(i - 1, j), t2 = T(i - 2, j);
            if (t1 < Mid && t2 >= Mid) { Px[n + 1] = double(i - 1) - (t1 - Mid) / (t1 - t2); break; }
        }
    }
    const double y0 = begin, y1 = begin + 1, y2 = begin + 2;
    double CaLeft = 0, CaRight = 0, CaCap = 0, XLeft = DY / 2, XRight = DY / 2, H = NAN, L = 0;
    int valid = 0;
    if (Px[0] != 0.0 && Px[1] != 0.0) {
        valid = 1;
        XLeft = Px[2] * ((y0 - y2) / (y1 - y2)) + Px[4] * ((y0 - y1) / (y2 - y1));
        XRight = Px[3] * ((y0 - y2) / (y1 - y2)) + Px[5] * ((y0 - y1) / (y2 - y1));
        double a = std::atan((y1 - y0) / (Px[2] - XLeft));
        if (a < 0) a += PI;
        CaLeft = a / PI * 180;
        a = std::atan(-(y1 - y0) / (Px[3] - XRight));
        if (a < 0) a += PI;
        CaRight = a / PI * 180;
        L = XRight - XLeft;
        const int ic = int((Px[4] + Px[5]) / 2 + 0.5);
        if (ic >= 0 && ic < DX)
            for (int j = begin; j < DZ - 2; ++j) {
                const double t = T(ic, j), t1 = T(ic, j + 1);
                if (t >= Mid && t1 < Mid) H = double(j) + (t - Mid) / (t - t1) - y0;
            }
        if (L <= 0) { L = Px[3] - Px[2]; H = H - 0.5; }
        if (XLeft > DY / 2 && XRight < DY / 2) { XLeft = DY / 2; XRight = DY / 2; }   // as FluidCuda.cu:672
        const double R = (H * H * 4 + L * L) / H / 8;
        CaCap = std::atan(L / (R - H) / 2);
        if (CaCap < 0) CaCap += PI;
        CaCap = CaCap / PI * 180;
    }
    out[0] = CaLeft; out[1] = CaRight; out[2] = CaCap; out[3] = XLeft; out[4] = XRight; out[5] = H; out[6] = L;
    out[7] = valid;
    return CPMP_OK;
}

// 3-D spherical cap on the bottom wall: base area = number of wall-adjacent columns
// whose first fluid layer is above the interface density, a = sqrt(area/pi);
// height H = highest interface crossing above the wall plane z = lo - 1/2
// (half-way bounce-back); R = (a^2+H^2)/(2H); angle = atan2(a, R-H).
int cpmp_cap_fit(cpmp_ctx *c, double rho_g, double rho_l, double *out) {
    if (!c || !out) return set_error(CPMP_EINVAL, "bad context or output");
    if (c->p.nwall_lo < 1) return set_error(CPMP_EINVAL, "no bottom wall");
    OBS_HIP(hipSetDevice(c->p.device));
    const double Mid = (rho_g + rho_l) / 2;
    unsigned nb = 0;
    OBS_HIP(launch_cap_columns(c, Mid, c->staging, &nb));
    std::vector<double> part(2 * (size_t)nb);
    OBS_HIP(hipMemcpyAsync(part.data(), c->staging, part.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    OBS_HIP(hipStreamSynchronize(c->stream));
    double area = 0, H = 0;
    for (unsigned b = 0; b < nb; ++b) {
        area += part[2 * b];
        H = std::max(H, part[2 * b + 1]);
    }
    const double a = std::sqrt(area / 3.14159265358979323846);
    double R = NAN, ang = NAN;
    if (H > 0) {
        R = (a * a + H * H) / (2 * H);
        ang = std::atan2(a, R - H) * 180.0 / 3.14159265358979323846;
    }
    out[0] = ang; out[1] = H; out[2] = a; out[3] = R;
    return CPMP_OK;
}

// Radius of the liquid cross-section of plane x (the neck of two coalescing
// droplets): r = sqrt(A / pi), A the plane's summed liquid fraction.
int cpmp_neck_radius(cpmp_ctx *c, int x, double rho_g, double rho_l, double *radius, double *area) {
    if (!c || !c->stream || !radius) return set_error(CPMP_EINVAL, "bad context or output");
    if (x < 0 || x >= c->p.nx) return set_error(CPMP_EINVAL, "plane x = %d outside the lattice", x);
    if (!(rho_l > rho_g)) return set_error(CPMP_EINVAL, "rho_l must exceed rho_g");
    OBS_HIP(hipSetDevice(c->p.device));
    unsigned nb = 0;
    OBS_HIP(launch_plane_fraction(c, x, rho_g, rho_l, c->staging, &nb));
    std::vector<double> part(nb);
    OBS_HIP(hipMemcpyAsync(part.data(), c->staging, part.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    OBS_HIP(hipStreamSynchronize(c->stream));
    double A = 0;
    for (unsigned b = 0; b < nb; ++b) A += part[b];
    *radius = std::sqrt(A / 3.14159265358979323846);
    if (area) *area = A;
    return CPMP_OK;
}

// One node's density or chemical potential (FP64), without copying the field.
int cpmp_probe(cpmp_ctx *c, int field, int x, int y, int z, double *value) {
    if (!c || !c->stream || !value) return set_error(CPMP_EINVAL, "bad context or output");
    if (field != CPMP_FIELD_RHO && field != CPMP_FIELD_MU) return set_error(CPMP_EINVAL, "field %d: rho or mu only", field);
    if (x < 0 || x >= c->p.nx || y < 0 || y >= c->p.ny || z < 0 || z >= c->p.nz)
        return set_error(CPMP_EINVAL, "node (%d,%d,%d) outside the lattice", x, y, z);
    OBS_HIP(hipSetDevice(c->p.device));
    const size_t I = ((size_t)x * c->p.ny + y) * c->p.nz + z;
    const char *src = (const char *)(field == CPMP_FIELD_RHO ? c->rho : c->mu) + I * c->scal_bytes;
    if (c->scal_bytes == sizeof(double)) {
        OBS_HIP(hipMemcpyAsync(value, src, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        OBS_HIP(hipStreamSynchronize(c->stream));
    } else {
        float v = 0;
        OBS_HIP(hipMemcpyAsync(&v, src, sizeof(float), hipMemcpyDeviceToHost, c->stream));
        OBS_HIP(hipStreamSynchronize(c->stream));
        *value = v;
    }
    return CPMP_OK;
}

// Laplace pressure jump K^2 [p(rho_in) - p(rho_out)] between two nodes, with the
// EOS pressure of EosPressure (FluidCuda.cu:1561-1573, all EOS of the build).
int cpmp_laplace_jump(cpmp_ctx *c, const int *in_xyz, const int *out_xyz, double *dp) {
    if (!c || !in_xyz || !out_xyz || !dp) return set_error(CPMP_EINVAL, "bad arguments");
    double ri = 0, ro = 0;
    int rc = cpmp_probe(c, CPMP_FIELD_RHO, in_xyz[0], in_xyz[1], in_xyz[2], &ri);
    if (!rc) rc = cpmp_probe(c, CPMP_FIELD_RHO, out_xyz[0], out_xyz[1], out_xyz[2], &ro);
    if (rc) return rc;
    const cpmp_params &p = c->p;
    *dp = p.K * p.K * (pressure_eos(p.eos, p.A, p.B, p.T, ri) - pressure_eos(p.eos, p.A, p.B, p.T, ro));
    return CPMP_OK;
}

}  // extern "C"
