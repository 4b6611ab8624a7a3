// api.hip -- the C ABI of include/cpmp.h: context lifetime, parameters, EOS
// helpers, initial conditions, the time loop, state transfer, reductions,
// checkpoints and measurement.
#include "internal.h"
#include "lbm_math.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

namespace cpmp {

static thread_local std::string g_err;

int set_error(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

}  // namespace cpmp

using namespace cpmp;

#define HIPCHK(expr)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess) return set_error(CPMP_EHIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

namespace {

bool valid_ctx(cpmp_ctx *c) { return c != nullptr && c->stream != nullptr; }

// FP16 form: after an initial state is written, its density is also the
// reference density the populations were shifted by
hipError_t sync_reference_density(cpmp_ctx *c) {
    if (c->form != CPMP_FORM_FP16) return hipSuccess;
    return hipMemcpyAsync(c->rho_next, c->rho, c->N * c->scal_bytes, hipMemcpyDeviceToDevice, c->stream);
}

int pop_size(int form) { return fp64_storage(form) ? 8 : (form == CPMP_FORM_FP32 ? 4 : 2); }
int scal_size(int form) { return fp64_storage(form) ? 8 : 4; }

// interface runs of the chemical-potential march (FP32 storage forms): 64 * xc
// slots per wave, MT / 64 waves per workgroup
size_t band_slots(const cpmp_ctx *c) {
    const size_t xc = (size_t)c->dp.xc;
    const size_t blocks = (size_t)((c->p.nz + kTileZ - 1) / kTileZ) * ((c->p.ny + kTileY - 1) / kTileY) * ((c->p.nx + xc - 1) / xc);
    return blocks * kTileY * kTileZ * xc;
}
// overflow descriptors: a run of n entries posts ceil((n - kBandWave) / kBandDesc) <=
// n / kBandDesc + 1, so slots / kBandDesc + (runs) bound their number
size_t band_desc_slots(const cpmp_ctx *c) {
    const size_t runs = band_slots(c) / (64 * (size_t)c->dp.xc);
    return band_slots(c) / kBandDesc + runs + 1;
}

void fill_devparams(cpmp_ctx *c) {
    const cpmp_params &p = c->p;
    DevParams &d = c->dp;
    d.nx = p.nx; d.ny = p.ny; d.nz = p.nz;
    d.lo = p.nwall_lo; d.hi = p.nwall_hi;
    d.N = (long long)p.nx * p.ny * p.nz;
    d.eos = p.eos; d.mu_form = p.mu_form; d.lap_points = p.lap_points; d.wall_scheme = p.wall_scheme;
    d.qp_on = (p.K1 != 0.0 || p.K2 != 0.0) ? 1 : 0;
    d.A = p.A; d.B = p.B; d.T = p.T; d.K2s = p.K * p.K; d.Ka = p.Ka; d.K1 = p.K1; d.K2 = p.K2;
    for (int q = 0; q < 19; ++q) { d.ts[q] = p.Ts[q]; d.tsf[q] = (float)p.Ts[q]; }
    d.unit_rates = 1;
    for (int q = 0; q < 19; ++q)
        if (cpmp::unit_rate_row(q) && p.Ts[q] != 1.0) d.unit_rates = 0;
    d.rho0 = p.rho0;
    d.forcing = p.forcing;
    d.xc = march_chunk(p.nx, p.ny, p.nz, p.form);
    // interface band of the FP64 EOS chemical potential (the last error is kept)
    d.band_lo = 1.0f;
    d.band_hi = 0.0f;
    {
        const std::string saved = g_err;
        double g = 0, l = 0;
        if (cpmp_coexistence(p.eos, p.A, p.B, p.T, &g, &l) == CPMP_OK) {
            d.band_lo = (float)(g + 0.02 * (l - g));
            d.band_hi = (float)(l - 0.02 * (l - g));
        }
        g_err = saved;
    }
}

int check_params(const cpmp_params *p) {
    if (!p) return set_error(CPMP_EINVAL, "params is NULL");
    if (p->nx < 1 || p->ny < 1 || p->nz < 1) return set_error(CPMP_EINVAL, "domain %dx%dx%d", p->nx, p->ny, p->nz);
    if (19LL * p->nx * p->ny * p->nz >= (1LL << 31))   // kernels index populations with 32-bit offsets
        return set_error(CPMP_EINVAL, "domain %dx%dx%d too large: 19*N must be below 2^31", p->nx, p->ny, p->nz);
    if (p->nx > 65535) return set_error(CPMP_EINVAL, "nx=%d exceeds the launch grid limit 65535", p->nx);
    if (p->nwall_lo < 0 || p->nwall_hi < 0 || p->nwall_lo + p->nwall_hi >= p->nz)
        return set_error(CPMP_EINVAL, "wall layers %d+%d do not fit nz=%d", p->nwall_lo, p->nwall_hi, p->nz);
    if (p->nwall_lo + p->nwall_hi > 0 && p->nz - p->nwall_lo - p->nwall_hi < 4)
        return set_error(CPMP_EINVAL, "fewer than 4 fluid layers between the walls");
    if (p->eos < 0 || p->eos > 5) return set_error(CPMP_EINVAL, "eos id %d", p->eos);
    if (p->lap_points != 5 && p->lap_points != 7) return set_error(CPMP_EINVAL, "lap_points must be 5 or 7");
    if (p->form < 0 || p->form > 3) return set_error(CPMP_EINVAL, "form %d", p->form);
    if (p->wall_scheme < 0 || p->wall_scheme > 2) return set_error(CPMP_EINVAL, "wall_scheme %d", p->wall_scheme);
    if (p->forcing != 0 && p->forcing != 1) return set_error(CPMP_EINVAL, "forcing %d", p->forcing);
    if (!(p->B > 0) || !(p->T > 0)) return set_error(CPMP_EINVAL, "EOS constants B=%g T=%g", p->B, p->T);
    for (int q = 0; q < 19; ++q)
        if (!(p->Ts[q] >= 0 && p->Ts[q] < 2)) return set_error(CPMP_EINVAL, "Ts[%d]=%g outside [0,2)", q, p->Ts[q]);
    return CPMP_OK;
}

}  // namespace

extern "C" {

const char *cpmp_last_error(void) { return g_err.c_str(); }
const char *cpmp_version(void) { return "cpmp-mi355x 0.1 (gfx950)"; }

// ---------------------------------------------------------------------------
int cpmp_set_multiphase(int eos, double Tr, double *A, double *B, double *Tc, double *Rc) {
    // SetMultiphase, FluidCuda.cu:1318-1344
    if (!A || !B || !Tc || !Rc) return set_error(CPMP_EINVAL, "NULL output");
    if (!(Tr > 0)) return set_error(CPMP_EINVAL, "Tr=%g", Tr);
    double a = 2.0 / 49, b = 2.0 / 21, tc, rc;
    switch (eos) {
    case CPMP_EOS_VDW: a = 9.0 / 49; b = 2.0 / 21; tc = 4.0 / 7; rc = 7.0 / 2; break;
    case CPMP_EOS_CSE: a = 1.0; b = 1.0; tc = 0.09432870; rc = 0.13044388; break;
    case CPMP_EOS_RKE: tc = 0.196133; rc = 2.729171; a *= 1.0 / std::sqrt(tc * Tr); break;
    case CPMP_EOS_RKS: {
        tc = 0.086861; rc = 2.729171;
        const double w = 0.344, m = 1.0 + (0.480 + 1.574 * w - 0.176 * w * w) * (1.0 - std::sqrt(Tr));
        a *= m * m;
    } break;
    case CPMP_EOS_PRW: case CPMP_EOS_PRM: {
        tc = 0.072919; rc = 2.65730416;
        const double w = eos == CPMP_EOS_PRW ? 0.344 : 0.011;
        const double m = 1.0 + (0.37464 + 1.54226 * w - 0.26992 * w * w) * (1.0 - std::sqrt(Tr));
        a *= m * m;
    } break;
    default: return set_error(CPMP_EINVAL, "eos id %d", eos);
    }
    *A = a; *B = b; *Tc = tc; *Rc = rc;
    return CPMP_OK;
}

// DimConversion (FluidCuda.cu:474-486): lattice-to-physical unit factors from a
// length, density and viscosity pair.  out: DimLen, DimTime, DimMass, DimF, DimT,
// DimV, DimA, DimW, DimB.
int cpmp_dim_conversion(double length, double macro_length, double density, double macro_density, double viscosity,
                        double macro_viscosity, double *out) {
    if (!out || !(length > 0) || !(density > 0) || !(macro_viscosity > 0))
        return set_error(CPMP_EINVAL, "lengths, densities and the physical viscosity must be > 0");
    const double len = macro_length / length;
    const double tim = viscosity / macro_viscosity * len * len;
    const double mass = macro_density / density * len * len * len;
    out[0] = len;
    out[1] = tim;
    out[2] = mass;
    out[3] = mass * len / (tim * tim);
    out[4] = mass * len * len / (tim * tim);
    out[5] = len / tim;
    out[6] = len / (tim * tim);
    out[7] = 1.0 / tim;
    out[8] = 1.0 / (tim * tim);
    return CPMP_OK;
}

int cpmp_default_params(cpmp_params *p, int nx, int ny, int nz, int eos, double Tr) {
    // Initialize(), FluidCuda.cu:119-163
    if (!p) return set_error(CPMP_EINVAL, "params is NULL");
    std::memset(p, 0, sizeof *p);
    p->nx = nx; p->ny = ny; p->nz = nz;
    p->nwall_lo = p->nwall_hi = 3;
    p->eos = eos;
    p->mu_form = 0;
    p->lap_points = 5;
    p->wall_scheme = 1;
    p->form = CPMP_FORM_FP64;
    double A, B, Tc, Rc;
    int rc = cpmp_set_multiphase(eos, Tr, &A, &B, &Tc, &Rc);
    if (rc) return rc;
    p->A = A; p->B = B; p->T = Tr * Tc;
    p->K = 0.2; p->Ka = 0.001; p->K1 = 0; p->K2 = 0;
    const double tau = 0.8;
    for (int q = 0; q < 19; ++q) p->Ts[q] = 1.0;
    p->Ts[2] = p->Ts[4] = p->Ts[6] = p->Ts[8] = 1.1;
    p->Ts[9] = p->Ts[11] = p->Ts[13] = p->Ts[14] = p->Ts[15] = 1.0 / tau;
    p->rho0 = 0.0;
    return CPMP_OK;
}

double cpmp_eos_pressure(int eos, double A, double B, double T, double rho) { return pressure_eos(eos, A, B, T, rho); }
double cpmp_eos_mu(int eos, int mu_form, double A, double B, double T, double rho) {
    return mu_eos(eos, mu_form, A, B, T, rho);
}

int cpmp_coexistence(int eos, double A, double B, double T, double *rho_g, double *rho_l) {
    // Maxwell construction: p(g) = p(l), mu(g) = mu(l); Newton with Gibbs-Duhem
    // derivatives dp/drho = rho dmu/drho.
    if (!rho_g || !rho_l) return set_error(CPMP_EINVAL, "NULL output");
    double g = *rho_g, l = *rho_l;
    if (!(g > 0 && l > g)) {
        // start from the spinodal bracket: scan mu'(rho) for its sign changes
        const double hi = 0.999 / B;
        double s1 = -1, s2 = -1, prev = dmu_eos(eos, 0, A, B, T, 1e-6 * hi);
        const int n = 20000;
        for (int i = 1; i < n; ++i) {
            const double r = hi * i / n, d = dmu_eos(eos, 0, A, B, T, r);
            if (prev > 0 && d <= 0 && s1 < 0) s1 = r;
            if (prev <= 0 && d > 0 && s1 > 0 && s2 < 0) s2 = r;
            prev = d;
        }
        if (s1 < 0 || s2 < 0) return set_error(CPMP_EINVAL, "no two-phase region at T=%g", T);
        g = s1 * 0.3; l = std::min(s2 * 1.5, 0.5 * (s2 + hi));
    }
    for (int it = 0; it < 500; ++it) {
        const double pg = pressure_eos(eos, A, B, T, g), pl = pressure_eos(eos, A, B, T, l);
        const double mg = mu_eos(eos, 0, A, B, T, g), ml = mu_eos(eos, 0, A, B, T, l);
        const double dg_ = dmu_eos(eos, 0, A, B, T, g), dl_ = dmu_eos(eos, 0, A, B, T, l);
        const double a11 = g * dg_, a12 = -l * dl_, a21 = dg_, a22 = -dl_;
        const double r1 = pg - pl, r2 = mg - ml, det = a11 * a22 - a12 * a21;
        if (!(std::fabs(det) > 0) || !std::isfinite(det)) return set_error(CPMP_EINVAL, "Maxwell Newton singular");
        const double dg = (r1 * a22 - a12 * r2) / det, dl = (a11 * r2 - a21 * r1) / det;
        double sc = 1.0;
        while ((g - sc * dg <= 0 || l - sc * dl <= g - sc * dg || B * (l - sc * dl) >= 1.0) && sc > 1e-8) sc *= 0.5;
        g -= sc * dg;
        l -= sc * dl;
        if (std::fabs(dg) <= 1e-15 * g && std::fabs(dl) <= 1e-15 * l) break;
    }
    if (!(g > 0 && l > g * (1 + 1e-9))) return set_error(CPMP_EINVAL, "Maxwell construction collapsed");
    *rho_g = g; *rho_l = l;
    return CPMP_OK;
}

int cpmp_cocurve_lookup(const char *path, double Tr, double *g, double *l) {
    // FluidCuda.cu:1356-1384 (rows "Tr rhoG/rhoc rhoL/rhoc"; match within 1e-12), but a
    // missing file or row is an error rather than zero densities.
    if (!path || !g || !l) return set_error(CPMP_EINVAL, "NULL argument");
    std::ifstream in(path);
    if (!in.is_open()) return set_error(CPMP_EIO, "cannot open %s", path);
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream iss(line);
        double t, a, b;
        if (!(iss >> t >> a >> b)) continue;
        if (std::fabs(t - Tr) < 1e-12) { *g = a; *l = b; return CPMP_OK; }
    }
    return set_error(CPMP_ENOTFOUND, "Tr=%.12g not in %s", Tr, path);
}

// ---------------------------------------------------------------------------
int cpmp_create(const cpmp_params *p, cpmp_ctx **out) {
    if (!out) return set_error(CPMP_EINVAL, "out is NULL");
    *out = nullptr;
    int rc = check_params(p);
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return set_error(CPMP_ENODEV, "no HIP device");
    if (p->device < 0 || p->device >= ndev) return set_error(CPMP_EINVAL, "device %d of %d", p->device, ndev);
    HIPCHK(hipSetDevice(p->device));
    cpmp_ctx *c = new cpmp_ctx();
    c->p = *p;
    c->form = p->form;
    c->N = (int64_t)p->nx * p->ny * p->nz;
    c->NXY = (int64_t)p->nx * p->ny;
    c->pop_bytes = pop_size(p->form);
    c->scal_bytes = scal_size(p->form);
    fill_devparams(c);
    c->collide_v1 = node_collide(p->nx, p->ny, p->nz, p->form);
    auto fail = [&](int code, const char *what) {
        cpmp_destroy(c);
        return set_error(code, "allocation of %s failed", what);
    };
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { delete c; return set_error(CPMP_EHIP, "stream"); }
    const size_t popset = (size_t)c->N * 19 * c->pop_bytes;
    if (hipMalloc(&c->f[0], popset) != hipSuccess) return fail(CPMP_ENOMEM, "populations A");
    if (hipMalloc(&c->f[1], popset) != hipSuccess) return fail(CPMP_ENOMEM, "populations B");
    if (hipMalloc(&c->rho, c->N * c->scal_bytes) != hipSuccess) return fail(CPMP_ENOMEM, "rho");
    if (hipMalloc(&c->mu, c->N * c->scal_bytes) != hipSuccess) return fail(CPMP_ENOMEM, "mu");
    const char *fuse = std::getenv("CPMP_FUSE");
    size_t ring_elems = 0, chunk_elems = 0;
    // every storage form sums the density inside the march collide where the tiles
    // cover the lattice (not with the one-thread-per-node collide of small FP64 lattices)
    if (!c->collide_v1 && !(fuse && std::strcmp(fuse, "0") == 0) &&
        fuse_geometry(p->nx, p->ny, p->nz, c->dp.xc, &ring_elems, &chunk_elems)) {
        if (hipMalloc(&c->rho_next, c->N * c->scal_bytes) != hipSuccess) return fail(CPMP_ENOMEM, "next rho");
        if (hipMalloc(&c->rbuf, ring_elems * c->scal_bytes) != hipSuccess) return fail(CPMP_ENOMEM, "tile rings");
        if (hipMalloc(&c->cbuf, chunk_elems * c->scal_bytes) != hipSuccess) return fail(CPMP_ENOMEM, "chunk planes");
        c->fuse_bytes = ((size_t)c->N + ring_elems + chunk_elems) * c->scal_bytes;
        // the fused path stores populations tile-contiguous (pop_site)
        c->dp.tiled = 1;
        c->dp.nty = p->ny / kTileY;
        c->dp.ntz = p->nz / kTileZ;
    }
    // the FP16 form double-buffers the density on every lattice: the previous
    // density is the reference its populations are shifted by
    if (c->form == CPMP_FORM_FP16 && !c->rho_next) {
        if (hipMalloc(&c->rho_next, c->N * c->scal_bytes) != hipSuccess) return fail(CPMP_ENOMEM, "previous rho");
        c->fuse_bytes = (size_t)c->N * c->scal_bytes;
    }
    if (!fp64_storage(c->form)) {
        if (hipMalloc((void **)&c->band, band_slots(c) * sizeof(uint4)) != hipSuccess)
            return fail(CPMP_ENOMEM, "interface runs");
        // descriptors, then the two counters
        if (hipMalloc((void **)&c->band_ovf, (band_desc_slots(c) + 1) * sizeof(uint2)) != hipSuccess)
            return fail(CPMP_ENOMEM, "interface overflow");
        c->band_cnt = (unsigned *)(c->band_ovf + band_desc_slots(c));
        c->band_cap = (unsigned)band_desc_slots(c);
    }
    if (hipMalloc((void **)&c->mv, 3 * c->N * sizeof(double)) != hipSuccess) return fail(CPMP_ENOMEM, "MV");
    if (hipMalloc((void **)&c->mu_w, 2 * c->NXY * sizeof(double)) != hipSuccess) return fail(CPMP_ENOMEM, "wall mu");
    if (hipMalloc((void **)&c->red, reduce_scratch_doubles() * sizeof(double)) != hipSuccess) return fail(CPMP_ENOMEM, "reduce");
    if (hipMalloc((void **)&c->staging, c->N * sizeof(double)) != hipSuccess) return fail(CPMP_ENOMEM, "staging");
    if (hipMemsetAsync(c->f[0], 0, popset, c->stream) != hipSuccess || hipMemsetAsync(c->f[1], 0, popset, c->stream) != hipSuccess ||
        hipMemsetAsync(c->rho, 0, c->N * c->scal_bytes, c->stream) != hipSuccess ||
        hipMemsetAsync(c->mu, 0, c->N * c->scal_bytes, c->stream) != hipSuccess ||
        (c->rho_next && hipMemsetAsync(c->rho_next, 0, c->N * c->scal_bytes, c->stream) != hipSuccess) ||
        (c->band_cnt && hipMemsetAsync(c->band_cnt, 0, 2 * sizeof(unsigned), c->stream) != hipSuccess) ||
        hipMemsetAsync(c->mv, 0, 3 * c->N * sizeof(double), c->stream) != hipSuccess ||
        hipMemsetAsync(c->mu_w, 0, 2 * c->NXY * sizeof(double), c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return fail(CPMP_EHIP, "initial clear");
    *out = c;
    return CPMP_OK;
}

int cpmp_destroy(cpmp_ctx *c) {
    if (!c) return CPMP_OK;
    if (c->stream) { hipSetDevice(c->p.device); hipStreamSynchronize(c->stream); }
    for (void *ptr : {c->f[0], c->f[1], c->rho, c->rho_next, c->rbuf, c->cbuf, c->mu, (void *)c->band,
                      (void *)c->band_ovf, (void *)c->mv, (void *)c->mu_w, (void *)c->red, (void *)c->staging})
        if (ptr) hipFree(ptr);
    if (c->timers) {
        for (auto &e : c->timers->ev) if (e) hipEventDestroy(e);
        delete c->timers;
    }
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return CPMP_OK;
}

int cpmp_get_params(cpmp_ctx *c, cpmp_params *out) {
    if (!valid_ctx(c) || !out) return set_error(CPMP_EINVAL, "bad context or output");
    *out = c->p;
    return CPMP_OK;
}

int cpmp_memory(cpmp_ctx *c, int64_t *bytes) {
    if (!valid_ctx(c) || !bytes) return set_error(CPMP_EINVAL, "bad context or output");
    *bytes = (int64_t)(2 * (size_t)c->N * 19 * c->pop_bytes + 2 * c->N * c->scal_bytes + 4 * c->N * sizeof(double)
                       + 2 * c->NXY * sizeof(double) + reduce_scratch_doubles() * sizeof(double) + c->fuse_bytes
                       + (c->band ? band_slots(c) * sizeof(uint4) + (band_desc_slots(c) + 1) * sizeof(uint2) : 0));
    return CPMP_OK;
}

// ---------------------------------------------------------------------------
int cpmp_init_reference(cpmp_ctx *c, double radius, double width, double rg, double rl, double solid) {
    if (!valid_ctx(c)) return set_error(CPMP_EINVAL, "bad context");
    if (!(width > 0)) return set_error(CPMP_EINVAL, "width must be > 0");
    HIPCHK(hipSetDevice(c->p.device));
    HIPCHK(launch_init_reference(c, radius, width, rg, rl, solid));
    HIPCHK(sync_reference_density(c));
    HIPCHK(hipMemsetAsync(c->mv, 0, 3 * c->N * sizeof(double), c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->step = 0;
    c->mv_valid = false;
    return CPMP_OK;
}

int cpmp_init_splash(cpmp_ctx *c, double radius, double width, double film_z, double rg, double rl) {
    if (!valid_ctx(c)) return set_error(CPMP_EINVAL, "bad context");
    if (!(width > 0)) return set_error(CPMP_EINVAL, "width must be > 0");
    HIPCHK(hipSetDevice(c->p.device));
    HIPCHK(launch_init_splash(c, radius, width, film_z, rg, rl));
    HIPCHK(sync_reference_density(c));
    HIPCHK(hipMemsetAsync(c->mv, 0, 3 * c->N * sizeof(double), c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->step = 0;
    c->mv_valid = false;
    return CPMP_OK;
}

int cpmp_init_spheres(cpmp_ctx *c, const cpmp_init_desc *d) {
    if (!valid_ctx(c) || !d) return set_error(CPMP_EINVAL, "bad context or descriptor");
    if (d->nspheres < 0 || d->nspheres > 8) return set_error(CPMP_EINVAL, "nspheres %d", d->nspheres);
    if (!(d->width > 0)) return set_error(CPMP_EINVAL, "width must be > 0");
    HIPCHK(hipSetDevice(c->p.device));
    HIPCHK(launch_init_spheres(c, d));
    HIPCHK(sync_reference_density(c));
    HIPCHK(hipMemsetAsync(c->mv, 0, 3 * c->N * sizeof(double), c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->step = 0;
    c->mv_valid = false;
    return CPMP_OK;
}

int cpmp_init_from_rho(cpmp_ctx *c, const double *rho) {
    if (!valid_ctx(c) || !rho) return set_error(CPMP_EINVAL, "bad context or input");
    HIPCHK(hipSetDevice(c->p.device));
    HIPCHK(hipMemcpyAsync(c->staging, rho, c->N * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(launch_scalar_from_double(c, c->staging, c->rho));
    HIPCHK(sync_reference_density(c));
    HIPCHK(launch_feq_from_rho(c));
    HIPCHK(hipMemsetAsync(c->mv, 0, 3 * c->N * sizeof(double), c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->step = 0;
    c->mv_valid = false;
    return CPMP_OK;
}

int cpmp_set_wall_mu(cpmp_ctx *c, const double *lo, const double *hi, int64_t n) {
    if (!valid_ctx(c) || !lo) return set_error(CPMP_EINVAL, "bad context or input");
    if (n != 1 && n != c->NXY) return set_error(CPMP_EINVAL, "wall field has %lld values, expected 1 or %lld",
                                                 (long long)n, (long long)c->NXY);
    std::vector<double> h(2 * c->NXY);
    for (int64_t i = 0; i < c->NXY; ++i) {
        h[i] = lo[n == 1 ? 0 : i];
        h[c->NXY + i] = (hi ? hi : lo)[n == 1 ? 0 : i];
    }
    HIPCHK(hipSetDevice(c->p.device));
    HIPCHK(hipMemcpyAsync(c->mu_w, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return CPMP_OK;
}

// ---------------------------------------------------------------------------
int cpmp_step(cpmp_ctx *c, int nsteps) {
    if (!valid_ctx(c)) return set_error(CPMP_EINVAL, "bad context");
    if (nsteps < 0) return set_error(CPMP_EINVAL, "nsteps %d", nsteps);
    if (nsteps == 0) return CPMP_OK;
    HIPCHK(hipSetDevice(c->p.device));
    HIPCHK(launch_steps(c, nsteps));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->mv_valid = true;
    return CPMP_OK;
}

int cpmp_sync(cpmp_ctx *c) {
    if (!valid_ctx(c)) return set_error(CPMP_EINVAL, "bad context");
    HIPCHK(hipSetDevice(c->p.device));
    HIPCHK(hipStreamSynchronize(c->stream));
    return CPMP_OK;
}

int64_t cpmp_step_count(cpmp_ctx *c) { return valid_ctx(c) ? c->step : -1; }

// ---------------------------------------------------------------------------
int cpmp_get_field(cpmp_ctx *c, int field, double *host) {
    if (!valid_ctx(c) || !host) return set_error(CPMP_EINVAL, "bad context or output");
    HIPCHK(hipSetDevice(c->p.device));
    const size_t nb = c->N * sizeof(double);
    switch (field) {
    case CPMP_FIELD_TYPE:
        HIPCHK(launch_type(c, c->staging));
        HIPCHK(hipMemcpyAsync(host, c->staging, nb, hipMemcpyDeviceToHost, c->stream));
        break;
    case CPMP_FIELD_RHO:
    case CPMP_FIELD_MU:
        HIPCHK(launch_scalar_to_double(c, field == CPMP_FIELD_RHO ? c->rho : c->mu, c->staging));
        HIPCHK(hipMemcpyAsync(host, c->staging, nb, hipMemcpyDeviceToHost, c->stream));
        break;
    case CPMP_FIELD_F:
        for (int q = 0; q < 19; ++q) {
            HIPCHK(launch_pop_to_double(c, q, c->f[c->cur], c->staging));
            HIPCHK(hipMemcpyAsync(host + (size_t)q * c->N, c->staging, nb, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
        }
        break;
    case CPMP_FIELD_VEL:
    case CPMP_FIELD_FORCE: {
        double *tmp = nullptr;
        HIPCHK(hipMalloc((void **)&tmp, 3 * nb));
        hipError_t e;
        if (field == CPMP_FIELD_VEL) {
            e = launch_velocity(c, tmp);
        } else {
            e = launch_wall(c);
            if (e == hipSuccess) e = launch_mu(c);
            if (e == hipSuccess) e = launch_force(c, tmp);
        }
        if (e == hipSuccess) e = hipMemcpyAsync(host, tmp, 3 * nb, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        hipFree(tmp);
        HIPCHK(e);
    } break;
    case CPMP_FIELD_MV:
        HIPCHK(hipMemcpyAsync(host, c->mv, 3 * nb, hipMemcpyDeviceToHost, c->stream));
        break;
    default:
        return set_error(CPMP_EINVAL, "field %d", field);
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return CPMP_OK;
}

int cpmp_set_field(cpmp_ctx *c, int field, const double *host) {
    if (!valid_ctx(c) || !host) return set_error(CPMP_EINVAL, "bad context or input");
    HIPCHK(hipSetDevice(c->p.device));
    const size_t nb = c->N * sizeof(double);
    switch (field) {
    case CPMP_FIELD_RHO:
    case CPMP_FIELD_MU:
        HIPCHK(hipMemcpyAsync(c->staging, host, nb, hipMemcpyHostToDevice, c->stream));
        HIPCHK(launch_scalar_from_double(c, c->staging, field == CPMP_FIELD_RHO ? c->rho : c->mu));
        break;
    case CPMP_FIELD_F:
        for (int q = 0; q < 19; ++q) {
            HIPCHK(hipMemcpyAsync(c->staging, host + (size_t)q * c->N, nb, hipMemcpyHostToDevice, c->stream));
            HIPCHK(launch_pop_from_double(c, q, c->staging, c->f[c->cur]));
            HIPCHK(hipStreamSynchronize(c->stream));
        }
        break;
    default:
        return set_error(CPMP_EINVAL, "field %d cannot be set", field);
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return CPMP_OK;
}

// ---------------------------------------------------------------------------
int cpmp_reduce(cpmp_ctx *c, cpmp_stats *out) {
    if (!valid_ctx(c) || !out) return set_error(CPMP_EINVAL, "bad context or output");
    HIPCHK(hipSetDevice(c->p.device));
    double v[12];
    HIPCHK(launch_reduce(c, v));
    return stats_from_reduction(c, v, out);
}

}  // extern "C"

int cpmp::stats_from_reduction(cpmp_ctx *c, const double *v, cpmp_stats *out) {
    out->mass = v[0];
    out->bad_count = (int64_t)v[1];
    out->first_bad = v[1] > 0 ? (int64_t)v[2] : -1;
    out->rho_min = v[3];
    out->rho_max = v[4];
    out->max_speed = std::sqrt(v[5]);
    out->momentum[0] = v[6]; out->momentum[1] = v[7]; out->momentum[2] = v[8];
    out->kinetic = v[9];
    if (out->bad_count > 0) {
        const int64_t I = out->first_bad;
        const int z = (int)(I % c->p.nz), y = (int)((I / c->p.nz) % c->p.ny), x = (int)(I / ((int64_t)c->p.ny * c->p.nz));
        return set_error(CPMP_ENAN, "bad density at (%d,%d,%d) step %lld (%lld bad nodes)", x, y, z,
                         (long long)c->step, (long long)out->bad_count);
    }
    return CPMP_OK;
}

extern "C" {

// ---------------------------------------------------------------------------
// Checkpoint: magic, version, params, step, cur populations in storage form (in
// the reference node order), rho, wall fields, and in the FP16 form the reference
// density its populations are shifted by.
static const char kMagic[8] = {'C', 'P', 'M', 'P', 'C', 'K', 'P', '1'};

int cpmp_checkpoint_save(cpmp_ctx *c, const char *path) {
    if (!valid_ctx(c) || !path) return set_error(CPMP_EINVAL, "bad context or path");
    HIPCHK(hipSetDevice(c->p.device));
    const size_t popset = (size_t)c->N * 19 * c->pop_bytes, sc = (size_t)c->N * c->scal_bytes;
    const size_t refb = c->form == CPMP_FORM_FP16 ? sc : 0;   // FP16: the reference density
    std::vector<char> buf(popset + sc + 2 * c->NXY * sizeof(double) + refb);
    // populations in the reference node order whatever the device layout
    const void *pops = c->f[c->cur];
    if (c->dp.tiled) {
        HIPCHK(launch_pop_relayout(c, c->f[c->cur], c->f[c->cur ^ 1], false));
        pops = c->f[c->cur ^ 1];
    }
    HIPCHK(hipMemcpyAsync(buf.data(), pops, popset, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(buf.data() + popset, c->rho, sc, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(buf.data() + popset + sc, c->mu_w, 2 * c->NXY * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (refb)
        HIPCHK(hipMemcpyAsync(buf.data() + popset + sc + 2 * c->NXY * sizeof(double), c->rho_next, refb,
                              hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    FILE *fp = std::fopen(path, "wb");
    if (!fp) return set_error(CPMP_EIO, "cannot write %s", path);
    const uint64_t nbytes = buf.size();
    bool ok = std::fwrite(kMagic, 1, 8, fp) == 8 && std::fwrite(&c->p, sizeof c->p, 1, fp) == 1 &&
              std::fwrite(&c->step, sizeof c->step, 1, fp) == 1 && std::fwrite(&nbytes, sizeof nbytes, 1, fp) == 1 &&
              std::fwrite(buf.data(), 1, buf.size(), fp) == buf.size();
    ok = (std::fclose(fp) == 0) && ok;
    return ok ? CPMP_OK : set_error(CPMP_EIO, "short write to %s", path);
}

int cpmp_checkpoint_load(cpmp_ctx *c, const char *path) {
    if (!valid_ctx(c) || !path) return set_error(CPMP_EINVAL, "bad context or path");
    FILE *fp = std::fopen(path, "rb");
    if (!fp) return set_error(CPMP_EIO, "cannot open %s", path);
    char magic[8];
    cpmp_params p;
    int64_t step;
    uint64_t nbytes;
    const size_t popset = (size_t)c->N * 19 * c->pop_bytes, sc = (size_t)c->N * c->scal_bytes;
    const size_t refb = c->form == CPMP_FORM_FP16 ? sc : 0;
    const size_t want = popset + sc + 2 * c->NXY * sizeof(double) + refb;
    bool ok = std::fread(magic, 1, 8, fp) == 8 && std::memcmp(magic, kMagic, 8) == 0;
    if (!ok) { std::fclose(fp); return set_error(CPMP_EIO, "%s is not a checkpoint (bad magic)", path); }
    ok = std::fread(&p, sizeof p, 1, fp) == 1 && std::fread(&step, sizeof step, 1, fp) == 1 &&
         std::fread(&nbytes, sizeof nbytes, 1, fp) == 1;
    if (!ok || nbytes != want || p.nx != c->p.nx || p.ny != c->p.ny || p.nz != c->p.nz || p.form != c->p.form ||
        p.nwall_lo != c->p.nwall_lo || p.nwall_hi != c->p.nwall_hi) {
        std::fclose(fp);
        return set_error(CPMP_EIO, "%s does not match this context (domain, form or size)", path);
    }
    std::vector<char> buf(want);
    ok = std::fread(buf.data(), 1, want, fp) == want;
    std::fclose(fp);
    if (!ok) return set_error(CPMP_EIO, "%s is truncated", path);
    HIPCHK(hipSetDevice(c->p.device));
    if (c->dp.tiled) {
        HIPCHK(hipMemcpyAsync(c->f[c->cur ^ 1], buf.data(), popset, hipMemcpyHostToDevice, c->stream));
        HIPCHK(launch_pop_relayout(c, c->f[c->cur ^ 1], c->f[c->cur], true));
    } else {
        HIPCHK(hipMemcpyAsync(c->f[c->cur], buf.data(), popset, hipMemcpyHostToDevice, c->stream));
    }
    HIPCHK(hipMemcpyAsync(c->rho, buf.data() + popset, sc, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->mu_w, buf.data() + popset + sc, 2 * c->NXY * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (refb)
        HIPCHK(hipMemcpyAsync(c->rho_next, buf.data() + popset + sc + 2 * c->NXY * sizeof(double), refb,
                              hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->step = step;
    c->mv_valid = false;
    return CPMP_OK;
}

// ---------------------------------------------------------------------------
int cpmp_time_steps(cpmp_ctx *c, int nsteps, double *ms) {
    if (!valid_ctx(c) || !ms || nsteps < 1) return set_error(CPMP_EINVAL, "bad arguments");
    HIPCHK(hipSetDevice(c->p.device));
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipEventRecord(a, c->stream));
    HIPCHK(launch_steps(c, nsteps));
    HIPCHK(hipEventRecord(b, c->stream));
    HIPCHK(hipEventSynchronize(b));
    float t = 0;
    HIPCHK(hipEventElapsedTime(&t, a, b));
    hipEventDestroy(a);
    hipEventDestroy(b);
    *ms = t;
    c->mv_valid = true;
    return CPMP_OK;
}

int cpmp_kernel_times(cpmp_ctx *c, int nsteps, double *ms, int *launches) {
    if (!valid_ctx(c) || !ms || nsteps < 1) return set_error(CPMP_EINVAL, "bad arguments");
    HIPCHK(hipSetDevice(c->p.device));
    // each stage timed with its own event pair on the context's stream
    const int passes = std::max(c->p.nwall_lo, c->p.nwall_hi);
    hipEvent_t ev[8];
    for (auto &e : ev) HIPCHK(hipEventCreate(&e));
    double acc[CPMP_K_COUNT] = {0, 0, 0, 0};
    // the last step's collide also writes the MV field (a slower instantiation): the
    // collide average covers the steps before it, unless there is only one
    const int ncol = nsteps > 1 ? nsteps - 1 : 1;
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int n = 0; n < nsteps; ++n) {
        HIPCHK(hipEventRecord(ev[0], c->stream));
        HIPCHK(launch_wall(c));
        HIPCHK(hipEventRecord(ev[1], c->stream));
        HIPCHK(hipEventRecord(ev[2], c->stream));
        HIPCHK(launch_mu(c));
        HIPCHK(hipEventRecord(ev[3], c->stream));
        HIPCHK(hipEventRecord(ev[4], c->stream));
        HIPCHK(launch_collide(c, n == nsteps - 1));
        HIPCHK(hipEventRecord(ev[5], c->stream));
        HIPCHK(hipEventRecord(ev[6], c->stream));
        HIPCHK(launch_macro(c));
        HIPCHK(hipEventRecord(ev[7], c->stream));
        HIPCHK(hipEventSynchronize(ev[7]));
        for (int k = 0; k < CPMP_K_COUNT; ++k) {
            if (k == CPMP_K_COLLIDE && n >= ncol) continue;
            float t = 0;
            HIPCHK(hipEventElapsedTime(&t, ev[2 * k], ev[2 * k + 1]));
            acc[k] += t;
        }
        ++c->step;
    }
    for (auto &e : ev) hipEventDestroy(e);
    const int per[CPMP_K_COUNT] = {passes, 1, 1, 1};
    for (int k = 0; k < CPMP_K_COUNT; ++k) {
        ms[k] = per[k] ? acc[k] / (k == CPMP_K_COLLIDE ? ncol : nsteps) / per[k] : 0.0;
        if (launches) launches[k] = per[k];
    }
    c->mv_valid = true;
    return CPMP_OK;
}

double cpmp_algorithmic_bytes_per_node(cpmp_ctx *c) {
    return valid_ctx(c) ? 2.0 * 19.0 * (double)c->pop_bytes : 0.0;
}

// ---------------------------------------------------------------------------
// CPU evaluation of the device collision code (parity self-test without a GPU)
int cpmp_selftest_collide(const double *f, double rho, const double *F, const double *ts, const double *qp,
                          double *out, double *u) {
    if (!f || !F || !ts || !out || !u) return set_error(CPMP_EINVAL, "NULL argument");
    host_collide_fp64(f, rho, F, ts, qp, out, u, 0);
    return CPMP_OK;
}
int cpmp_selftest_collide_guo(const double *f, double rho, const double *F, const double *ts, const double *qp,
                              double *out, double *u) {
    if (!f || !F || !ts || !out || !u) return set_error(CPMP_EINVAL, "NULL argument");
    host_collide_fp64(f, rho, F, ts, qp, out, u, 1);
    return CPMP_OK;
}

}  // extern "C"
