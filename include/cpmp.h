/*
 * cpmp.h -- C ABI of the MI355X (gfx950) chemical-potential multiphase D3Q19
 * MRT lattice-Boltzmann step.
 *
 * The reference (zhou-yiling/3DCPMPmodel) has no library interface: its host
 * code owns file-scope device pointers, sets __constant__ symbols and launches
 * five kernels per step from main().  Each entry point below replaces one of
 * those pieces; the comment names the reference lines it stands in for.
 *
 * Conventions
 *   - Every function returns 0 on success and a negative CPMP_E* code on
 *     failure; cpmp_last_error() gives the message of the last failure on the
 *     calling thread.
 *   - Host arrays use the reference layout I(x,y,z) = (x*DY + y)*DZ + z
 *     (Common.h:164); populations are SoA, f[q*N + I] (FluidCuda.cu:238).
 *     Host values are FP64 whatever the device storage form.
 *   - The context owns all device memory (hipMalloc; no managed memory) and
 *     one HIP stream.  Calls are synchronous on return.  One context per host
 *     thread; contexts on different threads are independent.
 */
#ifndef CPMP_H
#define CPMP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ---- */
#define CPMP_OK        0
#define CPMP_EINVAL   -1   /* bad argument */
#define CPMP_EHIP     -2   /* HIP runtime error */
#define CPMP_ENOMEM   -3   /* device or host allocation failed */
#define CPMP_EIO      -4   /* file missing / unreadable / malformed */
#define CPMP_ENAN     -5   /* NaN or negative density found (CalcMacroCPU, FluidCuda.cu:1632) */
#define CPMP_ENOTFOUND -6  /* Tr not present in a coexistence table */
#define CPMP_ENODEV   -7   /* no HIP device */

/* ---- EOS ids (model families of Common.h:54-59, FluidCuda.cu:1318-1344) ---- */
#define CPMP_EOS_VDW 0
#define CPMP_EOS_CSE 1   /* Carnahan-Starling */
#define CPMP_EOS_RKE 2   /* Redlich-Kwong */
#define CPMP_EOS_RKS 3   /* Redlich-Kwong-Soave */
#define CPMP_EOS_PRW 4   /* Peng-Robinson, omega = 0.344 */
#define CPMP_EOS_PRM 5   /* Peng-Robinson, omega = 0.011 */

/* ---- storage / arithmetic forms ---- */
#define CPMP_FORM_FP64 0  /* FP64 populations, FP64 arithmetic */
#define CPMP_FORM_FP32 1  /* FP32 populations, FP32 arithmetic, FP64 EOS mu and mass sums */
#define CPMP_FORM_FP16 2  /* FP16 populations stored as f_i - w_i*rho_ref, FP32 arithmetic;
                             rho_ref = density the pulling node had in the step that wrote
                             the value (the reference shift w_i*rho0, made per node) */
#define CPMP_FORM_FP32A 3 /* FP64 populations, rho and mu (the EOS mu in FP64 everywhere),
                             FP32 collision, force and streaming arithmetic */

/* ---- field ids for get/set ---- */
#define CPMP_FIELD_TYPE  0  /* 1 fluid, 2 solid (Common.h:37-38); get only */
#define CPMP_FIELD_RHO   1  /* Dens */
#define CPMP_FIELD_MU    2  /* Pote */
#define CPMP_FIELD_F     3  /* Dist, 19*N; a solid node's 19 values are never pulled and carry no state */
#define CPMP_FIELD_VEL   4  /* Vx,Vy,Vz = j/rho of the current populations, 3*N; get only */
#define CPMP_FIELD_MV    5  /* MVx,MVy,MVz of the last step, 3*N; get only */
#define CPMP_FIELD_FORCE 6  /* Fx,Fy,Fz of the current state, 3*N; get only */

/* Parameters: the Common.h domain + Initialize()/SetMultiphase() values
   (Common.h:25-27, FluidCuda.cu:119-163, 1314-1353). */
typedef struct {
    int nx, ny, nz;            /* DX, DY, DZ */
    int nwall_lo, nwall_hi;    /* solid layers at the z ends; the reference has 3 and 3
                                  (FluidCuda.cu:204); 0 and 0 = fully periodic */
    int eos;                   /* CPMP_EOS_* */
    int mu_form;               /* 0: the EOS's own closed-form mu; 1: the PR expression of
                                  FluidCuda.cu:453 for every EOS, as the reference kernel does */
    int lap_points;            /* 5: GradD5 (PTCALCFIVEPOINT), 7: GradD7 (PTCALCSEVENPOINT) */
    int wall_scheme;           /* 1: layered (COMPLEXCHEMBOUNDARY), 0: SIMPLECHEMBOUNDARY,
                                  2: ChemBoundaryComplex (FluidCuda.cu:403-429): solid layer d
                                  (LEVELd of Common.h:40-42, d = distance to the fluid) is the
                                  D3Q19-weighted average of its neighbours in layer d-1 */
    int form;                  /* CPMP_FORM_* */
    int device;                /* HIP device ordinal */
    double A, B, T;            /* EOS constants (SetMultiphase; T = Tr*Tc) */
    double K, Ka, K1, K2;      /* EOS scale, kappa, surface-tension tuning (FluidCuda.cu:124-127) */
    double Ts[19];             /* MRT relaxation rates (FluidCuda.cu:154-157) */
    double rho0;               /* not used: the FP16 form takes its shift density per node */
    int forcing;               /* 0: exact-difference force term (FluidCuda.cu:326-341),
                                  1: Guo forcing in MRT form (no adjoint) */
} cpmp_params;

/* Generic initial condition: up to 8 spheres with tanh interfaces, optional
   counter-based noise.  rho = rhoG + (rhoL-rhoG)*max_s phi_s, phi_s =
   (1 - tanh(2(r_s - R_s)/width))/2 (a bubble when invert != 0), then
   rho += noise_amp * U(seed, node) with U uniform in [-1,1). */
typedef struct {
    int nspheres;
    double cx[8], cy[8], cz[8], radius[8];
    double width;
    double rho_g, rho_l;
    int invert;
    double noise_amp;
    uint64_t seed;
    double solid_rho;          /* rho = mu value written into solid nodes at init */
} cpmp_init_desc;

/* FP64 reductions of CalcMacroCPU (FluidCuda.cu:1626-1657) on the device. */
typedef struct {
    double mass;               /* sum of rho over fluid nodes */
    double max_speed;          /* max |MV| over fluid nodes of the last step */
    int64_t bad_count;         /* NaN or negative densities */
    int64_t first_bad;         /* lowest node index with a bad density, -1 if none */
    double rho_min, rho_max;   /* over fluid nodes */
    double momentum[3];        /* sum of rho*MV over fluid nodes */
    double kinetic;            /* sum of rho*|MV|^2/2 over fluid nodes */
} cpmp_stats;

/* Adjoint of a droplet-shape objective (no reference counterpart). */
typedef struct {
    int nsteps;                /* steps of the forward run, starting from the current state */
    int checkpoint_every;      /* steps between stored population checkpoints (>=1) */
    const double *target;      /* rho target, N values (reference layout) */
    /* Device bytes for the tape (rho, u, grad mu, grad rho per node and step, 80 B) of the
       trailing steps recorded during the first forward run; those steps are not
       recomputed.  The steps before them are covered by checkpoints every
       checkpoint_every steps and recomputed segment by segment.  0: checkpoints and
       recompute for every step but those of the last segment, which the working tape
       holds during the first run; < 0: half of the free device memory. */
    int64_t tape_budget_bytes;
    /* Limit on the device bytes the call allocates (0: the free device memory less
       2 GiB).  A segment whose tape does not fit is recomputed in sub-segments:
       a first pass stores sub-checkpoints, then each sub-segment is taped in turn
       (each of its steps is recomputed twice).  The result does not depend on the
       schedule. */
    int64_t memory_limit_bytes;
} cpmp_adj_desc;

/* The schedule and memory of one adjoint call (cpmp_adjoint_plan). */
typedef struct {
    int64_t device_bytes;      /* device memory the call allocates besides the context */
    int taped_steps;           /* trailing steps taped during the first forward run */
    int segments;              /* checkpoint segments before them */
    int sub_length;            /* steps per taped sub-segment (= segment length when it fits) */
    int sub_checkpoints;       /* extra states stored per segment for the sub-segments */
    int64_t recomputed_steps;  /* forward steps run again during the reverse sweep */
    /* 1: each segment is recomputed into a second tape on a second stream while the
       reverse sweep of the segment after it runs (two tapes in device_bytes); 0: one
       tape, recompute and reverse sweep in turn.  The result does not depend on it. */
    int overlapped;
} cpmp_adj_plan;

typedef struct cpmp_ctx cpmp_ctx;

/* ---- host-side EOS / parameter helpers ---- */
/* SetMultiphase constants (FluidCuda.cu:1318-1344): A, B, Tc, Rc for an EOS at Tr. */
int cpmp_set_multiphase(int eos, double Tr, double *A, double *B, double *Tc, double *Rc);
/* Fill *p with the reference defaults of Initialize() (FluidCuda.cu:119-163) for
   the given domain, EOS and Tr. */
int cpmp_default_params(cpmp_params *p, int nx, int ny, int nz, int eos, double Tr);
/* DimConversion (FluidCuda.cu:474-486): unit factors from (lattice, physical) length,
   density and viscosity.  out[9]: DimLen, DimTime, DimMass, DimF, DimT, DimV, DimA,
   DimW, DimB. */
int cpmp_dim_conversion(double length, double macro_length, double density, double macro_density,
                        double viscosity, double macro_viscosity, double *out);
/* EOS pressure (FluidCuda.cu:1561-1573, extended to all EOS) and chemical potential
   (FluidCuda.cu:453 for PR) of the bulk, without the K^2 factor. */
double cpmp_eos_pressure(int eos, double A, double B, double T, double rho);
double cpmp_eos_mu(int eos, int mu_form, double A, double B, double T, double rho);
/* Coexistence densities by Maxwell construction (equal p, equal mu); starting
   values are taken from *rho_g, *rho_l when both are > 0. */
int cpmp_coexistence(int eos, double A, double B, double T, double *rho_g, double *rho_l);
/* Coexistence densities from a CoCurve table file (FluidCuda.cu:1356-1384):
   returns rho/rho_c of the row whose Tr matches within 1e-12; CPMP_ENOTFOUND
   if no row matches, CPMP_EIO if the file cannot be read. */
int cpmp_cocurve_lookup(const char *path, double Tr, double *rhog_reduced, double *rhol_reduced);

/* ---- context (replaces CudaInitialize / CudaFree, FluidCuda.cu:1398-1485) ---- */
int cpmp_create(const cpmp_params *p, cpmp_ctx **out);
int cpmp_destroy(cpmp_ctx *ctx);
int cpmp_get_params(cpmp_ctx *ctx, cpmp_params *out);
/* device memory of the context in bytes */
int cpmp_memory(cpmp_ctx *ctx, int64_t *bytes);

/* ---- initial conditions (SetFlowField, FluidCuda.cu:199-241) ---- */
/* CONTACTANGLE branch: sphere of `radius` centred at (DX/2, DY/2, radius+2), tanh
   width `width`, r evaluated with single-precision sqrt as the reference does. */
int cpmp_init_reference(cpmp_ctx *ctx, double radius, double width, double rho_g, double rho_l,
                        double solid_value);
/* DROPLETSPLASH branch (FluidCuda.cu:219-233): rho = rho_l + (rho_l-rho_g)/2 *
   (tanh((radius-r)*2/width) + tanh((film_z-k)*2/width)), r from the box centre in
   single precision; the reference fixes radius 50, width 10, film_z 32.  Solid rho and
   mu are 0. */
int cpmp_init_splash(cpmp_ctx *ctx, double radius, double width, double film_z, double rho_g, double rho_l);
int cpmp_init_spheres(cpmp_ctx *ctx, const cpmp_init_desc *d);
/* rho from the host; populations = Feq(rho, 0). */
int cpmp_init_from_rho(cpmp_ctx *ctx, const double *rho);

/* Wall chemical potential (the BasePt of FluidCuda.cu:387, 1268), generalised to a
   field over the nx*ny wall columns; hi may be NULL (same as lo).  n must be
   nx*ny, or 1 for a scalar. */
int cpmp_set_wall_mu(cpmp_ctx *ctx, const double *lo, const double *hi, int64_t n);

/* ---- the step (FluidCuda.cu:1277-1292) ---- */
int cpmp_step(cpmp_ctx *ctx, int nsteps);
int cpmp_sync(cpmp_ctx *ctx);
int64_t cpmp_step_count(cpmp_ctx *ctx);

/* ---- state transfer ---- */
int cpmp_get_field(cpmp_ctx *ctx, int field, double *host);
int cpmp_set_field(cpmp_ctx *ctx, int field, const double *host);   /* RHO, MU or F */

/* ---- diagnostics ---- */
int cpmp_reduce(cpmp_ctx *ctx, cpmp_stats *out);
/* cpmp_reduce without context scratch (no counterpart in the reference): for callers
   that cannot allocate device memory around the call (code under HIP graph capture,
   tight serving loops, harnesses that forbid hipMalloc).  The same kernels and
   results as cpmp_reduce, bit for bit; the reduction partials live in a static
   device array of 513 x 12 doubles (49 248 bytes) per GPU, held until the process
   ends, which the fixed 512-block grid fits for any lattice.  The call allocates
   nothing.  Calls that overlap on one GPU share that array and race. */
int cpmp_reduce_noalloc(cpmp_ctx *ctx, cpmp_stats *out);
/* ContactAngle(side=0, tier) (FluidCuda.cu:557-689) on the XZ plane y=tier, with the
   interface density (rho_g+rho_l)/2.  out[8]: CaLeft, CaRight, CaCap (degrees),
   XLeft, XRight, H, L, valid. */
int cpmp_contact_angle(cpmp_ctx *ctx, int tier, double rho_g, double rho_l, double *out);
/* 3-D spherical-cap fit of the rho=(rho_g+rho_l)/2 surface above the bottom wall.
   out[4]: angle (degrees), cap height, base radius, sphere radius. */
int cpmp_cap_fit(cpmp_ctx *ctx, double rho_g, double rho_l, double *out);
/* Density (CPMP_FIELD_RHO) or chemical potential (CPMP_FIELD_MU) of one node, FP64,
   without copying the field (the reference reads managed Dens / Pote elements). */
int cpmp_probe(cpmp_ctx *ctx, int field, int x, int y, int z, double *value);
/* Neck radius of two coalescing droplets (no reference counterpart): in plane x,
   A = sum over fluid nodes of clamp((rho - rho_g) / (rho_l - rho_g), 0, 1) and
   radius = sqrt(A / pi).  area may be NULL. */
int cpmp_neck_radius(cpmp_ctx *ctx, int x, double rho_g, double rho_l, double *radius, double *area);
/* Laplace pressure jump K^2 [p(rho(in)) - p(rho(out))] between two nodes, with the
   EOS pressure of EosPressure (FluidCuda.cu:1561-1573) for the context's EOS. */
int cpmp_laplace_jump(cpmp_ctx *ctx, const int *in_xyz, const int *out_xyz, double *dp);

/* ---- checkpoint / resume (no reference counterpart) ---- */
int cpmp_checkpoint_save(cpmp_ctx *ctx, const char *path);
int cpmp_checkpoint_load(cpmp_ctx *ctx, const char *path);

/* ---- Tecplot binary output (ordered_tec.cpp write_plt, #!TDV112) ---- */
/* Writes a one-zone PLT with variables i, j, k (as the reference's SaveTask) and the
   named fields (CPMP_FIELD_RHO / MU / VEL components).  decimate = 1 keeps every
   node; the reference's SaveTask uses 5 with a linear-index stride
   (FluidCuda.cu:466-472), reproduced when decimate_linear != 0. */
int cpmp_write_plt(cpmp_ctx *ctx, const char *path, const char *title, int decimate, int decimate_linear);

/* ---- discrete adjoint (no reference counterpart) ---- */
/* J = 1/2 sum_fluid (rho^N - target)^2 after d->nsteps steps from the current state.
   Returns dJ/d mu_w(bottom), dJ/d mu_w(top) (nx*ny each, may be NULL) and dJ/dKa.
   The state is left at step N of the forward run.  FP64 and FP32 forms: the forward
   run (and its recomputation) is the context's own step; the reverse sweep computes
   in FP64 on the tape, so in the FP32 form it is the FP64 adjoint evaluated along the
   FP32 trajectory (its population cotangents are stored in FP32 between steps). */
int cpmp_adjoint(cpmp_ctx *ctx, const cpmp_adj_desc *d, double *J, double *g_mu_lo,
                 double *g_mu_hi, double *g_ka);
/* The schedule cpmp_adjoint would take for d, without running it (CPMP_ENOMEM when
   no schedule fits the memory limit). */
int cpmp_adjoint_plan(cpmp_ctx *ctx, const cpmp_adj_desc *d, cpmp_adj_plan *plan);

/* ---- measurement ---- */
/* Average device time per launch (ms) of each kernel over nsteps steps, measured
   with HIP events on the context's stream (the collide over the first nsteps - 1
   steps: the last one also writes the MV field).  names: CPMP_K_* order below. */
#define CPMP_K_WALL    0
#define CPMP_K_MU      1
#define CPMP_K_COLLIDE 2
#define CPMP_K_MACRO   3
#define CPMP_K_COUNT   4
int cpmp_kernel_times(cpmp_ctx *ctx, int nsteps, double *ms_per_launch, int *launches_per_step);
/* Elapsed device time (ms) of nsteps whole steps on the context's stream. */
int cpmp_time_steps(cpmp_ctx *ctx, int nsteps, double *ms);
/* Algorithmic bytes per node per step of the current form (2*19*sizeof(population)). */
double cpmp_algorithmic_bytes_per_node(cpmp_ctx *ctx);

const char *cpmp_last_error(void);
const char *cpmp_version(void);

/* ---- host self-test (no reference counterpart) ----
   Runs the device collision code of one fluid node on the CPU in FP64
   (GlobalCollide's per-node arithmetic, FluidCuda.cu:316-344, without the
   stream): out = post-collision populations incl. the exact-difference force
   term, u3 = j(f)/rho used by it.  qp may be NULL (K1 = K2 = 0). */
int cpmp_selftest_collide(const double *f19, double rho, const double *F3, const double *ts19,
                          const double *qp19, double *out19, double *u3);
/* The same with the Guo force term of cpmp_params.forcing = 1 instead of the
   exact-difference term; u3 = j(f)/rho. */
int cpmp_selftest_collide_guo(const double *f19, double rho, const double *F3, const double *ts19,
                              const double *qp19, double *out19, double *u3);

#ifdef __cplusplus
}
#endif
#endif
