/*
 * cpmp_oracle.h -- CPU FP64 restatement of the chemical-potential multiphase
 * D3Q19 MRT lattice-Boltzmann step of zhou-yiling/3DCPMPmodel.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness checker for the
 * HIP implementation in 3dcpmpmodel_amd/ and the timed CPU baseline.  Only the
 * test suite (tests/), __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.  The product library never links, loads or calls it.
 *
 * Every function follows the reference line by line; the citation next to each
 * names the file:line it restates.  Arrays use the reference layout
 * I(x,y,z) = (x*DY + y)*DZ + z  (Common.h:164), populations SoA f[q*N + I].
 *
 * Pinning status of each function is listed in cpmp_oracle.cpp's header.
 */
#ifndef CPMP_ORACLE_H
#define CPMP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* EOS ids, order of FluidCuda.cu:1318-1344 cases (SetMultiphase). */
enum { ORC_VDW = 0, ORC_CSE = 1, ORC_RKE = 2, ORC_RKS = 3, ORC_PRW = 4, ORC_PRM = 5 };

typedef struct {
    int nx, ny, nz;          /* DX, DY, DZ (Common.h:25-27) */
    int nwall_lo, nwall_hi;  /* solid layers at z=0.. and z=DZ-1.. (FluidCuda.cu:204 uses 3,3) */
    int eos;                 /* ORC_* */
    int mu_form;             /* 0: each EOS's own closed-form mu; 1: the PR expression of
                                FluidCuda.cu:453 for every EOS (what the reference kernel does) */
    int lap_points;          /* 5: GradD5 (PTCALCFIVEPOINT), 7: GradD7 (PTCALCSEVENPOINT) */
    int wall_scheme;         /* 1: COMPLEXCHEMBOUNDARY (layered), 0: SIMPLECHEMBOUNDARY,
                                2: ChemBoundaryComplex (FluidCuda.cu:403-429, levels of
                                Common.h:40-42) */
    double A, B, T;          /* EOS constants after SetMultiphase (FluidCuda.cu:1321-1353) */
    double K, Ka, K1, K2;    /* FluidCuda.cu:124-127 */
    double Ts[19];           /* relaxation rates (FluidCuda.cu:154-157) */
    int forcing;             /* 0: exact-difference term (FluidCuda.cu:326-341), 1: Guo */
} orc_params;

typedef struct orc_state orc_state;

/* ---- EOS (FluidCuda.cu:1314-1394, 453, 1561-1573) ---- */
int    orc_set_multiphase(int eos, double Tr, double *A, double *B, double *Tc, double *Rc);
double orc_pressure(int eos, double A, double B, double T, double rho);
double orc_mu_eos(int eos, int mu_form, double A, double B, double T, double rho);
int    orc_maxwell(int eos, double A, double B, double T, double *rho_g, double *rho_l);

/* ---- lattice tables (Common.h:94-146) ---- */
void   orc_tables(int *ex, int *ey, int *ez, int *re, double *w, double *M, double *R);

/* ---- state ---- */
orc_state *orc_create(const orc_params *p);
void   orc_destroy(orc_state *s);
/* field ids for get/set */
enum { ORC_TYPE = 0, ORC_RHO = 1, ORC_MU = 2, ORC_F = 3, ORC_VEL = 4, ORC_MV = 5, ORC_FORCE = 6 };
int    orc_get(orc_state *s, int field, double *out);
int    orc_set(orc_state *s, int field, const double *in);
/* wall chemical potential, bottom / top fields of nx*ny (BasePt, FluidCuda.cu:387) */
int    orc_set_wall_mu(orc_state *s, const double *mu_lo, const double *mu_hi);

/* SetFlowField, CONTACTANGLE branch (FluidCuda.cu:199-218, 236-239) */
int    orc_init_reference(orc_state *s, double radius, double width, double den_g, double den_l,
                          double basept_init);
/* Types from the wall layers, given rho everywhere, f = Feq(rho, 0)  (FluidCuda.cu:236-239) */
int    orc_init_from_rho(orc_state *s, const double *rho);

/* One or more time steps in the order of FluidCuda.cu:1277-1292 */
int    orc_step(orc_state *s, int nsteps);
/* the individual stages, for stage-level parity tests */
int    orc_stage(orc_state *s, int stage);   /* 0 wall, 1 mu, 2 force, 3 collide+stream+swap, 4 macro */

/* CalcMacroCPU (FluidCuda.cu:1626-1657): out[0]=mass, [1]=nan/neg count, [2]=max speed */
int    orc_reduce(orc_state *s, double *out);

/* ContactAngle(side=0, tier) (FluidCuda.cu:557-689) on rho of the state, rows
   offset by the bottom wall depth (the reference's rows for nwall_lo = 3); -1 when
   nwall_lo < 1.  out: CaLeft, CaRight, CaCap, XLeft, XRight, H, L, valid */
int    orc_contact_angle(const orc_params *p, const double *rho, double den_g, double den_l,
                         int tier, double *out);

/* SplitMix64 output function: the generator's next value for state x, i.e. the mix
   of x + 0x9E3779B97F4A7C15 (Steele, Lea, Flood 2014) */
uint64_t orc_splitmix64(uint64_t x);
/* SplitMix64-based uniform noise in [-1,1) keyed by (seed, index) */
double orc_noise(uint64_t seed, uint64_t index);

/* ---- discrete adjoint (reverse mode of orc_step) ---- */
/* J = 1/2 sum_fluid (rho^N - target)^2 after nsteps from the state's current contents.
   g_mu_lo: dJ/d mu_w(bottom) [nx*ny], g_mu_hi: dJ/d mu_w(top) [nx*ny], g_ka: dJ/dKa.
   The forward trajectory is stored in memory (oracle sizes are small).  -1 for the
   Guo force term and for wall scheme 2, which it does not cover. */
int    orc_adjoint(orc_state *s, int nsteps, const double *target, double *J,
                   double *g_mu_lo, double *g_mu_hi, double *g_ka);

#ifdef __cplusplus
}
#endif
#endif
