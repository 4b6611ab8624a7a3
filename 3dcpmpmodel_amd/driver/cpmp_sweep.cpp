// cpmp_sweep -- the reference program's driver (main, FluidCuda.cu:1255-1304) on the
// C ABI of libcpmp: one droplet set up once (SetFlowField, CONTACTANGLE or
// DROPLETSPLASH branch), then a sweep of the wall chemical potential BasePt with
// the state carried over from one value to the next.  Every show_step steps it
// prints the ShowData row (FluidCuda.cu:1659-1728) and, at the end of each BasePt,
// appends the SaveContactAngle line (FluidCuda.cu:1730-1771) to
//   <out>/CP<EOS>_MRT_Tau=<tau>_FIELD<DX>x<DY>x<DZ>_ContactAngle_R<R>_Tr<Tr>.txt
//
//   cpmp_sweep [--dims DX DY DZ] [--eos PRW|PRM|VDW|RKE|RKS|CSE] [--tr Tr]
//              [--radius R] [--width W] [--basept FROM TO STEP] [--steps AllStep]
//              [--show STEP] [--form fp64|fp32|fp16] [--splash] [--out DIR]
//              [--cocurve FILE]
//
// Without --cocurve the coexistence densities come from the Maxwell construction
// (cpmp_coexistence); with it, from the reference's CoCurve table (times Rc).
#include "cpmp.h"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

namespace {

const char *eos_name(int eos) {
    switch (eos) {
    case CPMP_EOS_VDW: return "VDW";
    case CPMP_EOS_CSE: return "CSE";
    case CPMP_EOS_RKE: return "RKE";
    case CPMP_EOS_RKS: return "RKS";
    case CPMP_EOS_PRW: return "PRW";
    default: return "PRM";
    }
}

int eos_id(const char *s) {
    const char *n[] = {"VDW", "CSE", "RKE", "RKS", "PRW", "PRM"};
    const int id[] = {CPMP_EOS_VDW, CPMP_EOS_CSE, CPMP_EOS_RKE, CPMP_EOS_RKS, CPMP_EOS_PRW, CPMP_EOS_PRM};
    for (int k = 0; k < 6; ++k)
        if (std::strcmp(s, n[k]) == 0) return id[k];
    return -1;
}

int fail(const char *what) {
    std::fprintf(stderr, "cpmp_sweep: %s: %s\n", what, cpmp_last_error());
    return 1;
}

}  // namespace

int main(int argc, char **argv) {
    int DX = 128, DY = 128, DZ = 96, eos = CPMP_EOS_PRW, form = CPMP_FORM_FP64, all_steps = 10000, show = 1000;
    double Tr = 0.7, radius = 30, width = 5, bp_from = -0.08, bp_to = -0.015, bp_step = 0.005;
    bool splash = false;
    std::string out = ".", cocurve;
    for (int a = 1; a < argc; ++a) {
        const std::string k = argv[a];
        auto need = [&](int n) {
            if (a + n >= argc) { std::fprintf(stderr, "cpmp_sweep: %s needs %d value(s)\n", k.c_str(), n); std::exit(2); }
        };
        if (k == "--dims") { need(3); DX = std::atoi(argv[++a]); DY = std::atoi(argv[++a]); DZ = std::atoi(argv[++a]); }
        else if (k == "--eos") { need(1); eos = eos_id(argv[++a]); if (eos < 0) { std::fprintf(stderr, "cpmp_sweep: unknown EOS\n"); return 2; } }
        else if (k == "--tr") { need(1); Tr = std::atof(argv[++a]); }
        else if (k == "--radius") { need(1); radius = std::atof(argv[++a]); }
        else if (k == "--width") { need(1); width = std::atof(argv[++a]); }
        else if (k == "--basept") { need(3); bp_from = std::atof(argv[++a]); bp_to = std::atof(argv[++a]); bp_step = std::atof(argv[++a]); }
        else if (k == "--steps") { need(1); all_steps = std::atoi(argv[++a]); }
        else if (k == "--show") { need(1); show = std::atoi(argv[++a]); }
        else if (k == "--form") {
            need(1);
            const std::string f = argv[++a];
            form = f == "fp32" ? CPMP_FORM_FP32 : f == "fp16" ? CPMP_FORM_FP16 : CPMP_FORM_FP64;
        }
        else if (k == "--splash") splash = true;
        else if (k == "--out") { need(1); out = argv[++a]; }
        else if (k == "--cocurve") { need(1); cocurve = argv[++a]; }
        else { std::fprintf(stderr, "cpmp_sweep: unknown option %s\n", k.c_str()); return 2; }
    }
    if (show < 1 || all_steps < 1 || bp_step == 0) { std::fprintf(stderr, "cpmp_sweep: bad step settings\n"); return 2; }

    // Initialize() + SetMultiphase() (FluidCuda.cu:119-163, 1314-1394)
    cpmp_params p;
    if (cpmp_default_params(&p, DX, DY, DZ, eos, Tr)) return fail("parameters");
    p.form = form;
    double A, B, Tc, Rc, rhoG = 0, rhoL = 0;
    if (cpmp_set_multiphase(eos, Tr, &A, &B, &Tc, &Rc)) return fail("SetMultiphase");
    if (!cocurve.empty()) {
        if (cpmp_cocurve_lookup(cocurve.c_str(), Tr, &rhoG, &rhoL)) return fail("CoCurve");
        rhoG *= Rc;
        rhoL *= Rc;
    } else if (cpmp_coexistence(eos, A, B, Tr * Tc, &rhoG, &rhoL)) {
        return fail("coexistence");
    }

    // CudaInitialize() + SetFlowField (FluidCuda.cu:1260-1262)
    cpmp_ctx *ctx = nullptr;
    if (cpmp_create(&p, &ctx)) return fail("create");
    const int rc_init = splash ? cpmp_init_splash(ctx, radius, width, 32.0, rhoG, rhoL)
                               : cpmp_init_reference(ctx, radius, width, rhoG, rhoL, bp_from);
    if (rc_init) return fail("SetFlowField");

    char fname[512];
    std::snprintf(fname, sizeof fname, "%s/CP%s_MRT_Tau=%2.1f_FIELD%dx%dx%d_ContactAngle_R%3.1f_Tr%3.1f.txt", out.c_str(),
                  eos_name(eos), 1.0 / p.Ts[9], DX, DY, DZ, radius, Tr);
    int No = 0;
    double ca[8] = {0, 0, 0, 0, 0, 0, 0, 0};   // the last ContactAngle result, as the reference's globals
    auto t_last = std::chrono::steady_clock::now();
    for (double BasePt = bp_from; bp_step > 0 ? BasePt <= bp_to + 1e-12 : BasePt >= bp_to - 1e-12; BasePt += bp_step, ++No) {
        std::printf("Now the BasePt = %g\n", BasePt);
        if (cpmp_set_wall_mu(ctx, &BasePt, &BasePt, 1)) return fail("wall mu");
        for (int done = 0; done <= all_steps; done += show) {
            if (done > 0 && cpmp_step(ctx, show)) return fail("step");
            cpmp_stats s;
            if (cpmp_reduce(ctx, &s)) return fail("CalcMacroCPU");
            if (done > 0 && cpmp_contact_angle(ctx, DY / 2, rhoG, rhoL, ca)) return fail("ContactAngle");
            if (done == 0) {   // ShowData header (NowStep == 0)
                std::printf("Multiphase: CP%s    DX=%d    DY=%d    DZ=%d    Tau=%g    Tr=%g    DenG=%g    DenL=%g\n",
                            eos_name(eos), DX, DY, DZ, 1.0 / p.Ts[9], Tr, rhoG / Rc, rhoL / Rc);
                std::printf("Radius=%g    Width=%g    Ka=%g    K=%g    BasePt=%g    AllStep=%d\n", radius, width, p.Ka, p.K,
                            BasePt, all_steps);
                std::printf("NowStep    Mass    Den(DX/2 , 0 , 3)    Den(DX/2, DY/2/8, 3)    Den(DX/2, DY/2/4, 3)    MaxSpeed"
                            "    CaLeft    CaRight    CaCap    XLeft    XRight    StepTime\n");
            }
            double d0, d1, d2;
            if (cpmp_probe(ctx, CPMP_FIELD_RHO, DX / 2, 0, 3, &d0) || cpmp_probe(ctx, CPMP_FIELD_RHO, DX / 2, DY / 2 / 8, 3, &d1) ||
                cpmp_probe(ctx, CPMP_FIELD_RHO, DX / 2, DY / 2 / 4, 3, &d2))
                return fail("probe");
            const auto now = std::chrono::steady_clock::now();
            const double dt = std::chrono::duration<double>(now - t_last).count();
            t_last = now;
            std::printf("%9d    %.12f    %.16f    %.16f    %.16f    %.6f    %.6f    %.6f    %.6f    %.6f    %.6f    %.6f\n",
                        done, s.mass, d0 / Rc, d1 / Rc, d2 / Rc, s.max_speed, ca[0], ca[1], ca[2], ca[3], ca[4], dt);
            std::fflush(stdout);
        }
        // SaveContactAngle (header once per file, then one line per BasePt)
        FILE *fp = std::fopen(fname, "a");
        if (!fp) { std::fprintf(stderr, "cpmp_sweep: cannot write %s\n", fname); return 1; }
        if (No == 0) std::fprintf(fp, "Droprad  BasePt  NowStep  CaLeft  CaRight  CaCap  XLeft  XRight  \n");
        std::fprintf(fp, "%g  %g  %d  %g  %g  %g  %g  %g  \n", radius, BasePt, all_steps, ca[0], ca[1], ca[2], ca[3], ca[4]);
        std::fclose(fp);
    }
    cpmp_destroy(ctx);
    return 0;
}
