// plt_ref_driver.cpp -- writes a density field through the reference's own
// ordered_tec writer (ordered_tec.cpp, compiled from the reference checkout),
// the way SaveTask does (FluidCuda.cu:901-927).  Test infrastructure: the
// output is compared byte for byte with the library's cpmp_write_plt.
//
//   plt_ref rho.bin nx ny nz decimate outdir name
//   (decimate > 1 applies SaveTask's linear-index stride, FluidCuda.cu:466-472)
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "ordered_tec.h"

int main(int argc, char **argv) {
    if (argc != 8) { std::fprintf(stderr, "usage: plt_ref rho.bin nx ny nz decimate outdir name\n"); return 2; }
    const int nx = std::atoi(argv[2]), ny = std::atoi(argv[3]), nz = std::atoi(argv[4]), dr = std::atoi(argv[5]);
    const long n = (long)nx * ny * nz;
    std::vector<double> rho(n);
    FILE *fp = std::fopen(argv[1], "rb");
    if (!fp || std::fread(rho.data(), 8, n, fp) != (size_t)n) { std::fprintf(stderr, "bad input\n"); return 2; }
    std::fclose(fp);
    const int lx = nx / dr, ly = ny / dr, lz = nz / dr;
    const long m = (long)lx * ly * lz;
    std::vector<short> mi(m), mj(m), mk(m);
    std::vector<double> d(m);
    for (int i = 0; i < lx; ++i)
        for (int j = 0; j < ly; ++j)
            for (int k = 0; k < lz; ++k) {
                const long t = ((long)i * ly + j) * lz + k;
                mi[t] = i; mj[t] = j; mk[t] = k;
                d[t] = rho[dr == 1 ? t : (((long)i * ny + j) * nz + k) * dr];
            }
    using namespace ORDERED_TEC;
    TEC_FILE tec_file(argv[7], argv[6]);
    tec_file.Title = "Model3D";
    tec_file.Variables.push_back("i");
    tec_file.Variables.push_back("j");
    tec_file.Variables.push_back("k");
    tec_file.Variables.push_back("Density");
    tec_file.Zones.push_back(TEC_ZONE(argv[7]));
    tec_file.Zones[0].Max[0] = lz;
    tec_file.Zones[0].Max[1] = ly;
    tec_file.Zones[0].Max[2] = lx;
    tec_file.Zones[0].Data.push_back(TEC_DATA(mi.data()));
    tec_file.Zones[0].Data.push_back(TEC_DATA(mj.data()));
    tec_file.Zones[0].Data.push_back(TEC_DATA(mk.data()));
    tec_file.Zones[0].Data.push_back(TEC_DATA(d.data()));
    tec_file.write_plt(false);
    return 0;
}
