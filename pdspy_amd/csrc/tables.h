// tables.h -- the thermal Monte Carlo's host-built tables (tables.cpp): plain
// C++ with no HIP dependency, so the builder also compiles into host-only
// checks.
#pragma once
#include <string>
#include <vector>

#define RT_MAX_GROUP 4   // distinct absorption opacities per model
#define RT_NGUIDE 128    // guide-table entries per CDF row
#define RT_MRW_NY 2048   // entries of the diffusion escape-law inverse table
#define RT_MRW_LMAX 18.0f  // its logit range: ln(xi/(1-xi)) in [-LMAX, LMAX]

struct HostTables {
    int ngroup = 0;
    std::vector<int> grp;    // species -> group
    std::vector<int> grep;   // group -> representative species
    std::vector<double> nu, dnu;
    std::vector<double> kmax;        // per group
    std::vector<double> lnq;         // [G][ntemp]
    std::vector<float> cdf;          // [G][ntemp][nlam]
    std::vector<unsigned char> guide;  // [G][ntemp][NGUIDE]
    std::vector<float> scdf;         // [nstar][nlam]
    std::vector<unsigned char> sguide; // [nstar][NGUIDE]
    std::vector<float> spick;        // [nstar]
    std::vector<float> kross;        // [S][ntemp] Rosseland-mean extinction (MRW)
    std::vector<float> mrw_y;        // [RT_MRW_NY] inverse of the escape law
    std::vector<float> kplanck;      // [G][ntemp] Planck-mean absorption (MRW)
    double lum = 0;
    double lnt0 = 0, dlnt = 0;
};

// what the tables are built from: nlam wavelengths (um), nspec species'
// absorption and scattering opacities [nspec][nlam] (cm^2/g), the temperature
// grid exp(linspace(ln temp0, ln temp1, ntemp)) and nstar blackbody stars
// [nstar][5] (radius cm, T K, x, y, z)
struct TableSpec {
    int nlam, nspec, ntemp, nstar;
    double temp0, temp1;
    const double *lam, *kabs, *ksca, *stars;
};

int build_host_tables(const TableSpec &s, HostTables &t, std::string &err);
// only the parts the device builder needs from the host: frequency grid,
// absorption groups, their largest kabs and the temperature grid
int host_table_groups(const TableSpec &s, HostTables &t, std::string &err);
