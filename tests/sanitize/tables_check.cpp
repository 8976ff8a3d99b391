// Drives the library's host table builder (pdspy_amd/csrc/tables.cpp,
// build_tables) on the opacities of the benchmark models and on random
// power-law and tabulated dust, under AddressSanitizer and UBSan, and checks
// the invariants the thermal kernel relies on: CDF rows non-decreasing and
// ending at 1, guide entries pointing at the first bin whose CDF reaches the
// entry's threshold, finite ln q rising with T, positive mean opacities,
// star CDFs and pick fractions.  Prints "tables_check ok" and exits 0, or
// names the first violation and exits 1.
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <random>
#include <string>
#include <vector>

#include "../../pdspy_amd/csrc/tables.h"

struct Model {  // the inputs build_host_tables reads
    int nlam = 0, nspec = 0, ntemp = 1000, nstar = 0;
    double temp0 = 0.01, temp1 = 1e5;
    std::vector<double> lam, kabs, ksca, stars;
    HostTables tab;
};

static int fail(const char *what, int a, int b) {
    std::printf("tables_check FAILED: %s (%d, %d)\n", what, a, b);
    return 1;
}

static int check(Model &m) {
    std::string err;
    TableSpec sp{m.nlam, m.nspec, m.ntemp, m.nstar, m.temp0, m.temp1,
                 m.lam.data(), m.kabs.data(), m.ksca.data(), m.stars.data()};
    int rc = build_host_tables(sp, m.tab, err);
    if (rc) return fail(("build_tables: " + err).c_str(), rc, 0);
    const HostTables &t = m.tab;
    const int nl = m.nlam, nT = m.ntemp, G = t.ngroup;
    for (int q = 0; q < G; ++q) {
        for (int it = 0; it < nT; ++it) {
            const float *c = &t.cdf[((size_t)q * nT + it) * nl];
            const unsigned char *gd = &t.guide[((size_t)q * nT + it) * RT_NGUIDE];
            for (int i = 1; i < nl; ++i)
                if (!(c[i] >= c[i - 1])) return fail("cdf decreasing", q, it);
            if (c[nl - 1] != 1.0f) return fail("cdf end != 1", q, it);
            for (int k = 0; k < RT_NGUIDE; ++k) {
                float thr = (float)k / (float)RT_NGUIDE;
                int i = gd[k];
                if (i >= nl) return fail("guide out of range", q, k);
                if (c[i] < thr && i != nl - 1) return fail("guide below threshold", q, k);
                if (i > 0 && c[i - 1] >= thr) return fail("guide not the first bin", q, k);
            }
            double lq = t.lnq[(size_t)q * nT + it];
            if (!std::isfinite(lq)) return fail("ln q not finite", q, it);
            if (it > 0 && !(lq > t.lnq[(size_t)q * nT + it - 1])) return fail("ln q not rising", q, it);
            if (!(t.kplanck[(size_t)q * nT + it] > 0)) return fail("Planck mean <= 0", q, it);
        }
    }
    for (int s = 0; s < m.nspec; ++s)
        for (int it = 0; it < nT; ++it)
            if (!(t.kross[(size_t)s * nT + it] > 0)) return fail("Rosseland mean <= 0", s, it);
    for (int k = 0; k < m.nstar; ++k) {
        const float *c = &t.scdf[(size_t)k * nl];
        for (int i = 1; i < nl; ++i)
            if (!(c[i] >= c[i - 1])) return fail("star cdf decreasing", k, i);
        if (!(t.spick[k] > 0 && t.spick[k] <= 1)) return fail("star pick", k, 0);
    }
    if (!(t.lum > 0)) return fail("luminosity", 0, 0);
    if ((int)t.mrw_y.size() != RT_MRW_NY) return fail("escape table size", (int)t.mrw_y.size(), 0);
    return 0;
}

int main() {
    std::mt19937_64 rng(7);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    int cases = 0;
    for (int trial = 0; trial < 24; ++trial) {
        Model m;
        m.nlam = trial % 3 == 0 ? 100 : 20 + (int)(U(rng) * 200);
        m.nspec = 1 + trial % 8;
        m.ntemp = trial % 2 ? 1000 : 64;
        m.temp0 = 0.01;
        m.temp1 = 1e5;
        m.lam.resize(m.nlam);
        for (int i = 0; i < m.nlam; ++i)
            m.lam[i] = std::pow(10.0, -1.0 + 5.0 * i / (m.nlam - 1));
        m.kabs.resize((size_t)m.nspec * m.nlam);
        m.ksca.resize((size_t)m.nspec * m.nlam);
        const int ngrp = 1 + trial % RT_MAX_GROUP;
        for (int s = 0; s < m.nspec; ++s) {
            const int q = s % ngrp;  // species s has the opacity of group q
            const double beta = 0.3 + 0.4 * q, k0 = 1e3 * (1.0 + q);
            for (int i = 0; i < m.nlam; ++i) {
                double lam = m.lam[i];
                double ka = lam > 1.0 ? k0 * std::pow(lam, -beta) : k0;
                if (trial % 5 == 4 && i % 7 == 3) ka = 0.0;  // bins without absorption
                m.kabs[(size_t)s * m.nlam + i] = ka;
                m.ksca[(size_t)s * m.nlam + i] = ka * (0.2 + 0.8 * U(rng));
            }
        }
        m.nstar = 1 + trial % 3;
        m.stars.resize(5 * (size_t)m.nstar);
        for (int k = 0; k < m.nstar; ++k) {
            m.stars[5 * k] = 6.96e10 * (0.5 + 3 * U(rng));
            m.stars[5 * k + 1] = 3000.0 + 7000.0 * U(rng);
            m.stars[5 * k + 2] = m.stars[5 * k + 3] = m.stars[5 * k + 4] = 0.0;
        }
        if (check(m)) return 1;
        ++cases;
    }
    std::printf("tables_check ok (%d models)\n", cases);
    return 0;
}
