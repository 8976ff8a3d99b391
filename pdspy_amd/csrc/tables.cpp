// tables.cpp -- FP64 host construction of the thermal-MC tables.
//
// Frequency grid nu_i = c / lam_i with trapezoid weights dnu_i; Planck function
// from pdspy/misc/B_nu.py:4-6 and its temperature derivative (dB_nu.py:4-7);
// emission per gram q(T) = 4 pi sum_i kabs_i B_i(T) dnu_i on the log-spaced
// temperature grid of RADMC-3D's (ntemp, temp0, temp1) options
// (pdspy/radmc3d/write.py:47-52); Bjorkman & Wood re-emission CDFs
// P_i(T) ~ kabs_i dB_i/dT dnu_i; blackbody star spectra (stars.inp with -T,
// write.py:148-172) with L = 4 pi R^2 * pi sum B_i dnu_i.  Everything is built
// in log space so no row underflows at 0.01 K or overflows at 1e5 K.
#include <cmath>
#include <cstring>

#include <algorithm>

#include "../../include/pdspy_amd.h"
#include "tables.h"

namespace {

const double kH = 6.6260755e-27, kC = 2.99792458e10, kK = 1.380658e-16;
const double kPi = 3.14159265358979323846;

double ln_expm1(double x) { return x > 30.0 ? x + std::log1p(-std::exp(-x)) : std::log(std::expm1(x)); }

// ln B_nu(T)
double lnB(double nu, double T) {
    double x = kH * nu / (kK * T);
    return std::log(2.0 * kH / (kC * kC)) + 3.0 * std::log(nu) - ln_expm1(x);
}

// ln dB_nu/dT = ln(2 h^2 nu^4 / (c^2 k T^2)) - x - 2 ln(1 - e^-x)
double lndB(double nu, double T) {
    double x = kH * nu / (kK * T);
    return std::log(2.0 * kH * kH / (kC * kC * kK)) + 4.0 * std::log(nu) - 2.0 * std::log(T) - x -
           2.0 * std::log(-std::expm1(-x));
}

// normalised cumulative distribution from log weights; empty rows emit in the
// longest-wavelength bin
void make_cdf(const std::vector<double> &lw, float *cdf, unsigned char *guide, int ng) {
    int n = (int)lw.size();
    double mx = -INFINITY;
    for (double v : lw) mx = std::max(mx, v);
    std::vector<double> acc(n);
    if (mx == -INFINITY) {
        for (int i = 0; i < n; ++i) acc[i] = (i == n - 1) ? 1.0 : 0.0;
    } else {
        double s = 0;
        for (int i = 0; i < n; ++i) {
            s += std::exp(lw[i] - mx);
            acc[i] = s;
        }
        for (int i = 0; i < n; ++i) acc[i] /= s;
        acc[n - 1] = 1.0;
    }
    for (int i = 0; i < n; ++i) cdf[i] = (float)acc[i];
    cdf[n - 1] = 1.0f;
    // guide[k] = smallest i with cdf[i] >= k/ng (in the FP32 values the kernel sees)
    int i = 0;
    for (int k = 0; k < ng; ++k) {
        float thr = (float)k / (float)ng;
        while (i < n - 1 && cdf[i] < thr) ++i;
        guide[k] = (unsigned char)i;
    }
}

// modified random walk: y(xi) inverting the survival probability of a packet
// diffusing from the centre of a sphere, P(y) = 2 sum_n (-1)^(n+1)
// exp(-n^2 pi^2 y) (y = D t / R0^2).  The same for every model: built once per
// process (about 1e7 exponentials).
const std::vector<float> &mrw_escape_table() {
    static const std::vector<float> table = [] {
        auto surv = [](double y) {
            if (y < 0.1) {  // Jacobi dual form, fast for small y
                double s = 0;
                for (int mm = -40; mm <= 40; ++mm) s += std::exp(-(mm + 0.5) * (mm + 0.5) / y);
                return 1.0 - s / std::sqrt(kPi * y);
            }
            double s = 0;
            for (int n = 1; n <= 60; ++n)
                s += ((n & 1) ? 2.0 : -2.0) * std::exp(-(double)n * n * kPi * kPi * y);
            return s;
        };
        // tabulated on a uniform grid of logit(xi) = ln(xi / (1 - xi)), in which y
        // is close to linear at both ends (y ~ (ln 2 - ln xi)/pi^2 as xi -> 0 and
        // y ~ 1 / (4 ln(1/(1-xi))) as xi -> 1)
        std::vector<float> y(RT_MRW_NY, 0.f);
        for (int k = 0; k < RT_MRW_NY; ++k) {
            double L = -RT_MRW_LMAX + 2.0 * RT_MRW_LMAX * k / (RT_MRW_NY - 1);
            double xi = 1.0 / (1.0 + std::exp(-L));
            double lo = std::log(1e-9), hi = std::log(60.0);
            for (int it = 0; it < 90; ++it) {
                double mid = 0.5 * (lo + hi);
                if (surv(std::exp(mid)) > xi)
                    lo = mid;
                else
                    hi = mid;
            }
            y[k] = (float)std::exp(0.5 * (lo + hi));
        }
        return y;
    }();
    return table;
}

}  // namespace

int host_table_groups(const TableSpec &sp, HostTables &t, std::string &err) {
    const TableSpec *m = &sp;
    const int nl = m->nlam, ns = m->nspec, nT = m->ntemp;
    t.nu.resize(nl);
    t.dnu.resize(nl);
    for (int i = 0; i < nl; ++i) t.nu[i] = kC / (m->lam[i] * 1e-4);
    if (nl == 1) {
        t.dnu[0] = 1.0;
    } else {
        for (int i = 0; i < nl; ++i) {
            int a = i > 0 ? i - 1 : 0, b = i < nl - 1 ? i + 1 : nl - 1;
            t.dnu[i] = 0.5 * std::fabs(t.nu[b] - t.nu[a]);
        }
    }
    // species sharing an absorption opacity share a temperature (group)
    t.grp.assign(ns, 0);
    t.grep.clear();
    for (int s = 0; s < ns; ++s) {
        int found = -1;
        for (size_t q = 0; q < t.grep.size() && found < 0; ++q)
            if (!std::memcmp(&m->kabs[(size_t)t.grep[q] * nl], &m->kabs[(size_t)s * nl],
                             sizeof(double) * nl))
                found = (int)q;
        if (found < 0) {
            found = (int)t.grep.size();
            t.grep.push_back(s);
        }
        t.grp[s] = found;
    }
    t.ngroup = (int)t.grep.size();
    if (t.ngroup > RT_MAX_GROUP) {
        err = "more distinct absorption opacities than RT_MAX_GROUP";
        return RT_E_LIMIT;
    }
    const int G = t.ngroup;
    t.lnt0 = std::log(m->temp0);
    t.dlnt = (std::log(m->temp1) - t.lnt0) / (nT - 1);
    t.kmax.assign(G, 1.0);
    for (int q = 0; q < G; ++q) {
        const double *ka = &m->kabs[(size_t)t.grep[q] * nl];
        double km = 0;
        for (int i = 0; i < nl; ++i) km = std::max(km, ka[i]);
        t.kmax[q] = km > 0 ? km : 1.0;
    }
    return 0;
}

int build_host_tables(const TableSpec &sp, HostTables &t, std::string &err) {
    const TableSpec *m = &sp;
    const int nl = m->nlam, ns = m->nspec, nT = m->ntemp;
    int rc = host_table_groups(sp, t, err);
    if (rc) return rc;
    const int G = t.ngroup;
    t.lnq.assign((size_t)G * nT, 0.0);
    t.cdf.assign((size_t)G * nT * nl, 0.f);
    t.guide.assign((size_t)G * nT * RT_NGUIDE, 0);
    t.kross.assign((size_t)ns * nT, 0.f);
    t.kplanck.assign((size_t)G * nT, 0.f);
    // per-temperature quantities shared by every group and species, computed
    // once: ln B_nu(T), ln dB_nu/dT, and the Planck and Rosseland weights
    // exp(ln dnu + ln B - max) and exp(ln dnu + ln dB/dT - max) with their sums
    std::vector<double> ldnu(nl);
    for (int i = 0; i < nl; ++i) ldnu[i] = std::log(t.dnu[i]);
    const size_t nTl = (size_t)nT * nl;
    std::vector<double> LB(nTl), LD(nTl), Bw(nTl), Dw(nTl), wbT(nT), wdT(nT);
    for (int it = 0; it < nT; ++it) {
        const double T = std::exp(t.lnt0 + it * t.dlnt);
        double *lb = &LB[(size_t)it * nl], *ld = &LD[(size_t)it * nl];
        double *bw = &Bw[(size_t)it * nl], *dw = &Dw[(size_t)it * nl];
        for (int i = 0; i < nl; ++i) {
            lb[i] = lnB(t.nu[i], T);
            ld[i] = lndB(t.nu[i], T);
        }
        double mb = -INFINITY, md = -INFINITY;
        for (int i = 0; i < nl; ++i) {
            mb = std::max(mb, ldnu[i] + lb[i]);
            md = std::max(md, ldnu[i] + ld[i]);
        }
        double wb = 0, wd = 0;
        for (int i = 0; i < nl; ++i) {
            bw[i] = std::exp(ldnu[i] + lb[i] - mb);
            wb += bw[i];
            dw[i] = std::exp(ldnu[i] + ld[i] - md);
            wd += dw[i];
        }
        wbT[it] = wb;
        wdT[it] = wd;
    }
    std::vector<double> lw(nl), lwb(nl), lkd(nl);
    for (int q = 0; q < G; ++q) {
        const int s = t.grep[q];
        const double *ka = &m->kabs[(size_t)s * nl];
        for (int i = 0; i < nl; ++i) lkd[i] = ka[i] > 0 ? std::log(ka[i] * t.dnu[i]) : -INFINITY;
        for (int it = 0; it < nT; ++it) {
            const double *lb = &LB[(size_t)it * nl], *ld = &LD[(size_t)it * nl];
            double mx = -INFINITY;
            for (int i = 0; i < nl; ++i) {
                lwb[i] = ka[i] > 0 ? lkd[i] + lb[i] : -INFINITY;
                mx = std::max(mx, lwb[i]);
            }
            double s1 = 0;
            for (int i = 0; i < nl; ++i) s1 += std::exp(lwb[i] - mx);
            t.lnq[(size_t)q * nT + it] = std::log(4.0 * kPi) + mx + std::log(s1);
            for (int i = 0; i < nl; ++i) lw[i] = ka[i] > 0 ? lkd[i] + ld[i] : -INFINITY;
            make_cdf(lw, &t.cdf[((size_t)q * nT + it) * nl],
                     &t.guide[((size_t)q * nT + it) * RT_NGUIDE], RT_NGUIDE);
            // Planck-mean absorption at T
            const double *bw = &Bw[(size_t)it * nl];
            double wkb = 0;
            for (int i = 0; i < nl; ++i) wkb += bw[i] * ka[i];
            t.kplanck[(size_t)q * nT + it] = (float)(wkb / wbT[it]);
        }
    }
    // Rosseland-mean extinction per species (kabs + ksca), weights dB/dT dnu
    for (int s = 0; s < ns; ++s) {
        const double *ka = &m->kabs[(size_t)s * nl];
        const double *ks = &m->ksca[(size_t)s * nl];
        for (int it = 0; it < nT; ++it) {
            const double *dw = &Dw[(size_t)it * nl];
            double wdk = 0;
            for (int i = 0; i < nl; ++i) {
                double ke = ka[i] + ks[i];
                if (ke > 0) wdk += dw[i] / ke;
            }
            t.kross[(size_t)s * nT + it] = (float)(wdk > 0 ? wdT[it] / wdk : 0.0);
        }
    }
    t.mrw_y = mrw_escape_table();
    // stars
    t.scdf.assign((size_t)std::max(1, m->nstar) * nl, 0.f);
    t.sguide.assign((size_t)std::max(1, m->nstar) * RT_NGUIDE, 0);
    t.spick.assign(std::max(1, m->nstar), 1.f);
    double ltot = 0;
    std::vector<double> lum(m->nstar);
    for (int k = 0; k < m->nstar; ++k) {
        double R = m->stars[5 * k], Ts = m->stars[5 * k + 1];
        double mx = -INFINITY;
        for (int i = 0; i < nl; ++i) {
            lw[i] = std::log(t.dnu[i]) + lnB(t.nu[i], Ts);
            mx = std::max(mx, lw[i]);
        }
        double s1 = 0;
        for (int i = 0; i < nl; ++i) s1 += std::exp(lw[i] - mx);
        lum[k] = 4.0 * kPi * R * R * kPi * std::exp(mx) * s1;
        ltot += lum[k];
        make_cdf(lw, &t.scdf[(size_t)k * nl], &t.sguide[(size_t)k * RT_NGUIDE], RT_NGUIDE);
    }
    double acc = 0;
    for (int k = 0; k < m->nstar; ++k) {
        acc += lum[k];
        t.spick[k] = (float)(acc / ltot);
    }
    if (m->nstar > 0) t.spick[m->nstar - 1] = 1.f;
    t.lum = ltot;
    return 0;
}
