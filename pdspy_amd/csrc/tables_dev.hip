// tables_dev.hip -- the thermal Monte Carlo's opacity-dependent tables built on
// the device, for opacity updates inside a sampler step (pdspy's fits draw new
// dust per walker, run_disk_model.py:79-85): per absorption group and
// temperature row the emission per gram ln q(T) = ln(4 pi sum kabs B dnu), the
// Bjorkman & Wood re-emission CDF of kabs dB/dT dnu with its guide table and
// the Planck-mean absorption; per species the Rosseland-mean extinction.  One
// lane per row, in FP64, with the operations and summation order of the host
// builder (tables.cpp), so the FP32 tables agree with the host's to rounding.
#include <cmath>

#include "rt_internal.h"

namespace {

__device__ const double kHd = 6.6260755e-27, kCd = 2.99792458e10, kKd = 1.380658e-16;

__device__ __forceinline__ double ln_expm1_d(double x) {
    return x > 30.0 ? x + log1p(-exp(-x)) : log(expm1(x));
}

// ln B_nu(T) (pdspy/misc/B_nu.py:4-6)
__device__ __forceinline__ double lnB_d(double nu, double T) {
    const double x = kHd * nu / (kKd * T);
    return log(2.0 * kHd / (kCd * kCd)) + 3.0 * log(nu) - ln_expm1_d(x);
}

// ln |dB_nu/dT| (pdspy/misc/dB_nu.py:4-7)
__device__ __forceinline__ double lndB_d(double nu, double T) {
    const double x = kHd * nu / (kKd * T);
    return log(2.0 * kHd * kHd / (kCd * kCd * kKd)) + 4.0 * log(nu) - 2.0 * log(T) - x -
           2.0 * log(-expm1(-x));
}

}  // namespace

// in: [nu | dnu | ln dnu] (3 nl), [G][nl] the groups' kabs, [S][nl] kabs + ksca
__global__ void __launch_bounds__(64) k_tab_rows(int G, int S, int nT, int nl, double lnt0,
                                                 double dlnt, const double *in, double *lnq,
                                                 float *cdf, unsigned char *guide,
                                                 float *kplanck, float *kross) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const double *nu = in, *dnu = in + nl, *ldnu = in + 2 * nl;
    const double *gk = in + 3 * nl, *ke = gk + (size_t)G * nl;
    if (t < G * nT) {
        const int q = t / nT, it = t - q * nT;
        const double T = exp(lnt0 + it * dlnt);
        const double *ka = gk + (size_t)q * nl;
        // Planck weights and mean
        double mb = -INFINITY;
        for (int i = 0; i < nl; ++i) mb = fmax(mb, ldnu[i] + lnB_d(nu[i], T));
        double wb = 0, wkb = 0;
        for (int i = 0; i < nl; ++i) {
            const double bw = exp(ldnu[i] + lnB_d(nu[i], T) - mb);
            wb += bw;
            wkb += bw * ka[i];
        }
        kplanck[(size_t)q * nT + it] = (float)(wkb / wb);
        // emission per gram
        double mx = -INFINITY;
        for (int i = 0; i < nl; ++i)
            if (ka[i] > 0) mx = fmax(mx, log(ka[i] * dnu[i]) + lnB_d(nu[i], T));
        double s1 = 0;
        for (int i = 0; i < nl; ++i)
            s1 += ka[i] > 0 ? exp(log(ka[i] * dnu[i]) + lnB_d(nu[i], T) - mx) : 0.0;
        lnq[(size_t)q * nT + it] = log(4.0 * 3.14159265358979323846) + mx + log(s1);
        // re-emission CDF and its guide table
        float *c = cdf + ((size_t)q * nT + it) * nl;
        double md = -INFINITY;
        for (int i = 0; i < nl; ++i)
            if (ka[i] > 0) md = fmax(md, log(ka[i] * dnu[i]) + lndB_d(nu[i], T));
        if (md == -INFINITY) {
            for (int i = 0; i < nl; ++i) c[i] = (i == nl - 1) ? 1.f : 0.f;
        } else {
            double sd = 0;
            for (int i = 0; i < nl; ++i)
                sd += ka[i] > 0 ? exp(log(ka[i] * dnu[i]) + lndB_d(nu[i], T) - md) : 0.0;
            double acc = 0;
            for (int i = 0; i < nl; ++i) {
                acc += ka[i] > 0 ? exp(log(ka[i] * dnu[i]) + lndB_d(nu[i], T) - md) : 0.0;
                c[i] = (float)(acc / sd);
            }
        }
        c[nl - 1] = 1.f;
        unsigned char *gd = guide + ((size_t)q * nT + it) * RT_NGUIDE;
        int i = 0;
        for (int k = 0; k < RT_NGUIDE; ++k) {
            const float thr = (float)k / (float)RT_NGUIDE;
            while (i < nl - 1 && c[i] < thr) ++i;
            gd[k] = (unsigned char)i;
        }
        return;
    }
    const int r = t - G * nT;
    if (r >= S * nT) return;
    const int s = r / nT, it = r - s * nT;
    const double T = exp(lnt0 + it * dlnt);
    const double *kes = ke + (size_t)s * nl;
    double md = -INFINITY;
    for (int i = 0; i < nl; ++i) md = fmax(md, ldnu[i] + lndB_d(nu[i], T));
    double wd = 0, wdk = 0;
    for (int i = 0; i < nl; ++i) {
        const double dw = exp(ldnu[i] + lndB_d(nu[i], T) - md);
        wd += dw;
        if (kes[i] > 0) wdk += dw / kes[i];
    }
    kross[(size_t)s * nT + it] = (float)(wdk > 0 ? wd / wdk : 0.0);
}

// Uploads the inputs and enqueues the table build on the context's stream;
// m->tab holds the new groups (host_table_groups) and m->kabs / ksca the new
// opacities.
int tables_device_build(rt_model *m) {
    rt_ctx *ctx = m->ctx;
    const HostTables &t = m->tab;
    const int nl = m->nlam, S = m->nspec, G = t.ngroup, nT = m->ntemp;
    std::vector<double> in((size_t)(3 + G + S) * nl);
    for (int i = 0; i < nl; ++i) {
        in[i] = t.nu[i];
        in[nl + i] = t.dnu[i];
        in[2 * nl + i] = std::log(t.dnu[i]);
    }
    for (int q = 0; q < G; ++q)
        for (int i = 0; i < nl; ++i)
            in[(size_t)(3 + q) * nl + i] = m->kabs[(size_t)t.grep[q] * nl + i];
    for (int s = 0; s < S; ++s)
        for (int i = 0; i < nl; ++i)
            in[(size_t)(3 + G + s) * nl + i] = m->kabs[(size_t)s * nl + i] + m->ksca[(size_t)s * nl + i];
    if (!m->d_tabin || m->tabin_cap < in.size()) {
        if (m->d_tabin) (void)hipFree(m->d_tabin);
        m->d_tabin = nullptr;
        HIPCHK(ctx, hipMalloc(&m->d_tabin, in.size() * sizeof(double)));
        m->tabin_cap = in.size();
    }
    HIPCHK(ctx, hipMemcpy(m->d_tabin, in.data(), in.size() * sizeof(double), hipMemcpyHostToDevice));
    const int rows = (G + S) * nT;
    hipLaunchKernelGGL(k_tab_rows, dim3((rows + 63) / 64), dim3(64), 0, ctx->stream, G, S, nT, nl,
                       t.lnt0, t.dlnt, m->d_tabin, m->d_lnq, m->d_cdf, m->d_guide, m->d_kplanck,
                       m->d_kross);
    HIPCHK(ctx, hipGetLastError());
    return RT_OK;
}
