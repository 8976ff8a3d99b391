// tables_dev.hip -- the thermal Monte Carlo's opacity-dependent tables built on
// the device, for opacity updates inside a sampler step (pdspy's fits draw new
// dust per walker, run_disk_model.py:79-85): per absorption group and
// temperature row the emission per gram ln q(T) = ln(4 pi sum kabs B dnu), the
// Bjorkman & Wood re-emission CDF of kabs dB/dT dnu with its guide table and
// the Planck-mean absorption; per species the Rosseland-mean extinction.  As
// the host builder (tables.cpp) does, the quantities that depend only on the
// wavelength and temperature grids are made once per model and kept; an
// update then computes each row in FP64 with the host's operations and
// summation order, so the tables are the host's (to the last bit where the
// device's exp and log round as the host's do).
#include <algorithm>
#include <cmath>
#include <vector>

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

// The quantities that depend only on the wavelength and temperature grids, as
// the host builder computes them once per model: per temperature row it and
// wavelength i, LB = ln B_nu(T), LD = ln dB_nu/dT, the Planck and Rosseland
// weights Bw = exp(ln dnu + LB - max), Dw = exp(ln dnu + LD - max) and their
// sums wbT, wdT.  tT = [LB | LD | Bw | Dw] (nT nl each), then [wbT | wdT] (nT).
__global__ void __launch_bounds__(256) k_tab_T1(int nT, int nl, double lnt0, double dlnt,
                                                const double *in, double *tT) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)nT * nl) return;
    const int it = (int)(t / nl), i = (int)(t - (size_t)it * nl);
    const double T = exp(lnt0 + it * dlnt);
    const size_t n = (size_t)nT * nl;
    tT[t] = lnB_d(in[i], T);
    tT[n + t] = lndB_d(in[i], T);
}

__global__ void __launch_bounds__(64) k_tab_T2(int nT, int nl, const double *in, double *tT) {
    const int it = blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= nT) return;
    const size_t n = (size_t)nT * nl;
    const double *ldnu = in + 2 * nl;
    const double *lb = tT + (size_t)it * nl, *ld = lb + n;
    double *bw = tT + 2 * n + (size_t)it * nl, *dw = bw + n;
    double mb = -INFINITY, md = -INFINITY;
    for (int i = 0; i < nl; ++i) {
        mb = fmax(mb, ldnu[i] + lb[i]);
        md = fmax(md, ldnu[i] + ld[i]);
    }
    double wb = 0, wd = 0;
    for (int i = 0; i < nl; ++i) {
        bw[i] = exp(ldnu[i] + lb[i] - mb);
        wb += bw[i];
        dw[i] = exp(ldnu[i] + ld[i] - md);
        wd += dw[i];
    }
    tT[4 * n + it] = wb;
    tT[4 * n + nT + it] = wd;
}

#define TAB_ROWS 64           // table rows (consecutive temperatures) per workgroup
#define TAB_CH 32             // wavelengths per chunk
#define TAB_LD (TAB_CH + 1)   // LDS row stride (doubles)

// One workgroup of 256 per TAB_ROWS consecutive temperature rows of one group
// (emission, re-emission CDF and guide, Planck mean) or one species
// (Rosseland mean).  Per wavelength chunk all threads compute the terms into
// LDS; lane r of one wave then adds row r's terms in wavelength order, so every
// sum is the host builder's sequential sum and the tables equal its values.
// in: [nu | dnu | ln dnu] (3 nl), [G][nl] the groups' kabs, [S][nl] kabs +
// ksca, [G][nl] ln(kabs dnu) of the groups (-inf where kabs = 0).
__global__ void __launch_bounds__(256) k_tab_rows(int G, int S, int nT, int nl, const double *in,
                                                  const double *tT, double *lnq, float *cdf,
                                                  unsigned char *guide, float *kplanck,
                                                  float *kross) {
    __shared__ double b0[TAB_ROWS * TAB_LD], b1[TAB_ROWS * TAB_LD], b2[TAB_ROWS * TAB_LD];
    __shared__ float cb[TAB_ROWS * TAB_LD];
    __shared__ double smx[TAB_ROWS], smd[TAB_ROWS], ssd[TAB_ROWS];
    const int nbt = (nT + TAB_ROWS - 1) / TAB_ROWS;
    const int kind = blockIdx.x / nbt, it0 = (blockIdx.x - kind * nbt) * TAB_ROWS;
    const int nrow = min(TAB_ROWS, nT - it0);
    const int tid = threadIdx.x, w = tid >> 6, r = tid & 63;
    const bool lane_row = r < nrow;
    const size_t n = (size_t)nT * nl;
    const double *gk = in + 3 * nl, *ke = gk + (size_t)G * nl, *lk = ke + (size_t)S * nl;
    const int nch = (nl + TAB_CH - 1) / TAB_CH;
    if (kind >= G) {
        // Rosseland-mean extinction of species s: wdT / sum dw / (kabs + ksca)
        const int sp = kind - G;
        const double *kes = ke + (size_t)sp * nl, *dw = tT + 3 * n;
        double wdk = 0;
        for (int ch = 0; ch < nch; ++ch) {
            const int i0 = ch * TAB_CH, ni = min(TAB_CH, nl - i0);
            for (int e = tid; e < TAB_ROWS * TAB_CH; e += 256) {
                const int rr = e / TAB_CH, ii = e - rr * TAB_CH, i = i0 + ii;
                if (rr < nrow && ii < ni)
                    b0[rr * TAB_LD + ii] = kes[i] > 0 ? dw[(size_t)(it0 + rr) * nl + i] / kes[i] : 0.0;
            }
            __syncthreads();
            if (w == 0 && lane_row)
                for (int ii = 0; ii < ni; ++ii) wdk += b0[r * TAB_LD + ii];
            __syncthreads();
        }
        if (w == 0 && lane_row)
            kross[(size_t)sp * nT + it0 + r] = (float)(wdk > 0 ? tT[4 * n + nT + it0 + r] / wdk : 0.0);
        return;
    }
    const int q = kind;
    const double *ka = gk + (size_t)q * nl, *lkd = lk + (size_t)q * nl;
    const double *lb = tT, *ld = tT + n, *bw = tT + 2 * n;
    // pass 1: the largest exponent of the emission (wave 0) and re-emission
    // (wave 1) terms per row
    double mx = -INFINITY;
    for (int ch = 0; ch < nch; ++ch) {
        const int i0 = ch * TAB_CH, ni = min(TAB_CH, nl - i0);
        for (int e = tid; e < TAB_ROWS * TAB_CH; e += 256) {
            const int rr = e / TAB_CH, ii = e - rr * TAB_CH, i = i0 + ii;
            if (rr < nrow && ii < ni) {
                const size_t o = (size_t)(it0 + rr) * nl + i;
                const bool pos = ka[i] > 0;
                b0[rr * TAB_LD + ii] = pos ? lkd[i] + lb[o] : -INFINITY;
                b1[rr * TAB_LD + ii] = pos ? lkd[i] + ld[o] : -INFINITY;
            }
        }
        __syncthreads();
        if (w < 2 && lane_row) {
            const double *bb = w == 0 ? b0 : b1;
            for (int ii = 0; ii < ni; ++ii) mx = fmax(mx, bb[r * TAB_LD + ii]);
        }
        __syncthreads();
    }
    if (lane_row) {
        if (w == 0) smx[r] = mx;
        if (w == 1) smd[r] = mx;
    }
    __syncthreads();
    // pass 2: the sums of the emission (wave 0), re-emission (wave 1) and
    // Planck-mean (wave 2) terms
    double sum = 0;
    for (int ch = 0; ch < nch; ++ch) {
        const int i0 = ch * TAB_CH, ni = min(TAB_CH, nl - i0);
        for (int e = tid; e < TAB_ROWS * TAB_CH; e += 256) {
            const int rr = e / TAB_CH, ii = e - rr * TAB_CH, i = i0 + ii;
            if (rr < nrow && ii < ni) {
                const size_t o = (size_t)(it0 + rr) * nl + i;
                const bool pos = ka[i] > 0;
                b0[rr * TAB_LD + ii] = exp((pos ? lkd[i] + lb[o] : -INFINITY) - smx[rr]);
                b1[rr * TAB_LD + ii] = exp((pos ? lkd[i] + ld[o] : -INFINITY) - smd[rr]);
                b2[rr * TAB_LD + ii] = bw[o] * ka[i];
            }
        }
        __syncthreads();
        if (w < 3 && lane_row) {
            const double *bb = w == 0 ? b0 : (w == 1 ? b1 : b2);
            for (int ii = 0; ii < ni; ++ii) sum += bb[r * TAB_LD + ii];
        }
        __syncthreads();
    }
    if (lane_row) {
        const int it = it0 + r;
        if (w == 0) lnq[(size_t)q * nT + it] = log(4.0 * 3.14159265358979323846) + smx[r] + log(sum);
        if (w == 1) ssd[r] = sum;
        if (w == 2) kplanck[(size_t)q * nT + it] = (float)(sum / tT[4 * n + it]);
    }
    __syncthreads();
    // pass 3: the re-emission CDF (prefix sums over the total, wave 1), its
    // guide table (smallest i with cdf[i] >= k / RT_NGUIDE), coalesced stores
    double acc = 0;
    int kg = 0;
    unsigned char *gd = guide + ((size_t)q * nT + it0 + r) * RT_NGUIDE;
    for (int ch = 0; ch < nch; ++ch) {
        const int i0 = ch * TAB_CH, ni = min(TAB_CH, nl - i0);
        for (int e = tid; e < TAB_ROWS * TAB_CH; e += 256) {
            const int rr = e / TAB_CH, ii = e - rr * TAB_CH, i = i0 + ii;
            if (rr < nrow && ii < ni) {
                const size_t o = (size_t)(it0 + rr) * nl + i;
                b1[rr * TAB_LD + ii] = exp((ka[i] > 0 ? lkd[i] + ld[o] : -INFINITY) - smd[rr]);
            }
        }
        __syncthreads();
        if (w == 1 && lane_row) {
            const bool flat = smd[r] == -INFINITY;  // no re-emission: all at the last bin
            for (int ii = 0; ii < ni; ++ii) {
                const int i = i0 + ii;
                acc += b1[r * TAB_LD + ii];
                float cf = flat ? 0.f : (float)(acc / ssd[r]);
                if (i == nl - 1) cf = 1.f;
                cb[r * TAB_LD + ii] = cf;
                while (kg < RT_NGUIDE && cf >= (float)kg / (float)RT_NGUIDE) gd[kg++] = (unsigned char)i;
            }
        }
        __syncthreads();
        for (int e = tid; e < TAB_ROWS * TAB_CH; e += 256) {
            const int rr = e / TAB_CH, ii = e - rr * TAB_CH;
            if (rr < nrow && ii < ni)
                cdf[((size_t)q * nT + it0 + rr) * nl + i0 + ii] = cb[rr * TAB_LD + ii];
        }
        __syncthreads();
    }
}

struct GroupReps {
    int rep[RT_MAX_GROUP];  // the species whose kabs stands for each group
};

// The per-species arrays of an opacity update, from the staged FP64 values
// op = [kabs S nl | ksca S nl | g S nl | kmax G], with the host's formulas
// (the FP32 values are the same roundings, econv the same FP64 expression):
// kabs, ksca, g as FP32; the groups' kabs; 2^shift / kmax; the per-cell
// tally-to-energy factors kmax dr / (2^shift vol); and rows 3.. of the table
// build's inputs (the groups' kabs, kabs + ksca per species, ln(kabs dnu) of
// the groups).
__global__ void __launch_bounds__(256) k_opac_prep(int S, int G, int nl, int nr, int nc, int has_g,
                                                   GroupReps gr, const double *op,
                                                   const double *dr, const double *vol, float *ka,
                                                   float *ks, float *gg, float *gka, float *ginvk,
                                                   double *econv, double *tabin) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nk = (size_t)S * nl;
    const double *kabs = op, *ksca = op + nk, *g = op + 2 * nk, *kmax = op + 3 * nk;
    const double unit = ldexp(1.0, RT_TALLY_SHIFT);
    if (t < nk) {
        ka[t] = (float)kabs[t];
        ks[t] = (float)ksca[t];
        gg[t] = has_g ? (float)g[t] : 0.f;
        tabin[(size_t)(3 + G) * nl + t] = kabs[t] + ksca[t];
    }
    if (t < (size_t)G * nl) {
        const int q = (int)(t / nl), i = (int)(t - (size_t)q * nl);
        const double v = kabs[(size_t)gr.rep[q] * nl + i];
        gka[t] = (float)v;
        tabin[3 * (size_t)nl + t] = v;
        tabin[(size_t)(3 + G + S) * nl + t] = v > 0 ? log(v * tabin[nl + i]) : -INFINITY;
    }
    if (t < (size_t)G) ginvk[t] = (float)(unit / kmax[t]);
    if (t < (size_t)nc * G) {
        const int c = (int)(t / G), q = (int)(t - (size_t)c * G);
        econv[t] = kmax[q] * dr[c % nr] / (unit * vol[c]);
    }
}

// Enqueues an opacity update on the context's stream: m->tab holds the new
// groups and kmax (host_table_groups), m->kabs / ksca / gsca the new
// opacities.  One copy from a pinned staging buffer, then k_opac_prep and the
// table build; nothing waits for the device, so work already on the stream
// finishes with the old opacities and work enqueued after sees the new ones.
int opacity_device_update(rt_model *m) {
    rt_ctx *ctx = m->ctx;
    const HostTables &t = m->tab;
    const int nl = m->nlam, S = m->nspec, G = t.ngroup, nT = m->ntemp, nc = m->ncell;
    const size_t nk = (size_t)S * nl, nop = 3 * nk + RT_MAX_GROUP;
    if (G > RT_MAX_GROUP) return set_err(ctx, RT_E_LIMIT, "more than %d absorption groups", RT_MAX_GROUP);
    HIPCHK(ctx, hipSetDevice(ctx->device));
    // first update: buffers, and the inputs that stay (frequencies, cell widths
    // and volumes, the per-temperature quantities); each made once, so a call
    // that failed half way is completed by the next one
    if (!m->h_opin)
        HIPCHK(ctx, hipHostMalloc((void **)&m->h_opin, nop * sizeof(double), hipHostMallocDefault));
    if (!m->d_opin) HIPCHK(ctx, hipMalloc((void **)&m->d_opin, nop * sizeof(double)));
    if (!m->e_opin) HIPCHK(ctx, hipEventCreateWithFlags(&m->e_opin, hipEventDisableTiming));
    if (!m->d_tabin) {
        std::vector<double> in((size_t)(3 + 2 * RT_MAX_GROUP + S) * nl, 0.0);
        for (int i = 0; i < nl; ++i) {
            in[i] = t.nu[i];
            in[nl + i] = t.dnu[i];
            in[2 * nl + i] = std::log(t.dnu[i]);
        }
        double *d = nullptr;
        HIPCHK(ctx, hipMalloc((void **)&d, in.size() * sizeof(double)));
        HIPCHK(ctx, hipMemcpy(d, in.data(), in.size() * sizeof(double), hipMemcpyHostToDevice));
        m->d_tabin = d;
    }
    if (!m->d_tT) {
        const size_t nTl = (size_t)nT * nl;
        double *d = nullptr;
        HIPCHK(ctx, hipMalloc((void **)&d, (4 * nTl + 2 * (size_t)nT) * sizeof(double)));
        hipLaunchKernelGGL(k_tab_T1, dim3((unsigned)((nTl + 255) / 256)), dim3(256), 0, ctx->stream,
                           nT, nl, t.lnt0, t.dlnt, (const double *)m->d_tabin, d);
        hipLaunchKernelGGL(k_tab_T2, dim3((nT + 63) / 64), dim3(64), 0, ctx->stream, nT, nl,
                           (const double *)m->d_tabin, d);
        HIPCHK(ctx, hipGetLastError());
        m->d_tT = d;
    }
    if (!m->d_dr) {
        std::vector<double> dr(m->nr);
        for (int i = 0; i < m->nr; ++i) dr[i] = m->rw[i + 1] - m->rw[i];
        double *d = nullptr;
        HIPCHK(ctx, hipMalloc((void **)&d, m->nr * sizeof(double)));
        HIPCHK(ctx, hipMemcpy(d, dr.data(), m->nr * sizeof(double), hipMemcpyHostToDevice));
        m->d_dr = d;
    }
    if (!m->d_vol) {
        if (m->vol.size() != (size_t)nc) return set_err(ctx, RT_E_ARG, "model has no cell volumes");
        double *d = nullptr;
        HIPCHK(ctx, hipMalloc((void **)&d, sizeof(double) * nc));
        HIPCHK(ctx, hipMemcpy(d, m->vol.data(), sizeof(double) * nc, hipMemcpyHostToDevice));
        m->d_vol = d;
    }
    // the staging buffer is reused once the previous upload from it is done
    if (m->opin_busy) HIPCHK(ctx, hipEventSynchronize(m->e_opin));
    m->opin_busy = false;
    double *h = m->h_opin;
    std::copy(m->kabs.begin(), m->kabs.end(), h);
    std::copy(m->ksca.begin(), m->ksca.end(), h + nk);
    if (m->has_g) std::copy(m->gsca.begin(), m->gsca.end(), h + 2 * nk);
    for (int q = 0; q < G; ++q) h[3 * nk + q] = t.kmax[q];
    hipStream_t sm = ctx->stream;
    HIPCHK(ctx, hipMemcpyAsync(m->d_opin, h, (3 * nk + G) * sizeof(double), hipMemcpyHostToDevice, sm));
    HIPCHK(ctx, hipEventRecord(m->e_opin, sm));
    m->opin_busy = true;
    GroupReps gr{};
    for (int q = 0; q < G; ++q) gr.rep[q] = t.grep[q];
    const size_t nwork = std::max(nk, (size_t)nc * G);
    hipLaunchKernelGGL(k_opac_prep, dim3((unsigned)((nwork + 255) / 256)), dim3(256), 0, sm, S, G, nl,
                       m->nr, nc, m->has_g ? 1 : 0, gr, (const double *)m->d_opin,
                       (const double *)m->d_dr, (const double *)m->d_vol, m->d_ka, m->d_ks, m->d_g,
                       m->d_gka, m->d_ginvk, m->d_econv, m->d_tabin);
    HIPCHK(ctx, hipGetLastError());
    const int nbt = (nT + TAB_ROWS - 1) / TAB_ROWS;
    hipLaunchKernelGGL(k_tab_rows, dim3((G + S) * nbt), dim3(256), 0, sm, G, S, nT, nl,
                       (const double *)m->d_tabin, (const double *)m->d_tT, m->d_lnq, m->d_cdf,
                       m->d_guide, m->d_kplanck, m->d_kross);
    HIPCHK(ctx, hipGetLastError());
    return RT_OK;
}

// Releases the staging buffers (the device must be idle for this model).
void opacity_device_free(rt_model *m) {
    if (m->opin_busy) (void)hipEventSynchronize(m->e_opin);
    if (m->h_opin) (void)hipHostFree(m->h_opin);
    if (m->e_opin) (void)hipEventDestroy(m->e_opin);
    m->h_opin = nullptr;
    m->e_opin = nullptr;
    m->opin_busy = false;
}
