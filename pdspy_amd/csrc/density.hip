// Device-side density generation for the axisymmetric model components of
// pdspy's disk + envelope models, so that a sampler step can rebuild every
// model's densities without host round trips:
//
//  * flared power-law disk  -- pdspy/modeling/Disk.py:48-70 (surface density,
//    normalised by a 1000-point log trapezoid), :86-135 (rho = Sigma /
//    (sqrt(2 pi) h) exp(-z^2 / 2 h^2), h = h0 au (R/au)^plh);
//  * extended Ulrich envelope -- pdspy/modeling/UlrichEnvelopeExtended.py:37-110
//    (mu0 from the cubic mu0^3 - mu0 (1 - r/Rc) - mu r/Rc = 0 by Brent's method
//    with hyperbolic extrapolation, the midplane limits, the outflow cavity,
//    then the mass normalisation by a trapezoid rule over the cell centres).
//
// FP64 throughout; FMA contraction is off so each expression rounds as the
// host's numpy evaluation does.  The root finder follows the published
// algorithm of scipy.optimize.brenth (Brent 1973 with the hyperbolic
// extrapolation of Bus & Dekker 1975; xtol 2e-12, rtol 4 eps, 100 iterations)
// step for step, so in cells where the cubic has three real roots in [-1, 1]
// it returns the same root as the host model.
//
// Everything reads its parameters from device memory: the kernels can be
// captured once in a graph and replayed with new parameters uploaded in
// between.
#pragma clang fp contract(off)
#include "rt_internal.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace {

constexpr double AU_CM = 1.495978707e13;   // pdspy/constants/astronomy.py
constexpr double MSUN_G = 1.99e33;
constexpr double PI_PDSPY = 3.14159265;    // pdspy/constants/math.py:1
constexpr double PI_NP = 3.141592653589793;  // numpy's pi

__device__ double mu0_cubic(double x, double r, double mu, double rc) {
    return pow(x, 3.0) - x * (1.0 - r / rc) - mu * (r / rc);
}

// scipy.optimize.brenth(f, -1, 1) for f = mu0_cubic(., r, mu, rc)
__device__ double mu0_root(double r, double mu, double rc) {
    const double xtol = 2e-12, rtol = 4.0 * 2.220446049250313e-16;
    double xpre = -1.0, xcur = 1.0;
    double xblk = 0.0, fblk = 0.0, spre = 0.0, scur = 0.0;
    double fpre = mu0_cubic(xpre, r, mu, rc), fcur = mu0_cubic(xcur, r, mu, rc);
    if (fpre == 0.0) return xpre;
    if (fcur == 0.0) return xcur;
    if (signbit(fpre) == signbit(fcur)) return NAN;
    for (int i = 0; i < 100; ++i) {
        if (fpre != 0.0 && fcur != 0.0 && signbit(fpre) != signbit(fcur)) {
            xblk = xpre;
            fblk = fpre;
            spre = scur = xcur - xpre;
        }
        if (fabs(fblk) < fabs(fcur)) {
            xpre = xcur;
            xcur = xblk;
            xblk = xpre;
            fpre = fcur;
            fcur = fblk;
            fblk = fpre;
        }
        const double delta = (xtol + rtol * fabs(xcur)) / 2.0;
        const double sbis = (xblk - xcur) / 2.0;
        if (fcur == 0.0 || fabs(sbis) < delta) return xcur;
        if (fabs(spre) > delta && fabs(fcur) < fabs(fpre)) {
            double stry;
            if (xpre == xblk) {
                stry = -fcur * (xcur - xpre) / (fcur - fpre);  // secant
            } else {
                const double dpre = (fpre - fcur) / (xpre - xcur);
                const double dblk = (fblk - fcur) / (xblk - xcur);
                stry = -fcur * (fblk - fpre) / (fblk * dpre - fpre * dblk);  // hyperbolic
            }
            if (2.0 * fabs(stry) < fmin(fabs(spre), 3.0 * fabs(sbis) - delta)) {
                spre = scur;
                scur = stry;
            } else {
                spre = sbis;
                scur = sbis;
            }
        } else {
            spre = sbis;
            scur = sbis;
        }
        xpre = xcur;
        fpre = fcur;
        if (fabs(scur) > delta)
            xcur += scur;
        else
            xcur += (sbis > 0.0 ? delta : -delta);
        fcur = mu0_cubic(xcur, r, mu, rc);
    }
    return xcur;
}

// disk surface density before normalisation (Disk.py:48-70), r in au
__device__ double disk_sigma(const rt_density_comp &C, double r) {
    const double rin = C.rmin * AU_CM, rout = C.rmax * AU_CM;
    const double mass = C.mass * MSUN_G;
    const double gamma = C.plrho - C.plh;
    const double sigma0 = (2.0 - gamma) * mass / (2.0 * PI_PDSPY * pow(AU_CM, gamma)) /
                          (pow(rout, -gamma + 2.0) - pow(rin, -gamma + 2.0));
    double s = sigma0 * pow(r, -gamma);
    if ((r >= rout / AU_CM) != (r <= rin / AU_CM)) s = 0.0;
    return s;
}

// per disk component (one wave each): mass / (2 pi trapz(r Sigma, r)) on 1000
// log-spaced radii; the lanes share the points, the segment sums are reduced
// in a fixed order
#define RT_DISK_NPT 1000
__global__ void __launch_bounds__(64) k_disk_scale(const rt_density_comp *__restrict__ comps,
                                                   int ncomp, double *__restrict__ scale) {
    __shared__ double xs[RT_DISK_NPT], ys[RT_DISK_NPT];
    const int k = blockIdx.x, lane = threadIdx.x;
    if (k >= ncomp) return;
    const rt_density_comp C = comps[k];
    if (C.kind != RT_COMP_DISK) {
        if (lane == 0) scale[k] = 1.0;
        return;
    }
    const int n = RT_DISK_NPT;
    const double a = log10(C.rmin), b = log10(C.rmax);
    const double step = (b - a) / (double)(n - 1);
    for (int i = lane; i < n; i += 64) {
        const double e = i == n - 1 ? b : (double)i * step + a;
        const double r = pow(10.0, e);
        xs[i] = r * AU_CM;
        ys[i] = r * AU_CM * disk_sigma(C, r);
    }
    __syncthreads();
    double acc = 0.0;
    for (int i = lane + 1; i < n; i += 64) acc += (xs[i] - xs[i - 1]) * (ys[i] + ys[i - 1]) / 2.0;
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if (lane == 0) scale[k] = C.mass * MSUN_G / (2.0 * PI_NP * acc);
}

// component k's density at cell (ir, it) of the (r, theta) plane, envelopes
// before their mass normalisation; W[k][it][ir]
__global__ void k_comp_values(const rt_density_comp *__restrict__ comps, int ncomp,
                              const double *__restrict__ rc, const double *__restrict__ tc, int nr,
                              int nt, const double *__restrict__ scale, double *__restrict__ W) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nrt = (size_t)nr * nt;
    if (i >= nrt * (size_t)ncomp) return;
    const int k = (int)(i / nrt);
    const int ir = (int)(i % nrt) % nr, it = (int)((i % nrt) / nr);
    const rt_density_comp C = comps[k];
    const double r = rc[ir], th = tc[it];
    double v = 0.0;
    if (C.kind == RT_COMP_DISK) {
        const double R = r * sin(th), z = r * cos(th);
        const double sig = disk_sigma(C, R / AU_CM) * scale[k];
        const double h = C.h0 * AU_CM * pow(R / AU_CM, C.plh);
        v = sig / (sqrt(2.0 * PI_NP) * h) * exp(-0.5 * pow(z / h, 2.0));
    } else if (C.kind == RT_COMP_ULRICH_EXT) {
        const double rin = C.rmin * AU_CM, rout = C.rmax * AU_CM;
        const double rcent = C.rcent * AU_CM, cavz0 = C.zoffset * AU_CM;
        const double th_open = C.theta_open * PI_NP / 180.0;
        const double cava = (rout * sin(th_open / 2.0) / AU_CM / tan(th_open / 2.0) - cavz0 / AU_CM) *
                            pow(rout * sin(th_open / 2.0) / AU_CM, -C.cavpl);
        const double mu = cos(th), R = r * sin(th), z = r * cos(th);
        const double mu0 = mu0_root(r, mu, rcent);
        v = 1.0 * pow(r / rcent, -1.5) * pow(1.0 + mu / mu0, -0.5) *
            pow(mu / mu0 + 2.0 * pow(mu0, 2.0) * rcent / r, -1.0);
        if (fabs(mu) < 1.0e-10 && r < rcent)
            v = 1.0 * pow(r / rcent, -0.5) * pow(1.0 - r / rcent, -1.0) / 2.0;
        if (fabs(mu) < 1.0e-10 && r > rcent)
            v = 1.0 * pow(2.0 * r / rcent - 1.0, -0.5) * pow(r / rcent - 1.0, -1.0);
        if ((r >= rout) != (r <= rin)) v = 0.0;
        if (fabs(z) / AU_CM - cavz0 / AU_CM - cava * pow(R / AU_CM, C.cavpl) > 0.0) v *= C.cavrfact;
    }
    W[i] = v;
}

// envelopes: trapezoid over theta of rho r^2 sin(theta) at each radius
__global__ void k_env_rows(const rt_density_comp *__restrict__ comps, int ncomp,
                           const double *__restrict__ rc, const double *__restrict__ tc, int nr,
                           int nt, const double *__restrict__ W, double *__restrict__ rows) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nr * ncomp) return;
    const int k = i / nr, ir = i % nr;
    if (comps[k].kind != RT_COMP_ULRICH_EXT) return;
    const double *w = W + (size_t)k * nr * nt;
    const double r2 = pow(rc[ir], 2.0);
    double acc = 0.0;
    double y0 = w[ir] * r2 * sin(tc[0]);
    for (int it = 1; it < nt; ++it) {
        const double y = w[ir + (size_t)nr * it] * r2 * sin(tc[it]);
        acc += (tc[it] - tc[it - 1]) * (y + y0) / 2.0;
        y0 = y;
    }
    rows[i] = acc;
}

// envelopes: mass / (4 pi (or 2 pi) trapz over r); disks: 1
__global__ void k_env_norm(const rt_density_comp *__restrict__ comps, int ncomp,
                           const double *__restrict__ rc, int nr, double solid,
                           const double *__restrict__ rows, double *__restrict__ fac) {
    const int k = threadIdx.x;
    if (k >= ncomp) return;
    if (comps[k].kind != RT_COMP_ULRICH_EXT) {
        fac[k] = 1.0;
        return;
    }
    const double *y = rows + (size_t)k * nr;
    double acc = 0.0;
    for (int ir = 1; ir < nr; ++ir) acc += (rc[ir] - rc[ir - 1]) * (y[ir] + y[ir - 1]) / 2.0;
    fac[k] = comps[k].mass * MSUN_G / (solid * acc);
}

// species densities of every cell (r fastest) in the three device layouts the
// kernels read; components add into their species
__global__ void k_density_out(const rt_density_comp *__restrict__ comps, int ncomp, int nr,
                              int nt, int np, int S, const double *__restrict__ W,
                              const double *__restrict__ fac, float *__restrict__ rho,
                              float *__restrict__ rhoL, unsigned char *__restrict__ rhopos) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int nc = nr * nt * np;
    if (c >= nc) return;
    const size_t rt_ = (size_t)(c % (nr * nt));  // axisymmetric: the same for every phi
    for (int s = 0; s < S; ++s) {
        double v = 0.0;
        for (int k = 0; k < ncomp; ++k)
            if (comps[k].species == s) v += W[(size_t)k * nr * nt + rt_] * fac[k];
        rho[(size_t)s * nc + c] = (float)v;
        rhoL[(size_t)c * S + s] = (float)(v * RT_LEN_UNIT);
        rhopos[(size_t)s * nc + c] = v > 0.0 ? 1 : 0;
    }
}

inline unsigned nblk(size_t n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

// Checks the components, allocates the model's generator buffers (cell
// centres, per-component scratch) on first use.  Not capturable (allocates).
int density_prepare(rt_model *m, const rt_density_comp *h_comps, int ncomp) {
    rt_ctx *ctx = m->ctx;
    if (ncomp < 1 || ncomp > RT_MAX_COMP)
        return set_err(ctx, RT_E_ARG, "number of density components must be 1..%d", RT_MAX_COMP);
    if (!h_comps) return set_err(ctx, RT_E_ARG, "NULL density components");
    for (int k = 0; k < ncomp; ++k) {
        const rt_density_comp &C = h_comps[k];
        if (C.kind != RT_COMP_DISK && C.kind != RT_COMP_ULRICH_EXT)
            return set_err(ctx, RT_E_ARG, "component %d: unknown kind %d", k, C.kind);
        if (C.species < 0 || C.species >= m->nspec)
            return set_err(ctx, RT_E_ARG, "component %d: species %d not in 0..%d", k, C.species,
                           m->nspec - 1);
        const double v[] = {C.mass, C.rmin, C.rmax, C.plrho, C.h0, C.plh, C.rcent, C.cavpl,
                            C.cavrfact, C.theta_open, C.zoffset};
        for (double x : v)
            if (!std::isfinite(x)) return set_err(ctx, RT_E_ARG, "component %d: non-finite parameter", k);
        if (!(C.mass > 0) || !(C.rmin > 0) || !(C.rmax > C.rmin))
            return set_err(ctx, RT_E_ARG, "component %d: need mass > 0 and 0 < rmin < rmax", k);
        if (C.kind == RT_COMP_DISK && !(C.h0 > 0))
            return set_err(ctx, RT_E_ARG, "component %d: disk needs h0 > 0", k);
        if (C.kind == RT_COMP_ULRICH_EXT &&
            (!(C.rcent > 0) || !(C.theta_open > 0 && C.theta_open < 180) || !(C.cavrfact >= 0)))
            return set_err(ctx, RT_E_ARG,
                           "component %d: envelope needs rcent > 0, 0 < theta_open < 180, "
                           "cavrfact >= 0", k);
    }
    if (m->np < 1 || m->rw.size() != (size_t)m->nr + 1 || m->tw.size() != (size_t)m->nt + 1)
        return set_err(ctx, RT_E_ARG, "model has no host walls");
    if (!m->d_rc) {
        // cell centres as arithmetic midpoints of the walls (Grid.py:74-76), r in cm
        std::vector<double> rc(m->nr), tc(m->nt);
        for (int i = 0; i < m->nr; ++i) rc[i] = 0.5 * (m->rw[i] / RT_LEN_UNIT + m->rw[i + 1] / RT_LEN_UNIT) * RT_LEN_UNIT;
        for (int j = 0; j < m->nt; ++j) tc[j] = 0.5 * (m->tw[j] + m->tw[j + 1]);
        HIPCHK(ctx, hipMalloc((void **)&m->d_rc, sizeof(double) * m->nr));
        HIPCHK(ctx, hipMalloc((void **)&m->d_tc, sizeof(double) * m->nt));
        HIPCHK(ctx, hipMemcpy(m->d_rc, rc.data(), sizeof(double) * m->nr, hipMemcpyHostToDevice));
        HIPCHK(ctx, hipMemcpy(m->d_tc, tc.data(), sizeof(double) * m->nt, hipMemcpyHostToDevice));
        // UlrichEnvelopeExtended.py:100-106: 2 pi if the theta centres pass pi/2
        m->dens_solid = (tc[m->nt - 1] > PI_PDSPY / 2.0 ? 2.0 : 4.0) * PI_PDSPY;
        const size_t nrt = (size_t)m->nr * m->nt;
        HIPCHK(ctx, hipMalloc((void **)&m->d_dens_w, sizeof(double) * nrt * RT_MAX_COMP));
        HIPCHK(ctx, hipMalloc((void **)&m->d_dens_rows, sizeof(double) * m->nr * RT_MAX_COMP));
        HIPCHK(ctx, hipMalloc((void **)&m->d_dens_fac, sizeof(double) * 2 * RT_MAX_COMP));
        HIPCHK(ctx, hipMalloc((void **)&m->d_comps, sizeof(rt_density_comp) * RT_MAX_COMP));
    }
    return RT_OK;
}

// The generator's kernels on stream sm, reading ncomp components from d_comps
// (device memory).  Capturable.
void density_enqueue(rt_model *m, const rt_density_comp *d_comps, int ncomp, hipStream_t sm) {
    const size_t nrt = (size_t)m->nr * m->nt;
    double *scale = m->d_dens_fac, *fac = m->d_dens_fac + RT_MAX_COMP;
    hipLaunchKernelGGL(k_disk_scale, dim3(ncomp), dim3(64), 0, sm, d_comps, ncomp, scale);
    hipLaunchKernelGGL(k_comp_values, dim3(nblk(nrt * ncomp, 256)), dim3(256), 0, sm, d_comps,
                       ncomp, m->d_rc, m->d_tc, m->nr, m->nt, scale, m->d_dens_w);
    hipLaunchKernelGGL(k_env_rows, dim3(nblk((size_t)m->nr * ncomp, 64)), dim3(64), 0, sm,
                       d_comps, ncomp, m->d_rc, m->d_tc, m->nr, m->nt, m->d_dens_w,
                       m->d_dens_rows);
    hipLaunchKernelGGL(k_env_norm, dim3(1), dim3(64), 0, sm, d_comps, ncomp, m->d_rc, m->nr,
                       m->dens_solid, m->d_dens_rows, fac);
    hipLaunchKernelGGL(k_density_out, dim3(nblk(m->ncell, 256)), dim3(256), 0, sm, d_comps, ncomp,
                       m->nr, m->nt, m->np, m->nspec, m->d_dens_w, fac, m->d_rho, m->d_rhoL,
                       m->d_rhopos);
}

struct CompArgs {
    rt_density_comp c[RT_MAX_COMP];
};

// Stores components passed by value (kernel arguments are copied at launch, so
// the caller's array may change as soon as the launch returns).
__global__ void k_store_comps(CompArgs a, int ncomp, rt_density_comp *d_comps) {
    const int k = threadIdx.x;
    if (k < ncomp) d_comps[k] = a.c[k];
}

// One model, stream-ordered: the components go to the device as a kernel
// argument and the generator's kernels follow on the context's stream; nothing
// waits for the device.
int density_generate(rt_model *m, const rt_density_comp *h_comps, int ncomp) {
    rt_ctx *ctx = m->ctx;
    int rc = density_prepare(m, h_comps, ncomp);
    if (rc) return rc;
    hipStream_t sm = ctx->stream;
    CompArgs a{};
    std::copy(h_comps, h_comps + ncomp, a.c);
    hipLaunchKernelGGL(k_store_comps, dim3(1), dim3(RT_MAX_COMP), 0, sm, a, ncomp,
                       (rt_density_comp *)m->d_comps);
    density_enqueue(m, (const rt_density_comp *)m->d_comps, ncomp, sm);
    HIPCHK(ctx, hipGetLastError());
    m->rho_host_stale = true;
    m->have_T = false;
    return RT_OK;
}
