// Scattering source for images: the monochromatic Monte Carlo RADMC-3D runs
// at the image wavelength before ray tracing when the opacities scatter
// (nphot_scat packets; pdspy passes nphot to run_image, Model.py:515-534,
// write.py:34; run.py:157-289), for isotropic scattering.
//
// Sources: the point stars (L_nu = 4 pi^2 R^2 B_nu(T*)) and the thermal
// emission of the dust at the imaging temperature, L_nu,c = 4 pi V_c
// B_nu(T_c) sum_s rho_s kabs_s.  A packet of energy L/N starts at a star or
// uniformly inside a cell (r^3, cos theta, phi uniform; either hemisphere of a
// mirrored grid), travels with extinction rho (kabs + ksca), scatters
// isotropically with probability ksca / kext at each interaction and ends
// otherwise.  The mean intensity is the path-length estimate
// J_c = (L/N) sum l / (4 pi V_c), summed as u64 fixed point (l in units of
// 2^-24 of the cell's radial width) so the result does not depend on the
// order of the atomics.
//
// One packet per lane, grid-stride over packets: an albedo-weighted walk ends
// after a few interactions, so the lanes of a wave stay close to each other
// without the thermal kernel's refill machinery.  Random events (Philox key =
// seed, counter (k, packet)): k = 0 source (u0), position (u1 u2 u3),
// direction (u4 u5), tau (u6), hemisphere (u7); k >= 1: scatter or end (u0),
// direction (u1 u2), next tau (u3).
#include "rt_internal.h"

#include <cmath>
#include <vector>

using namespace rt;

namespace {

constexpr double PI_D = 3.141592653589793;
constexpr float TALLY_SCALE = 16777216.0f;  // 2^24

struct ScatKap {
    float a[RT_MAX_SPEC], s[RT_MAX_SPEC];  // kabs, ksca at the wavelength (cm^2/g)
};

struct ScatStars {
    double cum[RT_MAX_STAR];  // cumulative luminosity of the stars
    float pos[RT_MAX_STAR][3];  // au
};

struct ScatParams {
    DevGrid g;
    const float *rhoL;        // [cells][S] rho * au
    int S, nstar;
    ScatKap kap;
    ScatStars st;
    const double *cum;        // [cells] cumulative luminosity, stars first (last = total)
    const unsigned long long *seed;  // device word: Philox key (set per call or per step)
    uint64_t nphot;
    uint32_t max_events;
    const float *invdr;       // [nr] 1/dr (au^-1)
    unsigned long long *tally;  // [cells]
    unsigned long long *cnt;    // escape, absorbed, scattered, crossings, killed
};

__device__ __forceinline__ double planck_d(double nu, double T) {
    if (!(T > 0.0)) return 0.0;
    const double x = RT_H * nu / (RT_K * T);
    return 2.0 * RT_H * nu * nu * nu / (RT_C * RT_C) / expm1(x);
}

// dust luminosity of each cell at frequency nu (erg/s/Hz)
__global__ void k_scat_lum(const float *__restrict__ rho, int S, int nc, ScatKap kap,
                           const float *__restrict__ T, double nu, const double *__restrict__ vol,
                           double *__restrict__ lum) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    double a = 0.0;
    for (int s = 0; s < S; ++s) a += (double)rho[(size_t)s * nc + c] * (double)kap.a[s];
    lum[c] = 4.0 * PI_D * vol[c] * a * planck_d(nu, (double)T[c]);
}

// In-place inclusive prefix sum of n doubles plus `base`, in three kernels
// with a fixed partition (deterministic): block b owns a contiguous chunk and
// each of its threads a contiguous piece of that; the chunk totals are
// scanned by one block; then every block rewrites its chunk from its offset.
#define SCAN_BS 256
#define SCAN_MAXB 1024
__device__ __forceinline__ void scan_piece(int n, int nb, int b, int t, int &i0, int &i1) {
    const int chunk = (n + nb - 1) / nb;
    const int c0 = min(n, b * chunk), c1 = min(n, c0 + chunk);
    const int piece = (c1 - c0 + SCAN_BS - 1) / SCAN_BS;
    i0 = min(c1, c0 + t * piece);
    i1 = min(c1, i0 + piece);
}

// per block: the sums of its threads' pieces (part_t) and of its chunk
__global__ void __launch_bounds__(SCAN_BS) k_scan_sums(const double *__restrict__ x, int n,
                                                       double *__restrict__ bsum) {
    __shared__ double part[SCAN_BS];
    const int b = blockIdx.x, t = threadIdx.x, nb = gridDim.x;
    int i0, i1;
    scan_piece(n, nb, b, t, i0, i1);
    double s = 0.0;
    for (int i = i0; i < i1; ++i) s += x[i];
    part[t] = s;
    __syncthreads();
    if (t == 0) {
        double acc = 0.0;
        for (int k = 0; k < SCAN_BS; ++k) acc += part[k];
        bsum[b] = acc;
    }
}

// exclusive prefix of the chunk totals, starting at base
__global__ void k_scan_top(double *__restrict__ bsum, int nb, double base) {
    if (threadIdx.x != 0) return;
    double acc = base;
    for (int b = 0; b < nb; ++b) {
        const double v = bsum[b];
        bsum[b] = acc;
        acc += v;
    }
}

__global__ void __launch_bounds__(SCAN_BS) k_scan_apply(double *__restrict__ x, int n,
                                                        const double *__restrict__ boff) {
    __shared__ double part[SCAN_BS];
    const int b = blockIdx.x, t = threadIdx.x, nb = gridDim.x;
    int i0, i1;
    scan_piece(n, nb, b, t, i0, i1);
    double s = 0.0;
    for (int i = i0; i < i1; ++i) s += x[i];
    part[t] = s;
    __syncthreads();
    if (t == 0) {
        double acc = boff[b];
        for (int k = 0; k < SCAN_BS; ++k) {
            const double v = part[k];
            part[k] = acc;
            acc += v;
        }
    }
    __syncthreads();
    double acc = part[t];
    for (int i = i0; i < i1; ++i) {
        acc += x[i];
        x[i] = acc;
    }
}

// first index i with cum[i] > x (cum ascending); n - 1 if none
__device__ __forceinline__ int first_above(const double *cum, int n, double x) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (cum[mid] > x)
            hi = mid;
        else
            lo = mid + 1;
    }
    return lo;
}

__device__ __forceinline__ void tally_path(const ScatParams &P, int c, int ir, float len) {
    const unsigned long long v =
        (unsigned long long)rintf(len * P.invdr[ir] * TALLY_SCALE);
    if (v) atomicAdd(&P.tally[c], v);
}

__global__ void __launch_bounds__(256) k_scat_mc(ScatParams P) {
    const DevGrid &g = P.g;
    uint32_t c_esc = 0, c_abs = 0, c_sca = 0, c_cross = 0, c_kill = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const double ltot = P.cum[g.ncell - 1];
    const uint64_t seed = *P.seed;
    const uint64_t npk = ltot > 0.0 ? P.nphot : 0;
    for (uint64_t pid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; pid < npk;
         pid += stride) {
        float u[8];
        rng8(seed, pid, 0u, u);
        Pos p;
        const double x = (double)u[0] * ltot;
        int ks = -1;
        for (int k = 0; k < P.nstar; ++k)
            if (ks < 0 && x < P.st.cum[k]) ks = k;
        if (ks >= 0) {
            p.x = P.st.pos[ks][0];
            p.y = P.st.pos[ks][1];
            p.z = P.st.pos[ks][2];
        } else {
            const int c = first_above(P.cum, g.ncell, x);
            const int ir = c % g.nr, it = (c / g.nr) % g.nt, ip = c / (g.nr * g.nt);
            const double r1 = g.rwd[ir], r2 = g.rwd[ir + 1];
            const double r = cbrt(r1 * r1 * r1 + (double)u[1] * (r2 * r2 * r2 - r1 * r1 * r1));
            const double ct = g.ctd[it] + (double)u[2] * (g.ctd[it + 1] - g.ctd[it]);
            const double sth = sqrt(fmax(0.0, 1.0 - ct * ct));
            const double ph = (double)g.pw[ip] + (double)u[3] * (double)(g.pw[ip + 1] - g.pw[ip]);
            double sp, cp;
            sincos(ph, &sp, &cp);
            p.x = (float)(r * sth * cp);
            p.y = (float)(r * sth * sp);
            p.z = (float)(r * ct * ((g.mirror && u[7] < 0.5f) ? -1.0 : 1.0));
        }
        iso_dir(u[4], u[5], p.dx, p.dy, p.dz);
        locate(g, p);
        if (p.ir >= g.nr) {
            c_esc++;
            continue;
        }
        float tau = -__logf(u[6]);
        uint32_t k = 1, nev = 0, nzero = 0, nstep = 0;
        for (;;) {
            int kind;
            const float tw = next_wall(g, p, kind);
            float aabs = 0.f, asca = 0.f;
            int c = -1;
            if (p.ir >= 0) {
                c = cell_index(g, p);
                for (int s = 0; s < P.S; ++s) {
                    const float rl = P.rhoL[(size_t)c * P.S + s];
                    aabs = fmaf(rl, P.kap.a[s], aabs);
                    asca = fmaf(rl, P.kap.s[s], asca);
                }
            }
            const float aext = aabs + asca;
            if (aext * tw > tau) {
                const float sseg = tau / aext;
                tally_path(P, c, p.ir, sseg);
                p.x = fmaf(sseg, p.dx, p.x);
                p.y = fmaf(sseg, p.dy, p.y);
                p.z = fmaf(sseg, p.dz, p.z);
                float v[8];
                rng8(seed, pid, k++, v);
                nev++;
                if (v[0] * aext < asca) {
                    c_sca++;
                    iso_dir(v[1], v[2], p.dx, p.dy, p.dz);
                    tau = -__logf(v[3]);
                    nzero = 0;
                    if (P.max_events && nev >= P.max_events) {
                        c_kill++;
                        break;
                    }
                } else {
                    c_abs++;
                    break;
                }
            } else {
                if (c >= 0) tally_path(P, c, p.ir, tw);
                tau -= aext * tw;
                c_cross++;
                const float rcell = g.rw[p.ir + 1];
                if (++nstep > 100000000u) {  // last-resort cap on crossings
                    c_kill++;
                    break;
                }
                if (cross_wall(g, p, tw, kind)) {
                    c_esc++;
                    break;
                }
                if (unstick(g, p, tw, rcell, nzero) && p.ir >= g.nr) {
                    c_esc++;
                    break;
                }
            }
        }
    }
    // one atomic per counter per wave
    const uint32_t vals[5] = {c_esc, c_abs, c_sca, c_cross, c_kill};
#pragma unroll
    for (int q = 0; q < 5; ++q) {
        unsigned long long v = vals[q];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
        if (__lane_id() == 0 && v) atomicAdd(&P.cnt[q], v);
    }
}

// J_c = eps * l_c / (4 pi V_c), l in cm, eps = L / nphot
__global__ void k_scat_J(const unsigned long long *__restrict__ tally,
                         const float *__restrict__ invdr, const double *__restrict__ vol, int nr,
                         int nc, const double *__restrict__ cum, uint64_t nphot,
                         float *__restrict__ J) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const double ltot = cum[nc - 1];
    const double eps = (ltot > 0.0 && nphot > 0) ? ltot / (double)nphot : 0.0;
    const double len_cm =
        (double)tally[c] / (double)TALLY_SCALE / (double)invdr[c % nr] * RT_LEN_UNIT;
    J[c] = (float)(eps * len_cm / (4.0 * PI_D * vol[c]));
}

__global__ void k_zero_u64(unsigned long long *p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

inline unsigned nblk(size_t n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

// Allocates the model's scattering buffers (not capturable).
int scat_prepare(rt_model *m) {
    rt_ctx *ctx = m->ctx;
    if (m->vol.size() != (size_t)m->ncell) return set_err(ctx, RT_E_ARG, "model has no cell volumes");
    const int nc = m->ncell;
    if (!m->d_vol) {  // the device opacity update may have made it already
        HIPCHK(ctx, hipMalloc((void **)&m->d_vol, sizeof(double) * nc));
        HIPCHK(ctx, hipMemcpy(m->d_vol, m->vol.data(), sizeof(double) * nc, hipMemcpyHostToDevice));
    }
    if (!m->d_jtally) {
        HIPCHK(ctx, hipMalloc((void **)&m->d_jtally, sizeof(unsigned long long) * nc));
        HIPCHK(ctx, hipMalloc((void **)&m->d_scum, sizeof(double) * nc));
        HIPCHK(ctx, hipMalloc((void **)&m->d_scnt, sizeof(unsigned long long) * 8));
        HIPCHK(ctx, hipMalloc((void **)&m->d_sblk, sizeof(double) * SCAN_MAXB));
    }
    return RT_OK;
}

// The seed word the scattering kernels read (device memory).
unsigned long long *scat_seed_slot(rt_model *m) { return m->d_scnt + 7; }

// Source luminosities, prefix sum, Monte Carlo and J on stream sm, with the
// Philox key in scat_seed_slot(m).  No host synchronisation: capturable.
void scat_enqueue(rt_model *m, double lam_um, uint64_t nphot, const float *d_T, float *d_J,
                  hipStream_t sm) {
    const int nc = m->ncell;
    ScatParams P{};
    P.g = dev_grid(m);
    P.rhoL = m->d_rhoL;
    P.S = m->nspec;
    for (int s = 0; s < RT_MAX_SPEC; ++s) {
        P.kap.a[s] = s < m->nspec ? (float)kappa_at_lambda(m, m->kabs, s, lam_um) : 0.f;
        P.kap.s[s] = s < m->nspec ? (float)kappa_at_lambda(m, m->ksca, s, lam_um) : 0.f;
    }
    const double nu = RT_C / (lam_um * 1e-4);
    double lstar = 0.0;
    P.nstar = std::min(m->nstar, RT_MAX_STAR);
    for (int k = 0; k < P.nstar; ++k) {
        const double R = m->stars[5 * k], Ts = m->stars[5 * k + 1];
        const double x = RT_H * nu / (RT_K * Ts);
        lstar += 4.0 * PI_D * PI_D * R * R * (2.0 * RT_H * nu * nu * nu / (RT_C * RT_C) / std::expm1(x));
        P.st.cum[k] = lstar;
        for (int a = 0; a < 3; ++a) P.st.pos[k][a] = (float)(m->stars[5 * k + 2 + a] / RT_LEN_UNIT);
    }
    m->scat_lstar = lstar;
    hipLaunchKernelGGL(k_scat_lum, dim3(nblk(nc, 256)), dim3(256), 0, sm, m->d_rho, m->nspec, nc,
                       P.kap, d_T, nu, m->d_vol, m->d_scum);
    const int nb = (int)std::min<int64_t>(SCAN_MAXB, std::max<int64_t>(1, ((int64_t)nc + 4095) / 4096));
    hipLaunchKernelGGL(k_scan_sums, dim3(nb), dim3(SCAN_BS), 0, sm, (const double *)m->d_scum, nc,
                       m->d_sblk);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(64), 0, sm, m->d_sblk, nb, lstar);
    hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(SCAN_BS), 0, sm, m->d_scum, nc,
                       (const double *)m->d_sblk);
    P.cum = m->d_scum;
    P.seed = scat_seed_slot(m);
    P.nphot = nphot;
    P.max_events = 1000000u;
    P.invdr = m->d_invdr;
    P.tally = m->d_jtally;
    P.cnt = m->d_scnt;
    hipLaunchKernelGGL(k_zero_u64, dim3(nblk(nc, 256)), dim3(256), 0, sm, m->d_jtally, (size_t)nc);
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, sm, m->d_scnt, (size_t)5);
    const uint64_t blocks = std::min<uint64_t>((nphot + 255) / 256, (uint64_t)m->ctx->n_cu * 16);
    if (nphot > 0)
        hipLaunchKernelGGL(k_scat_mc, dim3((unsigned)std::max<uint64_t>(blocks, 1)), dim3(256), 0,
                           sm, P);
    hipLaunchKernelGGL(k_scat_J, dim3(nblk(nc, 256)), dim3(256), 0, sm, m->d_jtally, m->d_invdr,
                       m->d_vol, m->nr, nc, (const double *)m->d_scum, nphot, d_J);
}

int scat_source_run(rt_model *m, double lam_um, uint64_t nphot, uint64_t seed,
                    const float *d_T, float *d_J, rt_scat_stats *st) {
    rt_ctx *ctx = m->ctx;
    if (!(lam_um > 0)) return set_err(ctx, RT_E_ARG, "wavelength must be > 0");
    if (!d_T || !d_J) return set_err(ctx, RT_E_ARG, "NULL temperature or output pointer");
    int rc = scat_prepare(m);
    if (rc) return rc;
    hipStream_t sm = ctx->stream;
    unsigned long long key = seed;
    HIPCHK(ctx, hipMemcpyAsync(scat_seed_slot(m), &key, sizeof(key), hipMemcpyHostToDevice, sm));
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventRecord(e0, sm);
    scat_enqueue(m, lam_um, nphot, d_T, d_J, sm);
    (void)hipEventRecord(e1, sm);
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipStreamSynchronize(sm));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st) {
        unsigned long long cnt[8];
        double ltot = 0.0;
        HIPCHK(ctx, hipMemcpy(cnt, m->d_scnt, sizeof(cnt), hipMemcpyDeviceToHost));
        HIPCHK(ctx, hipMemcpy(&ltot, m->d_scum + (m->ncell - 1), sizeof(double),
                              hipMemcpyDeviceToHost));
        st->n_escape = cnt[0];
        st->n_abs = cnt[1];
        st->n_scat = cnt[2];
        st->n_cross = cnt[3];
        st->n_killed = cnt[4];
        st->lum_star = m->scat_lstar;
        st->lum_dust = ltot - m->scat_lstar;
        st->eps = (ltot > 0 && nphot > 0) ? ltot / (double)nphot : 0.0;
        st->kernel_ms = ms;
    }
    return RT_OK;
}
