// thermal.hip -- thermal Monte Carlo on gfx950: what RADMC-3D `mctherm` does for
// pdspy's Model.run_thermal (pdspy/modeling/Model.py:221-256, radmc3d/run.py:9-38).
//
// Algorithm (Bjorkman & Wood 2001 with a path-length absorbed-energy estimator):
//  * nphot packets of energy eps = L/nphot leave blackbody point stars; packet p
//    draws from Philox4x32-10 with key = seed, counter = (block, p, 0).
//  * flights of optical depth -ln(xi) through the spherical grid; every segment
//    of length s in cell c adds kabs_g(nu) * s to the cell's absorbed-energy
//    tally of absorption group g (u64 fixed point, unit 2^-28 kmax_g dr_c), so
//    the sum is exact and independent of execution order;
//  * at an interaction the packet scatters with probability alpha_sca/alpha_ext
//    (isotropic or Henyey-Greenstein) or is absorbed and re-emitted at once,
//    isotropically, with a frequency drawn from kabs dB/dT at the cell's current
//    temperature (T interpolated between the rows of the (ntemp, temp0, temp1)
//    table by a mixture draw);
//  * "current temperature" is the one implied by the tallies of all packets of
//    the earlier epochs (epoch k covers packets [N_k, N_k+1), N_k+1 = N_k +
//    max(first, N_k (growth-1))), which makes the result bit-reproducible for a
//    seed however the packets are scheduled.
//
// Execution: a persistent kernel, one packet per lane.  A lane whose packet has
// left the grid takes the next packet id at once (one atomic per wavefront per
// refill, ballot-compacted), so the 64 lanes of a wave stay busy until the
// epoch's packets run out.  Opacities live in LDS; densities ([cell][species],
// fp32) and tallies stay L2-resident for the grid sizes of interest.
#include <algorithm>
#include <cmath>
#include <cstdio>

#include "rt_internal.h"

using namespace rt;

struct ThermParams {
    DevGrid g;
    int nspec, ngroup, nlam, ntemp, nstar, hg;
    const float *rhoL;            // [cell][S]
    const float *ka, *ks, *gg;    // [S][nlam]
    const int *grp;               // [S]
    const float *gka;             // [G][nlam]
    const float *ginvk;           // [G]
    const float *invdr;           // [nr]
    const float *cdf;             // [G][ntemp][nlam]
    const unsigned char *guide;   // [G][ntemp][NG]
    const float *scdf;            // [nstar][nlam]
    const unsigned char *sguide;  // [nstar][NG]
    const float *spick;
    const float *spos;
    const float *xrow;            // [cell][G]
    const float *kross;           // [S][ntemp] Rosseland-mean extinction
    const float *kplanck;         // [G][ntemp] Planck-mean absorption
    const float *mrwy;            // [RT_MRW_NY] escape-law inverse
    unsigned long long *tally;    // [cell][G]
    unsigned int *visits;         // nullable
    unsigned long long *cnt;
    uint64_t pbegin, pend, seed, max_events;
    uint32_t max_steps;  // wall crossings per packet
    uint64_t max_ticks;
    int mrw, mrw_count;
    float mrw_tau, mrw_gamma;
};

// smallest i with u <= cdf[i], starting from the guide entry
__device__ __forceinline__ int sample_guided(const float *cdf, const unsigned char *guide,
                                             float u) {
    int i = guide[(int)(u * (float)RT_NGUIDE)];
    while (cdf[i] < u) ++i;
    return i;
}

__device__ __forceinline__ void flush_tally(const ThermParams &P, int c,
                                            unsigned long long (&acc)[RT_MAX_GROUP]) {
#pragma unroll
    for (int q = 0; q < RT_MAX_GROUP; ++q) {
        if (q < P.ngroup && acc[q]) {
            unsigned long long old = atomicAdd(&P.tally[(size_t)c * P.ngroup + q], acc[q]);
            if (old > (1ull << 62)) atomicOr(&P.cnt[CNT_OVERFLOW], 1ull);
            acc[q] = 0;
        }
    }
}

__global__ void __launch_bounds__(256) k_thermal(ThermParams P) {
    extern __shared__ float lds[];
    const int S = P.nspec, NL = P.nlam, G = P.ngroup;
    float *s_ka = lds;
    float *s_ks = s_ka + S * NL;
    float *s_gka = s_ks + S * NL;
    float *s_g = s_gka + G * NL;  // only when P.hg
    // grid walls in LDS: every step reads several of them
    const int NR = P.g.nr, NT = P.g.nt, NP = P.g.np;
    float *s_rw = s_g + (P.hg ? S * NL : 0);
    float *s_rw2 = s_rw + NR + 1;
    float *s_tw = s_rw2 + NR + 1;
    float *s_ct = s_tw + NT + 1;
    float *s_st = s_ct + NT + 1;
    float *s_pw = s_st + NT + 1;
    float *s_cp = s_pw + NP + 1;
    float *s_sp = s_cp + NP + 1;
    float *s_invdr = s_sp + NP + 1;
    for (int i = threadIdx.x; i < S * NL; i += blockDim.x) {
        s_ka[i] = P.ka[i];
        s_ks[i] = P.ks[i];
        if (P.hg) s_g[i] = P.gg[i];
    }
    for (int i = threadIdx.x; i < G * NL; i += blockDim.x) s_gka[i] = P.gka[i];
    for (int i = threadIdx.x; i <= NR; i += blockDim.x) {
        s_rw[i] = P.g.rw[i];
        s_rw2[i] = P.g.rw2[i];
        if (i < NR) s_invdr[i] = P.invdr[i];
    }
    for (int i = threadIdx.x; i <= NT; i += blockDim.x) {
        s_tw[i] = P.g.tw[i];
        s_ct[i] = P.g.ct[i];
        s_st[i] = P.g.st[i];
    }
    for (int i = threadIdx.x; i <= NP; i += blockDim.x) {
        s_pw[i] = P.g.pw[i];
        s_cp[i] = P.g.cp[i];
        s_sp[i] = P.g.sp[i];
    }
    __syncthreads();

    DevGrid g = P.g;
    g.rw = s_rw;
    g.rw2 = s_rw2;
    g.tw = s_tw;
    g.ct = s_ct;
    g.st = s_st;
    g.pw = s_pw;
    g.cp = s_cp;
    g.sp = s_sp;
    const uint64_t range = P.pend - P.pbegin;
    const int lane = __lane_id();
    const uint64_t t_start = wall_clock64();

    Pos p;
    Philox r;
    int ifr = 0;
    float tau = 0.f;
    uint32_t nev = 0, nstep = 0, nev_cell = 0;
    bool alive = false, exhausted = false;
    unsigned long long acc[RT_MAX_GROUP];
#pragma unroll
    for (int q = 0; q < RT_MAX_GROUP; ++q) acc[q] = 0;
    uint32_t n_cross = 0, n_abs = 0, n_scat = 0, n_esc = 0, n_kill = 0, n_ev = 0, n_mrw = 0;
    uint32_t max_pkt = 0;  // most steps of one packet on this lane

    for (uint32_t iter = 0;; ++iter) {
        // ---- refill dead lanes with new packets (one atomic per wave)
        if (!exhausted) {
            uint64_t need = __ballot(!alive);
            if (need) {
                int cnt = __popcll(need);
                int leader = __ffsll((unsigned long long)need) - 1;
                unsigned long long base = 0;
                if (lane == leader) base = atomicAdd(&P.cnt[CNT_PACKET], (unsigned long long)cnt);
                base = __shfl(base, leader);
                if (base + cnt >= range) exhausted = true;
                if (!alive) {
                    uint64_t rank = __popcll(need & ((1ull << lane) - 1ull));
                    uint64_t idx = base + rank;
                    if (idx < range) {
                        uint64_t pid = P.pbegin + idx;
                        max_pkt = max(max_pkt, nev + nstep);
                        rng_init(r, P.seed, pid, 0u);
                        float u = uni(r);
                        int ks = 0;
                        while (ks < P.nstar - 1 && u > P.spick[ks]) ++ks;
                        ifr = sample_guided(P.scdf + (size_t)ks * NL, P.sguide + (size_t)ks * RT_NGUIDE,
                                            uni(r));
                        iso_dir(r, p.dx, p.dy, p.dz);
                        p.x = P.spos[3 * ks];
                        p.y = P.spos[3 * ks + 1];
                        p.z = P.spos[3 * ks + 2];
                        locate(g, p);
                        nev = 0;
                        nstep = 0;
                        nev_cell = 0;
                        if (p.ir >= g.nr) {
                            n_esc++;
                        } else {
                            alive = true;
                            if (p.ir >= 0 && P.visits) atomicAdd(&P.visits[cell_index(g, p)], 1u);
                            tau = -__logf(uni(r));
                        }
                    }
                }
            }
        }
        if (!__any(alive)) break;

        // ---- watchdog (wave-uniform check)
        if ((iter & 1023u) == 0u) {
            bool stop = (wall_clock64() - t_start) > P.max_ticks ||
                        __atomic_load_n(&P.cnt[CNT_ABORT], __ATOMIC_RELAXED) != 0ull;
            if (stop) {
                if (lane == 0) atomicOr(&P.cnt[CNT_ABORT], 1ull);
                if (alive) {
                    if (p.ir >= 0) flush_tally(P, cell_index(g, p), acc);
                    n_kill++;
                    alive = false;
                }
                exhausted = true;
                break;
            }
        }
        if (!alive) continue;

        // ---- one step: to the next wall or to the next interaction
        int kind;
        float tw = next_wall(g, p, kind);
        float aabs = 0.f, asca = 0.f;
        int c = -1;
        if (p.ir >= 0) {
            c = cell_index(g, p);
            const float *rc = P.rhoL + (size_t)c * S;
            for (int s = 0; s < S; ++s) {
                float rho = rc[s];
                aabs = fmaf(rho, s_ka[s * NL + ifr], aabs);
                asca = fmaf(rho, s_ks[s * NL + ifr], asca);
            }
        }
        float aext = aabs + asca;
        if (aext * tw > tau) {
            // interaction inside the cell
            float sseg = tau / aext;
            float wf = s_invdr[p.ir] * sseg;
#pragma unroll
            for (int q = 0; q < RT_MAX_GROUP; ++q)
                if (q < G) acc[q] += (unsigned long long)rintf(s_gka[q * NL + ifr] * P.ginvk[q] * wf);
            p.x = fmaf(sseg, p.dx, p.x);
            p.y = fmaf(sseg, p.dy, p.y);
            p.z = fmaf(sseg, p.dz, p.z);
            n_ev++;
            nev++;
            nev_cell++;
            float ue = uni(r);
            const float *rc = P.rhoL + (size_t)c * S;
            if (ue * aext < asca) {
                n_scat++;
                if (P.hg) {
                    float us = uni(r) * asca, a = 0.f;
                    int sp = S - 1;
                    for (int s = 0; s < S; ++s) {
                        a = fmaf(rc[s], s_ks[s * NL + ifr], a);
                        if (us <= a) {
                            sp = s;
                            break;
                        }
                    }
                    hg_dir(r, s_g[sp * NL + ifr], p.dx, p.dy, p.dz);
                } else {
                    iso_dir(r, p.dx, p.dy, p.dz);
                }
            } else {
                n_abs++;
                float us = uni(r) * aabs, a = 0.f;
                int sp = S - 1;
                for (int s = 0; s < S; ++s) {
                    a = fmaf(rc[s], s_ka[s * NL + ifr], a);
                    if (us <= a) {
                        sp = s;
                        break;
                    }
                }
                int q = P.grp[sp];
                float x = P.xrow[(size_t)c * G + q];
                x = fminf(fmaxf(x, 0.f), (float)(P.ntemp - 1));
                int lo = min((int)x, P.ntemp - 2);
                float w = x - (float)lo;
                int row = lo + (uni(r) < w ? 1 : 0);
                size_t ro = (size_t)q * P.ntemp + row;
                ifr = sample_guided(P.cdf + ro * NL, P.guide + ro * RT_NGUIDE, uni(r));
                iso_dir(r, p.dx, p.dy, p.dz);
            }
            if (P.max_events && nev >= P.max_events) {
                flush_tally(P, c, acc);
                n_kill++;
                alive = false;
                continue;
            }
            if (P.mrw && nev_cell >= (uint32_t)P.mrw_count) {
                // modified random walk: jump to a sphere inside the cell when the
                // Rosseland optical depth to the nearest wall is large
                float aR = 0.f;
                for (int s = 0; s < S; ++s) {
                    float rho = rc[s];
                    if (rho > 0.f)
                        aR = fmaf(rho, trow(P.kross + (size_t)s * P.ntemp, P.ntemp,
                                            P.xrow[(size_t)c * G + P.grp[s]]), aR);
                }
                float dmin = wall_dist(g, p);
                float R0 = dmin - P.mrw_gamma / aR;
                if (aR * dmin >= P.mrw_tau && R0 > 0.f) {
                    float y = mrw_escape_y(P.mrwy, RT_MRW_NY, RT_MRW_LMAX, uni(r));
                    float ell = 3.f * aR * R0 * R0 * y;
                    float wl = s_invdr[p.ir] * ell;
                    float kP[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                    for (int q = 0; q < RT_MAX_GROUP; ++q)
                        if (q < G) {
                            kP[q] = trow(P.kplanck + (size_t)q * P.ntemp, P.ntemp,
                                         P.xrow[(size_t)c * G + q]);
                            acc[q] += (unsigned long long)rintf(kP[q] * P.ginvk[q] * wl);
                        }
                    float nx, ny, nz;
                    iso_dir(r, nx, ny, nz);
                    p.x = fmaf(R0, nx, p.x);
                    p.y = fmaf(R0, ny, p.y);
                    p.z = fmaf(R0, nz, p.z);
                    if (g.mirror) p.hemi = (p.z >= 0.f) ? 1 : -1;
                    float ct = sqrtf(uni(r));
                    rotate_dir(ct, uni(r), nx, ny, nz);
                    p.dx = nx;
                    p.dy = ny;
                    p.dz = nz;
                    float wsum = 0.f;
                    for (int s = 0; s < S; ++s) wsum = fmaf(rc[s], pick4(kP, P.grp[s]), wsum);
                    float us = uni(r) * wsum, a = 0.f;
                    int sp = S - 1;
                    for (int s = 0; s < S; ++s) {
                        a = fmaf(rc[s], pick4(kP, P.grp[s]), a);
                        if (us <= a) {
                            sp = s;
                            break;
                        }
                    }
                    int q = P.grp[sp];
                    float x = fminf(fmaxf(P.xrow[(size_t)c * G + q], 0.f), (float)(P.ntemp - 1));
                    int lo = min((int)x, P.ntemp - 2);
                    float w = x - (float)lo;
                    int row = lo + (uni(r) < w ? 1 : 0);
                    size_t ro = (size_t)q * P.ntemp + row;
                    ifr = sample_guided(P.cdf + ro * NL, P.guide + ro * RT_NGUIDE, uni(r));
                    n_mrw++;
                }
            }
            tau = -__logf(uni(r));
        } else {
            if (c >= 0) {
                float wf = s_invdr[p.ir] * tw;
#pragma unroll
                for (int q = 0; q < RT_MAX_GROUP; ++q)
                    if (q < G)
                        acc[q] += (unsigned long long)rintf(s_gka[q * NL + ifr] * P.ginvk[q] * wf);
                flush_tally(P, c, acc);
            }
            tau -= aext * tw;
            n_cross++;
            nev_cell = 0;
            if (++nstep > P.max_steps) {  // pathological zero-length wall ping-pong
                n_kill++;
                alive = false;
                continue;
            }
            if (cross_wall(g, p, tw, kind)) {
                max_pkt = max(max_pkt, nev + nstep);
                n_esc++;
                alive = false;
                continue;
            }
            if (p.ir >= 0 && P.visits) atomicAdd(&P.visits[cell_index(g, p)], 1u);
        }
    }
    // ---- statistics: wave-reduce and one atomic per counter per wave
    unsigned long long v[7] = {n_cross, n_abs, n_scat, n_esc, n_kill, n_ev, n_mrw};
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        unsigned long long x = v[k];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off);
        if (lane == 0 && x) atomicAdd(&P.cnt[CNT_CROSS + k], x);
    }
    unsigned long long mp = max_pkt;
    for (int off = 32; off > 0; off >>= 1) mp = max(mp, (unsigned long long)__shfl_down(mp, off));
    if (lane == 0 && mp) atomicMax(&P.cnt[CNT_MAXPKT], mp);
}

// lagged temperatures from the tallies: e = tally * econv * eps, log-log
// interpolation in ln q(T); xrow = fractional row of the T table (-1 if T=0)
__global__ void k_snapshot(const unsigned long long *tally, const double *econv, double eps,
                           const double *lnq, int ntemp, double lnt0, double dlnt, int ncg,
                           int G, float *xrow, float *Tg) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ncg) return;
    int q = i % G;
    double e = (double)tally[i] * econv[i] * eps;
    if (!(e > 0.0)) {
        xrow[i] = -1.f;
        Tg[i] = 0.f;
        return;
    }
    const double *lq = lnq + (size_t)q * ntemp;
    double le = log(e);
    int m;
    if (le <= lq[0])
        m = 0;
    else if (le >= lq[ntemp - 1])
        m = ntemp - 2;
    else {
        int lo = 0, hi = ntemp - 1;
        while (hi - lo > 1) {
            int mid = (lo + hi) >> 1;
            if (lq[mid] <= le)
                lo = mid;
            else
                hi = mid;
        }
        m = lo;
    }
    double x = m + (le - lq[m]) / (lq[m + 1] - lq[m]);
    xrow[i] = (float)x;
    Tg[i] = (float)exp(lnt0 + x * dlnt);
}

__global__ void k_fill_u64(unsigned long long *p, size_t n, unsigned long long v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill_u32(unsigned int *p, size_t n, unsigned int v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill_f32(float *p, size_t n, float v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// per-species output temperature: group temperature where rho_s > 0, else 0
__global__ void k_species_T(const float *Tg, const unsigned char *rhopos, const int *grp, int S,
                            int G, int ncell, float *Tout) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S * ncell) return;
    int s = i / ncell, c = i - s * ncell;
    Tout[i] = rhopos[i] ? Tg[(size_t)c * G + grp[s]] : 0.f;
}

static inline unsigned nblk(size_t n, int b) { return (unsigned)((n + b - 1) / b); }

int thermal_run(rt_model *m, uint64_t nphot, uint64_t seed, const rt_thermal_opts *o,
                rt_thermal_stats *st) {
    rt_ctx *ctx = m->ctx;
    hipStream_t sm = ctx->stream;
    rt_thermal_opts opt{};
    if (o) opt = *o;
    if (nphot == 0) return set_err(ctx, RT_E_ARG, "nphot must be > 0");
    if (m->nstar < 1) return set_err(ctx, RT_E_ARG, "model has no star");
    const uint64_t max_events = opt.max_events ? opt.max_events : 1000000ull;
    uint64_t first = opt.epoch_first ? opt.epoch_first : std::max<uint64_t>(1024, nphot / 4096);
    double growth = opt.epoch_growth > 1.0 ? opt.epoch_growth : 1.25;
    double max_s = opt.max_seconds > 0 ? opt.max_seconds : 60.0;
    int bs = opt.block_size > 0 ? opt.block_size : 256;
    if (bs % 64 || bs > 256) return set_err(ctx, RT_E_ARG, "block_size must be 64..256, multiple of 64");
    int bpc = opt.blocks_per_cu > 0 ? opt.blocks_per_cu : (2048 / bs);
    const int G = m->tab.ngroup;
    const size_t ncg = (size_t)m->ncell * G;
    const double eps = m->tab.lum / (double)nphot;

    // reset tallies, lagged temperatures, counters, visits
    hipLaunchKernelGGL(k_fill_u64, dim3(nblk(ncg, 256)), dim3(256), 0, sm, m->d_tally, ncg, 0ull);
    hipLaunchKernelGGL(k_fill_f32, dim3(nblk(ncg, 256)), dim3(256), 0, sm, m->d_xrow, ncg, -1.f);
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, sm, m->d_cnt, (size_t)CNT_N, 0ull);
    if (opt.count_visits)
        hipLaunchKernelGGL(k_fill_u32, dim3(nblk(m->ncell, 256)), dim3(256), 0, sm, m->d_visits,
                           (size_t)m->ncell, 0u);

    ThermParams P{};
    P.g = dev_grid(m);
    P.nspec = m->nspec;
    P.ngroup = G;
    P.nlam = m->nlam;
    P.ntemp = m->ntemp;
    P.nstar = m->nstar;
    P.hg = m->has_g ? 1 : 0;
    P.rhoL = m->d_rhoL;
    P.ka = m->d_ka;
    P.ks = m->d_ks;
    P.gg = m->d_g;
    P.grp = m->d_grp;
    P.gka = m->d_gka;
    P.ginvk = m->d_ginvk;
    P.invdr = m->d_invdr;
    P.cdf = m->d_cdf;
    P.guide = m->d_guide;
    P.scdf = m->d_scdf;
    P.sguide = m->d_sguide;
    P.spick = m->d_spick;
    P.spos = m->d_spos;
    P.xrow = m->d_xrow;
    P.kross = m->d_kross;
    P.mrwy = m->d_mrwy;
    P.kplanck = m->d_kplanck;
    P.tally = m->d_tally;
    P.visits = opt.count_visits ? m->d_visits : nullptr;
    P.cnt = m->d_cnt;
    P.seed = seed;
    P.max_events = max_events;
    P.max_steps = 100000000u;
    P.max_ticks = (uint64_t)(max_s * (double)ctx->wallclock_khz * 1000.0);
    P.mrw = opt.mrw;
    P.mrw_count = opt.mrw_count_trigger > 0 ? opt.mrw_count_trigger : 100;
    P.mrw_tau = (float)(opt.mrw_tauthres > 0 ? opt.mrw_tauthres : 10.0);
    P.mrw_gamma = (float)(opt.mrw_gamma > 0 ? opt.mrw_gamma : 2.0);
    size_t lds = sizeof(float) * ((size_t)m->nspec * m->nlam * (P.hg ? 3 : 2) + (size_t)G * m->nlam +
                                  2 * (size_t)(m->nr + 1) + (size_t)m->nr + 3 * (size_t)(m->nt + 1) +
                                  3 * (size_t)(m->np + 1));
    if (lds > 64 * 1024) return set_err(ctx, RT_E_LIMIT, "thermal kernel LDS %zu B > 64 KiB", lds);
    unsigned grid = (unsigned)(ctx->n_cu * bpc);

    std::vector<hipEvent_t> ev;
    auto mk = [&]() {
        hipEvent_t e;
        (void)hipEventCreate(&e);
        ev.push_back(e);
        return e;
    };
    hipEvent_t e_begin = mk();
    HIPCHK(ctx, hipEventRecord(e_begin, sm));
    std::vector<std::pair<hipEvent_t, hipEvent_t>> kev;
    uint64_t done = 0;
    int nep = 0;
    while (done < nphot) {
        uint64_t sz = (uint64_t)std::floor((double)done * (growth - 1.0));
        if (sz < first) sz = first;
        uint64_t end = std::min(nphot, done + sz);
        P.pbegin = done;
        P.pend = end;
        hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, sm, m->d_cnt, (size_t)1, 0ull);
        hipEvent_t a = mk(), b = mk();
        (void)hipEventRecord(a, sm);
        hipLaunchKernelGGL(k_thermal, dim3(grid), dim3(bs), lds, sm, P);
        (void)hipEventRecord(b, sm);
        kev.push_back({a, b});
        hipLaunchKernelGGL(k_snapshot, dim3(nblk(ncg, 256)), dim3(256), 0, sm, m->d_tally,
                           m->d_econv, eps, m->d_lnq, m->ntemp, m->tab.lnt0, m->tab.dlnt,
                           (int)ncg, G, m->d_xrow, m->d_Tg);
        HIPCHK(ctx, hipGetLastError());
        done = end;
        nep++;
    }
    hipLaunchKernelGGL(k_species_T, dim3(nblk((size_t)m->nspec * m->ncell, 256)), dim3(256), 0, sm,
                       m->d_Tg, m->d_rhopos, m->d_grp, m->nspec, G, m->ncell, m->d_Tout);
    hipEvent_t e_end = mk();
    HIPCHK(ctx, hipEventRecord(e_end, sm));
    HIPCHK(ctx, hipEventSynchronize(e_end));
    unsigned long long cnt[CNT_N];
    HIPCHK(ctx, hipMemcpy(cnt, m->d_cnt, sizeof(cnt), hipMemcpyDeviceToHost));
    float tot = 0, kms = 0;
    (void)hipEventElapsedTime(&tot, e_begin, e_end);
    float kmax = 0;
    for (auto &pr : kev) {
        float x = 0;
        (void)hipEventElapsedTime(&x, pr.first, pr.second);
        kms += x;
        kmax = std::max(kmax, x);
    }
    for (auto e : ev) (void)hipEventDestroy(e);
    m->eps_last = eps;
    m->have_T = true;
    if (st) {
        st->n_cross = cnt[CNT_CROSS];
        st->n_abs = cnt[CNT_ABS];
        st->n_scat = cnt[CNT_SCAT];
        st->n_escape = cnt[CNT_ESCAPE];
        st->n_killed = cnt[CNT_KILLED];
        st->n_events = cnt[CNT_EVENTS];
        st->n_mrw = cnt[CNT_MRW];
        st->eps = eps;
        st->lum = m->tab.lum;
        st->n_epochs = nep;
        st->watchdog_fired = cnt[CNT_ABORT] ? 1 : 0;
        st->kernel_ms = kms;
        st->total_ms = tot;
        st->max_epoch_ms = kmax;
        st->max_packet_steps = cnt[CNT_MAXPKT];
    }
    if (cnt[CNT_ABORT])
        return set_err(ctx, RT_E_WATCHDOG, "thermal kernel exceeded %.1f s; packets were killed",
                       max_s);
    if (cnt[CNT_OVERFLOW]) return set_err(ctx, RT_E_OVERFLOW, "fixed-point tally overflow");
    return RT_OK;
}
