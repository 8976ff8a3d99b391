// thermal.hip -- thermal Monte Carlo on gfx950: what RADMC-3D `mctherm` does for
// pdspy's Model.run_thermal (pdspy/modeling/Model.py:221-256, radmc3d/run.py:9-38).
//
// Algorithm (Bjorkman & Wood 2001 with a path-length absorbed-energy estimator):
//  * nphot packets of energy eps = L/nphot leave blackbody point stars; random
//    event k of packet p (emission, interaction, modified-random-walk step)
//    uses the Philox4x32-10 blocks with counter (k, p, 0|1) and key = seed.
//  * flights of optical depth -ln(xi) through the spherical grid; every segment
//    of length s in cell c adds kabs_g(nu) * s to the cell's absorbed-energy
//    tally of absorption group g (u64 fixed point, unit 2^-28 kmax_g dr_c), so
//    the sums are exact and independent of execution order;
//  * at an interaction the packet scatters with probability alpha_sca/alpha_ext
//    (isotropic or Henyey-Greenstein) or is absorbed and re-emitted at once,
//    isotropically, with a frequency drawn from kabs dB/dT at the cell's
//    temperature (interpolated between rows of the (ntemp, temp0, temp1) table
//    by a mixture draw); optionally the modified random walk replaces long
//    random walks deep inside optically thick cells;
//  * the packets are split into epochs [N_k, N_k+1), N_k+1 = N_k +
//    max(first, N_k (growth-1)); a packet of epoch e deposits into temperature
//    slot e + floor(interactions / slice) and re-emits with the temperatures
//    implied by the deposits of slots <= e-2.  The result is bit-reproducible
//    for a seed however packets are scheduled.
//
// Execution: ONE persistent launch per run.  Each lane carries one packet; a
// lane whose packet has finished takes the next one at once (one atomic per
// wavefront per refill, from a per-epoch counter).  Epoch e+2 may start as soon
// as slot e is complete -- every packet of epochs <= e has finished or moved on
// to a later slot -- and its snapshot is published, so a long random walk
// holds up the pipeline for at most one slice of interactions.
// Snapshots are computed cooperatively: any wave that finds the next epoch
// complete takes a 64-entry chunk, folds its working tallies into the
// cumulative ones and writes the temperature rows; the wave finishing the last
// chunk publishes it.
//
// Memory model: the L2 of each XCD is not coherent with the others, and an
// agent-scope acquire/release would invalidate/write back a whole L2.  So no
// such fence appears here: counters, flags and tallies are only touched by
// device-coherent atomics, each epoch's snapshot gets its own slot that is
// written once before its flag is raised and never rewritten during the run
// (so no cache can hold a stale copy); a release waits for the wave's
// outstanding memory operations (order_release), an acquire is a branch on
// the acquired value.
// Opacities and grid walls live in LDS; the current cell's densities and
// temperature rows stay in registers.
//
// Batches (BATCH instance): one launch serves many runs, each with its own
// parameter block, pipeline and tallies; a wavefront serves one run at a time
// and moves to another when its run has nothing to hand out.  A wave left with
// a few live packets while its run waits for a snapshot parks them in the
// run's pool (their whole state: position, direction, cell, Philox counter,
// frequency, remaining optical depth, event counts, slot) and moves on;
// waves of that run adopt parked packets into dead lanes.  A packet's result
// does not depend on which wave steps it, so every run's temperatures equal
// those of its plain run bit for bit.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "rt_internal.h"

using namespace rt;

#define RT_MAX_EPOCH 256

// RT_THERMAL_PROFILE=1 adds shader-clock accounting of the loop phases
#ifndef RT_THERMAL_PROFILE
#define RT_THERMAL_PROFILE 0
#endif
#define RT_NPHASE 8

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
    const float *kross;           // [S][ntemp] Rosseland-mean extinction
    const float *kplanck;         // [G][ntemp] Planck-mean absorption
    const float *mrwy;            // [RT_MRW_NY] escape-law inverse
    unsigned int *visits;         // nullable
    unsigned long long *cnt;
    uint64_t seed, max_events;
    uint32_t max_steps;  // wall crossings per packet
    int mrw, mrw_count;
    float mrw_tau, mrw_gamma;
    // epoch pipeline
    int nepoch, ncg, nchunk;
    const unsigned long long *ep_start;  // [E+1]
    unsigned long long *ep_claim;        // [E]
    unsigned long long *slot_start;      // [E] packets of epoch e that have started
    unsigned long long *slot_in;         // [E] packets that moved into slot e
    unsigned long long *slot_out;        // [E] packets that left slot e (moved on or ended)
    uint32_t slice;                      // interactions per slot
    unsigned int *snap_claim;            // [E]
    unsigned int *snap_done;             // [E]
    unsigned int *snap_ready;            // [E]
    unsigned int *next_snap;             // [1]
    unsigned long long *snap_time;       // [E] clock at each snapshot, then t_begin, t_end
    unsigned long long *ep_cur;          // [1] first epoch with unclaimed packets
    unsigned int *runs_done;             // [1] of the batch: runs with all snapshots out
    // batch only: pool of parked packets (a ring of park_cap states of
    // RT_PARK_WORDS words; flag = index + 1 once a state is written)
    unsigned int *park;
    unsigned int *park_flag;
    unsigned long long *park_head, *park_tail;
    uint32_t park_cap;
    unsigned long long *tslot;           // [E+1][slot] tallies by deposit slot (last: all later)
    unsigned long long *cum;             // [cell][G] tallies of all folded epochs
    float *xsnap;                        // [E][slot] temperature-row snapshots, one per epoch
    double *slot_sum;                    // [E+1] sum of all adds into each deposit slot
    size_t slot;                         // floats per snapshot slot (>= cells * G, aligned)
    const double *econv;                 // [cell][G]
    const double *lnq;                   // [G][ntemp]
    double eps, lnt0, dlnt;
};

// one launch over a batch of runs
struct BatchArgs {
    int nrun;                      // runs (parameter blocks passed beside, device memory)
    int lds_wave;                  // floats of LDS per wavefront
    uint64_t max_ticks;            // watchdog (wall-clock ticks)
    unsigned long long *abort;     // [1] watchdog flag
    unsigned int *runs_done;       // [1] runs whose last snapshot is published
    int *parked;                   // [1] packets parked in the runs' pools (batch)
};

#define RT_PARK_WORDS 22   // words of a parked packet's state
#ifndef RT_SNAP_POLL
#define RT_SNAP_POLL 64u  // a busy wave looks at its run's next snapshot every this many iterations
#endif
#ifndef RT_NUM_XCD
#define RT_NUM_XCD 8  // MI355X: 8 XCDs of 32 CUs, workgroups dealt to them in turn
#endif
#ifndef RT_PARK_MAXLIVE
#define RT_PARK_MAXLIVE 8  // a wave with at most this many live packets may park them
#endif

// ---------------------------------------------------------------- helpers
__device__ __forceinline__ unsigned ld_rlx32(const unsigned *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long ld_rlx64(const unsigned long long *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_rlx64(unsigned long long *p, unsigned long long v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_rlx32(unsigned *p, unsigned v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Ordering of the wave's device-coherent (agent-scope atomic) memory operations
// without L2 write-back or invalidation, which the data they order does not
// need.  Release: every earlier load, store and atomic of the wave has
// completed before any later one is issued -- a workgroup-scope fence alone is
// only a compiler barrier on gfx950 and emits no wait, so the wait is explicit.
// Acquire: a compiler barrier; the loads it orders are issued after a branch
// on the value that was acquired, so they follow it in hardware too.
__device__ __forceinline__ void order_release() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
}
__device__ __forceinline__ void order_acquire() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup"); }
__device__ __forceinline__ void count_up(unsigned long long *p) {
    order_release();
    __hip_atomic_fetch_add(p, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ unsigned uniform32(unsigned v) { return __builtin_amdgcn_readfirstlane(v); }
// value of lane `src` (wave-uniform src), known to the compiler to be uniform
__device__ __forceinline__ unsigned long long lane64(unsigned long long v, int src) {
    unsigned lo = __builtin_amdgcn_readlane((unsigned)v, src);
    unsigned hi = __builtin_amdgcn_readlane((unsigned)(v >> 32), src);
    return ((unsigned long long)hi << 32) | lo;
}

// smallest i with u <= cdf[i], starting from the guide entry
__device__ __forceinline__ int sample_guided(const float *cdf, const unsigned char *guide,
                                             float u) {
    int i = guide[(int)(u * (float)RT_NGUIDE)];
    while (cdf[i] < u) ++i;
    return i;
}

// a[i] for a register array without dynamic indexing (which would spill)
template <int N, typename T>
__device__ __forceinline__ T pickn(const T (&a)[N], int i) {
    T v = a[0];
#pragma unroll
    for (int k = 1; k < N; ++k) v = (i == k) ? a[k] : v;
    return v;
}

// fractional temperature row for an absorbed power per gram e (ln-q table)
__device__ __forceinline__ float xrow_of(const double *lq, int ntemp, double e) {
    if (!(e > 0.0)) return -1.f;
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
    return (float)(m + (le - lq[m]) / (lq[m + 1] - lq[m]));
}

// Adds the packet's pending deposits to the tallies of cell c with atomics
// whose result nobody waits for.  Overflow is ruled out by bounds instead of
// by the values the atomics find: every add stays below 2^62 (ovf), and the
// packet's adds into its current slot are summed (ssum, FP64) and added to the
// slot's total when the packet leaves the slot; a slot whose total of adds
// stays below 2^61 has no tally near 2^64 (checked after the launch).
template <int MG>
__device__ __forceinline__ void flush_tally(const ThermParams &P, unsigned long long *tb, int c,
                                            unsigned long long (&acc)[MG],
                                            unsigned long long &ovf, double &ssum) {
#pragma unroll
    for (int q = 0; q < MG; ++q) {
        if (q < P.ngroup && acc[q]) {
            ovf |= acc[q];
            ssum += (double)acc[q];
            __hip_atomic_fetch_add(&tb[(size_t)c * P.ngroup + q], acc[q], __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            acc[q] = 0;
        }
    }
}

// the packet leaves slot `slot`: its adds there join the slot's total
__device__ __forceinline__ void flush_ssum(const ThermParams &P, int slot, double &ssum) {
    if (ssum > 0.0) unsafeAtomicAdd(&P.slot_sum[slot], ssum);
    ssum = 0.0;
}

// One chunk of the next temperature snapshot, if its epoch is complete.
// Called with the whole wave active.  Returns true if the wave did work.
__device__ bool snapshot_help(const ThermParams &P, int lane) {
    unsigned k = uniform32(ld_rlx32(P.next_snap));
    if (k >= (unsigned)P.nepoch) return false;
    // slot k is complete when all of epoch k has started and every packet that
    // entered it has left (slots < k are complete: their snapshots are out)
    unsigned long long size = P.ep_start[k + 1] - P.ep_start[k];
    unsigned long long st = __shfl(ld_rlx64(&P.slot_start[k]), 0);
    if (st < size) return false;
    unsigned long long in = __shfl(ld_rlx64(&P.slot_in[k]), 0);
    unsigned long long out = __shfl(ld_rlx64(&P.slot_out[k]), 0);
    if (out < st + in) return false;
    unsigned ch = 0;
    if (lane == 0) ch = atomicAdd(&P.snap_claim[k], 1u);
    ch = uniform32(__shfl(ch, 0));
    if (ch >= (unsigned)P.nchunk) return false;
    order_acquire();
    int i = (int)ch * 64 + lane;
    if (i < P.ncg) {
        unsigned long long cv = ld_rlx64(P.cum + i) + ld_rlx64(P.tslot + (size_t)k * P.slot + i);
        st_rlx64(P.cum + i, cv);
        int q = i % P.ngroup;
        float x = xrow_of(P.lnq + (size_t)q * P.ntemp, P.ntemp, (double)cv * P.econv[i] * P.eps);
        __hip_atomic_store(P.xsnap + (size_t)k * P.slot + i, x, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
    order_release();
    if (lane == 0) {
        unsigned old = atomicAdd(&P.snap_done[k], 1u);
        if (old + 1u == (unsigned)P.nchunk) {
            order_release();
            st_rlx32(&P.snap_ready[k], 1u);
            st_rlx32(P.next_snap, k + 1u);
            st_rlx64(&P.snap_time[k], wall_clock64());
            if (k + 1u == (unsigned)P.nepoch) atomicAdd(P.runs_done, 1u);
        }
    }
    return true;
}

// Wave-level helpers of the batched loop --------------------------------------

// Stages run R's opacities and grid walls into this wave's LDS slice and
// points g and the table pointers at them.
__device__ __forceinline__ void stage_run(const ThermParams &R, float *wl, int lane, DevGrid &g,
                                          float *&s_ka, float *&s_ks, float *&s_gka,
                                          float *&s_g, float *&s_invdr) {
    const int S = R.nspec, NL = R.nlam, G = R.ngroup;
    const int NR = R.g.nr, NT = R.g.nt, NP = R.g.np;
    s_ka = wl;
    s_ks = s_ka + S * NL;
    s_gka = s_ks + S * NL;
    s_g = s_gka + G * NL;
    float *rw = s_g + (R.hg ? S * NL : 0);
    float *rw2 = rw + NR + 1;
    float *tw = rw2 + NR + 1;
    float *ct = tw + NT + 1;
    float *st = ct + NT + 1;
    float *pw = st + NT + 1;
    float *cp = pw + NP + 1;
    float *sp = cp + NP + 1;
    s_invdr = sp + NP + 1;
    for (int i = lane; i < S * NL; i += 64) {
        s_ka[i] = R.ka[i];
        s_ks[i] = R.ks[i];
        if (R.hg) s_g[i] = R.gg[i];
    }
    for (int i = lane; i < G * NL; i += 64) s_gka[i] = R.gka[i] * R.ginvk[i / NL];
    for (int i = lane; i <= NR; i += 64) {
        rw[i] = R.g.rw[i];
        rw2[i] = R.g.rw2[i];
        if (i < NR) s_invdr[i] = R.invdr[i];
    }
    for (int i = lane; i <= NT; i += 64) {
        tw[i] = R.g.tw[i];
        ct[i] = R.g.ct[i];
        st[i] = R.g.st[i];
    }
    for (int i = lane; i <= NP; i += 64) {
        pw[i] = R.g.pw[i];
        cp[i] = R.g.cp[i];
        sp[i] = R.g.sp[i];
    }
    // the wave reads what other lanes of it wrote
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    g = R.g;
    g.rw = rw;
    g.rw2 = rw2;
    g.tw = tw;
    g.ct = ct;
    g.st = st;
    g.pw = pw;
    g.cp = cp;
    g.sp = sp;
}

// First run, from `start` on, whose next unclaimed epoch may start now (its
// snapshot is published); -1 if none.  e_out: that epoch.
template <bool BATCH>
__device__ __forceinline__ const ThermParams &run_at(const ThermParams *__restrict__ runs,
                                                     const ThermParams &P1, int r) {
    return BATCH ? runs[r] : P1;
}

template <bool BATCH>
__device__ __forceinline__ bool pool_nonempty(const ThermParams &R) {
    if (!BATCH || !R.park) return false;
    return uniform32((unsigned)(ld_rlx64(R.park_tail) > ld_rlx64(R.park_head))) != 0u;
}

// whether run R's next unclaimed epoch may start now
__device__ __forceinline__ bool epoch_claimable(const ThermParams &R, unsigned &e) {
    e = uniform32((unsigned)ld_rlx64(R.ep_cur));
    return e < (unsigned)R.nepoch && (e < 2u || uniform32(ld_rlx32(&R.snap_ready[e - 2])) != 0u);
}

// another run (not `self`) with an epoch to hand out
template <bool BATCH>
__device__ __forceinline__ bool other_epoch_work(const BatchArgs &A,
                                                 const ThermParams *__restrict__ runs,
                                                 const ThermParams &P1, int self) {
    if (!BATCH) return false;
    for (int r = 0; r < A.nrun; ++r) {
        unsigned e;
        if (r != self && epoch_claimable(runs[r], e)) return true;
    }
    return false;
}

// First run, from `start` on, with work to hand out: an epoch that may start
// now or (batch) parked packets -- except the pool of run `skip`, whose packets
// the calling wave has just parked; -1 if none.  e_out: the run's next epoch.
template <bool BATCH>
__device__ __forceinline__ int select_run(const BatchArgs &A,
                                          const ThermParams *__restrict__ runs,
                                          const ThermParams &P1, int start, int skip, int &e_out) {
    if (!BATCH) {
        const unsigned e = uniform32((unsigned)ld_rlx64(P1.ep_cur));
        if (e < (unsigned)P1.nepoch &&
            (e < 2u || uniform32(ld_rlx32(&P1.snap_ready[e - 2])) != 0u)) {
            e_out = (int)e;
            return 0;
        }
        return -1;
    }
    int r = start;
    for (int k = 0; k < A.nrun; ++k) {
        const ThermParams &R = run_at<BATCH>(runs, P1, r);
        unsigned e;
        const bool ok = epoch_claimable(R, e);
        if (ok || (r != skip && pool_nonempty<BATCH>(R))) {
            e_out = (int)e;
            return r;
        }
        r = r + 1 == A.nrun ? 0 : r + 1;
    }
    return -1;
}

// parked packet state: device-coherent word stores and loads
__device__ __forceinline__ void park_put(unsigned *w, int k, unsigned v) {
    __hip_atomic_store(w + k, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned park_get(const unsigned *w, int k) {
    return __hip_atomic_load(w + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Adds a wave's counters for run R (one atomic per counter per wave) and
// clears them.
// The wave's event counts live in LDS, one row of 9 words per wave (8 counts
// and the largest packet's steps): counted per lane they held 9 VGPRs, which
// pushed the kernel past 128 into scratch -- and a launch that needs scratch
// held other streams' kernels back until its waves retired.  In a branch the
// first active lane adds the number of active lanes.
#define RT_CNT_WORDS 9
__device__ __forceinline__ void wave_count(unsigned *row, int k, int lane) {
    const uint64_t act = __ballot(1);
    if (lane == __ffsll((unsigned long long)act) - 1) row[k] += (unsigned)__popcll(act);
}

// adds the wave's counts to run R's counters and clears them (whole wave)
__device__ __forceinline__ void flush_counts(const ThermParams &R, int lane, unsigned *row,
                                             unsigned long long &ovf) {
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            if (row[k]) atomicAdd(&R.cnt[CNT_CROSS + k], (unsigned long long)row[k]);
            row[k] = 0u;
        }
        if (row[8]) atomicMax(&R.cnt[CNT_MAXPKT], (unsigned long long)row[8]);
        row[8] = 0u;
    }
    if (__any((ovf >> 62) != 0ull) && lane == 0) atomicOr(&R.cnt[CNT_OVERFLOW], 1ull);
    ovf = 0;
}

// waves per SIMD the register budget of the one-species variants is set for
// (4: 127 VGPRs, no spills; 5 spills about 30 VGPRs to scratch)
#ifndef RT_THERMAL_WAVES
#define RT_THERMAL_WAVES 4
#endif

// One persistent launch for a batch of runs (one model each).  Each wavefront
// serves one run at a time: its LDS slice holds that run's opacities and
// walls; when it has no packet left and the run has none to hand out (all
// claimed, or the next epoch waits for its snapshot) it moves to the next run
// that has.  MS, MG: compile-time bounds on species and absorption groups
// (register caches); two species in one absorption group fit 128 VGPRs (four
// waves per SIMD), up to four species get three waves' registers and more
// species two.  BATCH: the parameter blocks are read from `runs` in
// device memory; otherwise the single run's block P1 is a kernel argument.
template <int MS, int MG, bool BATCH>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MS <= 2 && MG <= 1 ? RT_THERMAL_WAVES : MS <= 4 ? 3 : 2)))
k_thermal(BatchArgs A, const ThermParams *__restrict__ runs, ThermParams P1) {
    extern __shared__ float lds[];
    const int lane = __lane_id();
    float *const wl = lds + (size_t)(threadIdx.x >> 6) * A.lds_wave;  // this wave's slice
    const uint64_t t_start = wall_clock64();
    if (blockIdx.x == 0 && threadIdx.x == 0)
        for (int r = 0; r < A.nrun; ++r)
            st_rlx64(&run_at<BATCH>(runs, P1, r).snap_time[RT_MAX_EPOCH], t_start);

    // the run this wave serves (wave-uniform; -1: none yet) and its staged data
    int m = -1;
    const ThermParams *Pp = BATCH ? runs : &P1;
#define P (*Pp)
    int S = 0, NL = 0, G = 0, E = 0;
    DevGrid g = P.g;  // LDS pointers set below, FP64 walls stay in global memory
    g.rw = g.rw2 = g.tw = g.ct = g.st = g.pw = g.cp = g.sp = wl;
    float *s_ka = wl, *s_ks = wl, *s_gka = wl, *s_g = wl, *s_invdr = wl;
    float ginv[MG];  // 2^shift / kmax per group
    int grpr[MS];    // absorption group of each species
#pragma unroll
    for (int q = 0; q < MG; ++q) ginv[q] = 0.f;
#pragma unroll
    for (int s = 0; s < MS; ++s) grpr[s] = 0;
    int scan = (int)((blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) % (unsigned)A.nrun);
    if (!BATCH) {
        // one run: staged once, before the loop
        m = 0;
        scan = 0;
        stage_run(P1, wl, lane, g, s_ka, s_ks, s_gka, s_g, s_invdr);
        S = P1.nspec;
        NL = P1.nlam;
        G = P1.ngroup;
        E = P1.nepoch;
#pragma unroll
        for (int q = 0; q < MG; ++q) ginv[q] = q < G ? P1.ginvk[q] : 0.f;
#pragma unroll
        for (int s = 0; s < MS; ++s) grpr[s] = s < S ? P1.grp[s] : 0;
    }

    Pos p;
    uint64_t pid = 0;   // packet id (Philox counter words 1, 2)
    uint32_t krng = 0;  // its next random event (Philox counter word 0)
    int ifr = 0;
    float tau = 0.f;
    uint32_t nev = 0, nstep = 0, nev_cell = 0, nzero = 0;
    uint32_t slice_left = 0;  // interactions until the packet moves to the next slot
    bool alive = false;
    int pe = 0;                                // epoch of this lane's packet
    int slot = 0;                              // its current deposit slot
    const float *xs = nullptr;                 // its temperature snapshot (null: T = 0)
    unsigned long long *tb = nullptr;          // its deposit slot's tallies
    int ccell = -1;                            // cell whose data sit in registers
    float aRc = -1.f;                          // its Rosseland extinction (-1: not yet)
    float rcache[MS];
    float xcache[MG];
#pragma unroll
    for (int s = 0; s < MS; ++s) rcache[s] = 0.f;
#pragma unroll
    for (int q = 0; q < MG; ++q) xcache[q] = -1.f;
    unsigned long long acc[MG];
    unsigned long long ovf = 0;  // OR of the adds (overflow check)
    double ssum = 0.0;           // the packet's adds into its current slot
#pragma unroll
    for (int q = 0; q < MG; ++q) acc[q] = 0;
    // per-run wave counters: crossings, absorptions, scatterings, escapes,
    // kills, interactions, MRW steps, nudges
    __shared__ unsigned s_cnt[4 * RT_CNT_WORDS];
    unsigned *const crow = s_cnt + (threadIdx.x >> 6) * RT_CNT_WORDS;
    if (lane < RT_CNT_WORDS) crow[lane] = 0u;
#define RT_COUNT(k) wave_count(crow, k, lane)
#define RT_MAXPKT(v) atomicMax(&crow[8], (unsigned)(v))
    int cur_e = 0;          // wave-uniform: epoch packets are claimed from
    bool epoch_ok = false;  // wave-uniform: its snapshot has been acquired
    uint64_t n_idle = 0;
    uint32_t idle_streak = 0;  // consecutive idle polls (back-off)
    uint32_t iter = 0;
    uint32_t last_adopt = 0;  // iteration of the wave's last adoption of parked packets
    int parked_from = -1;     // run whose pool the wave has just filled (not to be emptied by it)
    int prio = 0;  // wave priority set with s_setprio
#if RT_THERMAL_PROFILE
    // wave time by phase: refill, loop overhead, snapshot, idle, geometry,
    // crossing, interaction, modified random walk.  The wave's last time stamp
    // lives in LDS and the first active lane books each interval, so time
    // spent in divergent branches is counted once per wave.
    __shared__ unsigned long long s_tc[4];
    uint64_t tk[RT_NPHASE];
    for (int k = 0; k < RT_NPHASE; ++k) tk[k] = 0;
    const int wslot = threadIdx.x >> 6;
    if (lane == 0) s_tc[wslot] = clock64();
#define RT_PHASE(k)                                                      \
    do {                                                                 \
        uint64_t tn_ = clock64();                                        \
        if (lane == __ffsll((unsigned long long)__ballot(1)) - 1) {     \
            tk[k] += tn_ - s_tc[wslot];                                  \
            s_tc[wslot] = tn_;                                           \
        }                                                                \
    } while (0)
#else
#define RT_PHASE(k) \
    do {            \
    } while (0)
#endif

    for (;; ++iter) {
        RT_PHASE(1);
        // wave-uniform state, stated as such for the compiler (the loop is
        // re-entered from lane-divergent paths)
        m = (int)uniform32((unsigned)m);
        cur_e = (int)uniform32((unsigned)cur_e);
        epoch_ok = uniform32(epoch_ok ? 1u : 0u) != 0u;
        scan = (int)uniform32((unsigned)scan);
        S = (int)uniform32((unsigned)S);
        NL = (int)uniform32((unsigned)NL);
        G = (int)uniform32((unsigned)G);
        E = (int)uniform32((unsigned)E);
        Pp = BATCH ? runs + (m < 0 ? 0 : m) : &P1;
        const bool any0 = __any(alive);

        // ---- watchdog (wave-uniform check)
        if ((iter & 1023u) == 0u) {
            bool stop = (wall_clock64() - t_start) > A.max_ticks ||
                        uniform32((unsigned)ld_rlx64(A.abort)) != 0u;
            if (stop) {
                if (lane == 0) atomicOr(A.abort, 1ull);
                if (alive) {
                    if (p.ir >= 0) flush_tally(P, tb, cell_index(g, p), acc, ovf, ssum);
                    flush_ssum(P, slot, ssum);
                    RT_MAXPKT(nev + nstep);
                    RT_COUNT(4);
                    alive = false;
                }
                break;
            }
        }

        // ---- a wave without packets whose run has none to hand out moves on
        if (!any0) {
            bool can = m >= 0 && cur_e < E;
            if (can && !epoch_ok) {
                epoch_ok = cur_e < 2 || uniform32(ld_rlx32(&P.snap_ready[cur_e - 2])) != 0u;
                if (epoch_ok) order_acquire();
                can = epoch_ok;
            }
            if (!can && m >= 0 && m != parked_from) can = pool_nonempty<BATCH>(P);
            if (!can) {
                int e_new = 0;
                const int r = (int)uniform32(
                    (unsigned)select_run<BATCH>(A, runs, P1, scan, parked_from, e_new));
                e_new = (int)uniform32((unsigned)e_new);
                if (r >= 0) {
                    order_acquire();
                    if (BATCH && r != m) {
                        if (m >= 0) flush_counts(P, lane, crow, ovf);
                        m = r;
                        Pp = BATCH ? runs + m : &P1;
                        stage_run(P, wl, lane, g, s_ka, s_ks, s_gka, s_g, s_invdr);
                        S = P.nspec;
                        NL = P.nlam;
                        G = P.ngroup;
                        E = P.nepoch;
#pragma unroll
                        for (int q = 0; q < MG; ++q) ginv[q] = q < G ? P.ginvk[q] : 0.f;
#pragma unroll
                        for (int s = 0; s < MS; ++s) grpr[s] = s < S ? P.grp[s] : 0;
                        cur_e = e_new;
                    } else {
                        cur_e = max(cur_e, e_new);
                    }
                    // one run: select_run found its epoch claimable; batch: the run may
                    // have been picked for its pool, checked again before the next claim
                    epoch_ok = !BATCH;
                    scan = r;
                    parked_from = -1;
                    idle_streak = 0;
                } else {
                    // nothing to hand out anywhere: help a snapshot, finish or wait
                    if (uniform32(ld_rlx32(A.runs_done)) >= (unsigned)A.nrun &&
                        (!BATCH || uniform32((unsigned)__hip_atomic_load(
                                       A.parked, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) == 0u))
                        break;
                    const bool helped = snapshot_help(run_at<BATCH>(runs, P1, scan), lane);
                    // one run whose packets are all handed out: a wave without
                    // packets and without snapshot work leaves the launch, so a
                    // launch queued behind it (another run on another stream)
                    // takes over its CU while the last walks finish here.  The
                    // waves holding those walks stay and publish the remaining
                    // snapshots, which no packet of this run reads any more.
                    if (!BATCH && !helped &&
                        uniform32((unsigned)ld_rlx64(P1.ep_cur)) >= (unsigned)P1.nepoch)
                        break;
                    scan = scan + 1 == A.nrun ? 0 : scan + 1;
                    parked_from = -1;  // no other run has work: its own pool is fair game
                    n_idle++;
                    // back off while nothing turns up: every idle poll reads all
                    // runs' counters, load the remaining walks share
                    idle_streak = helped ? 0u : idle_streak + 1u;
                    if (idle_streak < 8u) {
                        __builtin_amdgcn_s_sleep(32);
                    } else {
                        const int naps = idle_streak < 64u ? 1 : 4;
                        for (int z = 0; z < naps; ++z) __builtin_amdgcn_s_sleep(127);
                    }
                    RT_PHASE(3);
                    continue;
                }
            }
        }

        // ---- help with the run's next temperature snapshot
        if ((iter & (RT_SNAP_POLL - 1u)) == 0u) {
            snapshot_help(P, lane);
            // a wave holding a long random walk is issued first on its SIMD:
            // those walks are what the end of a run waits for
            const bool l2 = __any(alive && nev > 200000u), l1 = __any(alive && nev > 20000u);
            const int want = l2 ? 2 : (l1 ? 1 : 0);
            if (want != prio) {
                if (want == 2)
                    __builtin_amdgcn_s_setprio(2);
                else if (want == 1)
                    __builtin_amdgcn_s_setprio(1);
                else
                    __builtin_amdgcn_s_setprio(0);
                prio = want;
            }
        }
        RT_PHASE(2);

        // ---- batch: a wave left with a few live packets and nothing to hand out
        // in its run parks them in the run's pool and moves on to another run
        // (a sparse wave costs as many issue slots as a full one)
        if (BATCH && (iter & 15u) == 0u && m >= 0 && P.park &&
            iter - last_adopt > 2048u) {
            const uint64_t live = __ballot(alive);
            const int n = __popcll(live);
            unsigned e_chk;
            if (n > 0 && n <= RT_PARK_MAXLIVE && !epoch_claimable(P, e_chk) &&
                !pool_nonempty<BATCH>(P) && other_epoch_work<BATCH>(A, runs, P1, m)) {
                const int leader = (int)uniform32((unsigned)(__ffsll((unsigned long long)live) - 1));
                unsigned long long base = 0, ok = 0;
                if (lane == leader) {
                    // counted first, so the count never misses a parked packet
                    __hip_atomic_fetch_add(A.parked, n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    for (int t = 0; t < 8 && !ok; ++t) {
                        unsigned long long tl = ld_rlx64(P.park_tail), hd = ld_rlx64(P.park_head);
                        if (tl + (unsigned long long)n - hd > P.park_cap) break;
                        unsigned long long ex = tl;
                        if (__hip_atomic_compare_exchange_strong(P.park_tail, &ex, tl + n,
                                                                 __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                                 __HIP_MEMORY_SCOPE_AGENT)) {
                            base = tl;
                            ok = 1;
                        }
                    }
                    if (!ok)
                        __hip_atomic_fetch_add(A.parked, -n, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT);
                    else
                        atomicAdd(&P.cnt[CNT_PARKED], (unsigned long long)n);
                }
                ok = lane64(ok, leader);
                base = lane64(base, leader);
                if (ok && alive) {
                    if (p.ir >= 0) flush_tally(P, tb, cell_index(g, p), acc, ovf, ssum);
                    flush_ssum(P, slot, ssum);
                    const unsigned long long idx = base + __popcll(live & ((1ull << lane) - 1ull));
                    const size_t s0 = (size_t)(idx % P.park_cap);
                    unsigned *w = P.park + s0 * RT_PARK_WORDS;
                    // the state last held here (a lap of the ring earlier) has been
                    // taken; its reader clears the flag once it has read it
                    bool free_ = false;
                    for (int wt = 0; wt < (1 << 22) && !free_; ++wt) {
                        free_ = ld_rlx32(&P.park_flag[s0]) == 0u;
                        if (!free_) __builtin_amdgcn_s_sleep(2);
                    }
                    if (!free_) atomicOr(A.abort, 1ull);  // (bounded) the run is invalid
                    park_put(w, 0, __float_as_uint(p.x));
                    park_put(w, 1, __float_as_uint(p.y));
                    park_put(w, 2, __float_as_uint(p.z));
                    park_put(w, 3, __float_as_uint(p.dx));
                    park_put(w, 4, __float_as_uint(p.dy));
                    park_put(w, 5, __float_as_uint(p.dz));
                    park_put(w, 6, (unsigned)p.ir);
                    park_put(w, 7, (unsigned)p.it);
                    park_put(w, 8, (unsigned)p.ip);
                    park_put(w, 9, (unsigned)p.hemi);
                    park_put(w, 10, (unsigned)pid);
                    park_put(w, 11, (unsigned)(pid >> 32));
                    park_put(w, 12, krng);
                    park_put(w, 13, (unsigned)ifr);
                    park_put(w, 14, __float_as_uint(tau));
                    park_put(w, 15, nev);
                    park_put(w, 16, nstep);
                    park_put(w, 17, nev_cell);
                    park_put(w, 18, nzero);
                    park_put(w, 19, slice_left);
                    park_put(w, 20, (unsigned)pe);
                    park_put(w, 21, (unsigned)slot);
                    order_release();
                    st_rlx32(&P.park_flag[s0], (unsigned)(idx + 1ull));
                    alive = false;
                }
                if (ok) {
                    parked_from = m;
                    RT_PHASE(0);
                    continue;  // empty now: the next iteration moves to another run
                }
            }
        }

        // ---- refill dead lanes with packets of the current epoch
        if (cur_e < E && !epoch_ok) {
            epoch_ok = cur_e < 2 || uniform32(ld_rlx32(&P.snap_ready[cur_e - 2])) != 0u;
            if (epoch_ok) order_acquire();
        }
        if (BATCH && !(cur_e < E && epoch_ok) && m >= 0 && m != parked_from && P.park &&
            pool_nonempty<BATCH>(P)) {
            // nothing to claim: take parked packets of this run into dead lanes
            const uint64_t need = __ballot(!alive);
            if (need) {
                const int cnt = __popcll(need);
                const int leader = (int)uniform32((unsigned)(__ffsll((unsigned long long)need) - 1));
                unsigned long long h = 0, got = 0;
                if (lane == leader) {
                    for (int t = 0; t < 8 && !got; ++t) {
                        unsigned long long hd = ld_rlx64(P.park_head), tl = ld_rlx64(P.park_tail);
                        if (tl <= hd) break;
                        unsigned long long want = std::min((unsigned long long)cnt, tl - hd);
                        unsigned long long ex = hd;
                        if (__hip_atomic_compare_exchange_strong(P.park_head, &ex, hd + want,
                                                                 __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                                 __HIP_MEMORY_SCOPE_AGENT)) {
                            h = hd;
                            got = want;
                        }
                    }
                }
                h = lane64(h, leader);
                got = lane64(got, leader);
                if (got) {
                    last_adopt = iter;
                    const unsigned long long r = __popcll(need & ((1ull << lane) - 1ull));
                    if (!alive && r < got) {
                        const unsigned long long idx = h + r;
                        const size_t s0 = (size_t)(idx % P.park_cap);
                        // the parker has reserved this state and writes it without
                        // waiting on anything: the wait ends (bounded all the same)
                        bool ready = false;
                        for (int wt = 0; wt < (1 << 22) && !ready; ++wt) {
                            ready = ld_rlx32(&P.park_flag[s0]) == (unsigned)(idx + 1ull);
                            if (!ready) __builtin_amdgcn_s_sleep(2);
                        }
                        order_acquire();
                        if (ready) {
                            const unsigned *w = P.park + s0 * RT_PARK_WORDS;
                            p.x = __uint_as_float(park_get(w, 0));
                            p.y = __uint_as_float(park_get(w, 1));
                            p.z = __uint_as_float(park_get(w, 2));
                            p.dx = __uint_as_float(park_get(w, 3));
                            p.dy = __uint_as_float(park_get(w, 4));
                            p.dz = __uint_as_float(park_get(w, 5));
                            p.ir = (int)park_get(w, 6);
                            p.it = (int)park_get(w, 7);
                            p.ip = (int)park_get(w, 8);
                            p.hemi = (int)park_get(w, 9);
                            pid = (unsigned long long)park_get(w, 10) |
                                  ((unsigned long long)park_get(w, 11) << 32);
                            krng = park_get(w, 12);
                            ifr = (int)park_get(w, 13);
                            tau = __uint_as_float(park_get(w, 14));
                            nev = park_get(w, 15);
                            nstep = park_get(w, 16);
                            nev_cell = park_get(w, 17);
                            nzero = park_get(w, 18);
                            slice_left = park_get(w, 19);
                            pe = (int)park_get(w, 20);
                            slot = (int)park_get(w, 21);
                            order_release();  // read before the slot is handed back
                            st_rlx32(&P.park_flag[s0], 0u);
                            xs = pe >= 2 ? P.xsnap + (size_t)(pe - 2) * P.slot : nullptr;
                            tb = P.tslot + (size_t)slot * P.slot;
                            ccell = -1;
                            aRc = -1.f;
                            alive = true;
                        } else {
                            atomicOr(A.abort, 1ull);  // a lost packet: the run is invalid
                        }
                    }
                    if (lane == leader)
                        __hip_atomic_fetch_add(A.parked, -(int)got, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT);
                }
            }
        }
        if (cur_e < E && epoch_ok) {
            uint64_t need = __ballot(!alive);
            if (need) {
                int cnt = __popcll(need);
                int leader = (int)uniform32((unsigned)(__ffsll((unsigned long long)need) - 1));
                unsigned long long base = 0;
                if (lane == leader) base = atomicAdd(&P.ep_claim[cur_e], (unsigned long long)cnt);
                base = lane64(base, leader);
                const unsigned long long e0 = P.ep_start[cur_e];
                const unsigned long long size = P.ep_start[cur_e + 1] - e0;
                if (!alive) {
                    uint64_t idx = base + __popcll(need & ((1ull << lane) - 1ull));
                    if (idx < size) {
                        pid = e0 + idx;
                        pe = cur_e;
                        slot = pe;
                        xs = pe >= 2 ? P.xsnap + (size_t)(pe - 2) * P.slot : nullptr;
                        tb = P.tslot + (size_t)pe * P.slot;
                        // emission (random event 0): star, frequency, direction, tau
                        float u[8];
                        rng8(P.seed, pid, 0u, u);
                        krng = 1;
                        int ks = 0;
                        while (ks < P.nstar - 1 && u[0] > P.spick[ks]) ++ks;
                        ifr = sample_guided(P.scdf + (size_t)ks * NL,
                                            P.sguide + (size_t)ks * RT_NGUIDE, u[1]);
                        iso_dir(u[2], u[3], p.dx, p.dy, p.dz);
                        p.x = P.spos[3 * ks];
                        p.y = P.spos[3 * ks + 1];
                        p.z = P.spos[3 * ks + 2];
                        locate(g, p);
                        nev = 0;
                        nstep = 0;
                        nev_cell = 0;
                        nzero = 0;
                        slice_left = P.slice;
                        ccell = -1;
                        if (p.ir >= g.nr) {
                            RT_COUNT(3);
                            count_up(&P.slot_out[pe]);
                        } else {
                            alive = true;
                            if (p.ir >= 0 && P.visits) atomicAdd(&P.visits[cell_index(g, p)], 1u);
                            tau = -__logf(u[4]);
                        }
                    }
                }
                {
                    unsigned long long nst =
                        base >= size ? 0ull : std::min((unsigned long long)cnt, size - base);
                    if (lane == leader && nst) {
                        order_release();
                        atomicAdd(&P.slot_start[cur_e], nst);
                    }
                }
                if (base + cnt >= size) {
                    cur_e++;
                    epoch_ok = false;
                    if (lane == leader) atomicMax(P.ep_cur, (unsigned long long)cur_e);
                }
            }
        }
        RT_PHASE(0);
        if (!alive) continue;

        // ---- one step: to the next wall or to the next interaction
        // (a new cell's data are requested first, the wall distances computed
        // while they arrive)
        const int c = p.ir >= 0 ? cell_index(g, p) : -1;
        if (c >= 0 && c != ccell) {
            const float *rc = P.rhoL + (size_t)c * S;
#pragma unroll
            for (int s = 0; s < MS; ++s) rcache[s] = s < S ? rc[s] : 0.f;
#pragma unroll
            for (int q = 0; q < MG; ++q)
                xcache[q] = (q < G && xs) ? xs[(size_t)c * G + q] : -1.f;
            ccell = c;
            aRc = -1.f;
        }
        int kind;
        float tw = next_wall(g, p, kind);
        float aabs = 0.f, asca = 0.f;
        if (c >= 0) {
#pragma unroll
            for (int s = 0; s < MS; ++s)
                if (s < S) {
                    aabs = fmaf(rcache[s], s_ka[s * NL + ifr], aabs);
                    asca = fmaf(rcache[s], s_ks[s * NL + ifr], asca);
                }
        }
        float aext = aabs + asca;
        RT_PHASE(4);
        if (aext * tw > tau) {
            // interaction inside the cell
            float sseg = tau / aext;
            float wf = s_invdr[p.ir] * sseg;
#pragma unroll
            for (int q = 0; q < MG; ++q)
                if (q < G) acc[q] += (unsigned long long)rintf(s_gka[q * NL + ifr] * wf);
            p.x = fmaf(sseg, p.dx, p.x);
            p.y = fmaf(sseg, p.dy, p.y);
            p.z = fmaf(sseg, p.dz, p.z);
            RT_COUNT(5);
            nev++;
            nev_cell++;
            nzero = 0;
            if (--slice_left == 0u && slot < E) {
                // next slot: earlier deposits are final for snapshots <= slot
                flush_tally(P, tb, c, acc, ovf, ssum);
                flush_ssum(P, slot, ssum);
                if (slot + 1 < E) count_up(&P.slot_in[slot + 1]);
                count_up(&P.slot_out[slot]);
                slot++;
                tb = P.tslot + (size_t)slot * P.slot;
            }
            if (slice_left == 0u) slice_left = P.slice;
            // interaction (one random event): u0 scatter or absorb, u1 species,
            // u2 temperature row, u3 frequency, u4 u5 direction, u6 next tau
            float u[8];
            rng8(P.seed, pid, krng++, u);
            // scattering (Henyey-Greenstein if the model has g) or absorption and
            // re-emission; the new direction is drawn once after the branch
            const bool scat = u[0] * aext < asca;
            bool use_hg = false;
            float ghg = 0.f;
            if (scat) {
                RT_COUNT(2);
                if (P.hg) {
                    float us = u[1] * asca, a = 0.f;
                    int sp = S - 1;
                    for (int s = 0; s < S; ++s) {
                        a = fmaf(pickn<MS>(rcache, s), s_ks[s * NL + ifr], a);
                        if (us <= a) {
                            sp = s;
                            break;
                        }
                    }
                    ghg = s_g[sp * NL + ifr];
                    use_hg = true;
                }
            } else {
                RT_COUNT(1);
                float us = u[1] * aabs, a = 0.f;
                int sp = S - 1;
                for (int s = 0; s < S; ++s) {
                    a = fmaf(pickn<MS>(rcache, s), s_ka[s * NL + ifr], a);
                    if (us <= a) {
                        sp = s;
                        break;
                    }
                }
                int q = pickn<MS>(grpr, sp);
                float x = fminf(fmaxf(pickn<MG>(xcache, q), 0.f), (float)(P.ntemp - 1));
                int lo = min((int)x, P.ntemp - 2);
                float w = x - (float)lo;
                int row = lo + (u[2] < w ? 1 : 0);
                size_t ro = (size_t)q * P.ntemp + row;
                ifr = sample_guided(P.cdf + ro * NL, P.guide + ro * RT_NGUIDE, u[3]);
            }
            if (use_hg)
                hg_dir(u[4], u[5], ghg, p.dx, p.dy, p.dz);
            else
                iso_dir(u[4], u[5], p.dx, p.dy, p.dz);
            tau = -__logf(u[6]);
            RT_PHASE(6);
            if (P.max_events && nev >= P.max_events) {
                flush_tally(P, tb, c, acc, ovf, ssum);
                flush_ssum(P, slot, ssum);
                RT_MAXPKT(nev + nstep);
                RT_COUNT(4);
                alive = false;
                if (slot < E) count_up(&P.slot_out[slot]);
                continue;
            }
            if (P.mrw && nev_cell >= (uint32_t)P.mrw_count) {
                // modified random walk: jump to a sphere inside the cell when the
                // Rosseland optical depth to the nearest wall is large
                if (aRc < 0.f) {
                    float a = 0.f;
                    for (int s = 0; s < S; ++s) {
                        float rho = pickn<MS>(rcache, s);
                        if (rho > 0.f)
                            a = fmaf(rho, trow(P.kross + (size_t)s * P.ntemp, P.ntemp,
                                               pickn<MG>(xcache, pickn<MS>(grpr, s))), a);
                    }
                    aRc = a;
                }
                const float aR = aRc;
                float dmin = wall_dist(g, p);
                float R0 = dmin - P.mrw_gamma / aR;
                if (aR * dmin >= P.mrw_tau && R0 > 0.f) {
                    // one random event: u0 path length, u1 u2 exit point, u3 u4
                    // exit direction, u5 species, u6 temperature row, u7 frequency
                    float v[8];
                    rng8(P.seed, pid, krng++, v);
                    float y = mrw_escape_y(P.mrwy, RT_MRW_NY, RT_MRW_LMAX, v[0]);
                    float ell = 3.f * aR * R0 * R0 * y;
                    float wl = s_invdr[p.ir] * ell;
                    float kP[MG];
#pragma unroll
                    for (int q = 0; q < MG; ++q) kP[q] = 0.f;
#pragma unroll
                    for (int q = 0; q < MG; ++q)
                        if (q < G) {
                            kP[q] = trow(P.kplanck + (size_t)q * P.ntemp, P.ntemp, xcache[q]);
                            acc[q] += (unsigned long long)rintf(kP[q] * ginv[q] * wl);
                        }
                    float nx, ny, nz;
                    iso_dir(v[1], v[2], nx, ny, nz);
                    p.x = fmaf(R0, nx, p.x);
                    p.y = fmaf(R0, ny, p.y);
                    p.z = fmaf(R0, nz, p.z);
                    if (g.mirror) p.hemi = (p.z >= 0.f) ? 1 : -1;
                    float ct = sqrtf(v[3]);
                    rotate_dir(ct, v[4], nx, ny, nz);
                    p.dx = nx;
                    p.dy = ny;
                    p.dz = nz;
                    float wsum = 0.f;
                    for (int s = 0; s < S; ++s) wsum = fmaf(pickn<MS>(rcache, s), pickn<MG>(kP, pickn<MS>(grpr, s)), wsum);
                    float us = v[5] * wsum, a = 0.f;
                    int sp = S - 1;
                    for (int s = 0; s < S; ++s) {
                        a = fmaf(pickn<MS>(rcache, s), pickn<MG>(kP, pickn<MS>(grpr, s)), a);
                        if (us <= a) {
                            sp = s;
                            break;
                        }
                    }
                    int q = pickn<MS>(grpr, sp);
                    float x = fminf(fmaxf(pickn<MG>(xcache, q), 0.f), (float)(P.ntemp - 1));
                    int lo = min((int)x, P.ntemp - 2);
                    float w = x - (float)lo;
                    int row = lo + (v[6] < w ? 1 : 0);
                    size_t ro = (size_t)q * P.ntemp + row;
                    ifr = sample_guided(P.cdf + ro * NL, P.guide + ro * RT_NGUIDE, v[7]);
                    RT_COUNT(6);
                }
            }
            RT_PHASE(7);
        } else {
            if (c >= 0) {
                float wf = s_invdr[p.ir] * tw;
#pragma unroll
                for (int q = 0; q < MG; ++q)
                    if (q < G)
                        acc[q] += (unsigned long long)rintf(s_gka[q * NL + ifr] * wf);
                flush_tally(P, tb, c, acc, ovf, ssum);
            }
            tau -= aext * tw;
            RT_COUNT(0);
            nev_cell = 0;
            bool end = false;
            const float rcell = g.rw[p.ir + 1];
            if (++nstep > P.max_steps) {  // last-resort cap on crossings
                RT_COUNT(4);
                end = true;
            } else if (cross_wall(g, p, tw, kind)) {
                RT_COUNT(3);
                end = true;
            } else if (unstick(g, p, tw, rcell, nzero)) {
                RT_COUNT(7);
                if (p.ir >= g.nr) {
                    RT_COUNT(3);
                    end = true;
                }
            }
            RT_PHASE(5);
            if (end) {
                flush_ssum(P, slot, ssum);
                RT_MAXPKT(nev + nstep);
                alive = false;
                if (slot < E) count_up(&P.slot_out[slot]);
                continue;
            }
            if (p.ir >= 0 && P.visits) atomicAdd(&P.visits[cell_index(g, p)], 1u);
        }
    }
    // ---- statistics of the last run served, batch counters
    if (m >= 0) flush_counts(P, lane, crow, ovf);
    if (lane == 0) {
        for (int r = 0; r < A.nrun; ++r)
            atomicMax(&run_at<BATCH>(runs, P1, r).snap_time[RT_MAX_EPOCH + 1],
                      (unsigned long long)wall_clock64());
        const ThermParams &R0 = run_at<BATCH>(runs, P1, 0);
        atomicAdd(&R0.cnt[CNT_WITER], (unsigned long long)iter + 1ull);
        if (n_idle) atomicAdd(&R0.cnt[CNT_IDLE], (unsigned long long)n_idle);
    }
#if RT_THERMAL_PROFILE
    for (int k = 0; k < RT_NPHASE; ++k) {
        unsigned long long x = tk[k];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off);
        if (lane == 0) atomicAdd(&run_at<BATCH>(runs, P1, 0).cnt[CNT_TK0 + k], x);
    }
#endif
#undef RT_PHASE
#undef RT_COUNT
#undef RT_MAXPKT
#undef P
}

// final temperatures from the cumulative tallies
__global__ void k_snapshot(const unsigned long long *tally, const double *econv, double eps,
                           const double *lnq, int ntemp, double lnt0, double dlnt, int ncg,
                           int G, float *xrow, float *Tg) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ncg) return;
    int q = i % G;
    float x = xrow_of(lnq + (size_t)q * ntemp, ntemp, (double)tally[i] * econv[i] * eps);
    xrow[i] = x;
    Tg[i] = x < 0.f ? 0.f : (float)exp(lnt0 + (double)x * dlnt);
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

// epoch boundaries: N_0 = 0, N_k+1 = min(nphot, N_k + max(first, floor(N_k (growth-1))))
static std::vector<unsigned long long> epoch_starts(uint64_t nphot, uint64_t first, double growth) {
    std::vector<unsigned long long> st{0};
    uint64_t done = 0;
    while (done < nphot) {
        uint64_t sz = (uint64_t)std::floor((double)done * (growth - 1.0));
        if (sz < first) sz = first;
        done = std::min(nphot, done + sz);
        st.push_back(done);
    }
    return st;
}

// pipeline counters and flags: u64 [ep_start E+1 | ep_claim | slot_start |
// slot_in | slot_out | snap_time | t_begin, t_end | ep_cur | abort | park_head,
// park_tail] then u32 [snap_claim | snap_done | snap_ready | next_snap, runs_done]
#define PIPE_U64 (6 * RT_MAX_EPOCH + 7)
#define PIPE_U32 (3 * RT_MAX_EPOCH + 4)
#define PIPE_F64 (RT_MAX_EPOCH + 1)  // after the u32 words (8-byte aligned): slot totals

struct PipeView {
    unsigned long long *ep_start, *ep_claim, *slot_start, *slot_in, *slot_out, *snap_time, *ep_cur,
        *abort, *park_head, *park_tail;
    unsigned int *u32base, *snap_claim, *snap_done, *snap_ready, *next_snap, *runs_done;
    double *slot_sum;
};

static PipeView pipe_view(void *d_pipe) {
    PipeView v;
    v.ep_start = (unsigned long long *)d_pipe;
    v.ep_claim = v.ep_start + RT_MAX_EPOCH + 1;
    v.slot_start = v.ep_claim + RT_MAX_EPOCH;
    v.slot_in = v.slot_start + RT_MAX_EPOCH;
    v.slot_out = v.slot_in + RT_MAX_EPOCH;
    v.snap_time = v.slot_out + RT_MAX_EPOCH;  // then t_begin, t_end
    v.ep_cur = v.snap_time + RT_MAX_EPOCH + 2;
    v.abort = v.ep_cur + 1;
    v.park_head = v.abort + 1;
    v.park_tail = v.park_head + 1;
    v.u32base = (unsigned int *)(v.ep_start + PIPE_U64);
    v.snap_claim = v.u32base;
    v.snap_done = v.snap_claim + RT_MAX_EPOCH;
    v.snap_ready = v.snap_done + RT_MAX_EPOCH;
    v.next_snap = v.snap_ready + RT_MAX_EPOCH;
    v.runs_done = v.next_snap + 1;
    v.slot_sum = (double *)(v.u32base + PIPE_U32);
    return v;
}

static int ensure_pipe_buffers(rt_model *m) {
    rt_ctx *ctx = m->ctx;
    if (!m->d_pipe)
        HIPCHK(ctx, hipMalloc(&m->d_pipe, sizeof(unsigned long long) * PIPE_U64 +
                                              sizeof(unsigned int) * PIPE_U32 +
                                              sizeof(double) * PIPE_F64));
    if (!m->d_run) HIPCHK(ctx, hipMalloc(&m->d_run, sizeof(ThermParams)));
    if (!m->h_run) m->h_run = std::calloc(1, sizeof(ThermParams));
    if (!m->h_run) return set_err(ctx, RT_E_NOMEM, "host allocation failed");
    return 0;
}

// one temperature snapshot and one tally array per slot (aligned to 4 KiB)
static size_t snap_slot(size_t ncg) { return (ncg + 1023) / 1024 * 1024; }

static int ensure_slot_buffers(rt_model *m, int E) {
    rt_ctx *ctx = m->ctx;
    const size_t sl = snap_slot((size_t)m->ncell * m->tab.ngroup);
    if (!m->d_xring || m->xring_cap < sl * (size_t)E) {
        if (m->d_xring) HIPCHK(ctx, hipFree(m->d_xring));
        m->d_xring = nullptr;
        HIPCHK(ctx, hipMalloc(&m->d_xring, sizeof(float) * sl * (size_t)E));
        m->xring_cap = sl * (size_t)E;
    }
    if (!m->d_tally2 || m->tslot_cap < sl * (size_t)(E + 1)) {
        if (m->d_tally2) HIPCHK(ctx, hipFree(m->d_tally2));
        m->d_tally2 = nullptr;
        HIPCHK(ctx, hipMalloc(&m->d_tally2, sizeof(unsigned long long) * sl * (size_t)(E + 1)));
        m->tslot_cap = sl * (size_t)(E + 1);
    }
    return 0;
}

// cumulative tallies plus the last slot (deposits after the last epoch's snapshot)
__global__ void k_add_u64(unsigned long long *dst, const unsigned long long *src, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] += src[i];
}

// LDS floats one wavefront needs for run m's opacities and walls
static int lds_floats(const rt_model *m) {
    const int S = m->nspec, NL = m->nlam, G = m->tab.ngroup;
    return S * NL * (m->has_g ? 3 : 2) + G * NL + 2 * (m->nr + 1) + m->nr + 3 * (m->nt + 1) +
           3 * (m->np + 1);
}

// Validates the options, lays out the epochs, allocates the run's buffers and
// fills its parameter block (host copy in m->h_run).
int thermal_plan(rt_model *m, uint64_t nphot, uint64_t seed, const rt_thermal_opts *o,
                 ThermPlan &rp) {
    rt_ctx *ctx = m->ctx;
    rt_thermal_opts opt{};
    if (o) opt = *o;
    if (nphot == 0) return set_err(ctx, RT_E_ARG, "nphot must be > 0");
    if (m->nstar < 1) return set_err(ctx, RT_E_ARG, "model has no star");
    const uint64_t max_events = opt.max_events ? opt.max_events : 1000000ull;
    uint64_t first = opt.epoch_first ? opt.epoch_first : std::max<uint64_t>(1024, nphot / 4096);
    double growth = opt.epoch_growth > 1.0 ? opt.epoch_growth : 1.25;
    const int G = m->tab.ngroup;
    const size_t ncg = (size_t)m->ncell * G;
    const double eps = m->tab.lum / (double)nphot;
    m->h_est = epoch_starts(nphot, first, growth);
    const int E = (int)m->h_est.size() - 1;
    if (E > RT_MAX_EPOCH)
        return set_err(ctx, RT_E_LIMIT, "%d epochs > %d: raise epoch_first or epoch_growth", E,
                       RT_MAX_EPOCH);
    if (ensure_pipe_buffers(m) || ensure_slot_buffers(m, E)) return RT_E_HIP;
    const uint64_t slice = opt.slice_events ? opt.slice_events : 10000ull;
    PipeView pv = pipe_view(m->d_pipe);

    ThermParams &P = *(ThermParams *)m->h_run;
    P = ThermParams{};
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
    P.kross = m->d_kross;
    P.kplanck = m->d_kplanck;
    P.mrwy = m->d_mrwy;
    P.visits = opt.count_visits ? m->d_visits : nullptr;
    P.cnt = m->d_cnt;
    P.seed = seed;
    P.max_events = max_events;
    P.max_steps = 100000000u;
    P.mrw = opt.mrw;
    P.mrw_count = opt.mrw_count_trigger > 0 ? opt.mrw_count_trigger : 100;
    P.mrw_tau = (float)(opt.mrw_tauthres > 0 ? opt.mrw_tauthres : 10.0);
    P.mrw_gamma = (float)(opt.mrw_gamma > 0 ? opt.mrw_gamma : 2.0);
    P.nepoch = E;
    P.ncg = (int)ncg;
    P.nchunk = (int)((ncg + 63) / 64);
    P.ep_start = pv.ep_start;
    P.ep_claim = pv.ep_claim;
    P.slot_start = pv.slot_start;
    P.slot_in = pv.slot_in;
    P.slot_out = pv.slot_out;
    P.slice = (uint32_t)std::min<uint64_t>(slice, 0x7fffffffull);
    P.snap_claim = pv.snap_claim;
    P.snap_done = pv.snap_done;
    P.snap_ready = pv.snap_ready;
    P.next_snap = pv.next_snap;
    P.snap_time = pv.snap_time;
    P.ep_cur = pv.ep_cur;
    P.runs_done = pv.runs_done;
    P.park_head = pv.park_head;  // the pool itself is a batch's (P.park stays null)
    P.park_tail = pv.park_tail;
    P.tslot = m->d_tally2;
    P.cum = m->d_tally;
    P.xsnap = m->d_xring;
    P.slot_sum = pv.slot_sum;
    P.slot = snap_slot(ncg);
    P.econv = m->d_econv;
    P.lnq = m->d_lnq;
    P.eps = eps;
    P.lnt0 = m->tab.lnt0;
    P.dlnt = m->tab.dlnt;

    rp.E = E;
    rp.eps = eps;
    rp.count_visits = opt.count_visits != 0;
    rp.lds_wave = lds_floats(m);
    return RT_OK;
}

// zero the run's tallies, counters and pipeline state (fill kernels: no memset
// nodes if the stream is being captured) and upload the epoch starts
int thermal_enqueue_reset(rt_model *m, const ThermPlan &rp, hipStream_t sm, bool upload_est) {
    rt_ctx *ctx = m->ctx;
    const int G = m->tab.ngroup;
    const size_t ncg = (size_t)m->ncell * G;
    const size_t sl = snap_slot(ncg);
    PipeView pv = pipe_view(m->d_pipe);
    hipLaunchKernelGGL(k_fill_u64, dim3(nblk(ncg, 256)), dim3(256), 0, sm, m->d_tally, ncg, 0ull);
    const size_t nts = sl * (size_t)(rp.E + 1);
    hipLaunchKernelGGL(k_fill_u64, dim3(nblk(nts, 256)), dim3(256), 0, sm, m->d_tally2, nts, 0ull);
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, sm, m->d_cnt, (size_t)CNT_N, 0ull);
    const size_t n64 = (size_t)(PIPE_U64 - (RT_MAX_EPOCH + 1));
    hipLaunchKernelGGL(k_fill_u64, dim3(nblk(n64, 256)), dim3(256), 0, sm, pv.ep_claim, n64, 0ull);
    hipLaunchKernelGGL(k_fill_u32, dim3(nblk(PIPE_U32, 256)), dim3(256), 0, sm, pv.u32base,
                       (size_t)PIPE_U32, 0u);
    hipLaunchKernelGGL(k_fill_u64, dim3(nblk(PIPE_F64, 256)), dim3(256), 0, sm,
                       (unsigned long long *)pv.slot_sum, (size_t)PIPE_F64, 0ull);
    if (upload_est)
        HIPCHK(ctx, hipMemcpyAsync(pv.ep_start, m->h_est.data(),
                                   sizeof(unsigned long long) * m->h_est.size(),
                                   hipMemcpyHostToDevice, sm));
    if (rp.count_visits)
        hipLaunchKernelGGL(k_fill_u32, dim3(nblk(m->ncell, 256)), dim3(256), 0, sm, m->d_visits,
                           (size_t)m->ncell, 0u);
    return RT_OK;
}

// final temperatures: cumulative tallies plus the last slot, rows, species
void thermal_enqueue_finish(rt_model *m, const ThermPlan &rp, hipStream_t sm) {
    const int G = m->tab.ngroup;
    const size_t ncg = (size_t)m->ncell * G;
    const size_t sl = snap_slot(ncg);
    hipLaunchKernelGGL(k_add_u64, dim3(nblk(ncg, 256)), dim3(256), 0, sm, m->d_tally,
                       m->d_tally2 + sl * (size_t)rp.E, ncg);
    hipLaunchKernelGGL(k_snapshot, dim3(nblk(ncg, 256)), dim3(256), 0, sm, m->d_tally, m->d_econv,
                       rp.eps, m->d_lnq, m->ntemp, m->tab.lnt0, m->tab.dlnt, (int)ncg, G, m->d_xrow,
                       m->d_Tg);
    hipLaunchKernelGGL(k_species_T, dim3(nblk((size_t)m->nspec * m->ncell, 256)), dim3(256), 0, sm,
                       m->d_Tg, m->d_rhopos, m->d_grp, m->nspec, G, m->ncell, m->d_Tout);
}

// The launch itself: kernel instance by the largest species / group count,
// LDS slices per wavefront, blocks per CU from the occupancy.
int thermal_enqueue_launch(rt_ctx *ctx, const void *d_runs, const void *h_run1, int nrun,
                           int lds_wave, int max_s, int max_g, const rt_thermal_opts *o,
                           unsigned long long *d_abort, unsigned int *d_done, int *d_parked,
                           hipStream_t sm) {
    rt_thermal_opts opt{};
    if (o) opt = *o;
    int bs = opt.block_size > 0 ? opt.block_size : 256;
    if (bs % 64 || bs > 256) return set_err(ctx, RT_E_ARG, "block_size must be 64..256, multiple of 64");
    double max_sec = opt.max_seconds > 0 ? opt.max_seconds : 60.0;
    size_t lds = sizeof(float) * (size_t)lds_wave * (size_t)(bs / 64);
    if (lds > 64 * 1024)
        return set_err(ctx, RT_E_LIMIT, "thermal kernel LDS %zu B > 64 KiB (grid too large for "
                                        "block_size %d)", lds, bs);
    // one run: its block is a kernel argument (h_run1); several: device array
    const bool batch = h_run1 == nullptr;
    void (*kern)(BatchArgs, const ThermParams *, ThermParams);
    if (max_s <= 2 && max_g <= 1)
        kern = batch ? k_thermal<2, 1, true> : k_thermal<2, 1, false>;
    else if (max_s <= 2 && max_g <= 2)  // pdspy's fits: disk and envelope dust differ
        kern = batch ? k_thermal<2, 2, true> : k_thermal<2, 2, false>;
    else if (max_s <= 4 && max_g <= 2)
        kern = batch ? k_thermal<4, 2, true> : k_thermal<4, 2, false>;
    else
        kern = batch ? k_thermal<RT_MAX_SPEC, RT_MAX_GROUP, true>
                     : k_thermal<RT_MAX_SPEC, RT_MAX_GROUP, false>;
    int bpc = opt.blocks_per_cu;
    if (bpc <= 0) {
        int occ = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, bs, lds) != hipSuccess || occ < 1)
            occ = 1;
        bpc = occ;
    }
    BatchArgs A{};
    A.nrun = nrun;
    A.lds_wave = lds_wave;
    A.max_ticks = (uint64_t)(max_sec * (double)ctx->wallclock_khz * 1000.0);
    A.abort = d_abort;
    A.runs_done = d_done;
    A.parked = d_parked;
    if (batch && !d_parked) return set_err(ctx, RT_E_ARG, "batch launch without a parked-packet counter");
    ThermParams P1{};
    if (!batch) P1 = *(const ThermParams *)h_run1;
    // free_cus_per_xcd: CUs of each XCD left free.  A persistent launch holds
    // every CU it gets until its last waves retire, and short kernels another
    // stream queues meanwhile (the next run's tally resets, the fold after a
    // run, an image) waited seconds behind it, so a second run could not start
    // while this one's long walks finished.  The dispatcher deals workgroup i
    // to XCD i mod 8; measured, a kernel of another queue ran at once only
    // with a whole CU free on every XCD (one free block slot per XCD did not
    // help).
    const int nxcd = RT_NUM_XCD;
    long nblocks = (long)ctx->n_cu * bpc;
    const int fc = opt.free_cus_per_xcd;
    if (fc > 0 && ctx->n_cu % nxcd == 0 && ctx->n_cu - nxcd * fc >= ctx->n_cu / 2)
        nblocks = (long)(ctx->n_cu - nxcd * fc) * bpc;
    hipLaunchKernelGGL(kern, dim3((unsigned)nblocks), dim3(bs), lds, sm, A,
                       (const ThermParams *)d_runs, P1);
    HIPCHK(ctx, hipGetLastError());
    return RT_OK;
}

// After the stream has drained: statistics, watchdog / overflow / pipeline
// checks and the timeline of run m.
int thermal_collect(rt_model *m, const ThermPlan &rp, bool aborted, float kms, float tot,
                    rt_thermal_stats *st) {
    rt_ctx *ctx = m->ctx;
    PipeView pv = pipe_view(m->d_pipe);
    unsigned long long cnt[CNT_N];
    HIPCHK(ctx, hipMemcpy(cnt, m->d_cnt, sizeof(cnt), hipMemcpyDeviceToHost));
    unsigned int nsnap = 0;
    HIPCHK(ctx, hipMemcpy(&nsnap, pv.next_snap, sizeof(nsnap), hipMemcpyDeviceToHost));
    const int E = rp.E;
    {
        std::vector<unsigned long long> tl(RT_MAX_EPOCH + 2);
        HIPCHK(ctx, hipMemcpy(tl.data(), pv.snap_time, sizeof(unsigned long long) * tl.size(),
                              hipMemcpyDeviceToHost));
        const double tick_ms = 1.0 / (double)ctx->wallclock_khz;
        const unsigned long long t0 = tl[RT_MAX_EPOCH];
        m->timeline.assign(E + 1, -1.0);
        for (int k = 0; k < E; ++k)
            if (tl[k]) m->timeline[k] = (double)(long long)(tl[k] - t0) * tick_ms;
        if (tl[RT_MAX_EPOCH + 1])
            m->timeline[E] = (double)(long long)(tl[RT_MAX_EPOCH + 1] - t0) * tick_ms;
    }
    m->eps_last = rp.eps;
    m->have_T = true;
    if (st) {
        st->n_cross = cnt[CNT_CROSS];
        st->n_abs = cnt[CNT_ABS];
        st->n_scat = cnt[CNT_SCAT];
        st->n_escape = cnt[CNT_ESCAPE];
        st->n_killed = cnt[CNT_KILLED];
        st->n_events = cnt[CNT_EVENTS];
        st->n_mrw = cnt[CNT_MRW];
        st->eps = rp.eps;
        st->lum = m->tab.lum;
        st->n_epochs = E;
        st->watchdog_fired = aborted ? 1 : 0;
        st->kernel_ms = kms;
        st->total_ms = tot;
        st->max_epoch_ms = kms;
        st->max_packet_steps = cnt[CNT_MAXPKT];
        st->wave_iters = cnt[CNT_WITER];
        st->n_nudged = cnt[CNT_NUDGE];
        st->idle_iters = cnt[CNT_IDLE];
        st->n_parked = cnt[CNT_PARKED];
        for (int k = 0; k < 8; ++k) st->phase_clocks[k] = cnt[CNT_TK0 + k];
    }
    if (aborted)
        return set_err(ctx, RT_E_WATCHDOG, "thermal kernel exceeded its time budget; packets were killed");
    // every slot's adds sum below 2^61: no tally of it reached 2^64 (wrapped)
    std::vector<double> ssum(E + 1);
    HIPCHK(ctx, hipMemcpy(ssum.data(), pv.slot_sum, sizeof(double) * (E + 1), hipMemcpyDeviceToHost));
    bool slot_ovf = false;
    for (double v : ssum) slot_ovf |= !(v < 0x1p61);
    if (cnt[CNT_OVERFLOW] || slot_ovf)
        return set_err(ctx, RT_E_OVERFLOW, "fixed-point tally overflow");
    if (nsnap != (unsigned)E)
        return set_err(ctx, RT_E_HIP, "thermal pipeline ended after %u of %d epochs", nsnap, E);
    return RT_OK;
}

int thermal_run(rt_model *m, uint64_t nphot, uint64_t seed, const rt_thermal_opts *o,
                rt_thermal_stats *st) {
    rt_ctx *ctx = m->ctx;
    hipStream_t sm = ctx->stream;
    ThermPlan rp{};
    int rc = thermal_plan(m, nphot, seed, o, rp);
    if (rc) return rc;
    rc = thermal_enqueue_reset(m, rp, sm, true);
    if (rc) return rc;
    PipeView pv = pipe_view(m->d_pipe);
    hipEvent_t e_begin, e_k0, e_k1, e_end;
    (void)hipEventCreate(&e_begin);
    (void)hipEventCreate(&e_k0);
    (void)hipEventCreate(&e_k1);
    (void)hipEventCreate(&e_end);
    HIPCHK(ctx, hipEventRecord(e_begin, sm));
    HIPCHK(ctx, hipEventRecord(e_k0, sm));
    rc = thermal_enqueue_launch(ctx, m->d_run, m->h_run, 1, rp.lds_wave, m->nspec, m->tab.ngroup,
                                o, pv.abort, pv.runs_done, nullptr, sm);
    if (rc) return rc;
    HIPCHK(ctx, hipEventRecord(e_k1, sm));
    thermal_enqueue_finish(m, rp, sm);
    HIPCHK(ctx, hipEventRecord(e_end, sm));
    HIPCHK(ctx, hipEventSynchronize(e_end));
    float tot = 0, kms = 0;
    (void)hipEventElapsedTime(&tot, e_begin, e_end);
    (void)hipEventElapsedTime(&kms, e_k0, e_k1);
    (void)hipEventDestroy(e_begin);
    (void)hipEventDestroy(e_k0);
    (void)hipEventDestroy(e_k1);
    (void)hipEventDestroy(e_end);
    unsigned long long ab = 0;
    HIPCHK(ctx, hipMemcpy(&ab, pv.abort, sizeof(ab), hipMemcpyDeviceToHost));
    return thermal_collect(m, rp, ab != 0, kms, tot, st);
}

size_t thermal_params_size() { return sizeof(ThermParams); }

// copies run m's planned parameter block (host) into dst_host
int thermal_params_copy(const rt_model *m, void *dst_host) {
    if (!m->h_run) return RT_E_ARG;
    std::memcpy(dst_host, m->h_run, sizeof(ThermParams));
    return RT_OK;
}

// points run m's parameter block at the batch's completion counter
void thermal_params_set_done(rt_model *m, unsigned int *d_done) {
    ((ThermParams *)m->h_run)->runs_done = d_done;
}

// gives run m's parameter block a pool for parked packets: `cap` states of
// RT_PARK_WORDS words and one flag word each (flags zero at the launch)
void thermal_params_set_pool(rt_model *m, unsigned int *park, unsigned int *flag, uint32_t cap) {
    ThermParams &P = *(ThermParams *)m->h_run;
    P.park = park;
    P.park_flag = flag;
    P.park_cap = cap;
}

size_t thermal_park_words() { return RT_PARK_WORDS; }

// where run m's epoch starts go (the reset of a batch step does not upload them)
PipeUpload thermal_epoch_starts(rt_model *m) {
    PipeUpload u;
    u.dst = pipe_view(m->d_pipe).ep_start;
    u.src = m->h_est.data();
    u.bytes = sizeof(unsigned long long) * m->h_est.size();
    return u;
}
