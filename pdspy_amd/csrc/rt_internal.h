// rt_internal.h -- host-side state behind the C ABI (include/pdspy_amd.h).
#pragma once
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "../../include/pdspy_amd.h"
#include "rt_device.h"
#include "tables.h"

#define RT_MAX_SPEC 8    // species per model
#define RT_MAX_LAM 255   // wavelength bins (guide tables are uint8)
#define RT_MAX_STAR 16   // point stars per model
#define RT_MAX_COMP 8    // density components of the device generator
#define RT_TALLY_SHIFT 28  // fixed-point tally unit: 2^-28 of kmax*dr
#define RT_LEN_UNIT 1.495978707e13  // device length unit: au in cm (pdspy AU)

#define RT_MAX_REFINE 24  // sub-pixel levels of the camera's recursive refinement
// control words of the sub-pixel pool: [0] nodes handed out, [1 + L] first node
// of level L (L = 0 .. RT_MAX_REFINE + 1: the last is the end of the deepest
// level), then the unsplit count (cumulative, cleared by rt_refine_stats) and
// the rays of the last image
#define RT_REF_UNSPLIT (RT_MAX_REFINE + 3)
#define RT_REF_RAYS (RT_MAX_REFINE + 4)
#define RT_REF_NCTL (RT_MAX_REFINE + 5)

// Node pool of the camera's sub-pixeling, shared by the context's models (their
// work is ordered on the context's stream): sky position, value, first child
// (-1: a leaf), cotangent
struct RefinePool {
    size_t cap = 0;
    float2 *xy = nullptr;
    float *val = nullptr;
    int *child = nullptr;
    float *gw = nullptr;
    unsigned *ctl = nullptr;
};

struct rt_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    std::mutex err_mu;  // models of one context may be updated from several host threads
    int n_cu = 0;
    int wallclock_khz = 0;
    RefinePool ref;
    // cells crossed by camera rays since the last read: [0] image rays
    // (sub-pixels included), [1] adjoint sweeps
    unsigned long long *d_icount = nullptr;
    ~rt_ctx();
};

// counters[] slots used by the thermal kernels
enum {
    CNT_PACKET = 0,
    CNT_CROSS,
    CNT_ABS,
    CNT_SCAT,
    CNT_ESCAPE,
    CNT_KILLED,
    CNT_EVENTS,
    CNT_MRW,
    CNT_NUDGE,
    CNT_ABORT,
    CNT_OVERFLOW,
    CNT_MAXPKT,
    CNT_WITER,     // loop iterations summed over wavefronts
    CNT_IDLE,      // of which with no packet to step
    CNT_PARKED,    // packets passed from a sparse wavefront to another (batch)
    CNT_TK0,       // 8 phase clocks (RT_THERMAL_PROFILE builds only)
    CNT_N = CNT_TK0 + 8
};



struct rt_model {
    rt_ctx *ctx = nullptr;
    int nr = 0, nt = 0, np = 0, ncell = 0, nspec = 0, nlam = 0, nstar = 0, mirror = 0;
    int ntemp = 1000;
    double temp0 = 0.01, temp1 = 1e5;
    bool has_g = false;
    std::vector<double> rw, tw, pw;              // walls (cm, rad, rad)
    std::vector<double> lam, kabs, ksca, gsca;   // host opacity tables
    std::vector<double> rho;                     // [S][cells] host copy
    std::vector<double> stars;                   // [nstar][5]
    std::vector<double> vol;                     // cell volumes (cm^3)
    HostTables tab;
    // device buffers
    float *d_rw = nullptr, *d_rw2 = nullptr, *d_tw = nullptr, *d_ct = nullptr, *d_st = nullptr;
    double *d_rwd = nullptr, *d_ctd = nullptr, *d_std = nullptr;
    float *d_pw = nullptr, *d_cp = nullptr, *d_sp = nullptr;
    float *d_rhoL = nullptr;      // [cells][S] rho * length unit
    float *d_rho = nullptr;       // [S][cells] rho (g/cm^3)
    unsigned char *d_rhopos = nullptr;  // [S][cells] 1 where the FP64 density is > 0
    float *d_ka = nullptr, *d_ks = nullptr, *d_g = nullptr;  // [S][nlam]
    int *d_grp = nullptr;
    float *d_gka = nullptr;       // [G][nlam]
    float *d_ginvk = nullptr;     // [G] 2^shift / kmax
    float *d_invdr = nullptr;     // [nr] 1/dr (au^-1)
    float *d_cdf = nullptr;
    unsigned char *d_guide = nullptr;
    double *d_lnq = nullptr;
    float *d_scdf = nullptr;
    unsigned char *d_sguide = nullptr;
    float *d_spick = nullptr;
    float *d_spos = nullptr;      // [nstar][3] au
    float *d_kross = nullptr, *d_kplanck = nullptr, *d_mrwy = nullptr;
    double *d_econv = nullptr;    // [cells][G] erg/s/g per tally unit (per packet energy 1)
    unsigned long long *d_tally = nullptr;  // [cells][G]
    float *d_xrow = nullptr;      // [cells][G] T-row coordinate of the lagged temperature
    float *d_Tg = nullptr;        // [cells][G] temperature (K)
    unsigned int *d_visits = nullptr;
    unsigned long long *d_cnt = nullptr;
    void *d_pipe = nullptr;                 // epoch pipeline counters and flags
    unsigned long long *d_tally2 = nullptr; // [epochs+1][slot] tallies by deposit slot
    size_t tslot_cap = 0;                   // entries allocated in d_tally2
    float *d_xring = nullptr;               // [epochs][slot] temperature-row snapshots
    size_t xring_cap = 0;                   // floats allocated in d_xring
    float *d_Tout = nullptr;      // [S][cells]
    // image scratch
    float *d_alpha = nullptr, *d_src = nullptr;
    float *d_csize = nullptr;     // [cell] cell size for the camera's sub-pixeling (au)
    double *d_tabin = nullptr;    // inputs of the device table build [3 + 2 G + S][nlam]
    double *d_tT = nullptr;       // its per-temperature quantities [4 ntemp nlam + 2 ntemp]
    // device opacity update: staged kabs, ksca, g [S][nlam] and kmax [G] (pinned
    // host copy, the device copy, the event of the last upload from it)
    double *h_opin = nullptr, *d_opin = nullptr;
    hipEvent_t e_opin = nullptr;
    bool opin_busy = false;
    double *d_dr = nullptr;       // [nr] radial cell widths (cm)
    double *d_galpha = nullptr, *d_gsrc = nullptr;  // adjoint sums, 64-bit fixed point
    double *d_qscale = nullptr;  // their units: [0] per unit dL/dalpha, [1] per unit dL/dS
    float *d_rowI = nullptr;      // per-ray intensity for the adjoint
    size_t rowI_cap = 0;
    // device density generator: cell centres, per-component scratch
    double *d_rc = nullptr, *d_tc = nullptr;      // r (cm), theta centres
    double *d_dens_w = nullptr;                   // [RT_MAX_COMP][nt][nr]
    double *d_dens_rows = nullptr;                // [RT_MAX_COMP][nr]
    double *d_dens_fac = nullptr;                 // [2][RT_MAX_COMP]
    void *d_comps = nullptr;                      // [RT_MAX_COMP] rt_density_comp
    double dens_solid = 0;                        // 4 pi (2 pi for full-sphere grids)
    bool rho_host_stale = false;                  // densities generated on the device
    // scattering source Monte Carlo
    double *d_vol = nullptr;                      // [cells] volumes (cm^3)
    unsigned long long *d_jtally = nullptr;       // [cells] path-length tallies
    double *d_scum = nullptr;                     // [cells] cumulative source luminosity
    unsigned long long *d_scnt = nullptr;         // [8] counters; [7] the Philox key
    double *d_sblk = nullptr;                     // [1024] prefix-sum chunk totals
    double scat_lstar = 0;                        // star luminosity of the last call
    void *h_run = nullptr;        // host parameter block of the thermal kernel (ThermParams)
    void *d_run = nullptr;        // its device copy
    std::vector<unsigned long long> h_est;  // epoch starts of the last planned run
    double eps_last = 0;          // packet energy of the last thermal run
    std::vector<double> timeline; // last run: ms from kernel start to each snapshot, then the end
    bool have_T = false;
    unsigned opac_version = 0;  // bumped by rt_model_set_opacity (captured graphs go stale)
};

int set_err(rt_ctx *ctx, int code, const char *fmt, ...);
#define HIPCHK(ctx, call)                                                                  \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            return set_err((ctx), RT_E_HIP, "%s failed: %s (%s:%d)", #call,                \
                           hipGetErrorString(e_), __FILE__, __LINE__);                     \
    } while (0)

// a planned thermal run (thermal.hip)
struct ThermPlan {
    int E = 0;            // epochs
    double eps = 0;       // packet energy
    bool count_visits = false;
    int lds_wave = 0;     // LDS floats per wavefront
};

// the opacity-dependent thermal-MC tables built by a kernel on the context's
// stream from m->kabs / ksca, m->tab holding the groups (tables_dev.hip)
int opacity_device_update(rt_model *m);
void opacity_device_free(rt_model *m);

// the model's thermal-MC tables from its host opacities and stars
inline int build_tables(rt_model *m, std::string &err) {
    TableSpec sp{m->nlam, m->nspec, m->ntemp, m->nstar, m->temp0, m->temp1,
                 m->lam.data(), m->kabs.data(), m->ksca.data(), m->stars.data()};
    return build_host_tables(sp, m->tab, err);
}
int thermal_plan(rt_model *m, uint64_t nphot, uint64_t seed, const rt_thermal_opts *o,
                 ThermPlan &rp);
// upload_est: also copy the epoch starts (a captured step has them uploaded once)
int thermal_enqueue_reset(rt_model *m, const ThermPlan &rp, hipStream_t sm, bool upload_est);
void thermal_enqueue_finish(rt_model *m, const ThermPlan &rp, hipStream_t sm);
// d_runs: device array of nrun parameter blocks, or (nrun = 1) h_run1: the
// single run's host block passed as a kernel argument
int thermal_enqueue_launch(rt_ctx *ctx, const void *d_runs, const void *h_run1, int nrun,
                           int lds_wave, int max_s, int max_g, const rt_thermal_opts *o,
                           unsigned long long *d_abort, unsigned int *d_done, int *d_parked,
                           hipStream_t sm);
int thermal_collect(rt_model *m, const ThermPlan &rp, bool aborted, float kms, float tot,
                    rt_thermal_stats *st);
size_t thermal_params_size();
int thermal_params_copy(const rt_model *m, void *dst_host);
void thermal_params_set_done(rt_model *m, unsigned int *d_done);
void thermal_params_set_pool(rt_model *m, unsigned int *park, unsigned int *flag, uint32_t cap);
size_t thermal_park_words();
struct PipeUpload {
    void *dst;
    const void *src;
    size_t bytes;
};
PipeUpload thermal_epoch_starts(rt_model *m);
int image_buffers(rt_model *m);
// the context's sub-pixel pool holds at least n nodes (synchronous; never
// inside a capture)
int refine_reserve(rt_ctx *ctx, size_t n);
// batches (batch.hip)
int batch_create(rt_ctx *ctx, rt_model *const *models, int n, rt_batch **out);
int batch_destroy(rt_batch *b);
int batch_thermal(rt_batch *b, const uint64_t *nphot, const uint64_t *seeds,
                  const rt_thermal_opts *o, rt_thermal_stats *stats);
int batch_step(rt_batch *b, const uint64_t *nphot, const uint64_t *seeds,
               const rt_thermal_opts *o, const rt_step_item *items, int use_graph,
               rt_thermal_stats *stats, double *step_ms);
rt::DevGrid dev_grid(const rt_model *m);
int density_prepare(rt_model *m, const rt_density_comp *h_comps, int ncomp);
void density_enqueue(rt_model *m, const rt_density_comp *d_comps, int ncomp, hipStream_t sm);
int density_generate(rt_model *m, const rt_density_comp *h_comps, int ncomp);
int scat_source_run(rt_model *m, double lam_um, uint64_t nphot, uint64_t seed,
                    const float *d_T, float *d_J, rt_scat_stats *st);
int scat_prepare(rt_model *m);
// opacity table `tab` ([species][lambda]) of species s at wavelength lam (um):
// the camera's log-log interpolation
double kappa_at_lambda(const rt_model *m, const std::vector<double> &tab, int s, double lam);
unsigned long long *scat_seed_slot(rt_model *m);
void scat_enqueue(rt_model *m, double lam_um, uint64_t nphot, const float *d_T, float *d_J,
                  hipStream_t sm);
int thermal_run(rt_model *m, uint64_t nphot, uint64_t seed, const rt_thermal_opts *o,
                rt_thermal_stats *st);
int image_run(rt_model *m, const rt_camera *cam, const float *d_T, const float *d_rho,
              float *d_img);
int image_adjoint_run(rt_model *m, const rt_camera *cam, const float *d_T, const float *d_rho,
                      const float *d_dimg, float *d_g_logrho, float *d_g_T);
int vis_run(rt_ctx *ctx, const float *d_img, int nx, int ny, double dxy, const double *d_u,
            const double *d_v, int K, double dra, double ddec, float *d_re, float *d_im);
int vis_adjoint_run(rt_ctx *ctx, const float *d_gre, const float *d_gim, int nx, int ny,
                    double dxy, const double *d_u, const double *d_v, int K, double dra,
                    double ddec, float *d_g_img);
// ws: nullable workspace of chisq_workspace(K) doubles (else allocated stream-ordered)
int chisq_run(rt_ctx *ctx, const float *dre, const float *dim, const float *w, const float *mre,
              const float *mim, int K, double *out, float *gre, float *gim, double *ws);
size_t chisq_workspace(int K);
// gridded sampler; work: nullable, vis_grid_workspace(nx, ny) bytes (else
// allocated stream-ordered); conv 0 = imtovis convention, 1 = pdspy's galario
// call (image[::-1], imaginary part negated)
int vis_grid_run(rt_ctx *ctx, const float *d_img, int nx, int ny, double dxy, const double *d_u,
                 const double *d_v, int K, double dra, double ddec, float *d_re, float *d_im,
                 void *d_work, int conv);
size_t vis_grid_workspace(int nx, int ny);
// the allocation-free gridded sampler: its scratch's device address for the
// current device (at context creation), and the run
int vis_scratch_init(rt_ctx *ctx, int device);
int vis_grid_noalloc_run(rt_ctx *ctx, const float *d_img, int nx, int ny, double dxy,
                         const double *d_u, const double *d_v, int K, double dra, double ddec,
                         float *d_re, float *d_im);
