// batch.hip -- one sampler step for many models: what pdspy's fitting loop does
// per walker (run_disk_model.py:31-511: thermal run, image, visibilities,
// chi^2), for a whole batch of models at once.
//
// * The thermal Monte Carlo of every model runs in ONE persistent launch
//   (k_thermal in batch mode): each wavefront serves one model at a time and
//   moves to another model's packets when its model has none to hand out, so
//   the epoch pipelines and the long random walks of different models overlap.
// * Then, per model: imaging temperature (density-weighted species mean,
//   Model.py:759-769), image, visibilities at the data's (u,v), chi^2 and its
//   gradient, and the adjoint back to d chi^2 / d log rho and d chi^2 / dT.
// * The whole step can be captured once as a hipGraph and replayed: the
//   parameter blocks and densities it reads live in device memory and may be
//   updated between replays; nothing in it allocates or synchronises.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "rt_internal.h"

// A chain's kernels are small (a 256x256 image is 256 blocks, one per CU), so
// chains of different models fill the GPU better side by side.
#define RT_CHAIN_STREAMS 4

struct rt_batch {
    rt_ctx *ctx = nullptr;
    std::vector<rt_model *> models;
    // thermal: parameter blocks of all runs (device array + host copies)
    void *d_runs = nullptr;
    std::vector<unsigned char> h_runs;
    std::vector<ThermPlan> plans;
    unsigned long long *d_abort = nullptr;  // [1]
    unsigned int *d_done = nullptr;         // [1]
    int *d_parked = nullptr;                // [1] packets in the pools
    // per model: a pool of parked packets (ring of park_cap states) and flags
    unsigned int *d_park = nullptr, *d_park_flag = nullptr;
    uint32_t park_cap = 8192;
    int lds_wave = 0, max_s = 0, max_g = 0;
    std::vector<uint64_t> plan_nphot;       // nphot the plans were made for
    rt_thermal_opts plan_opts{};
    bool planned = false;
    // step scratch per model
    std::vector<float *> d_Timg, d_imgs, d_re, d_im, d_gre, d_gim, d_gimg, d_J;
    std::vector<double *> d_ws;
    std::vector<int> K_cap, npix_cap;
    // captured step
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<rt_step_item> graph_items;
    std::vector<const void *> graph_bufs;  // model buffers the graph was captured with
    // timing; the step's own stream (captures need one that is not the null
    // stream) and the events that order it after / before the caller's stream
    hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr, e_in = nullptr, e_out = nullptr;
    // the per-model image chains run on RT_CHAIN_STREAMS streams forked from
    // the step's stream after the thermal launch and joined before its end
    hipStream_t chain_stream[RT_CHAIN_STREAMS] = {};
    hipEvent_t e_fork = nullptr, e_join[RT_CHAIN_STREAMS] = {};
    hipStream_t stream = nullptr;
};

// ---------------------------------------------------------------- kernels
// imaging temperature: sum_s rho_s T_s / sum_s rho_s, rho = 0 counted as 1e-30
__global__ void k_imaging_T(const float *rho, const float *T, int S, int ncell, float *Timg) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    float num = 0.f, den = 0.f;
    for (int s = 0; s < S; ++s) {
        float r = rho[(size_t)s * ncell + c];
        r = r == 0.f ? 1.0e-30f : r;
        num = fmaf(r, T[(size_t)s * ncell + c], num);
        den += r;
    }
    Timg[c] = num / den;
}

__global__ void k_scale_copy(const float *src, float a, size_t n, float *dst) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = a * src[i];
}

__global__ void k_fill_flags(unsigned long long *abort, unsigned int *done, int *parked,
                             unsigned int *park_flag, size_t nflag) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        abort[0] = 0ull;
        done[0] = 0u;
        parked[0] = 0;
    }
    if (i < nflag) park_flag[i] = 0u;
}

static inline unsigned nblk(size_t n, int b) { return (unsigned)((n + b - 1) / b); }

// ---------------------------------------------------------------- host
static int batch_free_graph(rt_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->graph_items.clear();
    b->graph_bufs.clear();
    return 0;
}

int batch_create(rt_ctx *ctx, rt_model *const *models, int n, rt_batch **out) {
    if (!ctx || !models || n < 1 || !out) return RT_E_ARG;
    for (int i = 0; i < n; ++i) {
        if (!models[i]) return set_err(ctx, RT_E_ARG, "model %d is NULL", i);
        if (models[i]->ctx != ctx) return set_err(ctx, RT_E_ARG, "model %d belongs to another context", i);
        for (int j = 0; j < i; ++j)
            if (models[j] == models[i]) return set_err(ctx, RT_E_ARG, "model %d appears twice", i);
    }
    rt_batch *b = new rt_batch();
    b->ctx = ctx;
    b->models.assign(models, models + n);
    const size_t psz = thermal_params_size();
    b->h_runs.assign(psz * (size_t)n, 0);
    hipError_t e = hipMalloc(&b->d_runs, psz * (size_t)n);
    if (e == hipSuccess) e = hipMalloc((void **)&b->d_abort, sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMalloc((void **)&b->d_done, sizeof(unsigned int));
    if (e == hipSuccess) e = hipMalloc((void **)&b->d_parked, sizeof(int));
    const size_t npark = (size_t)b->park_cap * (size_t)n;
    if (e == hipSuccess)
        e = hipMalloc((void **)&b->d_park, sizeof(unsigned int) * thermal_park_words() * npark);
    if (e == hipSuccess) e = hipMalloc((void **)&b->d_park_flag, sizeof(unsigned int) * npark);
    if (e == hipSuccess) e = hipEventCreate(&b->e0);
    if (e == hipSuccess) e = hipEventCreate(&b->e1);
    if (e == hipSuccess) e = hipEventCreate(&b->e2);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->e_in, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->e_out, hipEventDisableTiming);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->e_fork, hipEventDisableTiming);
    for (int k = 0; k < RT_CHAIN_STREAMS && e == hipSuccess; ++k) {
        e = hipStreamCreateWithFlags(&b->chain_stream[k], hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&b->e_join[k], hipEventDisableTiming);
    }
    if (e != hipSuccess) {
        batch_destroy(b);
        return set_err(ctx, RT_E_NOMEM, "batch allocation failed: %s", hipGetErrorString(e));
    }
    b->d_Timg.assign(n, nullptr);
    b->d_imgs.assign(n, nullptr);
    b->d_re.assign(n, nullptr);
    b->d_im.assign(n, nullptr);
    b->d_gre.assign(n, nullptr);
    b->d_gim.assign(n, nullptr);
    b->d_gimg.assign(n, nullptr);
    b->d_J.assign(n, nullptr);
    b->d_ws.assign(n, nullptr);
    b->K_cap.assign(n, 0);
    b->npix_cap.assign(n, 0);
    *out = b;
    return RT_OK;
}

int batch_destroy(rt_batch *b) {
    if (!b) return RT_E_ARG;
    batch_free_graph(b);
    (void)hipFree(b->d_runs);
    (void)hipFree(b->d_abort);
    (void)hipFree(b->d_done);
    (void)hipFree(b->d_parked);
    (void)hipFree(b->d_park);
    (void)hipFree(b->d_park_flag);
    for (size_t i = 0; i < b->models.size(); ++i) {
        (void)hipFree(b->d_Timg[i]);
        (void)hipFree(b->d_imgs[i]);
        (void)hipFree(b->d_re[i]);
        (void)hipFree(b->d_im[i]);
        (void)hipFree(b->d_gre[i]);
        (void)hipFree(b->d_gim[i]);
        (void)hipFree(b->d_gimg[i]);
        (void)hipFree(b->d_J[i]);
        (void)hipFree(b->d_ws[i]);
    }
    if (b->e0) (void)hipEventDestroy(b->e0);
    if (b->e1) (void)hipEventDestroy(b->e1);
    if (b->e2) (void)hipEventDestroy(b->e2);
    if (b->e_in) (void)hipEventDestroy(b->e_in);
    if (b->e_out) (void)hipEventDestroy(b->e_out);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    if (b->e_fork) (void)hipEventDestroy(b->e_fork);
    for (int k = 0; k < RT_CHAIN_STREAMS; ++k) {
        if (b->chain_stream[k]) (void)hipStreamDestroy(b->chain_stream[k]);
        if (b->e_join[k]) (void)hipEventDestroy(b->e_join[k]);
    }
    delete b;
    return RT_OK;
}

static bool same_opts(const rt_thermal_opts &a, const rt_thermal_opts &b) {
    return std::memcmp(&a, &b, sizeof(a)) == 0;
}

// Plans every run (allocations, epoch layout, parameter blocks), uploads the
// epoch starts and parameter blocks.  Synchronous; never inside a capture.
static int batch_plan(rt_batch *b, const uint64_t *nphot, const uint64_t *seeds,
                      const rt_thermal_opts *o) {
    rt_ctx *ctx = b->ctx;
    const int n = (int)b->models.size();
    rt_thermal_opts opt{};
    if (o) opt = *o;
    const size_t psz = thermal_params_size();
    bool replan = !b->planned || !same_opts(opt, b->plan_opts) ||
                  !std::equal(nphot, nphot + n, b->plan_nphot.begin(), b->plan_nphot.end());
    if (replan) {
        batch_free_graph(b);
        b->plans.assign(n, ThermPlan{});
        b->lds_wave = 0;
        b->max_s = 0;
        b->max_g = 0;
    }
    for (int i = 0; i < n; ++i) {
        rt_model *m = b->models[i];
        ThermPlan rp{};
        int rc = thermal_plan(m, nphot[i], seeds[i], &opt, rp);
        if (rc) return rc;
        thermal_params_set_done(m, b->d_done);
        if (!opt.no_park)
            thermal_params_set_pool(
                m, b->d_park + thermal_park_words() * (size_t)b->park_cap * (size_t)i,
                b->d_park_flag + (size_t)b->park_cap * (size_t)i, b->park_cap);
        thermal_params_copy(m, b->h_runs.data() + psz * (size_t)i);
        // the epoch starts every time: a single run of the model in between
        // may have left its own there
        PipeUpload up = thermal_epoch_starts(m);
        HIPCHK(ctx, hipMemcpy(up.dst, up.src, up.bytes, hipMemcpyHostToDevice));
        if (replan) {
            b->plans[i] = rp;
            b->lds_wave = std::max(b->lds_wave, rp.lds_wave);
            b->max_s = std::max(b->max_s, m->nspec);
            b->max_g = std::max(b->max_g, m->tab.ngroup);
        }
    }
    HIPCHK(ctx, hipMemcpy(b->d_runs, b->h_runs.data(), psz * (size_t)n, hipMemcpyHostToDevice));
    b->plan_nphot.assign(nphot, nphot + n);
    b->plan_opts = opt;
    b->planned = true;
    return RT_OK;
}

// the thermal part of a step: resets, one launch, finishes (capturable)
static int batch_enqueue_thermal(rt_batch *b, hipStream_t sm, bool timed) {
    rt_ctx *ctx = b->ctx;
    const int n = (int)b->models.size();
    for (int i = 0; i < n; ++i) {
        int rc = thermal_enqueue_reset(b->models[i], b->plans[i], sm, false);
        if (rc) return rc;
    }
    const size_t nflag = (size_t)b->park_cap * (size_t)n;
    hipLaunchKernelGGL(k_fill_flags, dim3(nblk(nflag, 256)), dim3(256), 0, sm, b->d_abort,
                       b->d_done, b->d_parked, b->d_park_flag, nflag);
    if (timed) HIPCHK(ctx, hipEventRecord(b->e0, sm));
    int rc = thermal_enqueue_launch(ctx, b->d_runs, nullptr, n, b->lds_wave, b->max_s, b->max_g,
                                    &b->plan_opts, b->d_abort, b->d_done, b->d_parked, sm);
    if (rc) return rc;
    if (timed) HIPCHK(ctx, hipEventRecord(b->e1, sm));
    for (int i = 0; i < n; ++i) thermal_enqueue_finish(b->models[i], b->plans[i], sm);
    return RT_OK;
}

static int batch_collect(rt_batch *b, float kms, float tot, rt_thermal_stats *stats) {
    rt_ctx *ctx = b->ctx;
    unsigned long long ab = 0;
    HIPCHK(ctx, hipMemcpy(&ab, b->d_abort, sizeof(ab), hipMemcpyDeviceToHost));
    int first_err = RT_OK;
    for (size_t i = 0; i < b->models.size(); ++i) {
        int rc = thermal_collect(b->models[i], b->plans[i], ab != 0, kms, tot,
                                 stats ? stats + i : nullptr);
        if (rc && first_err == RT_OK) first_err = rc;
    }
    return first_err;
}

int batch_thermal(rt_batch *b, const uint64_t *nphot, const uint64_t *seeds,
                  const rt_thermal_opts *o, rt_thermal_stats *stats) {
    rt_ctx *ctx = b->ctx;
    if (!nphot || !seeds) return set_err(ctx, RT_E_ARG, "nphot and seeds are required");
    int rc = batch_plan(b, nphot, seeds, o);
    if (rc) return rc;
    hipStream_t sm = ctx->stream;
    rc = batch_enqueue_thermal(b, sm, true);
    if (rc) return rc;
    HIPCHK(ctx, hipEventRecord(b->e2, sm));
    HIPCHK(ctx, hipEventSynchronize(b->e2));
    float kms = 0, tot = 0;
    (void)hipEventElapsedTime(&kms, b->e0, b->e1);
    (void)hipEventElapsedTime(&tot, b->e0, b->e2);
    return batch_collect(b, kms, tot, stats);
}

// scratch of model i for an item (allocated outside any capture)
static int batch_step_buffers(rt_batch *b, int i, const rt_step_item &it) {
    rt_ctx *ctx = b->ctx;
    rt_model *m = b->models[i];
    int rc = image_buffers(m);
    if (rc) return rc;
    if (!b->d_Timg[i]) HIPCHK(ctx, hipMalloc(&b->d_Timg[i], sizeof(float) * (size_t)m->ncell));
    if (it.nphot_scat) {
        int rc = scat_prepare(b->models[i]);
        if (rc) return rc;
        if (!b->d_J[i]) HIPCHK(ctx, hipMalloc(&b->d_J[i], sizeof(float) * b->models[i]->ncell));
    }
    const int npix = it.cam.nx * it.cam.ny;
    if (it.cam.nrefine > 0) {
        // sub-pixel pool for the captured image chain: room for the pixels'
        // rays and 15 times as many sub-pixels
        rc = refine_reserve(ctx, std::max(ctx->ref.cap, (size_t)16 * (size_t)(npix + 256) *
                                                            (size_t)std::max(1, it.cam.nsub * it.cam.nsub)));
        if (rc) return rc;
    }
    if (b->npix_cap[i] < npix) {
        (void)hipFree(b->d_imgs[i]);
        (void)hipFree(b->d_gimg[i]);
        b->d_imgs[i] = b->d_gimg[i] = nullptr;
        HIPCHK(ctx, hipMalloc(&b->d_imgs[i], sizeof(float) * (size_t)npix));
        HIPCHK(ctx, hipMalloc(&b->d_gimg[i], sizeof(float) * (size_t)npix));
        b->npix_cap[i] = npix;
    }
    if (b->K_cap[i] < it.K) {
        float **bufs[4] = {&b->d_re[i], &b->d_im[i], &b->d_gre[i], &b->d_gim[i]};
        for (float **p : bufs) {
            (void)hipFree(*p);
            *p = nullptr;
            HIPCHK(ctx, hipMalloc(p, sizeof(float) * (size_t)it.K));
        }
        (void)hipFree(b->d_ws[i]);
        b->d_ws[i] = nullptr;
        HIPCHK(ctx, hipMalloc(&b->d_ws[i], sizeof(double) * chisq_workspace(it.K)));
        b->K_cap[i] = it.K;
    }
    return RT_OK;
}

// image -> visibilities -> chi^2 -> adjoint of model i (capturable)
static int batch_enqueue_image_chain(rt_batch *b, int i, const rt_step_item &it, hipStream_t sm) {
    rt_ctx *ctx = b->ctx;
    rt_model *m = b->models[i];
    const int npix = it.cam.nx * it.cam.ny;
    hipLaunchKernelGGL(k_imaging_T, dim3(nblk(m->ncell, 256)), dim3(256), 0, sm, m->d_rho,
                       m->d_Tout, m->nspec, m->ncell, b->d_Timg[i]);
    // scattering source at the image wavelength (Philox key uploaded per step)
    rt_camera cam = it.cam;
    cam.d_jscat = nullptr;
    if (it.nphot_scat) {
        scat_enqueue(m, cam.lam_um, it.nphot_scat, b->d_Timg[i], b->d_J[i], sm);
        cam.d_jscat = b->d_J[i];
    }
    int rc = image_run(m, &cam, b->d_Timg[i], nullptr, it.d_img);
    if (rc) return rc;
    hipLaunchKernelGGL(k_scale_copy, dim3(nblk(npix, 256)), dim3(256), 0, sm, it.d_img,
                       (float)it.scale, (size_t)npix, b->d_imgs[i]);
    rc = vis_run(ctx, b->d_imgs[i], it.cam.nx, it.cam.ny, it.dxy, it.d_u, it.d_v, it.K, it.dra,
                 it.ddec, b->d_re[i], b->d_im[i]);
    if (rc) return rc;
    rc = chisq_run(ctx, it.d_re, it.d_im, it.d_w, b->d_re[i], b->d_im[i], it.K, it.d_chi2,
                   b->d_gre[i], b->d_gim[i], b->d_ws[i]);
    if (rc) return rc;
    if (!it.d_g_logrho || !it.d_g_T) return RT_OK;
    rc = vis_adjoint_run(ctx, b->d_gre[i], b->d_gim[i], it.cam.nx, it.cam.ny, it.dxy, it.d_u,
                         it.d_v, it.K, it.dra, it.ddec, b->d_gimg[i]);
    if (rc) return rc;
    hipLaunchKernelGGL(k_scale_copy, dim3(nblk(npix, 256)), dim3(256), 0, sm, b->d_gimg[i],
                       (float)it.scale, (size_t)npix, b->d_gimg[i]);
    return image_adjoint_run(m, &cam, b->d_Timg[i], nullptr, b->d_gimg[i], it.d_g_logrho,
                             it.d_g_T);
}

// device buffers of the models a captured step refers to: a replay is only
// valid while none of them has been reallocated
static std::vector<const void *> model_buffers(const rt_batch *b) {
    std::vector<const void *> v;
    for (const rt_model *m : b->models) {
        const void *p[] = {m->d_pipe, m->d_tally, m->d_tally2, m->d_xring, m->d_xrow, m->d_Tg,
                           m->d_Tout, m->d_rho, m->d_rhoL, m->d_alpha, m->d_src, m->d_galpha,
                           m->d_gsrc, m->d_qscale, m->d_cnt, m->d_visits, m->d_vol, m->d_jtally, m->d_scum,
                           m->d_scnt, m->d_sblk};
        v.insert(v.end(), p, p + sizeof(p) / sizeof(p[0]));
        // the image chains capture opacities at the camera wavelength by value
        v.push_back((const void *)(uintptr_t)m->opac_version);
        v.push_back(m->d_csize);
    }
    const RefinePool &P = b->ctx->ref;
    const void *q[] = {P.xy, P.val, P.child, P.gw, P.ctl};
    v.insert(v.end(), q, q + 5);
    return v;
}

static bool same_items(const std::vector<rt_step_item> &a, const rt_step_item *b, int n) {
    if ((int)a.size() != n) return false;
    return std::memcmp(a.data(), b, sizeof(rt_step_item) * (size_t)n) == 0;
}

static int batch_step_enqueue(rt_batch *b, const rt_step_item *items, int n, int use_graph);

int batch_step(rt_batch *b, const uint64_t *nphot, const uint64_t *seeds,
               const rt_thermal_opts *o, const rt_step_item *items, int use_graph,
               rt_thermal_stats *stats, double *step_ms) {
    rt_ctx *ctx = b->ctx;
    const int n = (int)b->models.size();
    if (!nphot || !seeds || !items) return set_err(ctx, RT_E_ARG, "nphot, seeds and items are required");
    for (int i = 0; i < n; ++i) {
        const rt_step_item &it = items[i];
        if (it.cam.mode != 0 || it.cam.nx < 1 || it.cam.ny < 1 || !it.d_img || it.K < 1 ||
            !it.d_u || !it.d_v || !it.d_re || !it.d_im || !it.d_w || !it.d_chi2 || !(it.dxy > 0) ||
            (it.d_g_logrho == nullptr) != (it.d_g_T == nullptr))
            return set_err(ctx, RT_E_ARG, "step item %d is invalid", i);
    }
    int rc = batch_plan(b, nphot, seeds, o);
    if (rc) return rc;
    for (int i = 0; i < n; ++i) {
        rc = batch_step_buffers(b, i, items[i]);
        if (rc) return rc;
        if (items[i].nphot_scat) {
            // the scattering Monte Carlo's key: the model's seed, decorrelated
            // from its thermal streams (same key, other counters would do too)
            const unsigned long long key = seeds[i] ^ 0x9E3779B97F4A7C15ull;
            HIPCHK(ctx, hipMemcpy(scat_seed_slot(b->models[i]), &key, sizeof(key),
                                  hipMemcpyHostToDevice));
        }
    }
    // the step runs on the batch's stream, ordered after the caller's stream;
    // the library's enqueue functions use ctx->stream, so it is switched for
    // the duration
    hipStream_t user = ctx->stream, sm = b->stream;
    HIPCHK(ctx, hipEventRecord(b->e_in, user));
    HIPCHK(ctx, hipStreamWaitEvent(sm, b->e_in, 0));
    ctx->stream = sm;
    rc = batch_step_enqueue(b, items, n, use_graph);
    ctx->stream = user;
    if (rc) return rc;
    HIPCHK(ctx, hipEventRecord(b->e_out, sm));
    HIPCHK(ctx, hipStreamWaitEvent(user, b->e_out, 0));
    HIPCHK(ctx, hipEventSynchronize(b->e2));
    float tot = 0;
    (void)hipEventElapsedTime(&tot, b->e0, b->e2);
    if (step_ms) *step_ms = tot;
    return batch_collect(b, -1.f, tot, stats);
}

// capture (once) and launch, or enqueue eagerly, on ctx->stream
// the image chains of all models: on the chain streams, forked from sm (the
// thermal launch before them) and joined back into it; on sm alone when an
// item sub-pixels (the refinement's node pool is the context's, one at a time)
static int batch_enqueue_chains(rt_batch *b, const rt_step_item *items, int n, hipStream_t sm) {
    rt_ctx *ctx = b->ctx;
    bool refine = false;
    for (int i = 0; i < n; ++i) refine = refine || items[i].cam.nrefine > 0;
    int rc = RT_OK;
    if (refine || n < 2) {
        for (int i = 0; i < n && rc == RT_OK; ++i) rc = batch_enqueue_image_chain(b, i, items[i], sm);
        return rc;
    }
    const int ns = std::min(n, RT_CHAIN_STREAMS);
    HIPCHK(ctx, hipEventRecord(b->e_fork, sm));
    for (int k = 0; k < ns; ++k) HIPCHK(ctx, hipStreamWaitEvent(b->chain_stream[k], b->e_fork, 0));
    // the chain's calls launch on the context's stream
    for (int i = 0; i < n && rc == RT_OK; ++i) {
        ctx->stream = b->chain_stream[i % ns];
        rc = batch_enqueue_image_chain(b, i, items[i], ctx->stream);
    }
    ctx->stream = sm;
    // joined even after an error, so a capture can end
    for (int k = 0; k < ns; ++k) {
        hipError_t e = hipEventRecord(b->e_join[k], b->chain_stream[k]);
        if (e == hipSuccess) e = hipStreamWaitEvent(sm, b->e_join[k], 0);
        if (e != hipSuccess && rc == RT_OK)
            rc = set_err(ctx, RT_E_HIP, "chain join: %s", hipGetErrorString(e));
    }
    return rc;
}

static int batch_step_enqueue(rt_batch *b, const rt_step_item *items, int n, int use_graph) {
    rt_ctx *ctx = b->ctx;
    hipStream_t sm = ctx->stream;
    int rc = RT_OK;
    if (use_graph && !(b->exec && same_items(b->graph_items, items, n) &&
                       b->graph_bufs == model_buffers(b))) {
        batch_free_graph(b);
        HIPCHK(ctx, hipStreamBeginCapture(sm, hipStreamCaptureModeThreadLocal));
        rc = batch_enqueue_thermal(b, sm, false);
        if (rc == RT_OK) rc = batch_enqueue_chains(b, items, n, sm);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(sm, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return set_err(ctx, RT_E_HIP, "graph capture failed: %s", hipGetErrorString(e));
        b->graph = g;
        e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            batch_free_graph(b);
            return set_err(ctx, RT_E_HIP, "graph instantiation failed: %s", hipGetErrorString(e));
        }
        b->graph_items.assign(items, items + n);
        b->graph_bufs = model_buffers(b);
    }
    HIPCHK(ctx, hipEventRecord(b->e0, sm));
    if (use_graph) {
        HIPCHK(ctx, hipGraphLaunch(b->exec, sm));
    } else {
        rc = batch_enqueue_thermal(b, sm, false);
        if (rc == RT_OK) rc = batch_enqueue_chains(b, items, n, sm);
        if (rc) return rc;
    }
    HIPCHK(ctx, hipEventRecord(b->e2, sm));
    return RT_OK;
}
