// abi.hip -- extern "C" entry points declared in include/pdspy_amd.h: context
// and model lifecycle, input validation and upload (the role of pdspy's
// Model.write_radmc3d, pdspy/modeling/Model.py:723-778), temperature read-back
// (the role of radmc3d/read.pyx:433-478), and dispatch to the kernels.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>

#include <cstdlib>

#include "rt_internal.h"

int set_err(rt_ctx *ctx, int code, const char *fmt, ...) {
    if (ctx) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        std::lock_guard<std::mutex> lk(ctx->err_mu);
        ctx->err = buf;
    }
    return code;
}

rt::DevGrid dev_grid(const rt_model *m) {
    rt::DevGrid g{};
    g.nr = m->nr;
    g.nt = m->nt;
    g.np = m->np;
    g.mirror = m->mirror;
    g.ncell = m->ncell;
    g.rw = m->d_rw;
    g.rw2 = m->d_rw2;
    g.rwd = m->d_rwd;
    g.tw = m->d_tw;
    g.ct = m->d_ct;
    g.st = m->d_st;
    g.ctd = m->d_ctd;
    g.std_ = m->d_std;
    g.pw = m->d_pw;
    g.cp = m->d_cp;
    g.sp = m->d_sp;
    return g;
}

template <typename T>
static int upload(rt_ctx *ctx, T **dst, const T *src, size_t n) {
    if (n == 0) n = 1;
    HIPCHK(ctx, hipMalloc((void **)dst, n * sizeof(T)));
    if (src) HIPCHK(ctx, hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice));
    return 0;
}

static bool finite_all(const double *p, size_t n) {
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(p[i])) return false;
    return true;
}

static void free_model(rt_model *m) {
    void *ptrs[] = {m->d_rw, m->d_rw2, m->d_tw, m->d_ct, m->d_st, m->d_rwd, m->d_ctd, m->d_std,
                    m->d_pw, m->d_cp, m->d_sp, m->d_rhoL, m->d_rho, m->d_rhopos, m->d_ka, m->d_ks, m->d_g,
                    m->d_grp, m->d_gka, m->d_ginvk, m->d_invdr, m->d_cdf, m->d_guide, m->d_lnq,
                    m->d_scdf, m->d_sguide, m->d_spick, m->d_spos, m->d_kross, m->d_kplanck, m->d_mrwy,
                    m->d_econv, m->d_tally, m->d_xrow, m->d_Tg, m->d_visits, m->d_cnt, m->d_Tout,
                    m->d_alpha, m->d_src, m->d_galpha, m->d_gsrc, m->d_qscale, m->d_rowI, m->d_pipe,
                    m->d_tally2,
                    m->d_csize, m->d_tabin, m->d_tT, m->d_opin, m->d_dr,
                    m->d_xring, m->d_run, m->d_rc, m->d_tc, m->d_dens_w, m->d_dens_rows,
                    m->d_dens_fac, m->d_comps, m->d_vol, m->d_jtally, m->d_scum, m->d_scnt, m->d_sblk};
    opacity_device_free(m);
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    std::free(m->h_run);
    m->h_run = nullptr;
}

static int upload_density(rt_model *m) {
    rt_ctx *ctx = m->ctx;
    const int S = m->nspec, nc = m->ncell;
    std::vector<float> rhoL((size_t)nc * S), rho((size_t)S * nc);
    std::vector<unsigned char> pos((size_t)S * nc);
    for (int s = 0; s < S; ++s)
        for (int c = 0; c < nc; ++c) {
            double r = m->rho[(size_t)s * nc + c];
            rho[(size_t)s * nc + c] = (float)r;
            rhoL[(size_t)c * S + s] = (float)(r * RT_LEN_UNIT);
            // the species' temperature is reported wherever its FP64 density is
            // positive, also where the FP32 copy underflows to 0
            pos[(size_t)s * nc + c] = r > 0 ? 1 : 0;
        }
    HIPCHK(ctx, hipMemcpy(m->d_rhopos, pos.data(), pos.size(), hipMemcpyHostToDevice));
    HIPCHK(ctx, hipMemcpy(m->d_rhoL, rhoL.data(), rhoL.size() * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(ctx, hipMemcpy(m->d_rho, rho.data(), rho.size() * sizeof(float), hipMemcpyHostToDevice));
    return 0;
}

// device copies of the opacities and of what derives from them besides the
// emission tables: FP32 kabs/ksca/g [S][nlam], the group representatives'
// kabs [G][nlam], 2^shift / kmax [G], and the per-cell tally-to-energy factors
// [cells][G] (these need the cell volumes)
struct OpacArrays {
    std::vector<float> ka, ks, gg, gka, ginvk;
    std::vector<double> econv;
};

static void opacity_arrays(const rt_model *m, OpacArrays &o) {
    const size_t nk = (size_t)m->nspec * m->nlam;
    const int G = m->tab.ngroup, nc = m->ncell;
    o.ka.assign(nk, 0.f);
    o.ks.assign(nk, 0.f);
    o.gg.assign(nk, 0.f);
    for (size_t i = 0; i < nk; ++i) {
        o.ka[i] = (float)m->kabs[i];
        o.ks[i] = (float)m->ksca[i];
        if (m->has_g) o.gg[i] = (float)m->gsca[i];
    }
    o.gka.assign((size_t)G * m->nlam, 0.f);
    for (int q = 0; q < G; ++q)
        for (int i = 0; i < m->nlam; ++i)
            o.gka[(size_t)q * m->nlam + i] = o.ka[(size_t)m->tab.grep[q] * m->nlam + i];
    o.ginvk.assign(G, 0.f);
    for (int q = 0; q < G; ++q) o.ginvk[q] = (float)(std::ldexp(1.0, RT_TALLY_SHIFT) / m->tab.kmax[q]);
    o.econv.assign((size_t)nc * G, 0.0);
    for (int c = 0; c < nc; ++c) {
        int i = c % m->nr;
        double dr_cm = m->rw[i + 1] - m->rw[i];
        for (int q = 0; q < G; ++q)
            o.econv[(size_t)c * G + q] =
                m->tab.kmax[q] * dr_cm / (std::ldexp(1.0, RT_TALLY_SHIFT) * m->vol[c]);
    }
}

static int create_model(rt_ctx *ctx, const rt_model_desc *d, rt_model *m) {
    if (!d) return set_err(ctx, RT_E_ARG, "NULL model description");
    if (d->nr < 1 || d->ntheta < 1 || d->nphi < 1 || d->nspec < 1 || d->nlam < 1)
        return set_err(ctx, RT_E_ARG, "grid, species and wavelength counts must be >= 1");
    if (d->nspec > RT_MAX_SPEC) return set_err(ctx, RT_E_LIMIT, "nspec > %d", RT_MAX_SPEC);
    if (d->nstar > RT_MAX_STAR) return set_err(ctx, RT_E_LIMIT, "nstar > %d", RT_MAX_STAR);
    if (d->nlam > RT_MAX_LAM) return set_err(ctx, RT_E_LIMIT, "nlam > %d", RT_MAX_LAM);
    if (!d->h_r_walls || !d->h_theta_walls || !d->h_phi_walls || !d->h_rho || !d->h_lam_um ||
        !d->h_kabs || !d->h_ksca || (d->nstar > 0 && !d->h_stars))
        return set_err(ctx, RT_E_ARG, "NULL input array");
    m->ctx = ctx;
    m->nr = d->nr;
    m->nt = d->ntheta;
    m->np = d->nphi;
    m->ncell = d->nr * d->ntheta * d->nphi;
    m->nspec = d->nspec;
    m->nlam = d->nlam;
    m->nstar = d->nstar;
    m->ntemp = d->ntemp > 0 ? d->ntemp : 1000;
    m->temp0 = d->temp0 > 0 ? d->temp0 : 0.01;
    m->temp1 = d->temp1 > 0 ? d->temp1 : 1e5;
    if (m->ntemp < 2 || !(m->temp1 > m->temp0)) return set_err(ctx, RT_E_ARG, "bad temperature grid");
    m->rw.assign(d->h_r_walls, d->h_r_walls + m->nr + 1);
    m->tw.assign(d->h_theta_walls, d->h_theta_walls + m->nt + 1);
    m->pw.assign(d->h_phi_walls, d->h_phi_walls + m->np + 1);
    for (int i = 0; i < m->nr; ++i)
        if (!(m->rw[i] > 0) || !(m->rw[i + 1] > m->rw[i]))
            return set_err(ctx, RT_E_ARG, "radial walls must be positive and increasing");
    for (int i = 0; i < m->nt; ++i)
        if (!(m->tw[i + 1] > m->tw[i])) return set_err(ctx, RT_E_ARG, "theta walls must increase");
    for (int i = 0; i < m->np; ++i)
        if (!(m->pw[i + 1] > m->pw[i])) return set_err(ctx, RT_E_ARG, "phi walls must increase");
    if (std::fabs(m->tw[0]) > 1e-12) return set_err(ctx, RT_E_ARG, "theta walls must start at 0");
    if (std::fabs(m->tw[m->nt] - 0.5 * M_PI) < 1e-9)
        m->mirror = 1;
    else if (std::fabs(m->tw[m->nt] - M_PI) < 1e-9)
        m->mirror = 0;
    else
        return set_err(ctx, RT_E_ARG, "theta walls must end at pi/2 (mirror) or pi");
    if (std::fabs(m->pw[0]) > 1e-12 || std::fabs(m->pw[m->np] - 2 * M_PI) > 1e-9)
        return set_err(ctx, RT_E_ARG, "phi walls must span [0, 2 pi]");
    const size_t nrho = (size_t)m->nspec * m->ncell;
    if (!finite_all(d->h_rho, nrho)) return set_err(ctx, RT_E_ARG, "density contains NaN/Inf");
    for (size_t i = 0; i < nrho; ++i)
        if (d->h_rho[i] < 0) return set_err(ctx, RT_E_ARG, "negative density");
    m->rho.assign(d->h_rho, d->h_rho + nrho);
    m->lam.assign(d->h_lam_um, d->h_lam_um + m->nlam);
    for (int i = 0; i < m->nlam; ++i)
        if (!(m->lam[i] > 0) || (i > 0 && !(m->lam[i] > m->lam[i - 1])))
            return set_err(ctx, RT_E_ARG, "wavelengths must be positive and ascending");
    const size_t nk = (size_t)m->nspec * m->nlam;
    m->kabs.assign(d->h_kabs, d->h_kabs + nk);
    m->ksca.assign(d->h_ksca, d->h_ksca + nk);
    if (!finite_all(m->kabs.data(), nk) || !finite_all(m->ksca.data(), nk))
        return set_err(ctx, RT_E_ARG, "opacity contains NaN/Inf");
    for (size_t i = 0; i < nk; ++i)
        if (m->kabs[i] < 0 || m->ksca[i] < 0) return set_err(ctx, RT_E_ARG, "negative opacity");
    m->has_g = d->h_g != nullptr;
    if (m->has_g) {
        m->gsca.assign(d->h_g, d->h_g + nk);
        for (double gv : m->gsca)
            if (!(gv > -1.0 && gv < 1.0)) return set_err(ctx, RT_E_ARG, "HG g must be in (-1, 1)");
    }
    m->stars.assign(d->h_stars, d->h_stars + 5 * (size_t)m->nstar);
    for (int s = 0; s < m->nstar; ++s)
        if (!(m->stars[5 * s] > 0) || !(m->stars[5 * s + 1] > 0) || !finite_all(&m->stars[5 * s], 5))
            return set_err(ctx, RT_E_ARG, "star radius and temperature must be positive");
    std::string err;
    int rc = build_tables(m, err);
    if (rc) return set_err(ctx, rc, "%s", err.c_str());

    // geometry (lengths in au on the device)
    const double L = RT_LEN_UNIT;
    std::vector<float> rw(m->nr + 1), rw2(m->nr + 1), tw(m->nt + 1), ct(m->nt + 1), st(m->nt + 1);
    std::vector<double> rwd(m->nr + 1), ctd(m->nt + 1), std_(m->nt + 1);
    for (int i = 0; i <= m->nr; ++i) {
        rwd[i] = m->rw[i] / L;
        rw[i] = (float)rwd[i];
        rw2[i] = (float)(rwd[i] * rwd[i]);
    }
    for (int j = 0; j <= m->nt; ++j) {
        tw[j] = (float)m->tw[j];
        ctd[j] = std::cos(m->tw[j]);
        std_[j] = std::sin(m->tw[j]);
    }
    if (m->mirror) {
        ctd[m->nt] = 0.0;
        std_[m->nt] = 1.0;
    }
    for (int j = 0; j <= m->nt; ++j) {
        ct[j] = (float)ctd[j];
        st[j] = (float)std_[j];
    }
    std::vector<float> pw(m->np + 1), cp(m->np + 1), sp(m->np + 1);
    for (int k = 0; k <= m->np; ++k) {
        pw[k] = (float)m->pw[k];
        cp[k] = (float)std::cos(m->pw[k]);
        sp[k] = (float)std::sin(m->pw[k]);
    }
    if (upload(ctx, &m->d_rw, rw.data(), rw.size()) || upload(ctx, &m->d_rw2, rw2.data(), rw2.size()) ||
        upload(ctx, &m->d_rwd, rwd.data(), rwd.size()) || upload(ctx, &m->d_tw, tw.data(), tw.size()) ||
        upload(ctx, &m->d_ct, ct.data(), ct.size()) || upload(ctx, &m->d_st, st.data(), st.size()) ||
        upload(ctx, &m->d_ctd, ctd.data(), ctd.size()) || upload(ctx, &m->d_std, std_.data(), std_.size()) ||
        upload(ctx, &m->d_pw, pw.data(), pw.size()) || upload(ctx, &m->d_cp, cp.data(), cp.size()) ||
        upload(ctx, &m->d_sp, sp.data(), sp.size()))
        return RT_E_HIP;

    // cell volumes (mirror doubles them) and per-cell energy conversion
    const int G = m->tab.ngroup, nc = m->ncell;
    m->vol.resize(nc);
    for (int k = 0; k < m->np; ++k)
        for (int j = 0; j < m->nt; ++j)
            for (int i = 0; i < m->nr; ++i) {
                double r1 = m->rw[i], r2 = m->rw[i + 1];
                double vr = (r2 - r1) * (r2 * r2 + r1 * r2 + r1 * r1) / 3.0;
                double vt = ctd[j] - ctd[j + 1];
                m->vol[i + m->nr * (j + m->nt * k)] = vr * vt * (m->pw[k + 1] - m->pw[k]) * (m->mirror ? 2.0 : 1.0);
            }
    std::vector<float> invdr(m->nr);
    for (int i = 0; i < m->nr; ++i) invdr[i] = (float)(1.0 / (rwd[i + 1] - rwd[i]));
    // opacities and tables
    OpacArrays o;
    opacity_arrays(m, o);
    std::vector<float> spos(3 * std::max(1, m->nstar), 0.f);
    for (int s = 0; s < m->nstar; ++s)
        for (int a = 0; a < 3; ++a) spos[3 * s + a] = (float)(m->stars[5 * s + 2 + a] / L);
    const size_t ncg = (size_t)nc * G;
    if (upload<float>(ctx, &m->d_rhoL, nullptr, nrho) || upload<float>(ctx, &m->d_rho, nullptr, nrho) ||
        upload<unsigned char>(ctx, &m->d_rhopos, nullptr, nrho) ||
        upload(ctx, &m->d_ka, o.ka.data(), nk) || upload(ctx, &m->d_ks, o.ks.data(), nk) ||
        upload(ctx, &m->d_g, o.gg.data(), nk) || upload(ctx, &m->d_grp, m->tab.grp.data(), m->nspec) ||
        upload(ctx, &m->d_gka, o.gka.data(), o.gka.size()) ||
        upload(ctx, &m->d_ginvk, o.ginvk.data(), o.ginvk.size()) ||
        upload(ctx, &m->d_invdr, invdr.data(), invdr.size()) ||
        upload(ctx, &m->d_cdf, m->tab.cdf.data(), m->tab.cdf.size()) ||
        upload(ctx, &m->d_guide, m->tab.guide.data(), m->tab.guide.size()) ||
        upload(ctx, &m->d_lnq, m->tab.lnq.data(), m->tab.lnq.size()) ||
        upload(ctx, &m->d_scdf, m->tab.scdf.data(), m->tab.scdf.size()) ||
        upload(ctx, &m->d_sguide, m->tab.sguide.data(), m->tab.sguide.size()) ||
        upload(ctx, &m->d_spick, m->tab.spick.data(), m->tab.spick.size()) ||
        upload(ctx, &m->d_spos, spos.data(), spos.size()) ||
        upload(ctx, &m->d_kross, m->tab.kross.data(), m->tab.kross.size()) ||
        upload(ctx, &m->d_kplanck, m->tab.kplanck.data(), m->tab.kplanck.size()) ||
        upload(ctx, &m->d_mrwy, m->tab.mrw_y.data(), m->tab.mrw_y.size()) ||
        upload(ctx, &m->d_econv, o.econv.data(), o.econv.size()) ||
        upload<unsigned long long>(ctx, &m->d_tally, nullptr, ncg) ||
        upload<float>(ctx, &m->d_xrow, nullptr, ncg) || upload<float>(ctx, &m->d_Tg, nullptr, ncg) ||
        upload<unsigned int>(ctx, &m->d_visits, nullptr, nc) ||
        upload<unsigned long long>(ctx, &m->d_cnt, nullptr, CNT_N) ||
        upload<float>(ctx, &m->d_Tout, nullptr, nrho))
        return RT_E_HIP;
    HIPCHK(ctx, hipMemset(m->d_visits, 0, sizeof(unsigned int) * nc));
    HIPCHK(ctx, hipMemset(m->d_Tout, 0, sizeof(float) * nrho));
    HIPCHK(ctx, hipMemset(m->d_Tg, 0, sizeof(float) * ncg));
    return upload_density(m);
}

extern "C" {

int rt_create(int device, rt_ctx **out) {
    if (!out) return RT_E_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= device || device < 0) return RT_E_HIP;
    if (hipSetDevice(device) != hipSuccess) return RT_E_HIP;
    rt_ctx *c = new rt_ctx();
    c->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
        delete c;
        return RT_E_HIP;
    }
    c->n_cu = prop.multiProcessorCount;
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) != hipSuccess || khz <= 0)
        khz = 100000;
    c->wallclock_khz = khz;
    // the allocation-free sampler's scratch address on this device
    if (vis_scratch_init(c, device) != RT_OK) {
        delete c;
        return RT_E_HIP;
    }
    if (hipMalloc(&c->d_icount, 2 * sizeof(unsigned long long)) != hipSuccess ||
        hipMemset(c->d_icount, 0, 2 * sizeof(unsigned long long)) != hipSuccess) {
        delete c;
        return RT_E_HIP;
    }
    *out = c;
    return RT_OK;
}

rt_ctx::~rt_ctx() {
    void *p[] = {ref.xy, ref.val, ref.child, ref.gw, ref.ctl, d_icount};
    for (void *q : p)
        if (q) (void)hipFree(q);
}

int rt_destroy(rt_ctx *ctx) {
    delete ctx;
    return RT_OK;
}

int rt_image_crossings(rt_ctx *ctx, uint64_t *image, uint64_t *adjoint, int reset) {
    if (!ctx || !ctx->d_icount) return RT_E_ARG;
    unsigned long long c[2] = {0, 0};
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    HIPCHK(ctx, hipMemcpy(c, ctx->d_icount, sizeof(c), hipMemcpyDeviceToHost));
    if (reset) HIPCHK(ctx, hipMemset(ctx->d_icount, 0, sizeof(c)));
    if (image) *image = c[0];
    if (adjoint) *adjoint = c[1];
    return RT_OK;
}

int rt_refine_stats(rt_ctx *ctx, uint64_t *nrays, uint64_t *unsplit) {
    if (!ctx) return RT_E_ARG;
    unsigned c[RT_REF_NCTL] = {0};
    if (ctx->ref.ctl) {
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        HIPCHK(ctx, hipMemcpy(c, ctx->ref.ctl, sizeof(c), hipMemcpyDeviceToHost));
        unsigned z = 0;
        HIPCHK(ctx, hipMemcpy(ctx->ref.ctl + RT_REF_UNSPLIT, &z, sizeof(z), hipMemcpyHostToDevice));
    }
    if (nrays) *nrays = c[RT_REF_RAYS];
    if (unsplit) *unsplit = c[RT_REF_UNSPLIT];
    return RT_OK;
}

int rt_refine_reserve(rt_ctx *ctx, uint64_t n) {
    if (!ctx || n == 0) return RT_E_ARG;
    return refine_reserve(ctx, (size_t)n);
}

int rt_set_stream(rt_ctx *ctx, void *s) {
    if (!ctx) return RT_E_ARG;
    ctx->stream = (hipStream_t)s;
    return RT_OK;
}

// the message is copied under the context's lock into a buffer of the calling
// thread, so a message written meanwhile by another thread (opacity updates of
// several models run on host threads) never tears the one returned; the
// pointer stays valid until this thread's next call
const char *rt_last_error(const rt_ctx *ctx) {
    if (!ctx) return "no context";
    thread_local std::string copy;
    std::lock_guard<std::mutex> lk(const_cast<rt_ctx *>(ctx)->err_mu);
    copy = ctx->err;
    return copy.c_str();
}

int rt_device_info(const rt_ctx *ctx, int *n_cu, int *khz) {
    if (!ctx) return RT_E_ARG;
    if (n_cu) *n_cu = ctx->n_cu;
    if (khz) *khz = ctx->wallclock_khz;
    return RT_OK;
}

int rt_model_create(rt_ctx *ctx, const rt_model_desc *desc, rt_model **out) {
    if (!ctx || !out) return RT_E_ARG;
    *out = nullptr;
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, RT_E_HIP, "hipSetDevice failed");
    rt_model *m = new rt_model();
    int rc = create_model(ctx, desc, m);
    if (rc) {
        free_model(m);
        delete m;
        return rc;
    }
    *out = m;
    return RT_OK;
}

int rt_model_destroy(rt_model *m) {
    if (!m) return RT_E_ARG;
    free_model(m);
    delete m;
    return RT_OK;
}

int rt_model_ncells(const rt_model *m) { return m ? m->ncell : RT_E_ARG; }

int rt_model_set_density(rt_model *m, const double *h_rho) {
    if (!m || !h_rho) return RT_E_ARG;
    size_t n = (size_t)m->nspec * m->ncell;
    if (!finite_all(h_rho, n)) return set_err(m->ctx, RT_E_ARG, "density contains NaN/Inf");
    for (size_t i = 0; i < n; ++i)
        if (h_rho[i] < 0) return set_err(m->ctx, RT_E_ARG, "negative density");
    m->rho.assign(h_rho, h_rho + n);
    m->have_T = false;
    m->rho_host_stale = false;
    return upload_density(m);
}

static int set_opacity(rt_model *m, const double *h_kabs, const double *h_ksca,
                       const double *h_g, bool on_device) {
    if (!m || !h_kabs || !h_ksca) return RT_E_ARG;
    rt_ctx *ctx = m->ctx;
    if ((h_g != nullptr) != m->has_g)
        return set_err(ctx, RT_E_ARG, "g must be given exactly when the model was created with g");
    const size_t nk = (size_t)m->nspec * m->nlam;
    if (!finite_all(h_kabs, nk) || !finite_all(h_ksca, nk))
        return set_err(ctx, RT_E_ARG, "opacity contains NaN/Inf");
    for (size_t i = 0; i < nk; ++i)
        if (h_kabs[i] < 0 || h_ksca[i] < 0) return set_err(ctx, RT_E_ARG, "negative opacity");
    if (h_g)
        for (size_t i = 0; i < nk; ++i)
            if (!(h_g[i] > -1.0 && h_g[i] < 1.0)) return set_err(ctx, RT_E_ARG, "HG g must be in (-1, 1)");
    // the new tables must keep the device layout: same species grouping; the
    // old opacities (and, for the host build, the old host tables) are kept
    // for a rollback
    std::vector<double> old_kabs, old_ksca, old_g;
    old_kabs.swap(m->kabs);
    old_ksca.swap(m->ksca);
    old_g.swap(m->gsca);
    m->kabs.assign(h_kabs, h_kabs + nk);
    m->ksca.assign(h_ksca, h_ksca + nk);
    if (h_g) m->gsca.assign(h_g, h_g + nk);
    std::string err;
    int rc;
    if (on_device) {
        // groups, frequency grid and kmax on the host (the host's emission
        // tables are not rebuilt, so only these fields change); the rest on
        // the device
        TableSpec sp{m->nlam, m->nspec, m->ntemp, m->nstar, m->temp0, m->temp1,
                     m->lam.data(), m->kabs.data(), m->ksca.data(), m->stars.data()};
        HostTables g;
        rc = host_table_groups(sp, g, err);
        if (!rc && g.grp != m->tab.grp) {
            rc = RT_E_ARG;
            err = "the new opacities group the species differently (species with identical kabs "
                  "share tables); create a new model instead";
        }
        if (rc) {
            m->kabs.swap(old_kabs);
            m->ksca.swap(old_ksca);
            m->gsca.swap(old_g);
            return set_err(ctx, rc, "%s", err.c_str());
        }
        m->tab.nu.swap(g.nu);
        m->tab.dnu.swap(g.dnu);
        m->tab.grep.swap(g.grep);
        m->tab.kmax.swap(g.kmax);
        m->tab.lnt0 = g.lnt0;
        m->tab.dlnt = g.dlnt;
        // stream-ordered: one staged copy and kernels on the context's stream
        rc = opacity_device_update(m);
        if (rc) return rc;
        m->have_T = false;
        m->opac_version++;
        return RT_OK;
    }
    const HostTables old_tab = m->tab;
    rc = build_tables(m, err);
    if (!rc && m->tab.grp != old_tab.grp) {
        rc = RT_E_ARG;
        err = "the new opacities group the species differently (species with identical kabs "
              "share tables); create a new model instead";
    }
    if (!rc &&
        (m->tab.cdf.size() != old_tab.cdf.size() || m->tab.kross.size() != old_tab.kross.size())) {
        rc = RT_E_ARG;
        err = "table sizes changed";
    }
    if (rc) {
        m->kabs.swap(old_kabs);
        m->ksca.swap(old_ksca);
        m->gsca.swap(old_g);
        m->tab = old_tab;
        return set_err(ctx, rc, "%s", err.c_str());
    }
    {
        OpacArrays o;
        opacity_arrays(m, o);
        const HostTables &t = m->tab;
        HIPCHK(ctx, hipSetDevice(ctx->device));
        HIPCHK(ctx, hipMemcpy(m->d_ka, o.ka.data(), nk * sizeof(float), hipMemcpyHostToDevice));
        HIPCHK(ctx, hipMemcpy(m->d_ks, o.ks.data(), nk * sizeof(float), hipMemcpyHostToDevice));
        HIPCHK(ctx, hipMemcpy(m->d_g, o.gg.data(), nk * sizeof(float), hipMemcpyHostToDevice));
        HIPCHK(ctx, hipMemcpy(m->d_gka, o.gka.data(), o.gka.size() * sizeof(float), hipMemcpyHostToDevice));
        HIPCHK(ctx, hipMemcpy(m->d_ginvk, o.ginvk.data(), o.ginvk.size() * sizeof(float), hipMemcpyHostToDevice));
        HIPCHK(ctx, hipMemcpy(m->d_econv, o.econv.data(), o.econv.size() * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(ctx, hipMemcpy(m->d_cdf, t.cdf.data(), t.cdf.size() * sizeof(float), hipMemcpyHostToDevice));
        HIPCHK(ctx, hipMemcpy(m->d_guide, t.guide.data(), t.guide.size(), hipMemcpyHostToDevice));
        HIPCHK(ctx, hipMemcpy(m->d_lnq, t.lnq.data(), t.lnq.size() * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(ctx, hipMemcpy(m->d_kross, t.kross.data(), t.kross.size() * sizeof(float), hipMemcpyHostToDevice));
        HIPCHK(ctx, hipMemcpy(m->d_kplanck, t.kplanck.data(), t.kplanck.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    m->have_T = false;
    m->opac_version++;
    return RT_OK;
}

int rt_model_set_opacity(rt_model *m, const double *h_kabs, const double *h_ksca,
                         const double *h_g) {
    return set_opacity(m, h_kabs, h_ksca, h_g, false);
}

int rt_model_set_opacity_device(rt_model *m, const double *h_kabs, const double *h_ksca,
                                const double *h_g) {
    return set_opacity(m, h_kabs, h_ksca, h_g, true);
}

int rt_model_get_tables(rt_model *m, double *h_lnq, float *h_cdf, unsigned char *h_guide,
                        float *h_kplanck, float *h_kross) {
    if (!m) return RT_E_ARG;
    rt_ctx *ctx = m->ctx;
    const size_t G = (size_t)m->tab.ngroup, nT = (size_t)m->ntemp, nl = (size_t)m->nlam;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if (h_lnq) HIPCHK(ctx, hipMemcpy(h_lnq, m->d_lnq, G * nT * sizeof(double), hipMemcpyDeviceToHost));
    if (h_cdf) HIPCHK(ctx, hipMemcpy(h_cdf, m->d_cdf, G * nT * nl * sizeof(float), hipMemcpyDeviceToHost));
    if (h_guide) HIPCHK(ctx, hipMemcpy(h_guide, m->d_guide, G * nT * RT_NGUIDE, hipMemcpyDeviceToHost));
    if (h_kplanck) HIPCHK(ctx, hipMemcpy(h_kplanck, m->d_kplanck, G * nT * sizeof(float), hipMemcpyDeviceToHost));
    if (h_kross)
        HIPCHK(ctx, hipMemcpy(h_kross, m->d_kross, (size_t)m->nspec * nT * sizeof(float), hipMemcpyDeviceToHost));
    return RT_OK;
}

int rt_model_get_opacity(rt_model *m, float *h_ka, float *h_ks, float *h_g, float *h_gka,
                         float *h_ginvk, double *h_econv) {
    if (!m) return RT_E_ARG;
    rt_ctx *ctx = m->ctx;
    const size_t nk = (size_t)m->nspec * m->nlam, G = (size_t)m->tab.ngroup;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if (h_ka) HIPCHK(ctx, hipMemcpy(h_ka, m->d_ka, nk * sizeof(float), hipMemcpyDeviceToHost));
    if (h_ks) HIPCHK(ctx, hipMemcpy(h_ks, m->d_ks, nk * sizeof(float), hipMemcpyDeviceToHost));
    if (h_g) HIPCHK(ctx, hipMemcpy(h_g, m->d_g, nk * sizeof(float), hipMemcpyDeviceToHost));
    if (h_gka) HIPCHK(ctx, hipMemcpy(h_gka, m->d_gka, G * m->nlam * sizeof(float), hipMemcpyDeviceToHost));
    if (h_ginvk) HIPCHK(ctx, hipMemcpy(h_ginvk, m->d_ginvk, G * sizeof(float), hipMemcpyDeviceToHost));
    if (h_econv)
        HIPCHK(ctx, hipMemcpy(h_econv, m->d_econv, (size_t)m->ncell * G * sizeof(double), hipMemcpyDeviceToHost));
    return RT_OK;
}

int rt_model_ngroups(const rt_model *m) { return m ? m->tab.ngroup : RT_E_ARG; }

int rt_model_generate_density(rt_model *m, const rt_density_comp *h_comps, int ncomp) {
    if (!m) return RT_E_ARG;
    return density_generate(m, h_comps, ncomp);
}

int rt_scattering_source(rt_model *m, double lam_um, uint64_t nphot, uint64_t seed,
                         const float *d_T, float *d_J, rt_scat_stats *stats) {
    if (!m) return RT_E_ARG;
    return scat_source_run(m, lam_um, nphot, seed, d_T, d_J, stats);
}

int rt_get_density_device(rt_model *m, float *d_rho) {
    if (!m || !d_rho) return RT_E_ARG;
    rt_ctx *ctx = m->ctx;
    HIPCHK(ctx, hipMemcpyAsync(d_rho, m->d_rho, sizeof(float) * m->nspec * m->ncell,
                               hipMemcpyDeviceToDevice, ctx->stream));
    return RT_OK;
}

int rt_thermal(rt_model *m, uint64_t nphot, uint64_t seed, const rt_thermal_opts *opts,
               rt_thermal_stats *stats) {
    if (!m) return RT_E_ARG;
    return thermal_run(m, nphot, seed, opts, stats);
}

int rt_get_temperature(rt_model *m, double *h_T, uint64_t *h_visits) {
    if (!m || !h_T) return RT_E_ARG;
    rt_ctx *ctx = m->ctx;
    if (!m->have_T) return set_err(ctx, RT_E_ARG, "no thermal run on this model yet");
    size_t n = (size_t)m->nspec * m->ncell;
    std::vector<float> T(n);
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    HIPCHK(ctx, hipMemcpy(T.data(), m->d_Tout, n * sizeof(float), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; ++i) h_T[i] = T[i];
    if (h_visits) {
        std::vector<unsigned int> v(m->ncell);
        HIPCHK(ctx, hipMemcpy(v.data(), m->d_visits, m->ncell * sizeof(unsigned int), hipMemcpyDeviceToHost));
        for (int c = 0; c < m->ncell; ++c) h_visits[c] = v[c];
    }
    return RT_OK;
}

int rt_thermal_timeline(const rt_model *m, double *h_ms, int cap) {
    if (!m) return RT_E_ARG;
    int n = (int)m->timeline.size();
    if (h_ms)
        for (int i = 0; i < n && i < cap; ++i) h_ms[i] = m->timeline[i];
    return n;
}

int rt_get_temperature_device(rt_model *m, float *d_T) {
    if (!m || !d_T) return RT_E_ARG;
    rt_ctx *ctx = m->ctx;
    if (!m->have_T) return set_err(ctx, RT_E_ARG, "no thermal run on this model yet");
    HIPCHK(ctx, hipMemcpyAsync(d_T, m->d_Tout, sizeof(float) * m->nspec * m->ncell,
                               hipMemcpyDeviceToDevice, ctx->stream));
    return RT_OK;
}

int rt_image(rt_model *m, const rt_camera *cam, const float *d_T, const float *d_rho, float *d_img) {
    if (!m || !d_img) return RT_E_ARG;
    return image_run(m, cam, d_T, d_rho, d_img);
}

int rt_image_adjoint(rt_model *m, const rt_camera *cam, const float *d_T, const float *d_rho,
                     const float *d_dimg, float *d_g_logrho, float *d_g_T) {
    if (!m) return RT_E_ARG;
    return image_adjoint_run(m, cam, d_T, d_rho, d_dimg, d_g_logrho, d_g_T);
}

int rt_vis_dft(rt_ctx *ctx, const float *d_img, int nx, int ny, double dxy, const double *d_u,
               const double *d_v, int K, double dra, double ddec, float *d_re, float *d_im) {
    if (!ctx) return RT_E_ARG;
    return vis_run(ctx, d_img, nx, ny, dxy, d_u, d_v, K, dra, ddec, d_re, d_im);
}

int rt_vis_dft_adjoint(rt_ctx *ctx, const float *d_gre, const float *d_gim, int nx, int ny,
                       double dxy, const double *d_u, const double *d_v, int K, double dra,
                       double ddec, float *d_g_img) {
    if (!ctx) return RT_E_ARG;
    return vis_adjoint_run(ctx, d_gre, d_gim, nx, ny, dxy, d_u, d_v, K, dra, ddec, d_g_img);
}

int rt_vis_grid(rt_ctx *ctx, const float *d_img, int nx, int ny, double dxy, const double *d_u,
                const double *d_v, int K, double dra, double ddec, float *d_re, float *d_im,
                void *d_work) {
    if (!ctx) return RT_E_ARG;
    return vis_grid_run(ctx, d_img, nx, ny, dxy, d_u, d_v, K, dra, ddec, d_re, d_im, d_work, 0);
}

int rt_vis_grid_noalloc(rt_ctx *ctx, const float *d_img, int nx, int ny, double dxy,
                        const double *d_u, const double *d_v, int K, double dra, double ddec,
                        float *d_re, float *d_im) {
    if (!ctx) return RT_E_ARG;
    return vis_grid_noalloc_run(ctx, d_img, nx, ny, dxy, d_u, d_v, K, dra, ddec, d_re, d_im);
}

int rt_interpolate_galario(rt_ctx *ctx, const float *d_img, int nx, int ny, double dxy,
                           const double *d_u, const double *d_v, int K, double dra, double ddec,
                           float *d_re, float *d_im, void *d_work) {
    if (!ctx) return RT_E_ARG;
    return vis_grid_run(ctx, d_img, nx, ny, dxy, d_u, d_v, K, dra, ddec, d_re, d_im, d_work, 1);
}

size_t rt_vis_grid_workspace(int nx, int ny) { return vis_grid_workspace(nx, ny); }

int rt_chisq(rt_ctx *ctx, const float *d_dre, const float *d_dim, const float *d_w,
             const float *d_mre, const float *d_mim, int K, double *d_out, float *d_gre,
             float *d_gim) {
    if (!ctx) return RT_E_ARG;
    return chisq_run(ctx, d_dre, d_dim, d_w, d_mre, d_mim, K, d_out, d_gre, d_gim, nullptr);
}

// ---------------------------------------------------------------- batches
int rt_batch_create(rt_ctx *ctx, rt_model *const *models, int nmodel, rt_batch **out) {
    if (!ctx) return RT_E_ARG;
    return batch_create(ctx, models, nmodel, out);
}

int rt_batch_destroy(rt_batch *b) { return batch_destroy(b); }

int rt_batch_thermal(rt_batch *b, const uint64_t *nphot, const uint64_t *seeds,
                     const rt_thermal_opts *opts, rt_thermal_stats *stats) {
    if (!b) return RT_E_ARG;
    return batch_thermal(b, nphot, seeds, opts, stats);
}

int rt_batch_step(rt_batch *b, const uint64_t *nphot, const uint64_t *seeds,
                  const rt_thermal_opts *opts, const rt_step_item *items, int use_graph,
                  rt_thermal_stats *stats, double *step_ms) {
    if (!b) return RT_E_ARG;
    return batch_step(b, nphot, seeds, opts, items, use_graph, stats, step_ms);
}

}  // extern "C"
