/*
 * pdspy_amd.h -- C ABI of the MI355X (gfx950) continuum radiative-transfer
 * path that pdspy's Model drives.  It replaces the process-and-files boundary
 * pdspy uses today:
 *
 *   Model.write_radmc3d         pdspy/modeling/Model.py:723-778   -> rt_model_create
 *   radmc3d mctherm (subprocess) pdspy/radmc3d/run.py:9-38
 *     + read.dust_temperature    pdspy/radmc3d/read.pyx:433-478   -> rt_thermal,
 *                                                                    rt_get_temperature
 *   radmc3d image (os.system)    pdspy/radmc3d/run.py:157-289
 *     + read.image               pdspy/radmc3d/read.pyx:25-152    -> rt_image
 *   imtovis / interpolate_model  pdspy/imaging/imtovis.py:6-27,
 *                                pdspy/interferometry/interpolate_model.py:11-57
 *                                                                 -> rt_vis_dft
 *   chisq                        pdspy/interferometry/libinterferometry.pyx:610-633,
 *                                pdspy/utils/emcee.py:31-43      -> rt_chisq
 *
 * Conventions: plain C types only; every function returns 0 on success and a
 * negative RT_E* code on failure (rt_last_error gives the message).  Pointer
 * arguments named h_* are host memory, d_* are device memory of the context's
 * GPU; buffers are owned by the caller.  A context is bound to one device and
 * one HIP stream and is not thread-safe.
 */
#ifndef PDSPY_AMD_H
#define PDSPY_AMD_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    RT_OK = 0,
    RT_E_ARG = -1,      /* invalid argument or input data (NaN, bad walls ...) */
    RT_E_HIP = -2,      /* HIP runtime error */
    RT_E_NOMEM = -3,    /* device allocation failed */
    RT_E_WATCHDOG = -4, /* kernel time budget exceeded; packets were killed */
    RT_E_OVERFLOW = -5, /* a fixed-point energy tally overflowed */
    RT_E_LIMIT = -6     /* a compile-time limit (species, wavelengths) exceeded */
};

typedef struct rt_ctx rt_ctx;
typedef struct rt_model rt_model;

/* Model description, as pdspy writes it for RADMC-3D (Model.py:723-778):
 * walls in cm / rad / rad (amr_grid.inp), densities in g/cm^3 with r fastest
 * (dust_density.inp, write.py:259-264), opacities on the wavelength grid
 * (wavelength_micron.inp + dustkappa_*.inp), blackbody point stars
 * (stars.inp with -T, write.py:148-172).  theta walls must start at 0 and end
 * at pi/2 (equatorial mirror, as RADMC-3D grids from pdspy do) or pi;
 * phi walls span [0, 2 pi].                                                  */
typedef struct {
    int nr, ntheta, nphi;
    const double *h_r_walls;     /* nr+1, cm */
    const double *h_theta_walls; /* ntheta+1, rad */
    const double *h_phi_walls;   /* nphi+1, rad */
    int nspec;
    const double *h_rho; /* [nspec][nphi][ntheta][nr] g/cm^3 */
    int nlam;
    const double *h_lam_um; /* nlam, ascending */
    const double *h_kabs;   /* [nspec][nlam] cm^2/g */
    const double *h_ksca;   /* [nspec][nlam] cm^2/g */
    const double *h_g;      /* [nspec][nlam] HG asymmetry, NULL = isotropic */
    int nstar;
    const double *h_stars; /* [nstar][5] = R_cm, T_K, x_cm, y_cm, z_cm */
    int ntemp;             /* emission-table T grid (RADMC-3D ntemp), 0 = 1000 */
    double temp0, temp1;   /* its range in K (RADMC-3D temp0/temp1), 0 = 0.01/1e5 */
} rt_model_desc;

/* Thermal Monte Carlo options (the radmc3d.inp keys pdspy passes,
 * write.py:1-146, plus GPU launch controls).  Zero fields take defaults. */
typedef struct {
    uint64_t max_events;   /* per-packet interaction cap (default 1e6)         */
    uint64_t epoch_first;  /* packets in the first epoch (default max(1024,N/4096)) */
    double epoch_growth;   /* epoch k+1 adds max(first, N_done*(growth-1)); 1.25;
                              epoch e re-emits with the temperatures of slots <= e-2 */
    int count_visits;      /* 1: count packet entries per cell                  */
    int mrw;               /* 1: modified random walk (RADMC-3D modified_random_walk) */
    double mrw_tauthres;   /* MRW trigger: Rosseland tau to the nearest wall (10) */
    int mrw_count_trigger; /* MRW trigger: interactions in the cell first (100)  */
    double mrw_gamma;      /* MRW sphere: radius = dist_to_wall - gamma/alpha_R (2) */
    double max_seconds;    /* watchdog per launch (default 60 s)                 */
    int block_size;        /* threads per block (0 = 256)                         */
    int blocks_per_cu;     /* persistent blocks per CU (0 = auto)                 */
    uint64_t slice_events; /* a packet of epoch e deposits into temperature slot
                              e + floor(interactions / slice_events) (default 1e4) */
    int no_park;           /* batches: 1 keeps every packet on the wavefront that
                              started it (default 0: a wavefront left with a few
                              packets whose model has none to hand out passes them
                              to another wavefront of that model); the results
                              are the same either way */
    int free_cus_per_xcd;  /* CUs of each of the 8 XCDs the persistent launch
                              leaves free (default 0: it takes every CU).  With 1,
                              kernels on other streams -- another run's resets,
                              images -- run beside it at once instead of waiting
                              for its waves to retire, at 3 % of the slots; for
                              runs on several streams.  Results do not depend
                              on it */
} rt_thermal_opts;

typedef struct {
    uint64_t n_cross, n_abs, n_scat, n_escape, n_killed, n_events, n_mrw;
    double eps;       /* packet energy, erg/s */
    double lum;       /* total stellar luminosity on the wavelength grid, erg/s */
    int n_epochs;
    int watchdog_fired;
    double kernel_ms; /* device time of the MC launches (HIP events) */
    double total_ms;  /* device time of the whole run incl. temperature updates */
    double max_epoch_ms;         /* longest single MC launch */
    uint64_t max_packet_steps;   /* most steps (crossings + interactions) of one packet */
    uint64_t wave_iters;  /* loop iterations summed over wavefronts; lane utilisation is
                             (n_cross + n_events) / (64 wave_iters)                      */
    uint64_t idle_iters;  /* of which a wavefront had no packet to step                 */
    uint64_t n_nudged;    /* packets moved 2^-16 r along their ray after bouncing
                             between two cells at a wall without moving             */
    uint64_t phase_clocks[8]; /* shader clocks summed over wavefronts spent in: refill,
                                 loop overhead, snapshot, idle, geometry, crossing,
                                 interaction, modified random walk; zero unless the
                                 library is built with RT_THERMAL_PROFILE              */
    uint64_t n_parked;    /* batches: packets handed from a wavefront left with a few
                             of them to another wavefront of the same model        */
} rt_thermal_stats;

/* Camera (RADMC-3D `image` options pdspy passes, Model.py:515-534):
 * parallel rays, pixel (j,k) centred at ((k - nx/2) pix, (j - ny/2) pix) cm
 * (integer division; pdspy's zoomau centring), observer direction
 * (sin i sin phi, -sin i cos phi, cos i), image rotated by posang.           */
typedef struct {
    int nx, ny, nsub;   /* pixels and sub-rays per pixel side (>=1) */
    double pix_cm;      /* pixel size in the model plane: pixelsize["] * dpc * AU */
    double incl_deg, posang_deg, phi_deg;
    double lam_um;      /* image wavelength; opacity log-log interpolated */
    int mode;           /* 0 intensity erg/s/cm^2/Hz/sr, 1 optical depth (tracetau) */
    int incl_star;      /* 1: add stars as point sources (pdspy nostar=False) */
    double tau_stop;    /* stop a ray once tau exceeds this (0 = 40) */
    int tile_w, tile_h; /* ray tile of one thread block, a multiple of 64 pixels
                           up to 256 (0, 0 = 64 x 1); results do not depend on it */
    const float *d_jscat; /* scattering source: mean intensity J_nu[cells] at lam_um
                             from rt_scattering_source; the extinction is then
                             rho (kabs + ksca) and the source function
                             (rho kabs B + rho ksca J) / (rho (kabs + ksca))
                             (isotropic scattering).  NULL: emission only, as
                             RADMC-3D with scattering_mode_max = 0.  The
                             adjoint holds J fixed.                           */
    int nrefine;          /* recursive sub-pixeling, RADMC-3D's camera_nrrefine
                             (write.py:57-58): at most this many levels, capped
                             at 24; 0 = off (one ray per pixel / sub-pixel)   */
    double refine_crit;   /* camera_refine_criterion (write.py:59-61; 0 = 1.0):
                             a (sub)pixel of side d is split into 2 x 2 when its
                             central ray crosses a cell of size l with
                             d > refine_crit * l, l = min(dr, r_c dtheta,
                             r_c sin(theta_c) dphi) at the cell centre; the
                             pixel is the mean of its sub-pixels.  Sub-pixels
                             live in a node pool of the context, grown as
                             needed by calls outside a graph capture          */
} rt_camera;

/* Scattering source (RADMC-3D's monochromatic Monte Carlo before an image,
 * nphot_scat packets: run.py:157-289 with write.py:34): nphot packets from
 * the stars and from the dust's thermal emission at the imaging temperature
 * d_T[cells], isotropic scattering, absorption ends a packet; d_J[cells] =
 * mean intensity (erg/s/cm^2/Hz/sr) at lam_um.  Philox key = seed;
 * bit-reproducible.  Synchronous.                                            */
typedef struct {
    uint64_t n_escape, n_abs, n_scat, n_cross, n_killed;
    double lum_star, lum_dust; /* source luminosities L_nu (erg/s/Hz) */
    double eps;                /* packet energy (erg/s/Hz)            */
    double kernel_ms;
} rt_scat_stats;
int rt_scattering_source(rt_model *m, double lam_um, uint64_t nphot, uint64_t seed,
                         const float *d_T, float *d_J, rt_scat_stats *stats);

/* A context is one device and one stream; the models created on it run their
 * work on that stream.  Contexts of one device are independent: two contexts
 * on two streams, each with its own models, used from two host threads, run
 * their thermal launches concurrently (a single-run launch's waves leave as
 * soon as its packets are all handed out, so a launch queued on the other
 * stream takes over their CUs). */
int rt_create(int device, rt_ctx **out);
int rt_destroy(rt_ctx *ctx);
/* HIP stream (hipStream_t) for all later work; NULL = the null stream. */
int rt_set_stream(rt_ctx *ctx, void *hip_stream);
/* Message of the context's last error; a copy owned by the calling thread,
 * valid until that thread's next call. */
const char *rt_last_error(const rt_ctx *ctx);
/* Number of CUs of the context's device and its wall-clock rate (kHz). */
int rt_device_info(const rt_ctx *ctx, int *n_cu, int *wallclock_khz);

/* Rays traced by the context's last sub-pixelled image (all levels) and
 * sub-pixels left unsplit for want of pool space since the last call of this
 * function (a captured graph cannot grow the pool; eager calls grow it and
 * retrace, so for them this stays 0).  Synchronises the context's stream. */
int rt_refine_stats(rt_ctx *ctx, uint64_t *nrays, uint64_t *unsplit);
/* Grows the context's sub-pixel pool to at least n nodes (n > 0); for graph
 * captures that refine.  Synchronous. */
int rt_refine_reserve(rt_ctx *ctx, uint64_t n);
/* Cells crossed by the context's camera rays since the last reset: *image by
 * image rays (sub-pixels included), *adjoint by the adjoints' reverse sweeps
 * (either pointer may be NULL); reset != 0 zeroes the counts.  Instrumentation
 * with no counterpart in the reference (the images' bytes per ray for their
 * roofline).  Synchronises the context's stream. */
int rt_image_crossings(rt_ctx *ctx, uint64_t *image, uint64_t *adjoint, int reset);

/* Uploads grid, densities and opacities, builds the emission tables.
 * Replaces write_radmc3d's input files (Model.py:723-778). */
int rt_model_create(rt_ctx *ctx, const rt_model_desc *desc, rt_model **out);
int rt_model_destroy(rt_model *m);
int rt_model_ncells(const rt_model *m);
/* Replace the model's densities (host, same layout as rt_model_desc.h_rho).
 * This and rt_model_set_opacity copy synchronously: no work using the model
 * may be in flight on the context's stream. */
int rt_model_set_density(rt_model *m, const double *h_rho);
/* Replace the dust opacities (pdspy: a new Dust per model draw,
 * DustGenerator.__call__, dust/DustGenerator.py:71-97, as run_disk_model.py:79-85
 * does for every walker): h_kabs, h_ksca [nspec][nlam] in cm^2/g on the model's
 * wavelength grid, h_g likewise exactly when the model was created with g.  The
 * emission, Planck- and Rosseland-mean tables are rebuilt; the species grouping
 * (species with identical kabs share tables) must stay as it was, else RT_E_ARG
 * and the model keeps its old opacities.  A batch's captured step graph is
 * recaptured on its next step. */
int rt_model_set_opacity(rt_model *m, const double *h_kabs, const double *h_ksca,
                         const double *h_g);
/* The same, stream-ordered: the groups and kmax are set on the host, then one
 * copy from a pinned staging buffer and two kernels on the context's stream
 * make the per-species arrays (the host's values bit for bit) and the
 * emission, re-emission and mean-opacity tables (FP64, the host builder's
 * operations and order: the FP32 tables agree with it to rounding).  Returns
 * without waiting for the device: work already on the stream finishes with the
 * old opacities, work enqueued after it sees the new ones.  For the per-walker
 * dust of a sampler step. */
int rt_model_set_opacity_device(rt_model *m, const double *h_kabs, const double *h_ksca,
                                const double *h_g);
/* Copies the model's device tables to the host (synchronises the stream):
 * ln q(T) [G][ntemp] (double), re-emission CDFs [G][ntemp][nlam], their guide
 * tables [G][ntemp][128] (uint8), Planck-mean absorption [G][ntemp],
 * Rosseland-mean extinction [nspec][ntemp]; any pointer may be NULL.  G =
 * rt_model_ngroups. */
int rt_model_get_tables(rt_model *m, double *h_lnq, float *h_cdf, unsigned char *h_guide,
                        float *h_kplanck, float *h_kross);
int rt_model_ngroups(const rt_model *m);
/* Copies the model's per-species device opacities to the host (synchronises
 * the stream), as the photon kernels read them: kabs, ksca, g [nspec][nlam]
 * (float), the groups' kabs [G][nlam], 2^28 / kmax [G] (float), the per-cell
 * tally-to-energy factors [cells][G] (double); any pointer may be NULL.  No
 * reference counterpart (RADMC-3D keeps these in its own memory); for checks
 * of rt_model_set_opacity_device against the host build. */
int rt_model_get_opacity(rt_model *m, float *h_ka, float *h_ks, float *h_g, float *h_gka,
                         float *h_ginvk, double *h_econv);

/* Density components generated on the device (pdspy's model builders, which
 * run on the host there):
 *   RT_COMP_DISK        Disk.density, pdspy/modeling/Disk.py:48-135
 *                       (mass, rmin, rmax, plrho, h0, plh)
 *   RT_COMP_ULRICH_EXT  UlrichEnvelopeExtended.density,
 *                       pdspy/modeling/UlrichEnvelopeExtended.py:37-110
 *                       (mass, rmin, rmax, rcent, cavpl, cavrfact, theta_open,
 *                       zoffset)
 * Masses in Msun, lengths in au, angles in degrees, as pdspy's add_disk /
 * add_ulrichextended_envelope take them.  Each component adds into the
 * density of `species`; species no component names get zero.           */
enum { RT_COMP_DISK = 1, RT_COMP_ULRICH_EXT = 2 };
typedef struct {
    int kind;
    int species;
    double mass, rmin, rmax;
    double plrho, h0, plh;                               /* disk */
    double rcent, cavpl, cavrfact, theta_open, zoffset;  /* envelope */
} rt_density_comp;
/* Replaces the model's densities with the sum of up to 8 components computed
 * on the device (FP64; the envelope's mu0 roots by the same Brent iteration
 * as the host builder).  Stream-ordered on the context's stream (the
 * components travel as a kernel argument; h_comps may be reused on return);
 * the host copy of the densities is not updated (rt_get_density_device reads
 * the result). */
int rt_model_generate_density(rt_model *m, const rt_density_comp *h_comps, int ncomp);
/* Device copy of the model's densities as float [nspec][cells]. */
int rt_get_density_device(rt_model *m, float *d_rho);

/* Thermal MC (RADMC-3D `mctherm`, run.py:9-38): nphot packets, Philox key =
 * seed.  Temperatures stay on the device; bit-reproducible for a given seed. */
int rt_thermal(rt_model *m, uint64_t nphot, uint64_t seed, const rt_thermal_opts *opts,
               rt_thermal_stats *stats);
/* Per-species temperatures [nspec][cells] (r fastest, dust_temperature.dat
 * order, read.pyx:433-478) and visits[cells] (may be NULL). */
int rt_get_temperature(rt_model *m, double *h_T, uint64_t *h_visits);
/* Pipeline timeline of the last rt_thermal (no counterpart in RADMC-3D; for
 * tuning): ms from the start of the Monte Carlo launch to the publication of
 * each epoch's temperature snapshot (n_epochs values, -1 if never published),
 * then to the launch's end.  Writes min(count, cap) values, returns count. */
int rt_thermal_timeline(const rt_model *m, double *h_ms, int cap);
/* Device copy of the same temperatures as float [nspec][cells]. */
int rt_get_temperature_device(rt_model *m, float *d_T);

/* Formal-solution image at one wavelength (RADMC-3D `image`, run.py:157-289,
 * with scattering_mode_max=0).  d_T[cells]: temperature used for imaging (pdspy
 * writes the density-weighted species mean, Model.py:759-769); d_rho
 * [nspec][cells] densities or NULL for the model's; d_img[ny][nx] output. */
int rt_image(rt_model *m, const rt_camera *cam, const float *d_T, const float *d_rho,
             float *d_img);
/* Adjoint of rt_image: given d_dimg = dL/dimg, writes dL/dlog(rho)
 * [nspec][cells] and dL/dT [cells].  tau is recomputed, not stored. */
int rt_image_adjoint(rt_model *m, const rt_camera *cam, const float *d_T,
                     const float *d_rho, const float *d_dimg, float *d_g_logrho,
                     float *d_g_T);

/* Visibilities of an image at K (u,v) points (wavelengths):
 * V = sum_jk I[j][k] exp(-2 pi i (u x_k + v y_j)), x_k = (k - nx/2) dxy,
 * y_j = (j - ny/2) dxy (rad), phase centre shifted by (dra, ddec) rad.
 * On the FFT grid this equals imtovis (imtovis.py:13-24). */
int rt_vis_dft(rt_ctx *ctx, const float *d_img, int nx, int ny, double dxy,
               const double *d_u, const double *d_v, int K, double dra, double ddec,
               float *d_re, float *d_im);
/* Transpose: d_g_img[j][k] = dL/dI given dL/dRe, dL/dIm at each point. */
int rt_vis_dft_adjoint(rt_ctx *ctx, const float *d_gre, const float *d_gim, int nx,
                       int ny, double dxy, const double *d_u, const double *d_v, int K,
                       double dra, double ddec, float *d_g_img);
/* Gridded sampler, galario's method (galario's sampleImage, called at
 * interpolate_model.py:20-27; Tazzari et al. 2018, MNRAS 476, 4527): the exact
 * DFT above on the image's pixel-conjugate grid (u_a = (a - nx/2)/(nx dxy),
 * v_b = (b - ny/2)/(ny dxy)), bilinear interpolation at each (u,v), then the
 * phase-centre shift.  Equal to rt_vis_dft at grid nodes; a point without
 * four grid nodes around it gets NaN.  2 <= nx, ny <= 4096.  d_work: device
 * scratch of rt_vis_grid_workspace(nx, ny) bytes, or NULL (allocated and freed
 * stream-ordered). */
int rt_vis_grid(rt_ctx *ctx, const float *d_img, int nx, int ny, double dxy,
                const double *d_u, const double *d_v, int K, double dra, double ddec,
                float *d_re, float *d_im, void *d_work);
size_t rt_vis_grid_workspace(int nx, int ny);
/* rt_vis_grid without any allocation (no counterpart in the reference): for
 * callers that cannot allocate device memory -- code under stream capture,
 * tight sampler loops, harnesses that forbid hipMalloc.  Its workspace is a
 * static device array of the library (16.8 MB on each GPU it runs on, held
 * until the process ends), so nx, ny <= 1024 (RT_E_LIMIT beyond).  Same
 * arguments and results as rt_vis_grid with d_work.  Every call on one GPU
 * shares that array: calls that overlap on one GPU (two streams, two host
 * threads) race and must be serialised by the caller. */
int rt_vis_grid_noalloc(rt_ctx *ctx, const float *d_img, int nx, int ny, double dxy,
                        const double *d_u, const double *d_v, int K, double dra, double ddec,
                        float *d_re, float *d_im);
/* pdspy's interpolate_model(code="galario") for one channel of a pdspy Image
 * (interpolate_model.py:20-27, which calls galario's
 * sampleImage(image[::-1], dxy, u, v, dRA, dDec) and negates the imaginary
 * part): galario's sampling of the row-flipped image -- transform on the
 * pixel-conjugate grid with the phase origin at pixel (nx/2, ny/2), bilinear
 * interpolation, shift factor exp(+2 pi i (u dra + v ddec)) -- returned
 * complex-conjugated.  Equivalently: the gridded sampler of the unflipped
 * transform evaluated as G(-u, -v) exp(-2 pi i (u dra + v ddec)) with
 * G = transform of image[::-1]; on grid nodes this is
 * exp(-2 pi i (u dra + v (ddec + dxy))) times rt_vis_dft at (-u, v).  Same
 * arguments, limits and workspace as rt_vis_grid; dra, ddec in rad. */
int rt_interpolate_galario(rt_ctx *ctx, const float *d_img, int nx, int ny, double dxy,
                           const double *d_u, const double *d_v, int K, double dra,
                           double ddec, float *d_re, float *d_im, void *d_work);
/* chi^2 = sum w ((dre-mre)^2 + (dim-mim)^2) into d_out[0] (double), and
 * (optional, non-NULL) d_gre/d_gim = d chi^2 / d(mre, mim). */
int rt_chisq(rt_ctx *ctx, const float *d_dre, const float *d_dim, const float *d_w,
             const float *d_mre, const float *d_mim, int K, double *d_out, float *d_gre,
             float *d_gim);

/* ---- batches: one sampler step for many models (no counterpart in pdspy,
 * whose fitting loop runs one model per MPI rank: run_disk_model.py:31-511,
 * utils/emcee.py:183-203).  The models are created as above and keep their
 * own grids, densities (rt_model_set_density between steps) and outputs. */
typedef struct rt_batch rt_batch;

/* What one model contributes to a step: its image, the data it is fitted to
 * and where the results go (device buffers, caller-owned). */
typedef struct {
    rt_camera cam;                  /* image of the model (mode 0)                 */
    double dxy;                     /* pixel size in radians for the visibilities  */
    double scale;                   /* image unit -> visibility unit, e.g. Jy per
                                       pixel per erg/s/cm^2/Hz/sr                   */
    const double *d_u, *d_v;        /* [K] (u,v) in wavelengths                     */
    int K;
    double dra, ddec;               /* phase-centre shift (rad)                     */
    const float *d_re, *d_im, *d_w; /* [K] data visibilities and weights            */
    float *d_img;                   /* out [ny][nx] image (intensity)               */
    double *d_chi2;                 /* out [1] chi^2                                */
    float *d_g_logrho;              /* out [nspec][cells] d chi^2 / d log rho, and  */
    float *d_g_T;                   /* out [cells] d chi^2 / d T_img; both NULL:
                                       no adjoint                                   */
    uint64_t nphot_scat;            /* scattering-source packets before the image
                                       (pdspy's fits: run_image nphot=1e5,
                                       run_disk_model.py:301-315); 0: emission
                                       only.  The step computes J itself; cam's
                                       d_jscat is ignored.                         */
} rt_step_item;

int rt_batch_create(rt_ctx *ctx, rt_model *const *models, int nmodel, rt_batch **out);
int rt_batch_destroy(rt_batch *b);
/* Thermal MC of every model in ONE persistent launch: nphot[i] packets with
 * Philox key seeds[i], the same options for all; stats: nmodel entries or
 * NULL.  Each model's temperatures equal those of rt_thermal with the same
 * arguments, bit for bit. */
int rt_batch_thermal(rt_batch *b, const uint64_t *nphot, const uint64_t *seeds,
                     const rt_thermal_opts *opts, rt_thermal_stats *stats);
/* One sampler step: the batched thermal launch, then per model the imaging
 * temperature (density-weighted species mean, Model.py:759-769), image,
 * visibilities, chi^2 (+ gradient) and the adjoint.  use_graph = 1 captures
 * the step as a hipGraph on first use and replays it while nphot, opts and
 * items stay the same (seeds and densities may change).  step_ms: device
 * time of the step (may be NULL). */
int rt_batch_step(rt_batch *b, const uint64_t *nphot, const uint64_t *seeds,
                  const rt_thermal_opts *opts, const rt_step_item *items, int use_graph,
                  rt_thermal_stats *stats, double *step_ms);

#ifdef __cplusplus
}
#endif
#endif
