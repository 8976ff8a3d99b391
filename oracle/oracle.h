/*
 * oracle.h -- FP64 CPU restatement of the continuum radiative-transfer path
 * that pdspy's Model drives (run_thermal -> RADMC-3D mctherm, run_image ->
 * RADMC-3D camera, run_visibilities / interpolate_model -> DFT, chi^2).
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * or the timed CPU baseline.  The product (pdspy_amd) never links or calls it.
 *
 * The reference's arithmetic for this path is the external RADMC-3D executable
 * (pdspy/radmc3d/run.py:9-38 mctherm, :157-289 image), which is not vendored in
 * the reference and is not installed here; this file restates the published
 * algorithms it implements (Bjorkman & Wood 2001 immediate re-emission with a
 * path-length absorbed-energy estimator, Lucy 1999) under the option set pdspy
 * writes (pdspy/radmc3d/write.py:1-146) and the data layout pdspy reads back
 * (pdspy/radmc3d/read.pyx:433-478, pdspy/modeling/Model.py:249-254).
 */
#ifndef PDSPY_AMD_ORACLE_H
#define PDSPY_AMD_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Model description.  All arrays are host arrays, FP64.
 * rw[nr+1] radial walls in cm, tw[nt+1] polar walls in rad (tw[0]=0,
 * tw[nt]=pi/2 means equatorial mirror, tw[nt]=pi a full sphere),
 * pw[np+1] azimuthal walls in rad (pw[0]=0, pw[np]=2pi).
 * rho[nspec][np][nt][nr] g/cm^3 (r fastest: the order pdspy writes
 * dust_density.inp, write.py:259-264).
 * lam_um[nlam] ascending; kabs/ksca/gsca[nspec][nlam] cm^2/g (gsca may be NULL).
 * star[nstar][5] = {R_cm, T_K, x_cm, y_cm, z_cm}.                           */
typedef struct {
    int nr, nt, np;
    const double *rw, *tw, *pw;
    int nspec;
    const double *rho;
    int nlam;
    const double *lam_um;
    const double *kabs, *ksca, *gsca;
    int nstar;
    const double *star;
} or_model;

typedef struct {
    uint64_t nphot;
    uint64_t seed;
    int ntemp;            /* T grid of the emission tables (RADMC-3D ntemp)  */
    double temp0, temp1;  /* its range in K (RADMC-3D temp0 / temp1)        */
    uint64_t max_events;  /* per-packet interaction cap (0 = none)          */
    int lagged;           /* 0: immediate B&W; 1: epoch-lagged temperatures */
    uint64_t epoch_first; /* first epoch size in packets (lagged mode)      */
    double epoch_growth;  /* later epochs: max(first, N_done*(growth-1))    */
    int nthreads;         /* OpenMP threads: lagged mode 0 = all; immediate
                             mode runs free-running in parallel when > 1
                             (schedule-dependent result, for timing)        */
    int mrw;              /* modified random walk on/off                    */
    double mrw_tauthres;  /* Rosseland tau to the nearest wall to trigger   */
    int mrw_count;        /* interactions in the cell before it may trigger */
    double mrw_gamma;     /* sphere radius = d_min - gamma / alpha_R        */
    int lag;              /* lagged mode: epoch e re-emits with the snapshot of
                             slots <= e - lag (default 2)                   */
    uint64_t slice_events;/* lagged mode: a packet of epoch e deposits into
                             slot e + floor(events / slice_events) (capped at
                             the epoch count); 0 = whole packets in slot e  */
} or_thermal_opts;

typedef struct {
    uint64_t n_cross, n_abs, n_scat, n_escape, n_killed, n_events, n_mrw;
    double eps;           /* packet energy L_tot/nphot, erg/s               */
    double lum;           /* total stellar luminosity on the grid, erg/s    */
    double tally_tau_sum; /* sum over cells/groups of eps*tally/eps        */
    int n_epochs;
    uint64_t n_nudged;    /* packets moved on by the ping-pong guard        */
} or_thermal_stats;

/* Thermal Monte Carlo.  T_out[nspec][cells] K (r fastest), visits[cells]
 * (packet entries per cell, may be NULL).  Returns 0 or a negative code.     */
int or_thermal(const or_model *m, const or_thermal_opts *o, double *T_out,
               uint64_t *visits, or_thermal_stats *st);

/* Scattering source: mean intensity J_nu[cells] (erg/s/cm^2/Hz/sr) at lam_um
 * from a monochromatic Monte Carlo with isotropic scattering (sources: stars
 * and the dust's thermal emission at the imaging temperature T[cells]).
 * kap_abs / kap_sca[nspec]: opacities at lam_um (cm^2/g).                    */
typedef struct {
    uint64_t n_escape, n_abs, n_scat, n_cross, n_killed;
    double lum_star, lum_dust; /* source luminosities L_nu (erg/s/Hz)      */
    double eps;                /* packet energy (L_star + L_dust) / nphot   */
} or_scat_stats;
int or_scat_source(const or_model *m, const double *T, const double *kap_abs,
                   const double *kap_sca, double lam_um, uint64_t nphot, uint64_t seed,
                   uint64_t max_events, int nthreads, double *J, or_scat_stats *st);

/* Camera.  Parallel rays; pixel (j,k) centre at ((k - nx/2)*pix, (j - ny/2)*pix)
 * cm in the image plane (integer division, the centring pdspy's zoomau
 * produces, Model.py:522-527), rotated by posang.  incl/posang/phi in degrees.
 * kap[nspec] = absorption opacity at the image wavelength (cm^2/g),
 * T[cells] = the single temperature per cell used for imaging
 * (Model.py:759-769 averages species).  mode 0 = intensity
 * (erg/s/cm^2/Hz/sr), mode 1 = optical depth through the model (tracetau).
 * incl_star adds the point-source stars (as pdspy's nostar=False).          */
typedef struct {
    int nx, ny, nsub;
    double pix_cm;
    double incl, posang, phi;
    double lam_um;
    int mode;
    int incl_star;
    double tau_stop;
    /* recursive sub-pixeling (RADMC-3D camera_nrrefine /
     * camera_refine_criterion): a (sub)pixel of side d whose central ray
     * crosses a cell of size l with d > crit l is replaced by the mean of its
     * 2x2 sub-pixels, at most nrefine levels deep (0 = off); a smaller crit
     * refines further.  Cell size: the
     * least of dr, r_c dtheta and (3D) r_c sin(theta_c) dphi, r_c and theta_c
     * the cell's centre.                                                     */
    int nrefine;
    double refine_crit;
} or_camera;

int or_image(const or_model *m, const double *T, const double *kap,
             const or_camera *cam, double *img);
/* rays traced by the last or_image / or_image_scat call (sub-pixels included) */
long or_image_last_nrays(void);

/* Reverse-mode gradient of L = sum(dimg * img) w.r.t. log(rho[s][c]) and
 * T[c], FP64.  The star term is included when cam->incl_star.              */
int or_image_adjoint(const or_model *m, const double *T, const double *kap,
                     const or_camera *cam, const double *dimg,
                     double *g_logrho, double *g_T);

/* The same with a scattering source: extinction rho (kabs + ksca), source
 * (rho kabs B + rho ksca J) / (rho (kabs + ksca)); J[cells] from
 * or_scat_source, held fixed in the gradient.  ksca/J NULL = emission only. */
int or_image_scat(const or_model *m, const double *T, const double *kap,
                  const double *ksca, const double *J, const or_camera *cam, double *img);
int or_image_adjoint_scat(const or_model *m, const double *T, const double *kap,
                          const double *ksca, const double *J, const or_camera *cam,
                          const double *dimg, double *g_logrho, double *g_T);

/* Ray segments through the grid: origin o[3], direction d[3] (unit); writes up
 * to maxseg (cell, length) pairs in crossing order, cell = -1 for the inner
 * hole.  Returns the number of segments or a negative code.                 */
int or_trace_ray(const or_model *m, const double *o, const double *d,
                 int maxseg, int *cell, double *len, int *hemi);
/* the same, counting the ping-pong guard's moves in *nudged (nullable) */
int or_trace_ray_n(const or_model *m, const double *o, const double *d, int maxseg,
                   int *cell, double *len, int *hemi, int *nudged);
/* the same from a position inside the grid whose cell indices are given,
 * idx = (ir, it, ip, hemi), not located from the position (a test sets up a
 * position a hair past a wall with the index one cell behind it); -1 for
 * indices outside the grid */
int or_trace_ray_from(const or_model *m, const double *o, const double *d, const int *idx,
                      int maxseg, int *cell, double *len, int *hemi);

/* Direct DFT: V(u,v) = sum_jk I[j][k] exp(-2 pi i (u x_k + v y_j)) with
 * x_k=(k-nx/2)dxy, y_j=(j-ny/2)dxy (rad), times exp(-2 pi i (u dra + v ddec)). */
void or_vis_dft(const double *img, int nx, int ny, double dxy, const double *u,
                const double *v, int K, double dra, double ddec, double *re,
                double *im);

/* Transpose: g_img[j][k] = d/dI of sum_k (gre_k ReV_k + gim_k ImV_k).       */
void or_vis_dft_adjoint(const double *gre, const double *gim, int nx, int ny,
                        double dxy, const double *u, const double *v, int K,
                        double dra, double ddec, double *g_img);

/* chi^2 = sum w ((dre-mre)^2 + (dim-mim)^2) (libinterferometry.pyx:619-633). */
double or_chisq(const double *dre, const double *dim, const double *w,
                const double *mre, const double *mim, int K);

/* Philox4x32-10 (Salmon et al. 2011) one block, for known-answer tests.     */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                      uint32_t out[4]);

/* Tables the MC uses, exposed for closed-form tests.
 * nu[nlam], dnu[nlam] (trapezoid weights), Bsum(T) = sum_i B_i(T) dnu_i.    */
void or_freq_grid(const double *lam_um, int nlam, double *nu, double *dnu);
double or_planck(double nu, double T);
double or_dplanck_dT(double nu, double T);
/* q(T) = 4 pi sum_i kabs_i B_i(T) dnu_i  (erg/s/g) and its inverse.        */
double or_emis_per_gram(const double *kabs, const double *lam_um, int nlam,
                        double T);
/* Modified-random-walk escape law: survival probability of a diffusing packet
 * started at the centre of a sphere, P(y) = 2 sum_n (-1)^(n+1) exp(-n^2 pi^2 y)
 * with y = D t / R0^2, and its inverse. */
/* The thermal Monte Carlo's tables for a model (temperature grid
 * exp(linspace(ln temp0, ln temp1, ntemp))): ln q(T) per absorption group
 * [ngroup][ntemp], Planck-mean absorption [ngroup][ntemp], Rosseland-mean
 * extinction per species [nspec][ntemp], re-emission CDFs
 * [ngroup][ntemp][nlam], group of each species.  or_reemit_bin is the
 * re-emission draw the thermal MC makes (T, two uniforms -> frequency bin);
 * or_reemit_hist histograms 2n such draws from Philox uniforms. */
typedef struct or_tables or_tables;
or_tables *or_tables_create(const or_model *m, int ntemp, double temp0, double temp1);
void or_tables_free(or_tables *t);
void or_tables_shape(const or_tables *t, int *ngroup, int *ntemp, int *nlam);
void or_tables_get(const or_tables *t, double *lnq, double *kplanck, double *kross,
                   double *cdf, int *grp);
int or_reemit_bin(const or_tables *t, int q, double T, double u_row, double u_freq);
void or_reemit_hist(const or_tables *t, int q, double T, uint64_t n, uint64_t seed,
                    uint64_t *counts);
/* Test only: plants a known mistake in the tables / draw (see rt_oracle.c). */
void or_set_mutation(int code);

double or_mrw_survival(double y);
double or_mrw_invert(double xi);

#ifdef __cplusplus
}
#endif
#endif
