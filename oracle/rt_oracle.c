/*
 * rt_oracle.c -- FP64 CPU restatement of the continuum RT path (test
 * infrastructure only; see oracle.h for who may call it).
 *
 * What each part restates:
 *  - Planck function / derivative: pdspy/misc/B_nu.py:4-6, dB_nu.py:4-7, with
 *    pdspy's CGS constants (pdspy/constants/physics.py).
 *  - Grid, densities, walls and cell order: pdspy/modeling/Grid.py:71-80,
 *    Model.py:754-758, radmc3d/write.py:198-266 (r fastest, theta in
 *    [0, pi/2] mirrored at the equator).
 *  - Stars as blackbody point sources (write.py:148-172, -T => blackbody,
 *    istar_sphere unset => point source).
 *  - Thermal MC: what RADMC-3D `mctherm` computes (run.py:9-38): packets of
 *    equal energy L/N, tau = -ln(xi) flights, albedo-weighted scattering,
 *    Bjorkman & Wood (2001) immediate re-emission drawn from the
 *    kappa_nu dB_nu/dT spectrum tabulated on RADMC-3D's (ntemp, temp0, temp1)
 *    temperature grid (write.py:47-52), path-length (Lucy 1999) absorbed-energy
 *    estimator, and per-species temperatures written back as
 *    dust_temperature.dat (read.pyx:433-478, Model.py:249-254).
 *  - Camera: formal solution I = sum S (1-e^-dtau) e^-tau with S = B_nu(T) and
 *    absorption-only opacity (RADMC-3D image with scattering_mode_max=0),
 *    pixel centring of Model.py:522-527.
 *  - Visibilities: DFT equal to imtovis (imaging/imtovis.py:13-24) on the FFT
 *    grid and to an exact sampler at arbitrary (u,v) (interpolate_model.py:20-27).
 *  - chi^2: interferometry/libinterferometry.pyx:619-633.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* pdspy CGS constants (pdspy/constants/physics.py) */
#define C_H 6.6260755e-27
#define C_C 2.99792458e10
#define C_K 1.380658e-16
#define PI_D 3.14159265358979323846
#define TWO_PI_D 6.28318530717958647692

#define TALLY_SHIFT 28 /* fixed-point tally resolution: 2^-28 of kmax*dr */

/* Deliberate mistakes a test plants to show that a pin catches them (0 =
 * none; 1 = re-emission spectrum kappa B instead of kappa dB/dT; 2 =
 * re-emission temperature row shifted up by one; 3 = ping-pong guard off;
 * 4 = a wall crossing found just behind the position is dropped instead of
 * taken at once).
 * Never set outside such a test. */
static int g_mutation = 0;
void or_set_mutation(int code) { g_mutation = code; }

/* ------------------------------------------------------------------ Philox */
#define PH_M0 0xD2511F53u
#define PH_M1 0xCD9E8D57u
#define PH_W0 0x9E3779B9u
#define PH_W1 0xBB67AE85u

void or_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2],
                      uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += PH_W0;
            k1 += PH_W1;
        }
        uint64_t p0 = (uint64_t)PH_M0 * c0;
        uint64_t p1 = (uint64_t)PH_M1 * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

/* Random event k of packet pid (k = 0 at emission, then one per interaction
 * and one per modified-random-walk step) takes the 8 uniforms of the Philox
 * blocks with counter (k, pid_lo, pid_hi, 0) and (k, pid_lo, pid_hi, 1) under
 * key = seed; each uniform is ((x>>8)+0.5)*2^-24, exact in FP32. */
static void rng8(uint64_t seed, uint64_t pid, uint32_t k, double u[8]) {
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (uint32_t b = 0; b < 2; ++b) {
        uint32_t ctr[4] = {k, (uint32_t)pid, (uint32_t)(pid >> 32), b}, out[4];
        or_philox4x32_10(ctr, key, out);
        for (int i = 0; i < 4; ++i) u[4 * b + i] = ((double)(out[i] >> 8) + 0.5) * (1.0 / 16777216.0);
    }
}

/* --------------------------------------------------------------- Planck */
double or_planck(double nu, double T) {
    if (T <= 0) return 0.0;
    double x = C_H * nu / (C_K * T);
    if (x > 700.0) return 2.0 * C_H * nu * nu * nu / (C_C * C_C) * exp(-x);
    return 2.0 * C_H * nu * nu * nu / (C_C * C_C) / expm1(x);
}

double or_dplanck_dT(double nu, double T) {
    if (T <= 0) return 0.0;
    double x = C_H * nu / (C_K * T);
    double pre = 2.0 * C_H * C_H * nu * nu * nu * nu / (C_C * C_C * C_K * T * T);
    if (x > 700.0) return pre * exp(-x);
    double s = sinh(0.5 * x);
    return pre / (4.0 * s * s);
}

/* ln of the Planck function, safe for any x > 0 */
static double ln_planck(double nu, double T) {
    double x = C_H * nu / (C_K * T);
    double lnexpm1 = (x > 30.0) ? x + log1p(-exp(-x)) : log(expm1(x));
    return log(2.0 * C_H / (C_C * C_C)) + 3.0 * log(nu) - lnexpm1;
}

static double ln_dplanck(double nu, double T) {
    double x = C_H * nu / (C_K * T);
    double ln_pre = log(2.0 * C_H * C_H / (C_C * C_C * C_K)) + 4.0 * log(nu) -
                    2.0 * log(T);
    /* 1/(4 sinh^2(x/2)) = e^-x / (1-e^-x)^2 */
    double ln_f = -x - 2.0 * log(-expm1(-x));
    return ln_pre + ln_f;
}

void or_freq_grid(const double *lam_um, int nlam, double *nu, double *dnu) {
    for (int i = 0; i < nlam; ++i) nu[i] = C_C / (lam_um[i] * 1e-4);
    if (nlam == 1) {
        dnu[0] = 1.0;
        return;
    }
    for (int i = 0; i < nlam; ++i) {
        if (i == 0)
            dnu[i] = 0.5 * fabs(nu[1] - nu[0]);
        else if (i == nlam - 1)
            dnu[i] = 0.5 * fabs(nu[nlam - 1] - nu[nlam - 2]);
        else
            dnu[i] = 0.5 * fabs(nu[i + 1] - nu[i - 1]);
    }
}

double or_emis_per_gram(const double *kabs, const double *lam_um, int nlam,
                        double T) {
    double *nu = malloc(sizeof(double) * nlam), *dnu = malloc(sizeof(double) * nlam);
    or_freq_grid(lam_um, nlam, nu, dnu);
    double s = 0;
    for (int i = 0; i < nlam; ++i) s += kabs[i] * or_planck(nu[i], T) * dnu[i];
    free(nu);
    free(dnu);
    return 4.0 * PI_D * s;
}

/* ------------------------------------------------------------- geometry */
typedef struct {
    int nr, nt, np, mirror, ncell;
    const double *rw, *tw, *pw;
    double *ct, *st; /* cos/sin of theta walls */
    double *cp, *sp; /* cos/sin of phi walls */
    double *vol;     /* cell volumes (mirror doubles them) */
    double *dr;      /* radial cell width per ir */
} geom_t;

#define INF_D 1e300

enum { W_NONE = 0, W_ROUT, W_RIN, W_TLO, W_THI, W_PLO, W_PHI, W_HOLE };

typedef struct {
    double x, y, z, dx, dy, dz;
    int ir, it, ip; /* ir = -1 hole, ir = nr outside */
    int hemi;       /* +1 / -1 (mirror mode: which copy of the half-space) */
} pk_t;

static int geom_init(geom_t *g, const or_model *m) {
    memset(g, 0, sizeof(*g));
    g->nr = m->nr;
    g->nt = m->nt;
    g->np = m->np;
    g->rw = m->rw;
    g->tw = m->tw;
    g->pw = m->pw;
    if (m->nr < 1 || m->nt < 1 || m->np < 1) return -1;
    for (int i = 0; i < m->nr; ++i)
        if (!(m->rw[i + 1] > m->rw[i]) || !(m->rw[i] > 0)) return -2;
    for (int i = 0; i < m->nt; ++i)
        if (!(m->tw[i + 1] > m->tw[i])) return -2;
    for (int i = 0; i < m->np; ++i)
        if (!(m->pw[i + 1] > m->pw[i])) return -2;
    if (fabs(m->tw[0]) > 1e-12) return -3;
    if (fabs(m->tw[m->nt] - 0.5 * PI_D) < 1e-9)
        g->mirror = 1;
    else if (fabs(m->tw[m->nt] - PI_D) < 1e-9)
        g->mirror = 0;
    else
        return -3;
    if (fabs(m->pw[0]) > 1e-12 || fabs(m->pw[m->np] - TWO_PI_D) > 1e-9) return -4;
    g->ncell = m->nr * m->nt * m->np;
    g->ct = malloc(sizeof(double) * (m->nt + 1));
    g->st = malloc(sizeof(double) * (m->nt + 1));
    g->cp = malloc(sizeof(double) * (m->np + 1));
    g->sp = malloc(sizeof(double) * (m->np + 1));
    for (int j = 0; j <= m->nt; ++j) {
        g->ct[j] = cos(m->tw[j]);
        g->st[j] = sin(m->tw[j]);
    }
    if (g->mirror) {
        g->ct[m->nt] = 0.0;
        g->st[m->nt] = 1.0;
    }
    for (int k = 0; k <= m->np; ++k) {
        g->cp[k] = cos(m->pw[k]);
        g->sp[k] = sin(m->pw[k]);
    }
    g->vol = malloc(sizeof(double) * g->ncell);
    g->dr = malloc(sizeof(double) * m->nr);
    for (int i = 0; i < m->nr; ++i) g->dr[i] = m->rw[i + 1] - m->rw[i];
    for (int k = 0; k < m->np; ++k)
        for (int j = 0; j < m->nt; ++j)
            for (int i = 0; i < m->nr; ++i) {
                double r1 = m->rw[i], r2 = m->rw[i + 1];
                double dv_r = (r2 - r1) * (r2 * r2 + r1 * r2 + r1 * r1) / 3.0;
                double dv_t = g->ct[j] - g->ct[j + 1];
                double dv_p = m->pw[k + 1] - m->pw[k];
                g->vol[i + m->nr * (j + m->nt * k)] =
                    dv_r * dv_t * dv_p * (g->mirror ? 2.0 : 1.0);
            }
    return 0;
}

static void geom_free(geom_t *g) {
    free(g->ct);
    free(g->st);
    free(g->cp);
    free(g->sp);
    free(g->vol);
    free(g->dr);
}

static int bsearch_walls(const double *w, int n, double v) {
    /* index i in [0, n-1] with w[i] <= v < w[i+1], clamped */
    if (v <= w[0]) return 0;
    if (v >= w[n]) return n - 1;
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        int mid = (lo + hi) / 2;
        if (w[mid] <= v)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

static void locate_angles(const geom_t *g, pk_t *p) {
    double rho = sqrt(p->x * p->x + p->y * p->y);
    if (g->mirror) {
        p->hemi = (p->z >= 0) ? 1 : -1;
        double th = atan2(rho, fabs(p->z));
        p->it = bsearch_walls(g->tw, g->nt, th);
    } else {
        p->hemi = 1;
        double th = atan2(rho, p->z);
        p->it = bsearch_walls(g->tw, g->nt, th);
    }
    if (g->np > 1) {
        double ph = atan2(p->y, p->x);
        if (ph < 0) ph += TWO_PI_D;
        p->ip = bsearch_walls(g->pw, g->np, ph);
    } else {
        p->ip = 0;
    }
}

static void locate(const geom_t *g, pk_t *p) {
    double r = sqrt(p->x * p->x + p->y * p->y + p->z * p->z);
    if (r < g->rw[0]) {
        p->ir = -1;
        p->hemi = (g->mirror && p->z < 0) ? -1 : 1;
        return;
    }
    if (r >= g->rw[g->nr]) {
        p->ir = g->nr;
        return;
    }
    p->ir = bsearch_walls(g->rw, g->nr, r);
    locate_angles(g, p);
}

static double sph_out(double b, double rr, double R) {
    double c = rr - R * R;
    double D = b * b - c;
    if (D < 0) D = 0;
    double sq = sqrt(D);
    double t = (b <= 0) ? (-b + sq) : (-c / (b + sq));
    return t > 0 ? t : 0.0;
}

static double sph_in(double b, double rr, double R) {
    if (b >= 0) return INF_D;
    double c = rr - R * R;
    double D = b * b - c;
    if (D < 0) return INF_D;
    double t = c / (-b + sqrt(D));
    return t > 0 ? t : 0.0;
}

/* Exit through the cone at polar angle th (cos cth, sin sth) in the folded
 * frame.  req = -1: leaving towards smaller theta, +1: towards larger theta.
 * The cone function f = rho^2 cos^2 - z^2 sin^2 is quadratic in t; the root
 * with f' > 0 is t+ = (-B+sqrt(D))/A and with f' < 0 is t-; on the nappe
 * sign(z) = sign(cos th) the sign of d(theta)/dt is sign(f') * sign(cos th). */
static double cone_hit(double x, double y, double z, double dx, double dy,
                       double dz, double cth, double sth, int req) {
    double c2 = cth * cth, s2 = sth * sth;
    double A = (dx * dx + dy * dy) * c2 - dz * dz * s2;
    double B = (x * dx + y * dy) * c2 - z * dz * s2;
    double C = (x * x + y * y) * c2 - z * z * s2;
    /* a ray lying in the cone (direction and position on it, as a radial ray
     * through the origin along a wall angle): it never leaves through it */
    if (fabs(A) <= 1e-6 && fabs(C) <= 1e-6 * (x * x + y * y + z * z)) return INF_D;
    double D = B * B - A * C;
    /* sreq > 0: leaving through this wall makes f grow */
    double sreq = cth > 0 ? req : -req;
    double t;
    if (D < 0) {
        /* no crossing: f keeps the sign of C along the line; the far side's
         * sign means the position is already past the wall */
        if (g_mutation == 4 || !(C * sreq > 0)) return INF_D;
        t = 0;
    } else {
        double sq = sqrt(D);
        if (sreq > 0) {
            t = (B < 0) ? (-B + sq) / A : -C / (B + sq);
        } else {
            t = (B > 0) ? -(B + sq) / A : C / (-B + sq);
        }
    }
    /* exit root behind the position while the ray moves through the wall:
     * the position is a hair past the wall with the index one cell behind it
     * (two crossings closer than the arithmetic resolves): cross at once */
    if (!(t >= 0)) {
        if (g_mutation == 4 || !(B * sreq > 0)) return INF_D;
        t = 0;
    }
    if (!(t < INF_D)) return INF_D;
    double zh = z + t * dz;
    if (zh * cth < 0) return INF_D;
    return t;
}

static double phi_hit(double x, double y, double dx, double dy, double cph,
                      double sph, int req) {
    double g0 = -x * sph + y * cph;
    double g1 = -dx * sph + dy * cph;
    if (req < 0 ? !(g1 < 0) : !(g1 > 0)) return INF_D;
    double t = -g0 / g1;
    /* already past the wall, moving through it: cross at once */
    if (!(t >= 0)) {
        if (g_mutation == 4) return INF_D;
        t = 0;
    }
    double xh = x + t * dx, yh = y + t * dy;
    if (xh * cph + yh * sph < 0) return INF_D;
    return t;
}

/* distance to the next wall of the current cell and the wall's kind */
static double next_wall(const geom_t *g, const pk_t *p, int *kind) {
    double rr = p->x * p->x + p->y * p->y + p->z * p->z;
    double b = p->x * p->dx + p->y * p->dy + p->z * p->dz;
    if (p->ir < 0) {
        *kind = W_HOLE;
        return sph_out(b, rr, g->rw[0]);
    }
    double tbest = sph_out(b, rr, g->rw[p->ir + 1]);
    *kind = W_ROUT;
    double t = sph_in(b, rr, g->rw[p->ir]);
    if (t < tbest) {
        tbest = t;
        *kind = W_RIN;
    }
    double zf = p->hemi * p->z, dzf = p->hemi * p->dz;
    int j = p->it;
    if (j > 0) { /* lower-theta wall (not the axis) */
        if (!g->mirror && fabs(g->ct[j]) < 1e-15) {
            /* equatorial plane seen from the lower hemisphere: g = -z, need g'<0 */
            if (p->dz > 0) {
                t = -p->z / p->dz;
                if (t < 0) t = g_mutation == 4 ? INF_D : 0;
                if (t < tbest) {
                    tbest = t;
                    *kind = W_TLO;
                }
            }
        } else {
            t = cone_hit(p->x, p->y, zf, p->dx, p->dy, dzf, g->ct[j], g->st[j], -1);
            if (t < tbest) {
                tbest = t;
                *kind = W_TLO;
            }
        }
    }
    if (j + 1 < g->nt || g->mirror) {
        if (g->mirror && j + 1 == g->nt) {
            if (dzf < 0) {
                t = -zf / dzf;
                if (t < 0) t = 0;
                if (t < tbest) {
                    tbest = t;
                    *kind = W_THI;
                }
            }
        } else if (!g->mirror && fabs(g->ct[j + 1]) < 1e-15) {
            if (p->dz < 0) {
                t = -p->z / p->dz;
                if (t < 0) t = g_mutation == 4 ? INF_D : 0;
                if (t < tbest) {
                    tbest = t;
                    *kind = W_THI;
                }
            }
        } else {
            t = cone_hit(p->x, p->y, zf, p->dx, p->dy, dzf, g->ct[j + 1],
                         g->st[j + 1], +1);
            if (t < tbest) {
                tbest = t;
                *kind = W_THI;
            }
        }
    }
    if (g->np > 1) {
        int k = p->ip;
        t = phi_hit(p->x, p->y, p->dx, p->dy, g->cp[k], g->sp[k], -1);
        if (t < tbest) {
            tbest = t;
            *kind = W_PLO;
        }
        t = phi_hit(p->x, p->y, p->dx, p->dy, g->cp[k + 1], g->sp[k + 1], +1);
        if (t < tbest) {
            tbest = t;
            *kind = W_PHI;
        }
    }
    return tbest;
}

static void snap_radius(pk_t *p, double R) {
    double r = sqrt(p->x * p->x + p->y * p->y + p->z * p->z);
    if (r > 0) {
        double f = R / r;
        p->x *= f;
        p->y *= f;
        p->z *= f;
    }
}

/* move by t and cross the wall of kind k; returns 1 if the packet left the grid */
static int cross_wall(const geom_t *g, pk_t *p, double t, int kind) {
    p->x += t * p->dx;
    p->y += t * p->dy;
    p->z += t * p->dz;
    switch (kind) {
    case W_HOLE:
        snap_radius(p, g->rw[0]);
        p->ir = 0;
        locate_angles(g, p);
        return 0;
    case W_ROUT:
        p->ir++;
        if (p->ir >= g->nr) return 1;
        snap_radius(p, g->rw[p->ir]);
        return 0;
    case W_RIN:
        snap_radius(p, g->rw[p->ir]);
        p->ir--;
        return 0;
    case W_TLO:
        p->it--;
        return 0;
    case W_THI:
        if (g->mirror && p->it + 1 == g->nt) {
            p->z = 0.0;
            p->hemi = -p->hemi;
        } else {
            p->it++;
        }
        return 0;
    case W_PLO:
        p->ip = (p->ip - 1 + g->np) % g->np;
        return 0;
    case W_PHI:
        p->ip = (p->ip + 1) % g->np;
        return 0;
    }
    return 0;
}

static inline int cell_of(const geom_t *g, const pk_t *p) {
    return p->ir + g->nr * (p->it + g->nt * p->ip);
}

/* Ping-pong guard (the FP64 counterpart of the device's): after 8 consecutive
 * crossings shorter than 2^-44 r (r: outer radius of the cell the step began
 * in) the packet moves 2^-36 r along its ray and is located again.  Returns 1
 * when it moved the packet. */
static int unstick(const geom_t *g, pk_t *p, double tw, double rcell, int *nzero) {
    if (g_mutation == 3) return 0;
    if (tw > rcell * 0x1p-44) {
        *nzero = 0;
        return 0;
    }
    if (++*nzero < 8) return 0;
    *nzero = 0;
    double d = rcell * 0x1p-36;
    p->x += d * p->dx;
    p->y += d * p->dy;
    p->z += d * p->dz;
    locate(g, p);
    return 1;
}

/* walk from p (cell indices as given) until it leaves the grid, recording
 * up to maxseg segments; returns the number of segments */
static int walk_segments(const geom_t *g, pk_t *p, int maxseg, int *cell, double *len,
                         int *hemi, int *nudged) {
    int n = 0, nzero = 0;
    while (p->ir >= -1 && p->ir < g->nr) {
        int kind;
        double t = next_wall(g, p, &kind);
        double rcell = g->rw[p->ir + 1];
        if (n < maxseg) {
            cell[n] = (p->ir < 0) ? -1 : cell_of(g, p);
            len[n] = t;
            if (hemi) hemi[n] = p->hemi;
        }
        n++;
        if (cross_wall(g, p, t, kind)) break;
        if (unstick(g, p, t, rcell, &nzero) && nudged) ++*nudged;
        if (n > 10000000) break;
    }
    return n;
}

int or_trace_ray_n(const or_model *m, const double *o, const double *d, int maxseg,
                   int *cell, double *len, int *hemi, int *nudged) {
    geom_t g;
    int rc = geom_init(&g, m);
    if (rc) return rc;
    pk_t p;
    memset(&p, 0, sizeof(p));
    p.x = o[0];
    p.y = o[1];
    p.z = o[2];
    p.dx = d[0];
    p.dy = d[1];
    p.dz = d[2];
    double rr = o[0] * o[0] + o[1] * o[1] + o[2] * o[2];
    double R = g.rw[g.nr];
    int n = 0;
    if (rr >= R * R) {
        double b = o[0] * d[0] + o[1] * d[1] + o[2] * d[2];
        double D = b * b - (rr - R * R);
        if (D <= 0 || b >= 0) {
            geom_free(&g);
            return 0;
        }
        double t = -b - sqrt(D);
        p.x += t * p.dx;
        p.y += t * p.dy;
        p.z += t * p.dz;
        snap_radius(&p, R);
        p.ir = g.nr - 1;
        locate_angles(&g, &p);
    } else {
        locate(&g, &p);
    }
    n = walk_segments(&g, &p, maxseg, cell, len, hemi, nudged);
    geom_free(&g);
    return n;
}

int or_trace_ray_from(const or_model *m, const double *o, const double *d, const int *idx,
                      int maxseg, int *cell, double *len, int *hemi) {
    geom_t g;
    int rc = geom_init(&g, m);
    if (rc) return rc;
    if (idx[0] < -1 || idx[0] >= g.nr || idx[1] < 0 || idx[1] >= g.nt || idx[2] < 0 ||
        idx[2] >= g.np || (idx[3] != 1 && idx[3] != -1) || (!g.mirror && idx[3] != 1)) {
        geom_free(&g);
        return -1;
    }
    pk_t p;
    memset(&p, 0, sizeof(p));
    p.x = o[0];
    p.y = o[1];
    p.z = o[2];
    p.dx = d[0];
    p.dy = d[1];
    p.dz = d[2];
    p.ir = idx[0];
    p.it = idx[1];
    p.ip = idx[2];
    p.hemi = idx[3];
    int n = walk_segments(&g, &p, maxseg, cell, len, hemi, NULL);
    geom_free(&g);
    return n;
}

int or_trace_ray(const or_model *m, const double *o, const double *d, int maxseg,
                 int *cell, double *len, int *hemi) {
    return or_trace_ray_n(m, o, d, maxseg, cell, len, hemi, NULL);
}

/* ------------------------------------------------------------- tables */
typedef struct {
    int nlam, ntemp, ngroup, nspec, nstar;
    double temp0, temp1, lnt0, dlnt;
    double *nu, *dnu;
    int *grp;        /* species -> group */
    int *grep;       /* group -> representative species */
    double *kmax;    /* [ngroup] */
    double *lnq;     /* [ngroup][ntemp] */
    double *cdf;     /* [ngroup][ntemp][nlam] */
    double *scdf;    /* [nstar][nlam] */
    double *spick;   /* [nstar] cumulative luminosity fractions */
    double *kross;   /* [nspec][ntemp] Rosseland-mean extinction */
    double *kplanck; /* [ngroup][ntemp] Planck-mean absorption */
    double lum;
} tab_t;

static void tab_free(tab_t *t) {
    free(t->nu);
    free(t->dnu);
    free(t->grp);
    free(t->grep);
    free(t->kmax);
    free(t->lnq);
    free(t->cdf);
    free(t->scdf);
    free(t->spick);
    free(t->kross);
    free(t->kplanck);
}

static double logsumexp(const double *v, int n) {
    double mx = -INFINITY;
    for (int i = 0; i < n; ++i)
        if (v[i] > mx) mx = v[i];
    if (mx == -INFINITY) return -INFINITY;
    double s = 0;
    for (int i = 0; i < n; ++i) s += exp(v[i] - mx);
    return mx + log(s);
}

static void cdf_from_log(const double *lw, int n, double *cdf) {
    double mx = -INFINITY;
    for (int i = 0; i < n; ++i)
        if (lw[i] > mx) mx = lw[i];
    if (mx == -INFINITY) { /* nothing emits: put everything in the last bin */
        for (int i = 0; i < n; ++i) cdf[i] = (i == n - 1) ? 1.0 : 0.0;
        return;
    }
    double s = 0;
    for (int i = 0; i < n; ++i) {
        s += exp(lw[i] - mx);
        cdf[i] = s;
    }
    for (int i = 0; i < n; ++i) cdf[i] /= s;
    cdf[n - 1] = 1.0;
}

static int tab_init(tab_t *t, const or_model *m, const or_thermal_opts *o) {
    memset(t, 0, sizeof(*t));
    int nl = m->nlam, ns = m->nspec;
    t->nlam = nl;
    t->nspec = ns;
    t->nstar = m->nstar;
    t->ntemp = o->ntemp;
    t->temp0 = o->temp0;
    t->temp1 = o->temp1;
    if (t->ntemp < 2 || !(t->temp1 > t->temp0) || !(t->temp0 > 0)) return -10;
    t->lnt0 = log(t->temp0);
    t->dlnt = (log(t->temp1) - t->lnt0) / (t->ntemp - 1);
    t->nu = malloc(sizeof(double) * nl);
    t->dnu = malloc(sizeof(double) * nl);
    or_freq_grid(m->lam_um, nl, t->nu, t->dnu);
    /* groups: species with identical absorption opacity share temperatures */
    t->grp = malloc(sizeof(int) * ns);
    t->grep = malloc(sizeof(int) * ns);
    int ng = 0;
    for (int s = 0; s < ns; ++s) {
        int found = -1;
        for (int q = 0; q < ng && found < 0; ++q) {
            const double *a = m->kabs + (size_t)t->grep[q] * nl, *b = m->kabs + (size_t)s * nl;
            if (memcmp(a, b, sizeof(double) * nl) == 0) found = q;
        }
        if (found < 0) {
            t->grep[ng] = s;
            found = ng++;
        }
        t->grp[s] = found;
    }
    t->ngroup = ng;
    t->kmax = malloc(sizeof(double) * ng);
    t->lnq = malloc(sizeof(double) * ng * t->ntemp);
    t->cdf = malloc(sizeof(double) * (size_t)ng * t->ntemp * nl);
    double *lw = malloc(sizeof(double) * nl);
    for (int q = 0; q < ng; ++q) {
        const double *ka = m->kabs + (size_t)t->grep[q] * nl;
        double km = 0;
        for (int i = 0; i < nl; ++i)
            if (ka[i] > km) km = ka[i];
        if (!(km > 0)) km = 1.0;
        t->kmax[q] = km;
        for (int it = 0; it < t->ntemp; ++it) {
            double T = exp(t->lnt0 + it * t->dlnt);
            for (int i = 0; i < nl; ++i)
                lw[i] = (ka[i] > 0) ? log(ka[i] * t->dnu[i]) + ln_planck(t->nu[i], T)
                                    : -INFINITY;
            t->lnq[q * t->ntemp + it] = log(4.0 * PI_D) + logsumexp(lw, nl);
            for (int i = 0; i < nl; ++i)
                lw[i] = (ka[i] > 0) ? log(ka[i] * t->dnu[i]) +
                                          (g_mutation == 1 ? ln_planck(t->nu[i], T)
                                                           : ln_dplanck(t->nu[i], T))
                                    : -INFINITY;
            cdf_from_log(lw, nl, t->cdf + ((size_t)q * t->ntemp + it) * nl);
        }
    }
    /* Planck-mean absorption per group and Rosseland-mean extinction per
     * species on the same temperature grid (used by the modified random walk) */
    t->kross = malloc(sizeof(double) * ns * t->ntemp);
    t->kplanck = malloc(sizeof(double) * ng * t->ntemp);
    for (int it = 0; it < t->ntemp; ++it) {
        double T = exp(t->lnt0 + it * t->dlnt);
        double *lb = malloc(sizeof(double) * nl), *ld = malloc(sizeof(double) * nl);
        double mb = -INFINITY, md = -INFINITY;
        for (int i = 0; i < nl; ++i) {
            lb[i] = log(t->dnu[i]) + ln_planck(t->nu[i], T);
            ld[i] = log(t->dnu[i]) + ln_dplanck(t->nu[i], T);
            if (lb[i] > mb) mb = lb[i];
            if (ld[i] > md) md = ld[i];
        }
        for (int q = 0; q < ng; ++q) {
            const double *ka = m->kabs + (size_t)t->grep[q] * nl;
            double wb = 0, wkb = 0;
            for (int i = 0; i < nl; ++i) {
                double b = exp(lb[i] - mb);
                wb += b;
                wkb += b * ka[i];
            }
            t->kplanck[q * t->ntemp + it] = wkb / wb;
        }
        for (int s = 0; s < ns; ++s) {
            double wd = 0, wdk = 0;
            for (int i = 0; i < nl; ++i) {
                double d = exp(ld[i] - md);
                double ke = m->kabs[(size_t)s * nl + i] + m->ksca[(size_t)s * nl + i];
                wd += d;
                if (ke > 0) wdk += d / ke;
            }
            t->kross[s * t->ntemp + it] = wdk > 0 ? wd / wdk : 0.0;
        }
        free(lb);
        free(ld);
    }
    /* stars: blackbody point sources, L = 4 pi R^2 * pi * sum B dnu */
    t->scdf = malloc(sizeof(double) * (m->nstar > 0 ? m->nstar : 1) * nl);
    t->spick = malloc(sizeof(double) * (m->nstar > 0 ? m->nstar : 1));
    double ltot = 0;
    for (int k = 0; k < m->nstar; ++k) {
        double R = m->star[5 * k], Ts = m->star[5 * k + 1];
        double s = 0;
        for (int i = 0; i < nl; ++i) {
            double bd = or_planck(t->nu[i], Ts) * t->dnu[i];
            s += bd;
            lw[i] = (bd > 0) ? log(bd) : -INFINITY;
        }
        cdf_from_log(lw, nl, t->scdf + (size_t)k * nl);
        double L = 4.0 * PI_D * R * R * PI_D * s;
        ltot += L;
        t->spick[k] = ltot;
    }
    for (int k = 0; k < m->nstar; ++k) t->spick[k] /= ltot;
    if (m->nstar > 0) t->spick[m->nstar - 1] = 1.0;
    t->lum = ltot;
    free(lw);
    return 0;
}

/* T from absorbed power per gram e (erg/s/g), log-log interpolation of lnq */
static double temp_from_e(const tab_t *t, int q, double e) {
    if (!(e > 0)) return 0.0;
    const double *lq = t->lnq + q * t->ntemp;
    double le = log(e);
    int n = t->ntemp, m;
    if (le <= lq[0])
        m = 0;
    else if (le >= lq[n - 1])
        m = n - 2;
    else {
        int lo = 0, hi = n - 1;
        while (hi - lo > 1) {
            int mid = (lo + hi) / 2;
            if (lq[mid] <= le)
                lo = mid;
            else
                hi = mid;
        }
        m = lo;
    }
    double lt = t->lnt0 + (m + (le - lq[m]) / (lq[m + 1] - lq[m])) * t->dlnt;
    return exp(lt);
}

static int sample_cdf(const double *cdf, int n, double u) {
    int lo = -1, hi = n - 1; /* smallest i with u <= cdf[i] */
    while (hi - lo > 1) {
        int mid = (lo + hi) / 2;
        if (u <= cdf[mid])
            hi = mid;
        else
            lo = mid;
    }
    return hi;
}

/* isotropic direction: cos(theta) = 2 u1 - 1, azimuth 2 pi u2 */
/* Bjorkman & Wood re-emission frequency of absorption group q at temperature
 * T: the table row below or above T's position on the log temperature grid,
 * the upper one with probability equal to the fractional position (u_row),
 * then the inverse CDF of that row's kappa_nu dB_nu/dT dnu spectrum (u_freq) */
static int reemit_bin(const tab_t *t, int q, double T, double u_row, double u_freq) {
    double x = (T > 0) ? (log(T) - t->lnt0) / t->dlnt : 0.0;
    if (x < 0) x = 0;
    if (x > t->ntemp - 1) x = t->ntemp - 1;
    int lo = (int)x;
    if (lo > t->ntemp - 2) lo = t->ntemp - 2;
    double w = x - lo;
    int row = lo + (u_row < w ? 1 : 0);
    if (g_mutation == 2 && row < t->ntemp - 1) ++row;
    return sample_cdf(t->cdf + ((size_t)q * t->ntemp + row) * t->nlam, t->nlam, u_freq);
}

static void iso_dir(double u1, double u2, double *d) {
    double mu = 2.0 * u1 - 1.0;
    double ph = TWO_PI_D * u2;
    double s = sqrt(fmax(0.0, 1.0 - mu * mu));
    d[0] = s * cos(ph);
    d[1] = s * sin(ph);
    d[2] = mu;
}

/* Henyey-Greenstein scattering of d: cosine from u1, azimuth 2 pi u2 */
static void hg_dir(double u1, double u2, double g, double *d) {
    double u = u1;
    double ct;
    if (fabs(g) < 1e-6)
        ct = 2.0 * u - 1.0;
    else {
        double f = (1.0 - g * g) / (1.0 - g + 2.0 * g * u);
        ct = (1.0 + g * g - f * f) / (2.0 * g);
    }
    if (ct > 1) ct = 1;
    if (ct < -1) ct = -1;
    double ps = TWO_PI_D * u2;
    double st = sqrt(fmax(0.0, 1.0 - ct * ct));
    double cps = cos(ps), sps = sin(ps);
    double ux = d[0], uy = d[1], uz = d[2];
    if (fabs(uz) > 0.99999) {
        d[0] = st * cps;
        d[1] = st * sps;
        d[2] = (uz > 0 ? 1.0 : -1.0) * ct;
    } else {
        double s = sqrt(1.0 - uz * uz);
        d[0] = st * (ux * uz * cps - uy * sps) / s + ux * ct;
        d[1] = st * (uy * uz * cps + ux * sps) / s + uy * ct;
        d[2] = -st * cps * s + uz * ct;
    }
    double n = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    d[0] /= n;
    d[1] /= n;
    d[2] /= n;
}

/* ------------------------------------------------- modified random walk */
/* Survival probability of a packet diffusing from the centre of a sphere of
 * radius R0, y = D t / R0^2 (Fleck & Canfield 1984; Min et al. 2009):
 * P(y) = 2 sum_{n>=1} (-1)^(n+1) exp(-n^2 pi^2 y); for small y the Jacobi
 * dual 1 - (pi y)^-1/2 sum_m exp(-(m+1/2)^2 / y) converges faster. */
double or_mrw_survival(double y) {
    if (y <= 0) return 1.0;
    if (y < 0.1) {
        double s = 0;
        for (int m = -40; m <= 40; ++m) s += exp(-(m + 0.5) * (m + 0.5) / y);
        return 1.0 - s / sqrt(PI_D * y);
    }
    double s = 0;
    for (int n = 1; n <= 60; ++n) s += ((n & 1) ? 2.0 : -2.0) * exp(-(double)n * n * PI_D * PI_D * y);
    return s;
}

double or_mrw_invert(double xi) {
    double lo = log(1e-9), hi = log(60.0);
    if (xi >= or_mrw_survival(1e-9)) return 1e-9;
    if (xi <= or_mrw_survival(60.0)) return 60.0;
    for (int k = 0; k < 80; ++k) {
        double mid = 0.5 * (lo + hi);
        if (or_mrw_survival(exp(mid)) > xi)
            lo = mid;
        else
            hi = mid;
    }
    return exp(0.5 * (lo + hi));
}

/* perpendicular distance to the nearest wall of the current cell that bounds
 * the medium (the equatorial mirror and the polar axis are not walls) */
static double wall_dist(const geom_t *g, const pk_t *p) {
    /* cone at theta_j: r sin|theta - theta_j| = |rho cos theta_j - z sin theta_j|
     * while |theta - theta_j| <= pi/2 (else the axis, r); phi half-plane at
     * phi_k: |y cos phi_k - x sin phi_k| while |phi - phi_k| <= pi/2 (else rho) */
    double r = sqrt(p->x * p->x + p->y * p->y + p->z * p->z);
    double d = fmin(r - g->rw[p->ir], g->rw[p->ir + 1] - r);
    double rho = sqrt(p->x * p->x + p->y * p->y);
    double zf = g->mirror ? fabs(p->z) : p->z;
    int j = p->it;
    for (int w = j; w <= j + 1; ++w) {
        if (w == 0 || w == g->nt) continue;
        double sn = rho * g->ct[w] - zf * g->st[w];
        double cs = zf * g->ct[w] + rho * g->st[w];
        d = fmin(d, cs >= 0 ? fabs(sn) : r);
    }
    if (g->np > 1) {
        for (int k = p->ip; k <= p->ip + 1; ++k) {
            double sn = p->y * g->cp[k] - p->x * g->sp[k];
            double cs = p->x * g->cp[k] + p->y * g->sp[k];
            d = fmin(d, cs >= 0 ? fabs(sn) : rho);
        }
    }
    return d;
}

/* table value at temperature row coordinate x (linear between rows) */
static double trow(const double *tab, int ntemp, double T, double lnt0, double dlnt) {
    double x = (T > 0) ? (log(T) - lnt0) / dlnt : 0.0;
    if (x < 0) x = 0;
    if (x > ntemp - 1) x = ntemp - 1;
    int lo = (int)x;
    if (lo > ntemp - 2) lo = ntemp - 2;
    double w = x - lo;
    return (1 - w) * tab[lo] + w * tab[lo + 1];
}

/* ------------------------------------------------------------ thermal MC */
typedef struct {
    const or_model *m;
    const geom_t *g;
    const tab_t *t;
    const or_thermal_opts *o;
    uint64_t *tally;   /* [ncell][ngroup] fixed point: folded (cumulative) tallies */
    uint64_t *slots;   /* lagged mode: [nslot][ncell][ngroup] tallies by deposit slot */
    int nslot;         /* number of epochs + 1 (the last slot takes all later deposits) */
    double *W;         /* [ncell][ngroup] units per (kabs*s) */
    double *Tcur;      /* [ncell][ngroup] lagged temperatures */
    double *econv;     /* [ncell][ngroup] erg/s/g per tally unit */
    uint64_t *visits;
    int overflow;
} mc_t;

typedef struct {
    uint64_t n_cross, n_abs, n_scat, n_escape, n_killed, n_events, n_mrw, n_nudged;
} cnt_t;

static void deposit(mc_t *mc, uint64_t *tl, int c, int ifr, double s) {
    const tab_t *t = mc->t;
    int ng = t->ngroup;
    for (int q = 0; q < ng; ++q) {
        double ka = mc->m->kabs[(size_t)t->grep[q] * t->nlam + ifr];
        double v = ka * s * mc->W[(size_t)c * ng + q];
        uint64_t add = (uint64_t)llround(v);
        if (add) {
            uint64_t old;
#pragma omp atomic capture
            {
                old = tl[(size_t)c * ng + q];
                tl[(size_t)c * ng + q] += add;
            }
            if (old > ((uint64_t)1 << 62)) mc->overflow = 1;
        }
    }
}

/* add an explicit absorbed amount (kappa * path, cm^3/g) to group q of cell c */
static void deposit_q(mc_t *mc, uint64_t *tl, int c, int q, double kpath) {
    int ng = mc->t->ngroup;
    uint64_t add = (uint64_t)llround(kpath * mc->W[(size_t)c * ng + q]);
    if (add) {
        uint64_t old;
#pragma omp atomic capture
        {
            old = tl[(size_t)c * ng + q];
            tl[(size_t)c * ng + q] += add;
        }
        if (old > ((uint64_t)1 << 62)) mc->overflow = 1;
    }
}

/* the temperature re-emission sees: lagged snapshot or the running tally */
static double cur_T(mc_t *mc, int c, int q) {
    int ng = mc->t->ngroup;
    if (mc->o->lagged) return mc->Tcur[(size_t)c * ng + q];
    uint64_t tv;
#pragma omp atomic read
    tv = mc->tally[(size_t)c * ng + q];
    return temp_from_e(mc->t, q, (double)tv * mc->econv[(size_t)c * ng + q]);
}

/* rotate unit vector d by polar angle (cos ct) and azimuth ps about itself */
static void rotate_dir(double *d, double ct, double ps) {
    double st = sqrt(fmax(0.0, 1.0 - ct * ct));
    double cps = cos(ps), sps = sin(ps);
    double ux = d[0], uy = d[1], uz = d[2];
    if (fabs(uz) > 0.99999) {
        d[0] = st * cps;
        d[1] = st * sps;
        d[2] = (uz > 0 ? 1.0 : -1.0) * ct;
    } else {
        double s = sqrt(1.0 - uz * uz);
        d[0] = st * (ux * uz * cps - uy * sps) / s + ux * ct;
        d[1] = st * (uy * uz * cps + ux * sps) / s + uy * ct;
        d[2] = -st * cps * s + uz * ct;
    }
    double n = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    d[0] /= n;
    d[1] /= n;
    d[2] /= n;
}

/* one modified-random-walk step (Min et al. 2009; Robitaille 2010) for a
 * packet at an interaction point in cell c; returns 1 if it was taken */
static int mrw_step(mc_t *mc, uint64_t *tl, pk_t *p, int c, uint64_t pid, uint32_t *krng,
                    int *ifr) {
    const or_model *m = mc->m;
    const geom_t *g = mc->g;
    const tab_t *t = mc->t;
    const int ns = m->nspec, ng = t->ngroup;
    double aR = 0;
    for (int s = 0; s < ns; ++s) {
        double rho = m->rho[(size_t)s * g->ncell + c];
        if (rho > 0)
            aR += rho * trow(t->kross + (size_t)s * t->ntemp, t->ntemp, cur_T(mc, c, t->grp[s]),
                             t->lnt0, t->dlnt);
    }
    double dmin = wall_dist(g, p);
    if (!(aR * dmin >= mc->o->mrw_tauthres)) return 0;
    double R0 = dmin - mc->o->mrw_gamma / aR;
    if (!(R0 > 0)) return 0;
    /* one random event: u0 path length, u1 u2 exit point, u3 u4 exit direction,
     * u5 species, u6 temperature row, u7 frequency */
    double v[8];
    rng8(mc->o->seed, pid, (*krng)++, v);
    /* diffusion path length to the sphere: y = D t / R0^2, D = 1/(3 alpha_R) */
    double y = or_mrw_invert(v[0]);
    double ell = 3.0 * aR * R0 * R0 * y;
    double kP[8];
    for (int q = 0; q < ng; ++q) {
        kP[q] = trow(t->kplanck + (size_t)q * t->ntemp, t->ntemp, cur_T(mc, c, q), t->lnt0, t->dlnt);
        deposit_q(mc, tl, c, q, kP[q] * ell);
    }
    /* exit point uniform on the sphere, direction cosine-weighted outwards */
    double nrm[3];
    iso_dir(v[1], v[2], nrm);
    p->x += R0 * nrm[0];
    p->y += R0 * nrm[1];
    p->z += R0 * nrm[2];
    if (g->mirror) p->hemi = (p->z >= 0) ? 1 : -1;
    double ct = sqrt(v[3]);
    double ps = TWO_PI_D * v[4];
    rotate_dir(nrm, ct, ps);
    p->dx = nrm[0];
    p->dy = nrm[1];
    p->dz = nrm[2];
    /* new frequency: re-emission spectrum of a species picked by rho * kappa_P */
    double wsum = 0;
    for (int s = 0; s < ns; ++s) wsum += m->rho[(size_t)s * g->ncell + c] * kP[t->grp[s]];
    double us = v[5] * wsum, acc = 0;
    int sp = ns - 1;
    for (int s = 0; s < ns; ++s) {
        acc += m->rho[(size_t)s * g->ncell + c] * kP[t->grp[s]];
        if (us <= acc) {
            sp = s;
            break;
        }
    }
    int q = t->grp[sp];
    *ifr = reemit_bin(t, q, cur_T(mc, c, q), v[6], v[7]);
    return 1;
}

static void run_packet(mc_t *mc, uint64_t pid, int epoch, cnt_t *cnt) {
    const or_model *m = mc->m;
    const geom_t *g = mc->g;
    const tab_t *t = mc->t;
    const int nl = t->nlam, ns = m->nspec, ng = t->ngroup;
    /* emission (random event 0): u0 star, u1 frequency, u2 u3 direction, u4 tau */
    double u[8];
    rng8(mc->o->seed, pid, 0, u);
    uint32_t krng = 1;
    int ks = 0;
    while (ks < m->nstar - 1 && u[0] > t->spick[ks]) ks++;
    int ifr = sample_cdf(t->scdf + (size_t)ks * nl, nl, u[1]);
    double d[3];
    iso_dir(u[2], u[3], d);
    pk_t p;
    memset(&p, 0, sizeof(p));
    p.x = m->star[5 * ks + 2];
    p.y = m->star[5 * ks + 3];
    p.z = m->star[5 * ks + 4];
    p.dx = d[0];
    p.dy = d[1];
    p.dz = d[2];
    locate(g, &p);
    if (p.ir >= g->nr) {
        cnt->n_escape++;
        return;
    }
    if (p.ir >= 0 && mc->visits) {
#pragma omp atomic
        mc->visits[cell_of(g, &p)]++;
    }
    double tau = -log(u[4]);
    uint64_t nev = 0, nev_cell = 0;
    int nzero = 0;
    /* deposit slot: epoch + floor(events / slice) in lagged mode */
    const uint64_t J = mc->o->lagged ? mc->o->slice_events : 0;
    int slot = epoch;
    const size_t ncg = (size_t)g->ncell * ng;
    uint64_t *tl = mc->o->lagged ? mc->slots + (size_t)(slot < mc->nslot - 1 ? slot : mc->nslot - 1) * ncg
                                 : mc->tally;
    for (;;) {
        int kind;
        double tw = next_wall(g, &p, &kind);
        double aabs = 0, asca = 0;
        int c = -1;
        if (p.ir >= 0) {
            c = cell_of(g, &p);
            for (int s = 0; s < ns; ++s) {
                double rho = m->rho[(size_t)s * g->ncell + c];
                aabs += rho * m->kabs[(size_t)s * nl + ifr];
                asca += rho * m->ksca[(size_t)s * nl + ifr];
            }
        }
        double aext = aabs + asca;
        if (aext * tw > tau) {
            double s = tau / aext;
            deposit(mc, tl, c, ifr, s);
            p.x += s * p.dx;
            p.y += s * p.dy;
            p.z += s * p.dz;
            cnt->n_events++;
            nev++;
            nev_cell++;
            nzero = 0;
            if (J && nev % J == 0 && slot < mc->nslot - 1) {
                slot++;
                tl = mc->slots + (size_t)slot * ncg;
            }
            /* interaction (one random event): u0 scatter or absorb, u1 species,
             * u2 temperature row, u3 frequency, u4 u5 direction, u6 next tau */
            rng8(mc->o->seed, pid, krng++, u);
            if (u[0] * aext < asca) {
                cnt->n_scat++;
                int hg = 0;
                double gg = 0;
                if (m->gsca) {
                    double us = u[1] * asca, acc = 0;
                    int sp = ns - 1;
                    for (int s = 0; s < ns; ++s) {
                        acc += m->rho[(size_t)s * g->ncell + c] * m->ksca[(size_t)s * nl + ifr];
                        if (us <= acc) {
                            sp = s;
                            break;
                        }
                    }
                    gg = m->gsca[(size_t)sp * nl + ifr];
                    hg = 1;
                }
                if (hg) {
                    double dd[3] = {p.dx, p.dy, p.dz};
                    hg_dir(u[4], u[5], gg, dd);
                    p.dx = dd[0];
                    p.dy = dd[1];
                    p.dz = dd[2];
                } else {
                    iso_dir(u[4], u[5], d);
                    p.dx = d[0];
                    p.dy = d[1];
                    p.dz = d[2];
                }
            } else {
                cnt->n_abs++;
                double us = u[1] * aabs, acc = 0;
                int sp = ns - 1;
                for (int s = 0; s < ns; ++s) {
                    acc += m->rho[(size_t)s * g->ncell + c] * m->kabs[(size_t)s * nl + ifr];
                    if (us <= acc) {
                        sp = s;
                        break;
                    }
                }
                int q = t->grp[sp];
                double T;
                if (mc->o->lagged) {
                    T = mc->Tcur[(size_t)c * ng + q];
                } else {
                    uint64_t tv;
#pragma omp atomic read
                    tv = mc->tally[(size_t)c * ng + q];
                    T = temp_from_e(t, q, (double)tv * mc->econv[(size_t)c * ng + q]);
                }
                ifr = reemit_bin(t, q, T, u[2], u[3]);
                iso_dir(u[4], u[5], d);
                p.dx = d[0];
                p.dy = d[1];
                p.dz = d[2];
            }
            if (mc->o->max_events && nev >= mc->o->max_events) {
                cnt->n_killed++;
                return;
            }
            tau = -log(u[6]);
            if (mc->o->mrw && nev_cell >= (uint64_t)mc->o->mrw_count &&
                mrw_step(mc, tl, &p, c, pid, &krng, &ifr))
                cnt->n_mrw++;
        } else {
            if (c >= 0) deposit(mc, tl, c, ifr, tw);
            tau -= aext * tw;
            cnt->n_cross++;
            nev_cell = 0;
            double rcell = g->rw[p.ir + 1];
            if (cross_wall(g, &p, tw, kind)) {
                cnt->n_escape++;
                return;
            }
            if (unstick(g, &p, tw, rcell, &nzero)) {
                cnt->n_nudged++;
                if (p.ir >= g->nr) {
                    cnt->n_escape++;
                    return;
                }
            }
            if (p.ir >= 0 && mc->visits) {
#pragma omp atomic
                mc->visits[cell_of(g, &p)]++;
            }
        }
    }
}

static void snapshot(mc_t *mc) {
    int ng = mc->t->ngroup, nc = mc->g->ncell;
#pragma omp parallel for schedule(static)
    for (int c = 0; c < nc; ++c)
        for (int q = 0; q < ng; ++q) {
            size_t k = (size_t)c * ng + q;
            mc->Tcur[k] = temp_from_e(mc->t, q, (double)mc->tally[k] * mc->econv[k]);
        }
}

int or_thermal(const or_model *m, const or_thermal_opts *o, double *T_out,
               uint64_t *visits, or_thermal_stats *st) {
    geom_t g;
    int rc = geom_init(&g, m);
    if (rc) return rc;
    if (m->nstar < 1 || m->nspec < 1 || m->nlam < 1) {
        geom_free(&g);
        return -5;
    }
    tab_t t;
    rc = tab_init(&t, m, o);
    if (rc) {
        geom_free(&g);
        tab_free(&t);
        return rc;
    }
    int ng = t.ngroup, nc = g.ncell;
    mc_t mc;
    memset(&mc, 0, sizeof(mc));
    mc.m = m;
    mc.g = &g;
    mc.t = &t;
    mc.o = o;
    mc.tally = calloc((size_t)nc * ng, sizeof(uint64_t));
    mc.W = malloc(sizeof(double) * nc * ng);
    mc.econv = malloc(sizeof(double) * nc * ng);
    mc.Tcur = calloc((size_t)nc * ng, sizeof(double));
    mc.visits = visits;
    if (visits) memset(visits, 0, sizeof(uint64_t) * nc);
    double eps = t.lum / (double)o->nphot;
    for (int c = 0; c < nc; ++c) {
        int ir = c % g.nr;
        for (int q = 0; q < ng; ++q) {
            double W = ldexp(1.0, TALLY_SHIFT) / (t.kmax[q] * g.dr[ir]);
            mc.W[(size_t)c * ng + q] = W;
            mc.econv[(size_t)c * ng + q] = eps / (W * g.vol[c]);
        }
    }
    cnt_t tot;
    memset(&tot, 0, sizeof(tot));
    int nep = 0;
    if (!o->lagged && o->nthreads > 1) {
        /* free-running immediate re-emission on several threads (the way an
         * OpenMP mctherm runs): packets read the shared tallies as they go, so
         * the result depends on the schedule; used for timing */
#ifdef _OPENMP
        omp_set_num_threads(o->nthreads);
#endif
        uint64_t a = 0, b = 0, c2 = 0, d = 0, e2 = 0, f = 0, h = 0, nu = 0;
        int64_t hi = (int64_t)o->nphot;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : a, b, c2, d, e2, f, h, nu)
        for (int64_t p = 0; p < hi; ++p) {
            cnt_t cc;
            memset(&cc, 0, sizeof(cc));
            run_packet(&mc, (uint64_t)p, 0, &cc);
            a += cc.n_cross;
            b += cc.n_abs;
            c2 += cc.n_scat;
            d += cc.n_escape;
            e2 += cc.n_killed;
            f += cc.n_events;
            h += cc.n_mrw;
            nu += cc.n_nudged;
        }
        tot.n_cross = a;
        tot.n_abs = b;
        tot.n_scat = c2;
        tot.n_escape = d;
        tot.n_killed = e2;
        tot.n_events = f;
        tot.n_mrw = h;
        tot.n_nudged = nu;
        nep = 1;
    } else if (!o->lagged) {
        for (uint64_t p = 0; p < o->nphot; ++p) run_packet(&mc, p, 0, &tot);
        nep = 1;
    } else {
#ifdef _OPENMP
        if (o->nthreads > 0) omp_set_num_threads(o->nthreads);
#endif
        uint64_t done = 0;
        uint64_t first = o->epoch_first ? o->epoch_first : 1;
        /* count epochs, allocate one deposit slot per epoch plus a final one */
        {
            uint64_t dd = 0;
            int ne = 0;
            while (dd < o->nphot) {
                uint64_t sz = (uint64_t)floor((double)dd * (o->epoch_growth - 1.0));
                if (sz < first) sz = first;
                dd += sz;
                if (dd > o->nphot) dd = o->nphot;
                ne++;
            }
            mc.nslot = ne + 1;
            mc.slots = calloc((size_t)mc.nslot * nc * ng, sizeof(uint64_t));
        }
        /* ring of `lag` snapshots: before epoch e, slot e % lag holds the
         * temperatures of epochs <= e - lag (zeros for e < lag) */
        int lag = o->lag > 0 ? o->lag : 2;
        double *ring = calloc((size_t)lag * nc * ng, sizeof(double));
        double *Tfinal = mc.Tcur;
        int e = 0;
        while (done < o->nphot) {
            mc.Tcur = ring + (size_t)(e % lag) * nc * ng;
            uint64_t sz = (uint64_t)floor((double)done * (o->epoch_growth - 1.0));
            if (sz < first) sz = first;
            uint64_t end = done + sz;
            if (end > o->nphot) end = o->nphot;
            uint64_t a = 0, b = 0, c2 = 0, d = 0, e2 = 0, f = 0, h = 0, nu = 0;
            int64_t lo = (int64_t)done, hi = (int64_t)end;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : a, b, c2, d, e2, f, h, nu)
            for (int64_t p = lo; p < hi; ++p) {
                cnt_t cc;
                memset(&cc, 0, sizeof(cc));
                run_packet(&mc, (uint64_t)p, e, &cc);
                a += cc.n_cross;
                b += cc.n_abs;
                c2 += cc.n_scat;
                d += cc.n_escape;
                e2 += cc.n_killed;
                f += cc.n_events;
                h += cc.n_mrw;
                nu += cc.n_nudged;
            }
            tot.n_cross += a;
            tot.n_abs += b;
            tot.n_scat += c2;
            tot.n_escape += d;
            tot.n_killed += e2;
            tot.n_events += f;
            tot.n_mrw += h;
            tot.n_nudged += nu;
            done = end;
            nep++;
            /* fold deposit slot e (final now: it only receives deposits of
             * epochs <= e) and take S_e, into the ring entry epoch e + lag reads */
            {
                uint64_t *sl = mc.slots + (size_t)e * nc * ng;
                for (size_t i = 0; i < (size_t)nc * ng; ++i) mc.tally[i] += sl[i];
            }
            snapshot(&mc);
            e++;
        }
        for (int a = e; a < mc.nslot; ++a) {
            uint64_t *sl = mc.slots + (size_t)a * nc * ng;
            for (size_t i = 0; i < (size_t)nc * ng; ++i) mc.tally[i] += sl[i];
        }
        free(mc.slots);
        mc.Tcur = Tfinal;
        free(ring);
    }
    snapshot(&mc);
    /* expected number of absorptions implied by the path-length tallies:
     * sum_c sum_s rho_s * (kabs_s * path) */
    double tsum = 0;
    for (int c = 0; c < nc; ++c)
        for (int s = 0; s < m->nspec; ++s) {
            int q = t.grp[s];
            tsum += (double)mc.tally[(size_t)c * ng + q] / mc.W[(size_t)c * ng + q] *
                    m->rho[(size_t)s * nc + c];
        }
    for (int s = 0; s < m->nspec; ++s) {
        int q = t.grp[s];
        for (int c = 0; c < nc; ++c) {
            double rho = m->rho[(size_t)s * nc + c];
            T_out[(size_t)s * nc + c] = (rho > 0) ? mc.Tcur[(size_t)c * ng + q] : 0.0;
        }
    }
    if (st) {
        st->n_cross = tot.n_cross;
        st->n_abs = tot.n_abs;
        st->n_scat = tot.n_scat;
        st->n_escape = tot.n_escape;
        st->n_killed = tot.n_killed;
        st->n_events = tot.n_events;
        st->n_mrw = tot.n_mrw;
        st->eps = eps;
        st->lum = t.lum;
        st->tally_tau_sum = tsum;
        st->n_epochs = nep;
        st->n_nudged = tot.n_nudged;
    }
    int ovf = mc.overflow;
    free(mc.tally);
    free(mc.W);
    free(mc.econv);
    free(mc.Tcur);
    tab_free(&t);
    geom_free(&g);
    return ovf ? -20 : 0;
}

/* ------------------------------------------------ scattering source MC */
/* Monochromatic Monte Carlo for the mean intensity J_nu per cell at the image
 * wavelength, which RADMC-3D computes before each image when the opacities
 * carry scattering (nphot_scat packets, pdspy/radmc3d/write.py:34 and
 * Model.py:520; run.py:157-289), for isotropic scattering.  Sources: the
 * stars (L_nu = 4 pi^2 R^2 B_nu(T*)) and the thermal emission of the dust at
 * the imaging temperature (L_nu,c = 4 pi V_c B_nu(T_c) sum_s rho_s kabs_s).
 * Packets of energy L/N travel with extinction rho (kabs + ksca); at each
 * interaction they scatter isotropically with probability ksca / kext and
 * end otherwise.  Path-length estimator: J_c = (L/N) sum l / (4 pi V_c).
 * Random events per packet (Philox key = seed, counter (k, packet)):
 * k = 0: u0 source, u1 u2 u3 position in the cell (r^3, cos theta, phi
 * uniform), u4 u5 direction, u6 first tau, u7 hemisphere (mirrored grids);
 * k >= 1 at the k-th interaction: u0 scatter or end, u1 u2 new direction,
 * u3 next tau.                                                           */
/* ------------------------------------------------- tables, for the tests */
struct or_tables {
    tab_t t;
};

or_tables *or_tables_create(const or_model *m, int ntemp, double temp0, double temp1) {
    or_thermal_opts o;
    memset(&o, 0, sizeof(o));
    o.ntemp = ntemp;
    o.temp0 = temp0;
    o.temp1 = temp1;
    or_tables *h = calloc(1, sizeof(or_tables));
    if (!h) return NULL;
    if (tab_init(&h->t, m, &o) != 0) {
        tab_free(&h->t);
        free(h);
        return NULL;
    }
    return h;
}

void or_tables_free(or_tables *h) {
    if (!h) return;
    tab_free(&h->t);
    free(h);
}

void or_tables_shape(const or_tables *h, int *ngroup, int *ntemp, int *nlam) {
    *ngroup = h->t.ngroup;
    *ntemp = h->t.ntemp;
    *nlam = h->t.nlam;
}

void or_tables_get(const or_tables *h, double *lnq, double *kplanck, double *kross, double *cdf,
                   int *grp) {
    const tab_t *t = &h->t;
    if (lnq) memcpy(lnq, t->lnq, sizeof(double) * t->ngroup * t->ntemp);
    if (kplanck) memcpy(kplanck, t->kplanck, sizeof(double) * t->ngroup * t->ntemp);
    if (kross) memcpy(kross, t->kross, sizeof(double) * t->nspec * t->ntemp);
    if (cdf) memcpy(cdf, t->cdf, sizeof(double) * (size_t)t->ngroup * t->ntemp * t->nlam);
    if (grp) memcpy(grp, t->grp, sizeof(int) * t->nspec);
}

int or_reemit_bin(const or_tables *h, int q, double T, double u_row, double u_freq) {
    return reemit_bin(&h->t, q, T, u_row, u_freq);
}

void or_reemit_hist(const or_tables *h, int q, double T, uint64_t n, uint64_t seed,
                    uint64_t *counts) {
    memset(counts, 0, sizeof(uint64_t) * h->t.nlam);
    for (uint64_t k = 0; k < n; ++k) {
        double u[8];
        rng8(seed, k, 0, u);
        counts[reemit_bin(&h->t, q, T, u[0], u[1])]++;
        counts[reemit_bin(&h->t, q, T, u[2], u[3])]++;
    }
}

static int sample_cum(const double *cum, int n, double x) {
    /* first i with cum[i] > x (cum ascending, cum[n-1] = total) */
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi) / 2;
        if (cum[mid] > x)
            hi = mid;
        else
            lo = mid + 1;
    }
    return lo;
}

int or_scat_source(const or_model *m, const double *T, const double *kap_abs,
                   const double *kap_sca, double lam_um, uint64_t nphot, uint64_t seed,
                   uint64_t max_events, int nthreads, double *J, or_scat_stats *st) {
    geom_t g;
    int rc = geom_init(&g, m);
    if (rc) return rc;
    const int nc = g.ncell, ns = m->nspec;
    const double nu = C_C / (lam_um * 1e-4);
    /* source luminosities: stars first, then cells (cumulative) */
    double *cum = malloc(sizeof(double) * (size_t)(m->nstar + nc));
    double acc = 0;
    for (int k = 0; k < m->nstar; ++k) {
        double R = m->star[5 * k];
        acc += 4.0 * PI_D * PI_D * R * R * or_planck(nu, m->star[5 * k + 1]);
        cum[k] = acc;
    }
    const double lstar = acc;
    for (int c = 0; c < nc; ++c) {
        double a = 0;
        for (int sp = 0; sp < ns; ++sp) a += m->rho[(size_t)sp * nc + c] * kap_abs[sp];
        acc += 4.0 * PI_D * g.vol[c] * a * or_planck(nu, T[c]);
        cum[m->nstar + c] = acc;
    }
    const double ltot = acc;
    memset(J, 0, sizeof(double) * nc);
    or_scat_stats tot;
    memset(&tot, 0, sizeof(tot));
    tot.lum_star = lstar;
    tot.lum_dust = ltot - lstar;
    if (!(ltot > 0) || nphot == 0) {
        free(cum);
        geom_free(&g);
        if (st) *st = tot;
        return 0;
    }
    int nth = nthreads > 0 ? nthreads : 1;
#ifdef _OPENMP
    if (nthreads <= 0) nth = omp_get_max_threads();
#else
    nth = 1;
#endif
    double *part = calloc((size_t)nth * nc, sizeof(double));
#pragma omp parallel num_threads(nth)
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double *Jt = part + (size_t)tid * nc;
        uint64_t c_esc = 0, c_abs = 0, c_sca = 0, c_cross = 0, c_kill = 0;
#pragma omp for schedule(dynamic, 256)
        for (long long ip = 0; ip < (long long)nphot; ++ip) {
            const uint64_t pid = (uint64_t)ip;
            double u[8];
            rng8(seed, pid, 0, u);
            pk_t p;
            memset(&p, 0, sizeof(p));
            const int src = sample_cum(cum, m->nstar + nc, u[0] * ltot);
            if (src < m->nstar) {
                p.x = m->star[5 * src + 2];
                p.y = m->star[5 * src + 3];
                p.z = m->star[5 * src + 4];
            } else {
                const int c = src - m->nstar;
                const int ir = c % g.nr, it = (c / g.nr) % g.nt, iq = c / (g.nr * g.nt);
                const double r1 = m->rw[ir], r2 = m->rw[ir + 1];
                const double r = cbrt(r1 * r1 * r1 + u[1] * (r2 * r2 * r2 - r1 * r1 * r1));
                const double ct = g.ct[it] + u[2] * (g.ct[it + 1] - g.ct[it]);
                const double sth = sqrt(fmax(0.0, 1.0 - ct * ct));
                const double ph = m->pw[iq] + u[3] * (m->pw[iq + 1] - m->pw[iq]);
                p.x = r * sth * cos(ph);
                p.y = r * sth * sin(ph);
                p.z = r * ct * ((g.mirror && u[7] < 0.5) ? -1.0 : 1.0);
            }
            double d[3];
            iso_dir(u[4], u[5], d);
            p.dx = d[0];
            p.dy = d[1];
            p.dz = d[2];
            locate(&g, &p);
            if (p.ir >= g.nr) {
                c_esc++;
                continue;
            }
            double tau = -log(u[6]);
            uint32_t k = 1;
            int nzero = 0;
            uint64_t nev = 0;
            for (;;) {
                int kind;
                double tw = next_wall(&g, &p, &kind);
                double aabs = 0, asca = 0;
                int c = -1;
                if (p.ir >= 0) {
                    c = cell_of(&g, &p);
                    for (int sp = 0; sp < ns; ++sp) {
                        double rho = m->rho[(size_t)sp * nc + c];
                        aabs += rho * kap_abs[sp];
                        asca += rho * kap_sca[sp];
                    }
                }
                double aext = aabs + asca;
                if (aext * tw > tau) {
                    double sseg = tau / aext;
                    Jt[c] += sseg;
                    p.x += sseg * p.dx;
                    p.y += sseg * p.dy;
                    p.z += sseg * p.dz;
                    double v[8];
                    rng8(seed, pid, k++, v);
                    nev++;
                    if (v[0] * aext < asca) {
                        c_sca++;
                        iso_dir(v[1], v[2], d);
                        p.dx = d[0];
                        p.dy = d[1];
                        p.dz = d[2];
                        tau = -log(v[3]);
                        nzero = 0;
                        if (max_events && nev >= max_events) {
                            c_kill++;
                            break;
                        }
                    } else {
                        c_abs++;
                        break;
                    }
                } else {
                    if (c >= 0) Jt[c] += tw;
                    tau -= aext * tw;
                    c_cross++;
                    double rcell = g.rw[p.ir + 1];
                    if (cross_wall(&g, &p, tw, kind)) {
                        c_esc++;
                        break;
                    }
                    if (unstick(&g, &p, tw, rcell, &nzero) && p.ir >= g.nr) {
                        c_esc++;
                        break;
                    }
                }
            }
        }
#pragma omp critical
        {
            tot.n_escape += c_esc;
            tot.n_abs += c_abs;
            tot.n_scat += c_sca;
            tot.n_cross += c_cross;
            tot.n_killed += c_kill;
        }
    }
    const double eps = ltot / (double)nphot;
    for (int c = 0; c < nc; ++c) {
        double sum = 0;
        for (int t = 0; t < nth; ++t) sum += part[(size_t)t * nc + c];
        J[c] = eps * sum / (4.0 * PI_D * g.vol[c]);
    }
    tot.eps = eps;
    if (st) *st = tot;
    free(part);
    free(cum);
    geom_free(&g);
    return 0;
}

/* ---------------------------------------------------------------- camera */
typedef struct {
    double n[3], ex[3], ey[3];
    double cpa, spa;
} cam_frame_t;

/* observer direction n = (sin i sin phi, -sin i cos phi, cos i);
 * image axes ex = (cos phi, sin phi, 0), ey = n x ex (z projects upwards). */
static void cam_frame(const or_camera *cam, cam_frame_t *f) {
    double i = cam->incl * PI_D / 180.0, ph = cam->phi * PI_D / 180.0,
           pa = cam->posang * PI_D / 180.0;
    f->n[0] = sin(i) * sin(ph);
    f->n[1] = -sin(i) * cos(ph);
    f->n[2] = cos(i);
    f->ex[0] = cos(ph);
    f->ex[1] = sin(ph);
    f->ex[2] = 0.0;
    f->ey[0] = f->n[1] * f->ex[2] - f->n[2] * f->ex[1];
    f->ey[1] = f->n[2] * f->ex[0] - f->n[0] * f->ex[2];
    f->ey[2] = f->n[0] * f->ex[1] - f->n[1] * f->ex[0];
    f->cpa = cos(pa);
    f->spa = sin(pa);
}

typedef struct {
    int cell;
    double s, dtau, S, tau0;
} seg_t;

/* Integrate one ray entering from the observer side.  Collects segments when
 * segs != NULL.  Returns intensity (mode 0) or total tau (mode 1).         */
/* size of cell (ir, it, ip) for the camera's sub-pixeling: the least of its
 * radial width, r_c dtheta and (with phi cells) r_c sin(theta_c) dphi */
static double cell_size(const geom_t *g, int ir, int it, int ip) {
    double rc = 0.5 * (g->rw[ir] + g->rw[ir + 1]);
    double l = g->rw[ir + 1] - g->rw[ir];
    double lt = rc * (g->tw[it + 1] - g->tw[it]);
    if (lt < l) l = lt;
    if (g->np > 1) {
        double lp = rc * sin(0.5 * (g->tw[it] + g->tw[it + 1])) * (g->pw[ip + 1] - g->pw[ip]);
        if (lp < l) l = lp;
    }
    return l;
}

static double trace_cam_ray_l(const or_model *m, const geom_t *g, const double *T,
                              const double *kap, const double *ksca, const double *Jsc,
                              const or_camera *cam, const cam_frame_t *f, double X, double Y,
                              seg_t *segs, int maxseg, int *nseg, double *lmin);

static double trace_cam_ray(const or_model *m, const geom_t *g, const double *T,
                            const double *kap, const double *ksca, const double *Jsc,
                            const or_camera *cam, const cam_frame_t *f, double X, double Y,
                            seg_t *segs, int maxseg, int *nseg) {
    return trace_cam_ray_l(m, g, T, kap, ksca, Jsc, cam, f, X, Y, segs, maxseg, nseg, NULL);
}

/* the same, also returning the smallest cell size along the traced part of
 * the ray in *lmin (INF_D when it crosses no cell) */
static double trace_cam_ray_l(const or_model *m, const geom_t *g, const double *T,
                              const double *kap, const double *ksca, const double *Jsc,
                              const or_camera *cam, const cam_frame_t *f, double X, double Y,
                              seg_t *segs, int maxseg, int *nseg, double *lmin) {
    /* sky (X,Y) -> model image plane (x', y') rotated by -posang */
    double xp = X * f->cpa + Y * f->spa;
    double yp = -X * f->spa + Y * f->cpa;
    double R = g->rw[g->nr];
    double o[3], d[3];
    for (int a = 0; a < 3; ++a) {
        o[a] = xp * f->ex[a] + yp * f->ey[a] + 2.0 * R * f->n[a];
        d[a] = -f->n[a];
    }
    if (nseg) *nseg = 0;
    if (lmin) *lmin = INF_D;
    double rr = o[0] * o[0] + o[1] * o[1] + o[2] * o[2];
    double b = o[0] * d[0] + o[1] * d[1] + o[2] * d[2];
    double D = b * b - (rr - R * R);
    if (D <= 0) return 0.0;
    double tin = -b - sqrt(D);
    pk_t p;
    memset(&p, 0, sizeof(p));
    p.x = o[0] + tin * d[0];
    p.y = o[1] + tin * d[1];
    p.z = o[2] + tin * d[2];
    p.dx = d[0];
    p.dy = d[1];
    p.dz = d[2];
    snap_radius(&p, R);
    p.ir = g->nr - 1;
    locate_angles(g, &p);
    double I = 0, tau = 0;
    int ns = m->nspec, n = 0;
    double nu = C_C / (cam->lam_um * 1e-4);
    int nzero = 0;
    for (int guard = 0; guard < 100000000; ++guard) {
        int kind;
        double tw = next_wall(g, &p, &kind);
        double rcell = g->rw[p.ir + 1];
        if (p.ir >= 0) {
            int c = cell_of(g, &p);
            if (lmin) {
                double l = cell_size(g, p.ir, p.it, p.ip);
                if (l < *lmin) *lmin = l;
            }
            /* extinction and source function; with a scattering source
             * (isotropic): S = (rho kabs B + rho ksca J) / (rho (kabs + ksca)) */
            double aa = 0, as = 0;
            for (int s = 0; s < ns; ++s) {
                aa += m->rho[(size_t)s * g->ncell + c] * kap[s];
                if (Jsc) as += m->rho[(size_t)s * g->ncell + c] * ksca[s];
            }
            double a = aa + as;
            double dtau = a * tw;
            double B = or_planck(nu, T[c]);
            double S = Jsc ? (a > 0 ? (aa * B + as * Jsc[c]) / a : 0.0) : B;
            if (cam->mode == 0) I += exp(-tau) * (-expm1(-dtau)) * S;
            if (segs && n < maxseg) {
                segs[n].cell = c;
                segs[n].s = tw;
                segs[n].dtau = dtau;
                segs[n].S = S;
                segs[n].tau0 = tau;
            }
            n++;
            tau += dtau;
            if (cam->mode == 0 && tau > cam->tau_stop) break;
        }
        if (cross_wall(g, &p, tw, kind)) break;
        if (unstick(g, &p, tw, rcell, &nzero) && p.ir >= g->nr) break;
    }
    if (nseg) *nseg = n;
    return cam->mode == 0 ? I : tau;
}

/* optical depth from point q to the observer along +n */
static double star_tau(const or_model *m, const geom_t *g, const double *kap,
                       const double *ksca, const double *q, const double *n, seg_t *segs,
                       int maxseg, int *nseg) {
    pk_t p;
    memset(&p, 0, sizeof(p));
    p.x = q[0];
    p.y = q[1];
    p.z = q[2];
    p.dx = n[0];
    p.dy = n[1];
    p.dz = n[2];
    locate(g, &p);
    double tau = 0;
    int k = 0, nzero = 0;
    while (p.ir >= -1 && p.ir < g->nr) {
        int kind;
        double tw = next_wall(g, &p, &kind);
        double rcell = g->rw[p.ir + 1];
        if (p.ir >= 0) {
            int c = cell_of(g, &p);
            double a = 0;
            for (int s = 0; s < m->nspec; ++s)
                a += m->rho[(size_t)s * g->ncell + c] * (kap[s] + (ksca ? ksca[s] : 0.0));
            if (segs && k < maxseg) {
                segs[k].cell = c;
                segs[k].s = tw;
                segs[k].dtau = a * tw;
            }
            k++;
            tau += a * tw;
        }
        if (cross_wall(g, &p, tw, kind)) break;
        unstick(g, &p, tw, rcell, &nzero);
    }
    if (nseg) *nseg = k;
    return tau;
}

static int star_pixel(const or_camera *cam, const cam_frame_t *f, const double *q,
                      int *jj, int *kk) {
    double xp = q[0] * f->ex[0] + q[1] * f->ex[1] + q[2] * f->ex[2];
    double yp = q[0] * f->ey[0] + q[1] * f->ey[1] + q[2] * f->ey[2];
    /* model plane -> sky: rotate by +posang */
    double X = xp * f->cpa - yp * f->spa;
    double Y = xp * f->spa + yp * f->cpa;
    int k = (int)floor(X / cam->pix_cm + 0.5) + cam->nx / 2;
    int j = (int)floor(Y / cam->pix_cm + 0.5) + cam->ny / 2;
    if (k < 0 || k >= cam->nx || j < 0 || j >= cam->ny) return 0;
    *jj = j;
    *kk = k;
    return 1;
}

/* Value of the (sub)pixel of side d centred at sky (X, Y), level lev of the
 * recursive sub-pixeling: the central ray, or, when that ray crosses a cell of
 * size l with d > crit l and lev < nrefine, the mean of the 2x2 sub-pixels
 * taken in the order (-,-), (+,-), (-,+), (+,+) in (X, Y).  *nrays counts
 * the rays traced.                                                          */
static double subpixel_value(const or_model *m, const geom_t *g, const double *T,
                             const double *kap, const double *ksca, const double *Jsc,
                             const or_camera *cam, const cam_frame_t *f, double X, double Y,
                             double d, int lev, long *nrays) {
    double lmin;
    double I = trace_cam_ray_l(m, g, T, kap, ksca, Jsc, cam, f, X, Y, NULL, 0, NULL, &lmin);
    if (nrays) ++*nrays;
    if (lev >= cam->nrefine || !(d > cam->refine_crit * lmin)) return I;
    double h = 0.25 * d, acc = 0;
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            acc += subpixel_value(m, g, T, kap, ksca, Jsc, cam, f, X + (b ? h : -h),
                                  Y + (a ? h : -h), 0.5 * d, lev + 1, nrays);
    return 0.25 * acc;
}

/* rays traced by the last or_image / or_image_scat call */
static long g_last_nrays = 0;
long or_image_last_nrays(void) { return g_last_nrays; }

int or_image(const or_model *m, const double *T, const double *kap,
             const or_camera *cam, double *img) {
    return or_image_scat(m, T, kap, NULL, NULL, cam, img);
}

int or_image_scat(const or_model *m, const double *T, const double *kap,
                  const double *ksca, const double *Jsc, const or_camera *cam, double *img) {
    geom_t g;
    int rc = geom_init(&g, m);
    if (rc) return rc;
    cam_frame_t f;
    cam_frame(cam, &f);
    int nsub = cam->nsub > 0 ? cam->nsub : 1;
    const double dsub = cam->pix_cm / nsub;
    long nrays = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : nrays)
    for (int j = 0; j < cam->ny; ++j)
        for (int k = 0; k < cam->nx; ++k) {
            double acc = 0;
            for (int a = 0; a < nsub; ++a)
                for (int b = 0; b < nsub; ++b) {
                    double X = (k - cam->nx / 2 + ((b + 0.5) / nsub - 0.5)) * cam->pix_cm;
                    double Y = (j - cam->ny / 2 + ((a + 0.5) / nsub - 0.5)) * cam->pix_cm;
                    acc += subpixel_value(m, &g, T, kap, Jsc ? ksca : NULL, Jsc, cam, &f, X, Y,
                                          dsub, 0, &nrays);
                }
            img[(size_t)j * cam->nx + k] = acc / (nsub * nsub);
        }
    g_last_nrays = nrays;
    if (cam->incl_star && cam->mode == 0) {
        double nu = C_C / (cam->lam_um * 1e-4);
        for (int s = 0; s < m->nstar; ++s) {
            const double *q = m->star + 5 * s + 2;
            int jj, kk;
            if (!star_pixel(cam, &f, q, &jj, &kk)) continue;
            double R = m->star[5 * s], Ts = m->star[5 * s + 1];
            double F = PI_D * R * R * or_planck(nu, Ts) / (cam->pix_cm * cam->pix_cm);
            img[(size_t)jj * cam->nx + kk] +=
                F * exp(-star_tau(m, &g, kap, Jsc ? ksca : NULL, q, f.n, NULL, 0, NULL));
        }
    }
    geom_free(&g);
    return 0;
}

int or_image_adjoint(const or_model *m, const double *T, const double *kap,
                     const or_camera *cam, const double *dimg, double *g_logrho,
                     double *g_T) {
    return or_image_adjoint_scat(m, T, kap, NULL, NULL, cam, dimg, g_logrho, g_T);
}

/* With a scattering source J (held fixed): dtau = a s with a = sum rho (kabs +
 * ksca), S = (sum rho kabs B + sum rho ksca J) / a, so
 * dI/dlog rho_s = dI/ddtau s rho_s (kabs_s + ksca_s)
 *               + w (rho_s kabs_s B + rho_s ksca_s J - S rho_s (kabs_s + ksca_s)) / a
 * and dI/dT = w (sum rho kabs / a) dB/dT, w = e^-tau0 (1 - e^-dtau).  Without
 * J this reduces to the emission-only gradient.                            */
/* adds gI * dI/dlog rho and gI * dI/dT of the ray through sky (X, Y) */
static int adjoint_ray(const or_model *m, const geom_t *g, const double *T, const double *kap,
                       const double *ksca, const double *Jsc, const or_camera *cam,
                       const cam_frame_t *f, double X, double Y, double gI, seg_t *segs,
                       int maxseg, double *g_logrho, double *g_T) {
    const int nc = g->ncell, ns = m->nspec;
    const double nu = C_C / (cam->lam_um * 1e-4);
    int n;
    double I = trace_cam_ray(m, g, T, kap, Jsc ? ksca : NULL, Jsc, cam, f, X, Y, segs, maxseg, &n);
    if (n > maxseg) return -30;
    double cum = 0;
    for (int q = 0; q < n; ++q) {
        seg_t *sg = &segs[q];
        double att = exp(-sg->tau0);
        double contrib = att * (-expm1(-sg->dtau)) * sg->S;
        cum += contrib;
        double Rk = I - cum;
        double dI_ddtau = att * exp(-sg->dtau) * sg->S - Rk;
        double w = att * (-expm1(-sg->dtau));
        double a = 0, aa = 0;
        for (int s = 0; s < ns; ++s) {
            double rho = m->rho[(size_t)s * nc + sg->cell];
            aa += rho * kap[s];
            a += rho * (kap[s] + (Jsc ? ksca[s] : 0.0));
        }
        double B = or_planck(nu, T[sg->cell]);
        for (int s = 0; s < ns; ++s) {
            double rho = m->rho[(size_t)s * nc + sg->cell];
            double ks = Jsc ? ksca[s] : 0.0;
            double gS = Jsc && a > 0 ? (rho * kap[s] * B + rho * ks * Jsc[sg->cell] -
                                        sg->S * rho * (kap[s] + ks)) / a
                                     : 0.0;
            g_logrho[(size_t)s * nc + sg->cell] +=
                gI * (dI_ddtau * sg->s * rho * (kap[s] + ks) + w * gS);
        }
        double fa = Jsc ? (a > 0 ? aa / a : 0.0) : 1.0;
        g_T[sg->cell] += gI * w * fa * or_dplanck_dT(nu, T[sg->cell]);
    }
    return 0;
}

/* the adjoint of subpixel_value: the same refinement decisions, each leaf
 * ray weighted by gI / 4^(its depth) */
static int subpixel_adjoint(const or_model *m, const geom_t *g, const double *T,
                            const double *kap, const double *ksca, const double *Jsc,
                            const or_camera *cam, const cam_frame_t *f, double X, double Y,
                            double d, int lev, double gI, seg_t *segs, int maxseg,
                            double *g_logrho, double *g_T) {
    if (cam->nrefine > lev) {
        double lmin;
        trace_cam_ray_l(m, g, T, kap, Jsc ? ksca : NULL, Jsc, cam, f, X, Y, NULL, 0, NULL, &lmin);
        if (d > cam->refine_crit * lmin) {
            double h = 0.25 * d;
            for (int a = 0; a < 2; ++a)
                for (int b = 0; b < 2; ++b) {
                    int rc = subpixel_adjoint(m, g, T, kap, ksca, Jsc, cam, f, X + (b ? h : -h),
                                              Y + (a ? h : -h), 0.5 * d, lev + 1, 0.25 * gI,
                                              segs, maxseg, g_logrho, g_T);
                    if (rc) return rc;
                }
            return 0;
        }
    }
    return adjoint_ray(m, g, T, kap, ksca, Jsc, cam, f, X, Y, gI, segs, maxseg, g_logrho, g_T);
}

int or_image_adjoint_scat(const or_model *m, const double *T, const double *kap,
                          const double *ksca, const double *Jsc, const or_camera *cam,
                          const double *dimg, double *g_logrho, double *g_T) {
    geom_t g;
    int rc = geom_init(&g, m);
    if (rc) return rc;
    cam_frame_t f;
    cam_frame(cam, &f);
    int nc = g.ncell, ns = m->nspec;
    memset(g_logrho, 0, sizeof(double) * ns * nc);
    memset(g_T, 0, sizeof(double) * nc);
    int nsub = cam->nsub > 0 ? cam->nsub : 1;
    const double dsub = cam->pix_cm / nsub;
    int maxseg = 4 * (g.nr + g.nt + g.np) + 64;
    seg_t *segs = malloc(sizeof(seg_t) * maxseg);
    double nu = C_C / (cam->lam_um * 1e-4);
    for (int j = 0; j < cam->ny; ++j)
        for (int k = 0; k < cam->nx; ++k) {
            double gI = dimg[(size_t)j * cam->nx + k] / (nsub * nsub);
            if (gI == 0) continue;
            for (int a = 0; a < nsub; ++a)
                for (int b = 0; b < nsub; ++b) {
                    double X = (k - cam->nx / 2 + ((b + 0.5) / nsub - 0.5)) * cam->pix_cm;
                    double Y = (j - cam->ny / 2 + ((a + 0.5) / nsub - 0.5)) * cam->pix_cm;
                    rc = subpixel_adjoint(m, &g, T, kap, ksca, Jsc, cam, &f, X, Y, dsub, 0, gI,
                                          segs, maxseg, g_logrho, g_T);
                    if (rc) {
                        free(segs);
                        geom_free(&g);
                        return rc;
                    }
                }
        }
    if (cam->incl_star && cam->mode == 0) {
        for (int s = 0; s < m->nstar; ++s) {
            const double *q = m->star + 5 * s + 2;
            int jj, kk;
            if (!star_pixel(cam, &f, q, &jj, &kk)) continue;
            double gI = dimg[(size_t)jj * cam->nx + kk];
            double R = m->star[5 * s], Ts = m->star[5 * s + 1];
            double F = PI_D * R * R * or_planck(nu, Ts) / (cam->pix_cm * cam->pix_cm);
            int n;
            double tau = star_tau(m, &g, kap, Jsc ? ksca : NULL, q, f.n, segs, maxseg, &n);
            double Is = F * exp(-tau);
            for (int q2 = 0; q2 < n && q2 < maxseg; ++q2)
                for (int sp = 0; sp < ns; ++sp) {
                    double rho = m->rho[(size_t)sp * nc + segs[q2].cell];
                    g_logrho[(size_t)sp * nc + segs[q2].cell] +=
                        gI * (-Is) * segs[q2].s * rho * (kap[sp] + (Jsc ? ksca[sp] : 0.0));
                }
        }
    }
    free(segs);
    geom_free(&g);
    return 0;
}

/* ---------------------------------------------------------- visibilities */
void or_vis_dft(const double *img, int nx, int ny, double dxy, const double *u,
                const double *v, int K, double dra, double ddec, double *re,
                double *im) {
#pragma omp parallel
    {
        double *ax = malloc(sizeof(double) * 2 * nx), *by = malloc(sizeof(double) * 2 * ny);
#pragma omp for schedule(static)
        for (int q = 0; q < K; ++q) {
            for (int k = 0; k < nx; ++k) {
                double ph = -TWO_PI_D * u[q] * (k - nx / 2) * dxy;
                ax[2 * k] = cos(ph);
                ax[2 * k + 1] = sin(ph);
            }
            for (int j = 0; j < ny; ++j) {
                double ph = -TWO_PI_D * v[q] * (j - ny / 2) * dxy;
                by[2 * j] = cos(ph);
                by[2 * j + 1] = sin(ph);
            }
            double sr = 0, si = 0;
            for (int j = 0; j < ny; ++j) {
                double rr = 0, ri = 0;
                const double *row = img + (size_t)j * nx;
                for (int k = 0; k < nx; ++k) {
                    rr += row[k] * ax[2 * k];
                    ri += row[k] * ax[2 * k + 1];
                }
                sr += rr * by[2 * j] - ri * by[2 * j + 1];
                si += rr * by[2 * j + 1] + ri * by[2 * j];
            }
            double ph = -TWO_PI_D * (u[q] * dra + v[q] * ddec);
            double c = cos(ph), s = sin(ph);
            re[q] = sr * c - si * s;
            im[q] = sr * s + si * c;
        }
        free(ax);
        free(by);
    }
}

void or_vis_dft_adjoint(const double *gre, const double *gim, int nx, int ny,
                        double dxy, const double *u, const double *v, int K,
                        double dra, double ddec, double *g_img) {
    memset(g_img, 0, sizeof(double) * nx * ny);
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; ++j)
        for (int k = 0; k < nx; ++k) {
            double acc = 0;
            for (int q = 0; q < K; ++q) {
                double ph = -TWO_PI_D * (u[q] * ((k - nx / 2) * dxy + dra) +
                                         v[q] * ((j - ny / 2) * dxy + ddec));
                acc += gre[q] * cos(ph) + gim[q] * sin(ph);
            }
            g_img[(size_t)j * nx + k] = acc;
        }
}

double or_chisq(const double *dre, const double *dim, const double *w,
                const double *mre, const double *mim, int K) {
    double s = 0;
    for (int q = 0; q < K; ++q) {
        double a = dre[q] - mre[q], b = dim[q] - mim[q];
        s += (a * a + b * b) * w[q];
    }
    return s;
}
