// rt_device.h -- device helpers shared by the gfx950 kernels: Philox4x32-10,
// the spherical-grid traversal (FP32 with FP64 re-evaluation of near-tangent
// sphere/cone crossings), Planck functions and wave-level helpers.
//
// Geometry follows the grids pdspy writes for RADMC-3D (radmc3d/write.py:198-244):
// r walls (lengths in au on the device), theta walls in [0, pi/2] with an
// equatorial mirror (the lower half-space folds onto the upper one) or [0, pi],
// phi walls over [0, 2 pi].  Cells are indexed c = ir + nr*(it + nt*ip)
// (r fastest, dust_density.inp order, write.py:259-264).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define RT_INF 3.0e38f

namespace rt {

// ---------------------------------------------------------------- Philox
// Salmon, Moraes, Dror & Shaw, "Parallel random numbers: as easy as 1, 2, 3"
// (SC'11): 10 rounds of the 4x32 bijection, Weyl key schedule.

__device__ __forceinline__ void philox_block(uint32_t c0, uint32_t c1, uint32_t c2,
                                             uint32_t c3, uint32_t k0, uint32_t k1,
                                             uint32_t out[4]) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint32_t hi0 = __umulhi(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
        uint32_t hi1 = __umulhi(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
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

// uniform in (0,1) from the top 24 bits of a word, exact in FP32
__device__ __forceinline__ float u24(uint32_t x) {
    return ((float)(x >> 8) + 0.5f) * (1.0f / 16777216.0f);
}

// The random numbers of a packet come in events: event k (k = 0 at emission,
// then one per interaction and one per modified-random-walk step) uses the 8
// uniforms of the blocks with counter (k, pid_lo, pid_hi, 0) and
// (k, pid_lo, pid_hi, 1) under key = seed.  A fixed number of Philox blocks per
// event keeps the wavefront from re-running the generator at every draw.
__device__ __forceinline__ void rng8(uint64_t seed, uint64_t pid, uint32_t k, float u[8]) {
    uint32_t a[4], b[4];
    philox_block(k, (uint32_t)pid, (uint32_t)(pid >> 32), 0u, (uint32_t)seed,
                 (uint32_t)(seed >> 32), a);
    philox_block(k, (uint32_t)pid, (uint32_t)(pid >> 32), 1u, (uint32_t)seed,
                 (uint32_t)(seed >> 32), b);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        u[i] = u24(a[i]);
        u[4 + i] = u24(b[i]);
    }
}

// ---------------------------------------------------------------- grid
struct DevGrid {
    int nr, nt, np, mirror, ncell;
    const float *rw;    // nr+1 radial walls (au)
    const float *rw2;   // their squares
    const double *rwd;  // FP64 walls
    const float *tw;    // nt+1 theta walls
    const float *ct;    // cos of theta walls (ct[nt] = 0 exactly in mirror mode)
    const float *st;    // sin
    const double *ctd;  // FP64 cos/sin for the near-tangent re-evaluation
    const double *std_;
    const float *pw;    // np+1 phi walls
    const float *cp;
    const float *sp;
};

enum { W_NONE = 0, W_ROUT, W_RIN, W_TLO, W_THI, W_PLO, W_PHI, W_HOLE };

struct Pos {
    float x, y, z, dx, dy, dz;
    int ir, it, ip, hemi;  // ir = -1: inner hole; ir = nr: outside
};

__device__ __forceinline__ int cell_index(const DevGrid &g, const Pos &p) {
    return p.ir + g.nr * (p.it + g.nt * p.ip);
}

__device__ __forceinline__ int search_walls(const float *w, int n, float v) {
    // i in [0, n-1] with w[i] <= v < w[i+1], clamped
    if (v <= w[0]) return 0;
    if (v >= w[n]) return n - 1;
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (w[mid] <= v)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

__device__ __forceinline__ void locate_angles(const DevGrid &g, Pos &p) {
    float rho = sqrtf(p.x * p.x + p.y * p.y);
    if (g.mirror) {
        p.hemi = (p.z >= 0.f) ? 1 : -1;
        p.it = search_walls(g.tw, g.nt, atan2f(rho, fabsf(p.z)));
    } else {
        p.hemi = 1;
        p.it = search_walls(g.tw, g.nt, atan2f(rho, p.z));
    }
    if (g.np > 1) {
        float ph = atan2f(p.y, p.x);
        if (ph < 0.f) ph += 6.28318530718f;
        p.ip = search_walls(g.pw, g.np, ph);
    } else {
        p.ip = 0;
    }
}

// Cell of the position.  Every index is written once at the end (separate
// early returns writing different fields would keep Pos out of registers).
__device__ __forceinline__ void locate(const DevGrid &g, Pos &p) {
    float rr = p.x * p.x + p.y * p.y + p.z * p.z;
    const bool hole = rr < g.rw2[0], outside = !hole && rr >= g.rw2[g.nr];
    int ir;
    Pos q = p;
    q.it = 0;
    q.ip = 0;
    q.hemi = (g.mirror && p.z < 0.f) ? -1 : 1;
    if (hole) {
        ir = -1;
    } else if (outside) {
        ir = g.nr;
    } else {
        ir = search_walls(g.rw, g.nr, sqrtf(rr));
        locate_angles(g, q);
    }
    p.ir = ir;
    p.it = q.it;
    p.ip = q.ip;
    p.hemi = q.hemi;
}

__device__ __forceinline__ float sph_out(float b, float rr, float R2) {
    float c = rr - R2;
    float D = fmaxf(fmaf(b, b, -c), 0.f);
    float sq = sqrtf(D);
    float t = (b <= 0.f) ? (sq - b) : (-c / (b + sq));
    return fmaxf(t, 0.f);
}

// inner sphere; near-tangent cases (D small against b^2) are re-evaluated in
// FP64 from the FP32 position so the hit/miss decision is consistent.  A
// discriminant below -1e-5 (r^2 + R^2), twenty times its FP32 rounding bound,
// is a miss without that check (most rays moving inwards miss the inner
// sphere; one lane re-evaluating holds its whole wave).
__device__ __forceinline__ float sph_in(const Pos &p, float b, float rr, float R2, double Rd) {
    if (b >= 0.f) return RT_INF;
    float c = rr - R2;
    float D = fmaf(b, b, -c);
    if (D < -1e-5f * (rr + R2)) return RT_INF;
    if (D < 1e-4f * b * b) {
        double bd = (double)p.x * p.dx + (double)p.y * p.dy + (double)p.z * p.dz;
        double rrd = (double)p.x * p.x + (double)p.y * p.y + (double)p.z * p.z;
        double cd = rrd - Rd * Rd;
        double Dd = bd * bd - cd;
        if (Dd < 0.0) return RT_INF;
        double t = cd / (sqrt(Dd) - bd);
        return t > 0.0 ? (float)t : 0.f;
    }
    float t = c / (sqrtf(D) - b);
    return fmaxf(t, 0.f);
}

// Exit through the cone at polar angle th (folded frame): f = rho^2 cos^2 -
// z^2 sin^2; root with f' > 0 is (-B+sqrt(D))/A, with f' < 0 is (-B-sqrt(D))/A;
// on the nappe sign(z) = sign(cos th) the sign of dtheta/dt is
// sign(f') * sign(cos th).  req = -1 leaves towards smaller theta, +1 larger.
// A ray whose cell index is consistent with its position always meets its
// exit root ahead of it.  When a radial and a theta crossing lie within FP32
// resolution of each other, the ray can take the radial one first and stand a
// hair past the cone while the index still names the cell before it: the exit
// root is then just behind the position (the ray moving through the wall in
// the req direction), or, near tangency, missing while f already has the far
// side's sign.  The wall is then crossed at once (t = 0); rejecting it would
// leave the theta index behind for the rest of the ray.
__device__ __forceinline__ float cone_root(double A, double B, double C, double D,
                                           int want_plus) {
    double sq = sqrt(D);
    double t;
    if (want_plus)
        t = (B < 0) ? (-B + sq) / A : -C / (B + sq);
    else
        t = (B > 0) ? -(B + sq) / A : C / (-B + sq);
    return (float)t;
}

__device__ __forceinline__ float cone_hit(float x, float y, float z, float dx, float dy,
                                          float dz, float rr, float cth, float sth, double cthd,
                                          double sthd, int req) {
    float c2 = cth * cth, s2 = sth * sth;
    float A = (dx * dx + dy * dy) * c2 - dz * dz * s2;
    float B = (x * dx + y * dy) * c2 - z * dz * s2;
    float C = (x * x + y * y) * c2 - z * z * s2;
    // a ray lying in the cone (direction and position on it, as a radial ray
    // through the origin along a wall angle): it never leaves through it
    if (fabsf(A) <= 1e-6f && fabsf(C) <= 1e-6f * (x * x + y * y + z * z)) return RT_INF;
    float D = B * B - A * C;
    const float sreq = (float)(cth > 0.f ? req : -req);
    int want_plus = sreq > 0.f;
    float t;
    if (D < -1e-5f * rr) {
        // no crossing, clear of rounding (|A| <= 1, |B| <= r, |C| <= r^2, so D
        // is good to a few ulp of r^2): as the FP64 test below would find
        if (!(C * sreq > 0.f)) return RT_INF;
        t = 0.f;
    } else if (D < 1e-4f * B * B) {
        double c2d = cthd * cthd, s2d = sthd * sthd;
        double Ad = ((double)dx * dx + (double)dy * dy) * c2d - (double)dz * dz * s2d;
        double Bd = ((double)x * dx + (double)y * dy) * c2d - (double)z * dz * s2d;
        double Cd = ((double)x * x + (double)y * y) * c2d - (double)z * z * s2d;
        double Dd = Bd * Bd - Ad * Cd;
        if (Dd < 0.0) {
            // no crossing: f keeps the sign of C along the whole line
            if (!(Cd * sreq > 0.0)) return RT_INF;
            t = 0.f;
        } else {
            t = cone_root(Ad, Bd, Cd, Dd, want_plus);
        }
    } else {
        float sq = sqrtf(D);
        if (want_plus)
            t = (B < 0.f) ? (-B + sq) / A : -C / (B + sq);
        else
            t = (B > 0.f) ? -(B + sq) / A : C / (-B + sq);
    }
    if (!(t >= 0.f)) {
        if (!(B * sreq > 0.f)) return RT_INF;
        t = 0.f;
    }
    if (!(t < RT_INF)) return RT_INF;
    float zh = fmaf(t, dz, z);
    if (zh * cth < 0.f) return RT_INF;
    return t;
}

// Exit through the phi half-plane (cos ph, sin ph); a crossing just behind
// the position (the ray already past the wall, moving through it) is taken at
// once, as for the cones.
__device__ __forceinline__ float phi_hit(float x, float y, float dx, float dy, float cph,
                                         float sph, int req) {
    float g0 = -x * sph + y * cph;
    float g1 = -dx * sph + dy * cph;
    if (req < 0 ? !(g1 < 0.f) : !(g1 > 0.f)) return RT_INF;
    float t = fmaxf(-g0 / g1, 0.f);
    float xh = fmaf(t, dx, x), yh = fmaf(t, dy, y);
    if (xh * cph + yh * sph < 0.f) return RT_INF;
    return t;
}

// distance to the next wall of the current cell and which wall it is
__device__ __forceinline__ float next_wall(const DevGrid &g, const Pos &p, int &kind) {
    float rr = p.x * p.x + p.y * p.y + p.z * p.z;
    float b = p.x * p.dx + p.y * p.dy + p.z * p.dz;
    if (p.ir < 0) {
        kind = W_HOLE;
        return sph_out(b, rr, g.rw2[0]);
    }
    float tb = sph_out(b, rr, g.rw2[p.ir + 1]);
    kind = W_ROUT;
    float t = sph_in(p, b, rr, g.rw2[p.ir], g.rwd[p.ir]);
    if (t < tb) {
        tb = t;
        kind = W_RIN;
    }
    float zf = p.hemi > 0 ? p.z : -p.z, dzf = p.hemi > 0 ? p.dz : -p.dz;
    int j = p.it;
    if (j > 0) {
        if (!g.mirror && g.ct[j] == 0.f) {
            // the equatorial plane, left upwards (a crossing behind the
            // position is taken at once)
            if (p.dz > 0.f) {
                t = fmaxf(-p.z / p.dz, 0.f);
                if (t < tb) {
                    tb = t;
                    kind = W_TLO;
                }
            }
        } else {
            t = cone_hit(p.x, p.y, zf, p.dx, p.dy, dzf, rr, g.ct[j], g.st[j], g.ctd[j], g.std_[j],
                         -1);
            if (t < tb) {
                tb = t;
                kind = W_TLO;
            }
        }
    }
    if (g.mirror && j + 1 == g.nt) {
        if (dzf < 0.f) {
            t = fmaxf(-zf / dzf, 0.f);
            if (t < tb) {
                tb = t;
                kind = W_THI;
            }
        }
    } else if (j + 1 < g.nt) {
        if (!g.mirror && g.ct[j + 1] == 0.f) {
            if (p.dz < 0.f) {
                t = fmaxf(-p.z / p.dz, 0.f);
                if (t < tb) {
                    tb = t;
                    kind = W_THI;
                }
            }
        } else {
            t = cone_hit(p.x, p.y, zf, p.dx, p.dy, dzf, rr, g.ct[j + 1], g.st[j + 1],
                         g.ctd[j + 1], g.std_[j + 1], +1);
            if (t < tb) {
                tb = t;
                kind = W_THI;
            }
        }
    }
    if (g.np > 1) {
        int k = p.ip;
        t = phi_hit(p.x, p.y, p.dx, p.dy, g.cp[k], g.sp[k], -1);
        if (t < tb) {
            tb = t;
            kind = W_PLO;
        }
        t = phi_hit(p.x, p.y, p.dx, p.dy, g.cp[k + 1], g.sp[k + 1], +1);
        if (t < tb) {
            tb = t;
            kind = W_PHI;
        }
    }
    return tb;
}

__device__ __forceinline__ void snap_radius(Pos &p, float R) {
    float r = sqrtf(p.x * p.x + p.y * p.y + p.z * p.z);
    if (r > 0.f) {
        float f = R / r;
        p.x *= f;
        p.y *= f;
        p.z *= f;
    }
}

// move by t, cross the wall; returns true when the packet leaves the grid.
// The cell indices change by +-1 through selects (no branch writes through a
// pointer into the packet state, which would keep it out of registers).
__device__ __forceinline__ bool cross_wall(const DevGrid &g, Pos &p, float t, int kind) {
    p.x = fmaf(t, p.dx, p.x);
    p.y = fmaf(t, p.dy, p.y);
    p.z = fmaf(t, p.dz, p.z);
    if (kind == W_HOLE) {
        snap_radius(p, g.rw[0]);
        p.ir = 0;
        locate_angles(g, p);
        return false;
    }
    const bool mirror_hit = kind == W_THI && g.mirror && p.it + 1 == g.nt;
    const int dir_r = (kind == W_ROUT ? 1 : 0) - (kind == W_RIN ? 1 : 0);
    const int dir_t = (kind == W_THI && !mirror_hit ? 1 : 0) - (kind == W_TLO ? 1 : 0);
    const int dir_p = (kind == W_PHI ? 1 : 0) - (kind == W_PLO ? 1 : 0);
    p.ir += dir_r;
    if (p.ir >= g.nr) return true;
    // the wall crossed: inner wall of the new cell going out, outer going in
    if (dir_r != 0) snap_radius(p, g.rw[dir_r > 0 ? p.ir : p.ir + 1]);
    p.it += dir_t;
    if (mirror_hit) {
        p.z = 0.f;
        p.hemi = -p.hemi;
    }
    int ip = p.ip + dir_p;
    p.ip = ip < 0 ? g.np - 1 : (ip >= g.np ? 0 : ip);
    return false;
}

// Ping-pong guard.  A ray almost tangent to a sphere or cone can find the wall
// it has just crossed again at a distance below FP32 resolution and bounce
// between two cells without moving.  After 8 consecutive crossings shorter
// than 2^-20 r (r: outer radius of the cell the step started in) the position
// is advanced by 2^-16 r along the ray and its cell is found again from the
// position.  Returns true when it moved the packet.
__device__ __forceinline__ bool unstick(const DevGrid &g, Pos &p, float tw, float rcell,
                                        uint32_t &nzero) {
    if (tw > rcell * 9.5367431640625e-07f) {
        nzero = 0;
        return false;
    }
    if (++nzero < 8u) return false;
    nzero = 0;
    float d = rcell * 1.52587890625e-05f;
    p.x = fmaf(d, p.dx, p.x);
    p.y = fmaf(d, p.dy, p.y);
    p.z = fmaf(d, p.dz, p.z);
    locate(g, p);
    return true;
}

// ---------------------------------------------------------------- physics
// pdspy CGS constants (pdspy/constants/physics.py)
#define RT_H 6.6260755e-27
#define RT_C 2.99792458e10
#define RT_K 1.380658e-16

// isotropic direction: cos(theta) = 2 u1 - 1, azimuth 2 pi u2
__device__ __forceinline__ void iso_dir(float u1, float u2, float &dx, float &dy, float &dz) {
    float mu = 2.f * u1 - 1.f;
    float s, c;
    sincospif(2.f * u2, &s, &c);
    float st = sqrtf(fmaxf(0.f, 1.f - mu * mu));
    dx = st * c;
    dy = st * s;
    dz = mu;
}

// rotate the unit vector (dx,dy,dz) by polar angle (cosine ct) and azimuth
// 2 pi u2 about itself
__device__ __forceinline__ void rotate_dir(float ct, float u2, float &dx, float &dy, float &dz) {
    float sps, cps;
    sincospif(2.f * u2, &sps, &cps);
    float st = sqrtf(fmaxf(0.f, 1.f - ct * ct));
    float ux = dx, uy = dy, uz = dz;
    if (fabsf(uz) > 0.99999f) {
        dx = st * cps;
        dy = st * sps;
        dz = (uz > 0.f ? 1.f : -1.f) * ct;
    } else {
        float s = sqrtf(1.f - uz * uz);
        dx = st * (ux * uz * cps - uy * sps) / s + ux * ct;
        dy = st * (uy * uz * cps + ux * sps) / s + uy * ct;
        dz = -st * cps * s + uz * ct;
    }
    float n = rsqrtf(dx * dx + dy * dy + dz * dz);
    dx *= n;
    dy *= n;
    dz *= n;
}

// Henyey-Greenstein scattering about the current direction: cosine from u1,
// azimuth 2 pi u2
__device__ __forceinline__ void hg_dir(float u1, float u2, float g, float &dx, float &dy,
                                       float &dz) {
    float u = u1;
    float ct;
    if (fabsf(g) < 1e-6f) {
        ct = 2.f * u - 1.f;
    } else {
        float f = (1.f - g * g) / (1.f - g + 2.f * g * u);
        ct = (1.f + g * g - f * f) / (2.f * g);
    }
    ct = fminf(1.f, fmaxf(-1.f, ct));
    rotate_dir(ct, u2, dx, dy, dz);
}

// perpendicular distance to the nearest wall of the current cell that bounds
// the medium; the equatorial mirror and the polar axis are not walls.  Cone at
// theta_j: r sin|theta - theta_j| = |rho cos theta_j - z sin theta_j| while
// |theta - theta_j| <= pi/2 (else the axis, r); phi half-plane at phi_k:
// |y cos phi_k - x sin phi_k| while |phi - phi_k| <= pi/2 (else rho).
__device__ __forceinline__ float wall_dist(const DevGrid &g, const Pos &p) {
    float r = sqrtf(p.x * p.x + p.y * p.y + p.z * p.z);
    float d = fminf(r - g.rw[p.ir], g.rw[p.ir + 1] - r);
    float rho = sqrtf(p.x * p.x + p.y * p.y);
    float zf = g.mirror ? fabsf(p.z) : p.z;
    int j = p.it;
    if (j > 0) {
        float sn = rho * g.ct[j] - zf * g.st[j];
        float cs = zf * g.ct[j] + rho * g.st[j];
        d = fminf(d, cs >= 0.f ? fabsf(sn) : r);
    }
    if (j + 1 < g.nt) {
        float sn = rho * g.ct[j + 1] - zf * g.st[j + 1];
        float cs = zf * g.ct[j + 1] + rho * g.st[j + 1];
        d = fminf(d, cs >= 0.f ? fabsf(sn) : r);
    }
    if (g.np > 1) {
        int k = p.ip;
        float sn = p.y * g.cp[k] - p.x * g.sp[k];
        float cs = p.x * g.cp[k] + p.y * g.sp[k];
        d = fminf(d, cs >= 0.f ? fabsf(sn) : rho);
        sn = p.y * g.cp[k + 1] - p.x * g.sp[k + 1];
        cs = p.x * g.cp[k + 1] + p.y * g.sp[k + 1];
        d = fminf(d, cs >= 0.f ? fabsf(sn) : rho);
    }
    return d;
}

// table value at the fractional temperature-row coordinate x
__device__ __forceinline__ float trow(const float *tab, int ntemp, float x) {
    x = fminf(fmaxf(x, 0.f), (float)(ntemp - 1));
    int lo = min((int)x, ntemp - 2);
    float w = x - (float)lo;
    return fmaf(w, tab[lo + 1] - tab[lo], tab[lo]);
}

// diffusion escape law: y(xi) with P(y) = xi, tabulated on a uniform grid of
// logit(xi) over [-lmax, lmax] (24-bit uniforms stay inside it)
__device__ __forceinline__ float mrw_escape_y(const float *Y, int ny, float lmax, float xi) {
    float L = __logf(xi) - __logf(1.f - xi);
    float f = (L + lmax) * ((float)(ny - 1) / (2.f * lmax));
    f = fminf(fmaxf(f, 0.f), (float)(ny - 1));
    int k = min((int)f, ny - 2);
    float w = f - (float)k;
    return fmaf(w, Y[k + 1] - Y[k], Y[k]);
}

// a[i] for i < 4 without dynamic register indexing
__device__ __forceinline__ float pick4(const float (&a)[4], int i) {
    return i == 0 ? a[0] : i == 1 ? a[1] : i == 2 ? a[2] : a[3];
}

// B_nu(T) in FP32 from h nu/k (pdspy/misc/B_nu.py:4-6); pre = 2 h nu^3 / c^2
__device__ __forceinline__ float planck_f(float pre, float hnu_k, float T) {
    if (!(T > 0.f)) return 0.f;
    float x = hnu_k / T;
    if (x > 80.f) return pre * __expf(-x);
    return pre / expm1f(x);
}

// dB_nu/dT (pdspy/misc/dB_nu.py:4-7 up to sign) = pre * x e^x / (T (e^x-1)^2)
__device__ __forceinline__ float dplanck_f(float pre, float hnu_k, float T) {
    if (!(T > 0.f)) return 0.f;
    float x = hnu_k / T;
    if (x > 80.f) return pre * x * __expf(-x) / T;
    float e = expm1f(x);
    return pre * x * (e + 1.f) / (T * e * e);
}

// ---------------------------------------------------------------- waves
__device__ __forceinline__ int lane_id() { return __lane_id(); }

}  // namespace rt
