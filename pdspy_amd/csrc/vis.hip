// vis.hip -- model visibilities and chi^2 on gfx950.
//
// V(u,v) = sum_jk I[j][k] exp(-2 pi i (u x_k + v y_j)) with x_k = (k - nx/2) dxy,
// y_j = (j - ny/2) dxy, times the phase-centre shift exp(-2 pi i (u dra + v ddec)).
// On the pixel-conjugate grid this is pdspy's imtovis (imaging/imtovis.py:13-24,
// fftshift(fft2(ifftshift(I)))); at arbitrary (u,v) it is the exact sampler
// behind interpolate_model (interferometry/interpolate_model.py:20-27), and the
// gridded sampler further down samples it as galario does.
// chi^2 = sum w ((Re_d - Re_m)^2 + (Im_d - Im_m)^2) (libinterferometry.pyx:619-633).
//
// Forward: one (u,v) point per lane in each of four row groups of a
// workgroup; the image is staged through LDS in column strips and read as
// wave-wide broadcasts; the strip's column phasors sit in registers and the row
// phasor is one sincospi every 16 rows and a rotation by one row step in
// between, so the inner loop is two FMAs per pixel.  Adjoint: 16x16 tiles, four groups of lanes per tile taking alternate
// chunks of 64 points; per chunk the tile's column phasors (times the point's
// complex cotangent) and row phasors are built once in LDS, and the tile's
// update G += B^T Ha (16 x 16, 4 points deep) runs on v_mfma_f32_16x16x4f32.
#include <cmath>
#include <mutex>

#include "rt_internal.h"

#define VIS_BS 256
#define VIS_STRIP 32
#define VIS_MAXROWS 1024
#define VIS_G 4  // row groups per workgroup of the forward DFT

// NG groups of VIS_BS lanes per workgroup share the points and take a
// contiguous range of rows each; their sums are added in group order
template <int NG>
__global__ void __launch_bounds__(VIS_BS * NG) k_vis(const float *img, int nx, int ny,
                                                     const double *u, const double *v, int K,
                                                     double dxy, double dra, double ddec,
                                                     float *re, float *im) {
    constexpr int ng = NG;
    __shared__ float strip[VIS_MAXROWS * VIS_STRIP / 4];  // rows x 8 columns per pass (32 KB)
    const int g = NG == 1 ? 0 : (int)threadIdx.x / VIS_BS;
    const int lt = (int)threadIdx.x - g * VIS_BS;
    const int q = blockIdx.x * VIS_BS + lt;
    const bool act = q < K;
    const int rpg = NG == 1 ? ny : (ny + ng - 1) / ng;
    const int jb = NG == 1 ? 0 : min(ny, g * rpg), je = NG == 1 ? ny : min(ny, jb + rpg);
    float ud = 0.f, vd = 0.f;
    if (act) {
        ud = (float)(u[q] * dxy);
        vd = (float)(v[q] * dxy);
    }
    float sr = 0.f, si = 0.f;
    float ds, dc;  // one row step of the row phasor
    sincospif(-2.f * vd, &ds, &dc);
    // process columns in strips of 8 so that a strip of up to 1024 rows fits LDS
    const int W = 8;
    for (int k0 = 0; k0 < nx; k0 += W) {
        int wk = min(W, nx - k0);
        __syncthreads();
        for (int i = threadIdx.x; i < ny * W; i += VIS_BS * ng) {
            int j = i / W, kk = i - j * W;
            strip[i] = (kk < wk) ? img[(size_t)j * nx + k0 + kk] : 0.f;
        }
        __syncthreads();
        float ar[W], ai[W];
#pragma unroll
        for (int kk = 0; kk < W; ++kk) {
            float s, c;
            sincospif(-2.f * ud * (float)(k0 + kk - nx / 2), &s, &c);
            ar[kk] = c;
            ai[kk] = s;
        }
        // row phasors: exact every 16 rows from the group's first row, rotated
        // by one row step in between (FP32 drift below 2e-6 over 15 rotations)
        float bs = 0.f, bc = 1.f;
        for (int j = jb; j < je; ++j) {
            const float *row = strip + j * W;
            float rr = 0.f, ri = 0.f;
#pragma unroll
            for (int kk = 0; kk < W; ++kk) {
                rr = fmaf(row[kk], ar[kk], rr);
                ri = fmaf(row[kk], ai[kk], ri);
            }
            if (((j - jb) & 15) == 0) {
                sincospif(-2.f * vd * (float)(j - ny / 2), &bs, &bc);
            } else {
                const float nc = bc * dc - bs * ds;
                bs = fmaf(bc, ds, bs * dc);
                bc = nc;
            }
            sr += rr * bc - ri * bs;
            si += rr * bs + ri * bc;
        }
    }
    // the groups' sums through the strip buffer (free after the last strip)
    if (ng > 1) __syncthreads();
    float *red = strip;  // [2][VIS_G - 1][VIS_BS]
    if (g > 0) {
        red[(g - 1) * VIS_BS + lt] = sr;
        red[((VIS_G - 1) + g - 1) * VIS_BS + lt] = si;
    }
    if (ng > 1) __syncthreads();
    if (g == 0 && act) {
        for (int h = 0; h < ng - 1; ++h) {
            sr += red[h * VIS_BS + lt];
            si += red[((VIS_G - 1) + h) * VIS_BS + lt];
        }
        double ph = -2.0 * (u[q] * dra + v[q] * ddec);
        double s, c;
        sincospi(ph, &s, &c);
        re[q] = (float)(sr * c - si * s);
        im[q] = (float)(sr * s + si * c);
    }
}

#define ADJ_T 16
#define ADJ_CH 64
#define ADJ_G 4   // point groups per tile: each takes every ADJ_G-th chunk of points
#define ADJ_BS (ADJ_T * ADJ_T * ADJ_G)

typedef float vf4 __attribute__((ext_vector_type(4)));

// g[j][k] = Re sum_q H_q a_qk b_qj,  H_q = (gre_q - i gim_q) shift_q,
// a_qk = exp(-2 pi i u_q x_k), b_qj = exp(-2 pi i v_q y_j)
// One workgroup per 16x16 tile, ADJ_G groups of 256 lanes; group g takes the
// chunks g, g + ADJ_G, ... of the points, builds the chunk's phasor tables in
// LDS, and each of its four waves adds 16 of the chunk's points to its own copy
// of the tile on the matrix cores.  The 16 copies are added in wave order, so
// the result does not depend on scheduling.
__global__ void __launch_bounds__(ADJ_BS) k_vis_adj(const float *gre, const float *gim, int nx,
                                                    int ny, const double *u, const double *v,
                                                    int K, double dxy, double dra, double ddec,
                                                    float *gimg) {
    __shared__ float Ha_r[ADJ_G][ADJ_CH][ADJ_T], Ha_i[ADJ_G][ADJ_CH][ADJ_T];
    __shared__ float b_r[ADJ_G][ADJ_CH][ADJ_T], b_i[ADJ_G][ADJ_CH][ADJ_T];
    __shared__ float H_r[ADJ_G][ADJ_CH], H_i[ADJ_G][ADJ_CH];
    __shared__ float sa_c[ADJ_G][ADJ_CH], sa_s[ADJ_G][ADJ_CH], sb_c[ADJ_G][ADJ_CH],
        sb_s[ADJ_G][ADJ_CH];
    __shared__ float s_u[ADJ_G][ADJ_CH], s_v[ADJ_G][ADJ_CH];  // u dxy, v dxy
    const int g = threadIdx.x / (ADJ_T * ADJ_T), lt = threadIdx.x - g * (ADJ_T * ADJ_T);
    const int tx = lt % ADJ_T, ty = lt / ADJ_T;
    const int k0 = blockIdx.x * ADJ_T, j0 = blockIdx.y * ADJ_T;
    const int k = k0 + tx, j = j0 + ty;
    // matrix-core layout: wave wv of the group, lane = mq * 16 + mc
    const int wv = lt >> 6, mq = (lt >> 4) & 3, mc = lt & 15;
    vf4 acc4 = {0.f, 0.f, 0.f, 0.f};
    // loader: 4 threads per point, 4 consecutive columns / rows each
    const int lq = lt >> 2, ll = (lt & 3) * 4;
    const int nch = (K + ADJ_CH - 1) / ADJ_CH;
    // the setup lanes' point data, loaded one round ahead so that the loads
    // are in flight while the round before runs (the barrier waits for none)
    double pu = 0.0, pv = 0.0;
    float pgr = 0.f, pgi = 0.f;
    if (lt < ADJ_CH && g * ADJ_CH + lt < K) {
        const int q = g * ADJ_CH + lt;
        pu = u[q];
        pv = v[q];
        pgr = gre[q];
        pgi = gim[q];
    }
    for (int c0 = 0; c0 < nch; c0 += ADJ_G) {
        const int q0 = (c0 + g) * ADJ_CH;  // past K for some groups in the last round
        const double cu = pu, cv = pv;
        const float gr = pgr, gi = pgi;
        if (lt < ADJ_CH && q0 + ADJ_G * ADJ_CH + lt < K) {
            const int qn = q0 + ADJ_G * ADJ_CH + lt;
            pu = u[qn];
            pv = v[qn];
            pgr = gre[qn];
            pgi = gim[qn];
        }
        __syncthreads();
        // per point: the cotangent with the phase-centre shift and the
        // one-pixel steps of its column and row phasors
        if (lt < ADJ_CH) {
            const int q = q0 + lt;
            float hr = 0.f, hi = 0.f, s1 = 0.f, c1 = 1.f, s2 = 0.f, c2 = 1.f, su = 0.f, sv = 0.f;
            if (q < K) {
                double s, c;
                sincospi(-2.0 * (cu * dra + cv * ddec), &s, &c);
                // H = (gre - i gim) * (c + i s)
                hr = (float)(gr * c + gi * s);
                hi = (float)(gr * s - gi * c);
                su = (float)(cu * dxy);
                sv = (float)(cv * dxy);
                sincospif(-2.f * su, &s1, &c1);
                sincospif(-2.f * sv, &s2, &c2);
            }
            H_r[g][lt] = hr;
            H_i[g][lt] = hi;
            sa_c[g][lt] = c1;
            sa_s[g][lt] = s1;
            sb_c[g][lt] = c2;
            sb_s[g][lt] = s2;
            s_u[g][lt] = su;
            s_v[g][lt] = sv;
        }
        __syncthreads();
        {
            const int q = q0 + lq;
            float hr = 0.f, hi = 0.f, ar = 0.f, ai = 0.f, br = 0.f, bi = 0.f;
            float dac = 1.f, das = 0.f, dbc = 1.f, dbs = 0.f;
            if (q < K) {
                hr = H_r[g][lq];
                hi = H_i[g][lq];
                sincospif(-2.f * s_u[g][lq] * (float)(k0 + ll - nx / 2), &ai, &ar);
                sincospif(-2.f * s_v[g][lq] * (float)(j0 + ll - ny / 2), &bi, &br);
                dac = sa_c[g][lq];
                das = sa_s[g][lq];
                dbc = sb_c[g][lq];
                dbs = sb_s[g][lq];
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                Ha_r[g][lq][ll + e] = hr * ar - hi * ai;
                Ha_i[g][lq][ll + e] = hr * ai + hi * ar;
                b_r[g][lq][ll + e] = br;
                b_i[g][lq][ll + e] = bi;
                // next column / row: rotate by one pixel step
                const float nar = ar * dac - ai * das;
                ai = fmaf(ar, das, ai * dac);
                ar = nar;
                const float nbr = br * dbc - bi * dbs;
                bi = fmaf(br, dbs, bi * dbc);
                br = nbr;
            }
        }
        __syncthreads();
        // the tile's update by the wave's 16 points on the matrix cores:
        // G[j][k] += sum_q b_r[q][j] Ha_r[q][k] - b_i[q][j] Ha_i[q][k], four points
        // per v_mfma_f32_16x16x4f32 (A[j][q] from lane q*16 + j, B[q][k] from
        // lane q*16 + k)
#pragma unroll
        for (int st = 0; st < ADJ_CH / 16; ++st) {
            const int qq = wv * 16 + st * 4 + mq;
            acc4 = __builtin_amdgcn_mfma_f32_16x16x4f32(b_r[g][qq][mc], Ha_r[g][qq][mc], acc4, 0, 0, 0);
            acc4 = __builtin_amdgcn_mfma_f32_16x16x4f32(-b_i[g][qq][mc], Ha_i[g][qq][mc], acc4, 0, 0, 0);
        }
    }
    // the 16 waves' tiles (lane q*16 + k holds rows 4q..4q+3 of column k),
    // added in wave order; the phasor buffers are free now
    __syncthreads();
    float *red = &Ha_r[0][0][0];
    const int w16 = g * 4 + wv;
#pragma unroll
    for (int r = 0; r < 4; ++r) red[w16 * (ADJ_T * ADJ_T) + (4 * mq + r) * ADJ_T + mc] = acc4[r];
    __syncthreads();
    if (g == 0 && k < nx && j < ny) {
        float sum = red[lt];
        for (int h = 1; h < 4 * ADJ_G; ++h) sum += red[h * (ADJ_T * ADJ_T) + lt];
        gimg[(size_t)j * nx + k] = sum;
    }
}

// ---- gridded sampler: galario's method (Tazzari, Beaujean & Testi 2018,
// MNRAS 476, 4527: Fourier transform of the image on its pixel-conjugate grid,
// bilinear interpolation at each (u,v), then the phase-centre shift), which is
// what pdspy's interpolate_model calls (interpolate_model.py:20-27).
// F[b][a] = sum_jk I[j][k] w_nx((a - nx/2)(k - nx/2)) w_ny((b - ny/2)(j - ny/2)),
// w_n(r) = exp(-2 pi i r / n), is the exact DFT at u_a = (a - nx/2) / (nx dxy),
// v_b = (b - ny/2) / (ny dxy); it is computed as two passes of 1D transforms
// (rows, then columns) with integer phase indices into one twiddle table per
// axis, so every twiddle is exact to FP32 rounding.  Work:
// [A: ny x nx float2][F: ny x nx float2][twiddles: nx + ny float2].
#define GRID_BS 256
#define GRID_MAXN 4096

__global__ void __launch_bounds__(GRID_BS) k_grid_twiddles(int nx, int ny, float2 *tw) {
    const int i = blockIdx.x * GRID_BS + threadIdx.x;
    if (i >= nx + ny) return;
    const int n = i < nx ? nx : ny, r = i < nx ? i : i - nx;
    double s, c;
    sincospi(-2.0 * (double)r / (double)n, &s, &c);
    tw[i] = make_float2((float)c, (float)s);
}

// n-periodic phase index of k * x0 and of the step k, both in [0, n)
__device__ __forceinline__ void phase_index(int k, int x0, int n, int &r, int &dr) {
    long long a = ((long long)k * (long long)x0) % n;
    r = (int)(a < 0 ? a + n : a);
    const int d = k % n;
    dr = d < 0 ? d + n : d;
}

// rows: A[j][a] = sum_k I[j'][k] w_nx((a - nx/2)(k - nx/2)); block = 256 a's of
// row j, which reads image row j' = j, or j' = ny - 1 - j when flip is set (the
// transform of image[::-1])
__global__ void __launch_bounds__(GRID_BS) k_grid_rows(const float *img, const float2 *twx,
                                                       int nx, int ny, int flip, float2 *A) {
    __shared__ float row[GRID_MAXN];
    __shared__ float2 tw[GRID_MAXN];
    const int j = blockIdx.y;
    const int js = flip ? ny - 1 - j : j;
    for (int i = threadIdx.x; i < nx; i += GRID_BS) {
        row[i] = img[(size_t)js * nx + i];
        tw[i] = twx[i];
    }
    __syncthreads();
    const int a = blockIdx.x * GRID_BS + threadIdx.x;
    if (a >= nx) return;
    int r, dr;
    phase_index(a - nx / 2, -(nx / 2), nx, r, dr);
    float sr = 0.f, si = 0.f;
    for (int k = 0; k < nx; ++k) {
        const float p = row[k];
        const float2 w = tw[r];
        sr = fmaf(p, w.x, sr);
        si = fmaf(p, w.y, si);
        r += dr;
        if (r >= nx) r -= nx;
    }
    A[(size_t)j * nx + a] = make_float2(sr, si);
}

// columns: F[b][a] = sum_j A[j][a] w_ny((b - ny/2)(j - ny/2)); block = 256 a's x
// KB rows b, so each A element read serves KB outputs (KB = 1 for small grids,
// where the launch would otherwise hold fewer blocks than the GPU has CUs)
template <int GRID_KB>
__global__ void __launch_bounds__(GRID_BS) k_grid_cols(const float2 *A, const float2 *twy, int nx,
                                                       int ny, float2 *F) {
    __shared__ float2 tw[GRID_MAXN];
    for (int i = threadIdx.x; i < ny; i += GRID_BS) tw[i] = twy[i];
    __syncthreads();
    const int a = blockIdx.x * GRID_BS + threadIdx.x;
    const int b0 = blockIdx.y * GRID_KB;
    if (a >= nx) return;
    int r[GRID_KB], dr[GRID_KB];
    float fr[GRID_KB], fi[GRID_KB];
#pragma unroll
    for (int t = 0; t < GRID_KB; ++t) {
        phase_index(b0 + t - ny / 2, -(ny / 2), ny, r[t], dr[t]);
        fr[t] = 0.f;
        fi[t] = 0.f;
    }
    for (int j = 0; j < ny; ++j) {
        const float2 x = A[(size_t)j * nx + a];
#pragma unroll
        for (int t = 0; t < GRID_KB; ++t) {
            const float2 w = tw[r[t]];
            fr[t] = fmaf(x.x, w.x, fmaf(-x.y, w.y, fr[t]));
            fi[t] = fmaf(x.x, w.y, fmaf(x.y, w.x, fi[t]));
            r[t] += dr[t];
            if (r[t] >= ny) r[t] -= ny;
        }
    }
#pragma unroll
    for (int t = 0; t < GRID_KB; ++t)
        if (b0 + t < ny) F[(size_t)(b0 + t) * nx + a] = make_float2(fr[t], fi[t]);
}

// one (u,v) point per lane: bilinear interpolation between the four grid nodes
// around (usgn u, usgn v), then the phase-centre shift at (u, v); a point whose
// four nodes are not all on the grid gets NaN
__global__ void __launch_bounds__(GRID_BS) k_grid_sample(const float2 *F, int nx, int ny,
                                                         double dxy, const double *u,
                                                         const double *v, int K, double usgn,
                                                         double dra, double ddec, float *re,
                                                         float *im) {
    const int q = blockIdx.x * GRID_BS + threadIdx.x;
    if (q >= K) return;
    const double uq = u[q], vq = v[q];
    const double fa = usgn * uq * ((double)nx * dxy) + (double)(nx / 2);
    const double fb = usgn * vq * ((double)ny * dxy) + (double)(ny / 2);
    const double a0 = floor(fa), b0 = floor(fb);
    if (!(a0 >= 0.0 && a0 <= (double)(nx - 2) && b0 >= 0.0 && b0 <= (double)(ny - 2))) {
        re[q] = __builtin_nanf("");
        im[q] = __builtin_nanf("");
        return;
    }
    const size_t o = (size_t)b0 * nx + (size_t)a0;
    const float2 f00 = F[o], f01 = F[o + 1], f10 = F[o + nx], f11 = F[o + nx + 1];
    const double ta = fa - a0, tb = fb - b0;
    const double w00 = (1.0 - ta) * (1.0 - tb), w01 = ta * (1.0 - tb), w10 = (1.0 - ta) * tb,
                 w11 = ta * tb;
    const double sr = w00 * f00.x + w01 * f01.x + w10 * f10.x + w11 * f11.x;
    const double si = w00 * f00.y + w01 * f01.y + w10 * f10.y + w11 * f11.y;
    double s, c;
    sincospi(-2.0 * (uq * dra + vq * ddec), &s, &c);
    re[q] = (float)(sr * c - si * s);
    im[q] = (float)(sr * s + si * c);
}

size_t vis_grid_workspace(int nx, int ny) {
    return sizeof(float2) * (2 * (size_t)nx * ny + (size_t)nx + ny);
}

// Scratch of the allocation-free gridded sampler (rt_vis_grid_noalloc): the
// workspace of an image of up to VIS_NOALLOC_MAXN x VIS_NOALLOC_MAXN pixels
// (16.8 MB), one copy per GPU for the life of the process.  Host code never
// passes the array's name to a kernel: it is a host-side stand-in, and the
// device address of each GPU's copy is looked up once per device.
#define VIS_NOALLOC_MAXN 1024
static __device__ __align__(256) float2 scratch[2 * VIS_NOALLOC_MAXN * VIS_NOALLOC_MAXN +
                                                 2 * VIS_NOALLOC_MAXN];
#define VIS_MAX_DEVICES 64
static void *g_scratch_ptr[VIS_MAX_DEVICES];
static std::mutex g_scratch_mu;

// with `device` current: the device address of this GPU's scratch array
int vis_scratch_init(rt_ctx *ctx, int device) {
    if (device < 0 || device >= VIS_MAX_DEVICES)
        return set_err(ctx, RT_E_LIMIT, "device %d beyond the %d the library supports", device,
                       VIS_MAX_DEVICES);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return RT_OK;
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return set_err(ctx, RT_E_HIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
    g_scratch_ptr[device] = ptr;
    return RT_OK;
}

int vis_grid_noalloc_run(rt_ctx *ctx, const float *d_img, int nx, int ny, double dxy,
                         const double *d_u, const double *d_v, int K, double dra, double ddec,
                         float *d_re, float *d_im) {
    if (nx > VIS_NOALLOC_MAXN || ny > VIS_NOALLOC_MAXN)
        return set_err(ctx, RT_E_LIMIT, "rt_vis_grid_noalloc: %d x %d image exceeds its scratch "
                                        "(at most %d x %d pixels); use rt_vis_grid", nx, ny,
                       VIS_NOALLOC_MAXN, VIS_NOALLOC_MAXN);
    void *ptr = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (ctx->device >= 0 && ctx->device < VIS_MAX_DEVICES) ptr = g_scratch_ptr[ctx->device];
    }
    if (!ptr) return set_err(ctx, RT_E_HIP, "rt_vis_grid_noalloc: scratch of device %d not set up",
                             ctx->device);
    // every scratch word the kernels read is written first in the same call
    // (twiddles, row pass, column pass), so nothing needs clearing
    return vis_grid_run(ctx, d_img, nx, ny, dxy, d_u, d_v, K, dra, ddec, d_re, d_im, ptr, 0);
}

// conv 0: the imtovis convention above.  conv 1: pdspy's galario call
// (interpolate_model.py:23-27): galario's sampleImage of image[::-1] with the
// imaginary part negated.  galario samples G = transform of image[::-1] at
// (u, v) and multiplies by exp(+2 pi i (u dra + v ddec)); for a real image
// conj(G(u, v)) = G(-u, -v) holds node by node, hence also for the bilinear
// interpolant, so the negated imaginary part is G interpolated at (-u, -v)
// times exp(-2 pi i (u dra + v ddec)).
int vis_grid_run(rt_ctx *ctx, const float *d_img, int nx, int ny, double dxy, const double *d_u,
                 const double *d_v, int K, double dra, double ddec, float *d_re, float *d_im,
                 void *d_work, int conv) {
    if (nx < 2 || ny < 2 || nx > GRID_MAXN || ny > GRID_MAXN || K < 0 || !(dxy > 0))
        return set_err(ctx, RT_E_ARG, "vis grid: invalid sizes (2 <= nx, ny <= %d)", GRID_MAXN);
    if (K == 0) return RT_OK;
    if (!d_img || !d_u || !d_v || !d_re || !d_im)
        return set_err(ctx, RT_E_ARG, "vis grid: NULL image, (u,v) or output");
    float2 *ws = (float2 *)d_work;
    if (!ws) HIPCHK(ctx, hipMallocAsync((void **)&ws, vis_grid_workspace(nx, ny), ctx->stream));
    float2 *A = ws, *F = ws + (size_t)nx * ny, *tw = ws + 2 * (size_t)nx * ny;
    hipLaunchKernelGGL(k_grid_twiddles, dim3((nx + ny + GRID_BS - 1) / GRID_BS), dim3(GRID_BS), 0,
                       ctx->stream, nx, ny, tw);
    hipLaunchKernelGGL(k_grid_rows, dim3((nx + GRID_BS - 1) / GRID_BS, ny), dim3(GRID_BS), 0,
                       ctx->stream, d_img, tw, nx, ny, conv == 1 ? 1 : 0, A);
    const int bx = (nx + GRID_BS - 1) / GRID_BS;
    if (bx * ((ny + 7) / 8) >= 512)
        hipLaunchKernelGGL(k_grid_cols<8>, dim3(bx, (ny + 7) / 8), dim3(GRID_BS), 0, ctx->stream, A,
                           tw + nx, nx, ny, F);
    else if (bx * ((ny + 3) / 4) >= 512)
        hipLaunchKernelGGL(k_grid_cols<4>, dim3(bx, (ny + 3) / 4), dim3(GRID_BS), 0, ctx->stream, A,
                           tw + nx, nx, ny, F);
    else
        hipLaunchKernelGGL(k_grid_cols<1>, dim3(bx, ny), dim3(GRID_BS), 0, ctx->stream, A, tw + nx,
                           nx, ny, F);
    hipLaunchKernelGGL(k_grid_sample, dim3((K + GRID_BS - 1) / GRID_BS), dim3(GRID_BS), 0,
                       ctx->stream, F, nx, ny, dxy, d_u, d_v, K, conv == 1 ? -1.0 : 1.0, dra, ddec,
                       d_re, d_im);
    if (!d_work) HIPCHK(ctx, hipFreeAsync(ws, ctx->stream));
    HIPCHK(ctx, hipGetLastError());
    return RT_OK;
}

#define CHI_BS 256

__global__ void __launch_bounds__(CHI_BS) k_chisq(const float *dre, const float *dim, const float *w,
                                                  const float *mre, const float *mim, int K,
                                                  double *partial, float *gre, float *gim) {
    __shared__ double sh[CHI_BS];
    int q = blockIdx.x * CHI_BS + threadIdx.x;
    double x = 0.0;
    if (q < K) {
        double a = (double)dre[q] - mre[q], b = (double)dim[q] - mim[q];
        x = (a * a + b * b) * (double)w[q];
        if (gre) {
            gre[q] = (float)(-2.0 * w[q] * a);
            gim[q] = (float)(-2.0 * w[q] * b);
        }
    }
    sh[threadIdx.x] = x;
    __syncthreads();
    for (int s = CHI_BS / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

// fixed-order final sum (deterministic)
__global__ void __launch_bounds__(CHI_BS) k_sum(const double *partial, int n, double *out) {
    __shared__ double sh[CHI_BS];
    double x = 0.0;
    for (int i = threadIdx.x; i < n; i += CHI_BS) x += partial[i];
    sh[threadIdx.x] = x;
    __syncthreads();
    for (int s = CHI_BS / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0];
}

int vis_run(rt_ctx *ctx, const float *d_img, int nx, int ny, double dxy, const double *d_u,
            const double *d_v, int K, double dra, double ddec, float *d_re, float *d_im) {
    if (nx < 1 || ny < 1 || ny > VIS_MAXROWS || K < 0 || !(dxy > 0))
        return set_err(ctx, RT_E_ARG, "vis: invalid sizes (ny <= %d)", VIS_MAXROWS);
    if (K == 0) return RT_OK;
    if (!d_img || !d_u || !d_v || !d_re || !d_im)
        return set_err(ctx, RT_E_ARG, "vis: NULL image, (u,v) or output");
    // row groups only where the points alone leave the GPU short of waves
    const int nb = (K + VIS_BS - 1) / VIS_BS;
    if (nb < 2 * ctx->n_cu)
        hipLaunchKernelGGL(k_vis<VIS_G>, dim3(nb), dim3(VIS_BS * VIS_G), 0, ctx->stream, d_img, nx,
                           ny, d_u, d_v, K, dxy, dra, ddec, d_re, d_im);
    else
        hipLaunchKernelGGL(k_vis<1>, dim3(nb), dim3(VIS_BS), 0, ctx->stream, d_img, nx, ny, d_u,
                           d_v, K, dxy, dra, ddec, d_re, d_im);
    HIPCHK(ctx, hipGetLastError());
    return RT_OK;
}

int vis_adjoint_run(rt_ctx *ctx, const float *d_gre, const float *d_gim, int nx, int ny, double dxy,
                    const double *d_u, const double *d_v, int K, double dra, double ddec,
                    float *d_g_img) {
    if (nx < 1 || ny < 1 || K < 0 || !(dxy > 0)) return set_err(ctx, RT_E_ARG, "vis adjoint: invalid sizes");
    dim3 grid((nx + ADJ_T - 1) / ADJ_T, (ny + ADJ_T - 1) / ADJ_T);
    hipLaunchKernelGGL(k_vis_adj, grid, dim3(ADJ_BS), 0, ctx->stream, d_gre, d_gim, nx, ny,
                       d_u, d_v, K, dxy, dra, ddec, d_g_img);
    HIPCHK(ctx, hipGetLastError());
    return RT_OK;
}

size_t chisq_workspace(int K) { return (size_t)((K + CHI_BS - 1) / CHI_BS); }

int chisq_run(rt_ctx *ctx, const float *dre, const float *dim, const float *w, const float *mre,
              const float *mim, int K, double *out, float *gre, float *gim, double *ws) {
    if (K < 1 || !out) return set_err(ctx, RT_E_ARG, "chisq: invalid arguments");
    if ((gre == nullptr) != (gim == nullptr)) return set_err(ctx, RT_E_ARG, "chisq: gre/gim both or neither");
    int nb = (int)chisq_workspace(K);
    double *partial = ws;
    if (!partial) HIPCHK(ctx, hipMallocAsync((void **)&partial, sizeof(double) * nb, ctx->stream));
    hipLaunchKernelGGL(k_chisq, dim3(nb), dim3(CHI_BS), 0, ctx->stream, dre, dim, w, mre, mim, K,
                       partial, gre, gim);
    hipLaunchKernelGGL(k_sum, dim3(1), dim3(CHI_BS), 0, ctx->stream, partial, nb, out);
    if (!ws) HIPCHK(ctx, hipFreeAsync(partial, ctx->stream));
    HIPCHK(ctx, hipGetLastError());
    return RT_OK;
}
