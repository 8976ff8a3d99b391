// image.hip -- ray-traced images and their adjoint on gfx950: what RADMC-3D's
// camera computes for pdspy's Model.run_image (pdspy/modeling/Model.py:515-577,
// radmc3d/run.py:157-289), i.e. the formal solution
//     I = sum_k S_k (1 - exp(-dtau_k)) exp(-tau_k)
// along parallel rays, with S = B_nu(T) and absorption-only extinction
// (scattering_mode_max=0), or, given the scattering source's mean intensity J
// (scat.hip), extinction rho (kabs + ksca) and S = (rho kabs B + rho ksca J) /
// (rho kext) (isotropic scattering); plus blackbody point stars attenuated by
// the model (pdspy's nostar=False).  Pixel (j,k) is centred at
// ((k - nx/2) pix, (j - ny/2) pix) in the sky plane, the centring pdspy's
// zoomau produces (Model.py:522-527).
//
// Kernels: a per-cell prep pass (alpha and S per cell) so
// the ray loop reads 8 bytes per crossed cell; one ray per lane; the adjoint
// re-integrates each ray twice (first for I, then for dI/d dtau_k = e^-tau_k
// e^-dtau_k S_k - (I - I_<=k)) instead of storing tau, and accumulates
// dL/dalpha, dL/dS per cell with FP64 atomics; a post pass applies the chain
// rule to log(rho_s) and T.
#include <algorithm>
#include <cmath>

#include "rt_internal.h"

using namespace rt;

struct CamParams {
    DevGrid g;
    float n[3], ex[3], ey[3];
    float cpa, spa;
    float pix;       // au
    int nx, ny, nsub, mode;
    int tw, th, ntx;  // ray tile of a block: tw x th pixels, ntx tiles per image row
    unsigned hbits;   // adjoint LDS table: 2^hbits entries (>= 4 per ray of a block)
    float tau_stop;
    const float *alpha;  // [cell] per au
    const float *src;    // [cell]
    const float *csize;  // [cell] cell size for the sub-pixeling (au)
    unsigned long long *icount;  // cells crossed: [0] image rays, [1] adjoint sweeps
    const double *qs;  // adjoint: fixed-point units per unit of dL/dalpha, dL/dS
};

// adds the block's crossings n (summed over its threads) to *dst; every thread
// of the block calls it
__device__ __forceinline__ void count_crossings(unsigned long long *dst, uint32_t n) {
    __shared__ unsigned s_n;
    if (threadIdx.x == 0) s_n = 0u;
    __syncthreads();
    if (n) atomicAdd(&s_n, n);
    __syncthreads();
    if (threadIdx.x == 0 && s_n && dst) atomicAdd(dst, (unsigned long long)s_n);
}

// A camera ray is the line p0 + s d: p0 the point closest to the origin (in
// the sky plane through it), d = -n.  Marching the FP32 position from the
// outer sphere accumulates a transverse error of a few ulp of the outer radius
// (~1e-3 au at 5000 au); on a nearly radial ray (the central pixels) the theta
// crossings then move along the ray by that error times r/b, a sizeable part
// of a cell.  At every radial crossing the position is therefore set back on
// the exact line at s = -+sqrt(r_wall^2 - b^2) (FP64), which leaves only the
// error gathered since the last radial crossing, a few ulp of the local radius.
struct CamRay {
    float x0, y0, z0;  // p0 (au)
    double b2;         // |p0|^2
};

__device__ __forceinline__ bool ray_start(const CamParams &C, float X, float Y, Pos &p,
                                          CamRay &cr) {
    float xp = X * C.cpa + Y * C.spa;
    float yp = -X * C.spa + Y * C.cpa;
    cr.x0 = xp * C.ex[0] + yp * C.ey[0];
    cr.y0 = xp * C.ex[1] + yp * C.ey[1];
    cr.z0 = xp * C.ex[2] + yp * C.ey[2];
    cr.b2 = (double)cr.x0 * cr.x0 + (double)cr.y0 * cr.y0 + (double)cr.z0 * cr.z0;
    const double R = C.g.rwd[C.g.nr];
    if (cr.b2 >= R * R) return false;
    p.dx = -C.n[0];
    p.dy = -C.n[1];
    p.dz = -C.n[2];
    const float s = (float)-sqrt(R * R - cr.b2);
    p.x = fmaf(s, p.dx, cr.x0);
    p.y = fmaf(s, p.dy, cr.y0);
    p.z = fmaf(s, p.dz, cr.z0);
    p.ir = C.g.nr - 1;
    locate_angles(C.g, p);
    return true;
}

// cross_wall for a camera ray: radial crossings put the position on the exact
// line (see CamRay); theta, phi and mirror crossings as cross_wall
__device__ __forceinline__ bool cross_wall_cam(const DevGrid &g, Pos &p, float t, int kind,
                                               const CamRay &cr) {
    if (kind != W_RIN && kind != W_ROUT && kind != W_HOLE) return cross_wall(g, p, t, kind);
    const bool out = kind != W_RIN;
    const int ir = kind == W_HOLE ? 0 : p.ir + (out ? 1 : -1);
    p.ir = ir;
    if (ir >= g.nr) return true;
    const double rw = g.rwd[out ? ir : ir + 1];
    float s = (float)sqrt(fmax(rw * rw - cr.b2, 0.0));
    if (!out) s = -s;
    p.x = fmaf(s, p.dx, cr.x0);
    p.y = fmaf(s, p.dy, cr.y0);
    p.z = fmaf(s, p.dz, cr.z0);
    if (kind == W_HOLE) locate_angles(g, p);
    return false;
}

// formal solution along the ray from p; with LMIN also the smallest cell size
// on the part of the ray traced (RT_INF when it crosses no cell)
template <bool LMIN>
__device__ __forceinline__ float integrate_ray_l(const CamParams &C, Pos p, const CamRay &cr,
                                                 float &lmin, uint32_t &nseg) {
    float I = 0.f, tau = 0.f;
    uint32_t nzero = 0;
    for (int guard = 0; guard < (1 << 24); ++guard) {
        int kind;
        float tw = next_wall(C.g, p, kind);
        const float rcell = C.g.rw[p.ir + 1];
        if (p.ir >= 0) {
            ++nseg;
            int c = cell_index(C.g, p);
            if (LMIN) lmin = fminf(lmin, C.csize[c]);
            float dtau = C.alpha[c] * tw;
            if (C.mode == 0) {
                I = fmaf(__expf(-tau) * (-expm1f(-dtau)), C.src[c], I);
                tau += dtau;
                if (tau > C.tau_stop) break;
            } else {
                tau += dtau;
            }
        }
        if (cross_wall_cam(C.g, p, tw, kind, cr)) break;
        if (unstick(C.g, p, tw, rcell, nzero) && p.ir >= C.g.nr) break;
    }
    return C.mode == 0 ? I : tau;
}

__device__ __forceinline__ float integrate_ray(const CamParams &C, Pos p, const CamRay &cr,
                                              uint32_t &nseg) {
    float unused = RT_INF;
    return integrate_ray_l<false>(C, p, cr, unused, nseg);
}

// pixel of this lane: blocks cover tw x th tiles of the image, so the rays of
// a wavefront stay close together (they cross the same cells)
__device__ __forceinline__ bool tile_pixel(const CamParams &C, int &j, int &k) {
    const int tile = blockIdx.x, ty = tile / C.ntx, tx = tile - ty * C.ntx;
    const int t = threadIdx.x, dy = t / C.tw, dx = t - dy * C.tw;
    k = tx * C.tw + dx;
    j = ty * C.th + dy;
    return k < C.nx && j < C.ny;
}

__global__ void __launch_bounds__(256) k_image(CamParams C, float *img) {
    int j, k;
    uint32_t nseg = 0;
    if (tile_pixel(C, j, k)) {
        const int pix = j * C.nx + k;
        float acc = 0.f;
        for (int a = 0; a < C.nsub; ++a)
            for (int b = 0; b < C.nsub; ++b) {
                float X = ((float)(k - C.nx / 2) + ((b + 0.5f) / C.nsub - 0.5f)) * C.pix;
                float Y = ((float)(j - C.ny / 2) + ((a + 0.5f) / C.nsub - 0.5f)) * C.pix;
                Pos p;
                CamRay cr;
                if (ray_start(C, X, Y, p, cr)) acc += integrate_ray(C, p, cr, nseg);
            }
        img[pix] = acc / (float)(C.nsub * C.nsub);
    }
    count_crossings(C.icount, nseg);
}

// The adjoint's per-cell sums are 64-bit fixed-point integers (units in
// C.qs, set by k_adj_scale): integer adds commute, so the gradient does not
// depend on the order in which rays and blocks add to a cell -- it is
// bit-reproducible, as the thermal tallies are.  They are first gathered in a
// small open-addressing table in LDS: the rays of one block are neighbours on
// the sky and cross mostly the same cells, so each cell then costs one global
// atomic per block instead of one per ray.  Entries: 2^hbits keys (cell index,
// -1 = empty) and two sums; a contribution that finds no slot in 8 probes goes
// straight to global memory.
typedef unsigned long long u64;
__device__ __forceinline__ void adj_add(int c, long long va, long long vs, int *key, u64 *ga,
                                        u64 *gs, unsigned hbits, double *g_alpha,
                                        double *g_src) {
    const unsigned mask = (1u << hbits) - 1u;
    unsigned slot = ((unsigned)c * 2654435761u) >> (32u - hbits);
    for (int probe = 0; probe < 8; ++probe) {
        int cur = __hip_atomic_load(&key[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (cur == -1) {
            int exp = -1;
            __hip_atomic_compare_exchange_strong(&key[slot], &exp, c, __ATOMIC_RELAXED,
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            cur = exp == -1 ? c : exp;
        }
        if (cur == c) {
            __hip_atomic_fetch_add(&ga[slot], (u64)va, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
            __hip_atomic_fetch_add(&gs[slot], (u64)vs, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
            return;
        }
        slot = (slot + 1u) & mask;
    }
    atomicAdd((u64 *)&g_alpha[c], (u64)va);
    atomicAdd((u64 *)&g_src[c], (u64)vs);
}

// the block's LDS table into the global sums
__device__ __forceinline__ void adj_flush(const int *key, const u64 *ga, const u64 *gs,
                                          unsigned H, double *g_alpha, double *g_src) {
    for (unsigned i = threadIdx.x; i < H; i += blockDim.x) {
        const int c = key[i];
        if (c >= 0) {
            if (ga[i]) atomicAdd((u64 *)&g_alpha[c], ga[i]);
            if (gs[i]) atomicAdd((u64 *)&g_src[c], gs[i]);
        }
    }
}

// gI times the gradient of the ray through sky (X, Y) w.r.t. the per-cell
// alpha and S: the ray is integrated for I, then swept again front to back
// recomputing tau, with R_k = I - I_<=k (nothing is stored per segment)
__device__ __forceinline__ void adjoint_ray(const CamParams &C, float X, float Y, float gI,
                                            int *key, u64 *ga, u64 *gs, double *g_alpha,
                                            double *g_src, uint32_t &nseg) {
    Pos p0;
    CamRay cr;
    if (!ray_start(C, X, Y, p0, cr)) return;
    const double qa = C.qs[0], qsrc = C.qs[1];
    uint32_t nfwd = 0;
    float I = integrate_ray(C, p0, cr, nfwd);
    Pos p = p0;
    float tau = 0.f, cum = 0.f;
    uint32_t nzero = 0;
    for (int guard = 0; guard < (1 << 24); ++guard) {
        int kind;
        float tw = next_wall(C.g, p, kind);
        const float rcell = C.g.rw[p.ir + 1];
        if (p.ir >= 0) {
            ++nseg;
            int c = cell_index(C.g, p);
            float dtau = C.alpha[c] * tw;
            float att = __expf(-tau);
            float om = -expm1f(-dtau);
            float S = C.src[c];
            cum = fmaf(att * om, S, cum);
            float Rk = I - cum;
            float dI_ddtau = att * (1.f - om) * S - Rk;
            adj_add(c, __double2ll_rn((double)(gI * dI_ddtau * tw) * qa),
                    __double2ll_rn((double)(gI * att * om) * qsrc), key, ga, gs, C.hbits,
                    g_alpha, g_src);
            tau += dtau;
            if (tau > C.tau_stop) break;
        }
        if (cross_wall_cam(C.g, p, tw, kind, cr)) break;
        if (unstick(C.g, p, tw, rcell, nzero) && p.ir >= C.g.nr) break;
    }
}

__global__ void __launch_bounds__(256) k_image_adj(CamParams C, const float *dimg, double *g_alpha,
                                                   double *g_src) {
    extern __shared__ u64 adj_lds[];
    const unsigned H = 1u << C.hbits;
    u64 *ga = adj_lds, *gs = adj_lds + H;
    int *key = (int *)(adj_lds + 2 * H);
    for (unsigned i = threadIdx.x; i < H; i += blockDim.x) {
        key[i] = -1;
        ga[i] = 0ull;
        gs[i] = 0ull;
    }
    __syncthreads();
    int j, k;
    bool act = tile_pixel(C, j, k);
    float gI = 0.f;
    if (act) {
        gI = dimg[j * C.nx + k] / (float)(C.nsub * C.nsub);
        act = gI != 0.f;
    }
    uint32_t nseg = 0;
    for (int a = 0; act && a < C.nsub; ++a)
        for (int b = 0; b < C.nsub; ++b) {
            float X = ((float)(k - C.nx / 2) + ((b + 0.5f) / C.nsub - 0.5f)) * C.pix;
            float Y = ((float)(j - C.ny / 2) + ((a + 0.5f) / C.nsub - 0.5f)) * C.pix;
            adjoint_ray(C, X, Y, gI, key, ga, gs, g_alpha, g_src, nseg);
        }
    __syncthreads();
    adj_flush(key, ga, gs, H, g_alpha, g_src);
    count_crossings(C.icount ? C.icount + 1 : nullptr, nseg);
}

// ---------------------------------------------------------------- sub-pixels
// Recursive sub-pixeling (RADMC-3D's camera_nrrefine / camera_refine_criterion,
// as pdspy leaves them at RADMC-3D's defaults): a (sub)pixel of side d whose
// central ray crosses a cell of size l with d > crit l is replaced by the mean
// of its 2 x 2 sub-pixels.  Done breadth first over a node pool: every node of
// level L is traced by one lane, which hands its four children (at most
// RT_MAX_REFINE levels deep) their slots in level L + 1 with one atomic per
// wavefront; the values are then averaged bottom-up, level by level.  Siblings
// are summed in a fixed order, so the image does not depend on the order in
// which slots were handed out.  The adjoint hands each leaf ray the pixel's
// cotangent divided by 4 per level and runs the per-ray reverse sweep.
#define RT_REF_DEAD 1.0e30f  // sky position of a node with no ray (tile padding, no room)

struct RefArgs {
    float2 *xy;       // sky position of the node's central ray (au)
    float *val;       // its value (intensity or tau)
    int *child;       // first of its four children, -1: a leaf
    float *gw;        // adjoint: its cotangent
    unsigned *ctl;    // control words (RT_REF_*)
    unsigned cap;     // pool size (nodes)
    float crit;       // refine criterion
    int nref;         // levels that may split
    float d0;         // side of a level-0 node (au): pixel / nsub
    unsigned n0;      // level-0 nodes: every tile's pixels x nsub^2
    int ns2;          // nsub^2
};

// level-0 node i: tiles in order, the tile's pixels, then the pixel's nsub^2
// sub-rays, so the rays of a wavefront are neighbours on the sky
__device__ __forceinline__ bool node0_pixel(const CamParams &C, int ns2, unsigned i, int &j, int &k,
                                            int &a, int &b) {
    const unsigned tp = (unsigned)(C.tw * C.th);
    const unsigned sub = i % (unsigned)ns2, q = i / (unsigned)ns2;
    const unsigned t = q / tp, r = q - t * tp;
    const int ty = (int)t / C.ntx, tx = (int)t - ty * C.ntx;
    const int dy = (int)r / C.tw, dx = (int)r - dy * C.tw;
    k = tx * C.tw + dx;
    j = ty * C.th + dy;
    a = (int)sub / C.nsub;
    b = (int)sub - a * C.nsub;
    return k < C.nx && j < C.ny;
}

__device__ __forceinline__ unsigned pixel_node0(const CamParams &C, int ns2, int j, int k, int sub) {
    const int ty = j / C.th, tx = k / C.tw;
    const unsigned t = (unsigned)(ty * C.ntx + tx);
    const unsigned r = (unsigned)((j - ty * C.th) * C.tw + (k - tx * C.tw));
    return (t * (unsigned)(C.tw * C.th) + r) * (unsigned)ns2 + (unsigned)sub;
}

__global__ void k_ref_init(CamParams C, RefArgs R) {
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        R.ctl[0] = R.n0;
        R.ctl[1] = 0u;
        for (int L = 1; L <= RT_MAX_REFINE + 1; ++L) R.ctl[1 + L] = R.n0;
        R.ctl[RT_REF_RAYS] = 0u;
    }
    if (i >= R.n0) return;
    int j, k, a, b;
    float2 xy = make_float2(RT_REF_DEAD, RT_REF_DEAD);  // tile padding: no ray
    if (node0_pixel(C, R.ns2, i, j, k, a, b)) {
        xy.x = ((float)(k - C.nx / 2) + ((b + 0.5f) / C.nsub - 0.5f)) * C.pix;
        xy.y = ((float)(j - C.ny / 2) + ((a + 0.5f) / C.nsub - 0.5f)) * C.pix;
    }
    R.xy[i] = xy;
}

// traces level L: value of every node, children for those that split
__global__ void __launch_bounds__(256) k_ref_trace(CamParams C, RefArgs R, int L) {
    const unsigned start = R.ctl[1 + L], end = R.ctl[2 + L];
    const float d = ldexpf(R.d0, -L);
    const bool may_split = L < R.nref;
    const int lane = __lane_id();
    unsigned nrays = 0;
    uint32_t nseg = 0;
    for (unsigned base = start + blockIdx.x * blockDim.x; base < end;
         base += gridDim.x * blockDim.x) {
        const unsigned i = base + threadIdx.x;
        const bool act = i < end;
        float I = 0.f, lmin = RT_INF;
        float2 xy = make_float2(RT_REF_DEAD, RT_REF_DEAD);
        if (act) {
            xy = R.xy[i];
            Pos p;
            CamRay cr;
            if (xy.x < 0.5f * RT_REF_DEAD && ray_start(C, xy.x, xy.y, p, cr)) {
                I = integrate_ray_l<true>(C, p, cr, lmin, nseg);
                ++nrays;
            }
            R.val[i] = I;
        }
        const bool want = act && may_split && d > R.crit * lmin;
        const uint64_t mk = __ballot(want);
        int ch = -1;
        if (mk) {
            const int cnt = __popcll(mk);
            const int leader = __ffsll((unsigned long long)mk) - 1;
            // slots are only handed out while the pool has room (the count
            // stays within a wavefront's 256 slots of the pool size)
            unsigned b0 = R.cap;
            if (lane == leader && __hip_atomic_load(&R.ctl[0], __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_AGENT) < R.cap)
                b0 = atomicAdd(&R.ctl[0], 4u * (unsigned)cnt);
            b0 = __shfl(b0, leader);
            if (want) {
                const unsigned c0 = b0 + 4u * (unsigned)__popcll(mk & ((1ull << lane) - 1ull));
                if (b0 < R.cap && c0 + 4u <= R.cap) {
                    ch = (int)c0;
                    const float h = 0.25f * d;
                    R.xy[c0] = make_float2(xy.x - h, xy.y - h);
                    R.xy[c0 + 1] = make_float2(xy.x + h, xy.y - h);
                    R.xy[c0 + 2] = make_float2(xy.x - h, xy.y + h);
                    R.xy[c0 + 3] = make_float2(xy.x + h, xy.y + h);
                } else {
                    // no room: this node stays a leaf; slots of its that lie in
                    // the pool become rayless orphans of the next level
                    if (b0 < R.cap)
                        for (unsigned q = 0; q < 4u; ++q)
                            if (c0 + q < R.cap) {
                                R.xy[c0 + q] = make_float2(RT_REF_DEAD, RT_REF_DEAD);
                                R.gw[c0 + q] = 0.f;
                            }
                    atomicAdd(&R.ctl[RT_REF_UNSPLIT], 1u);
                }
            }
        }
        if (act) R.child[i] = ch;
    }
    for (int off = 32; off > 0; off >>= 1) nrays += __shfl_down(nrays, off);
    if (lane == 0 && nrays) atomicAdd(&R.ctl[RT_REF_RAYS], nrays);
    count_crossings(C.icount, nseg);
}

// level L + 1 ends where the slots handed out so far end
__global__ void k_ref_close(RefArgs R, int L) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const unsigned c = R.ctl[0];
        R.ctl[3 + L] = c < R.cap ? c : R.cap;
    }
}

__global__ void __launch_bounds__(256) k_ref_combine(RefArgs R, int L) {
    const unsigned start = R.ctl[1 + L], end = R.ctl[2 + L];
    for (unsigned i = start + blockIdx.x * blockDim.x + threadIdx.x; i < end;
         i += gridDim.x * blockDim.x) {
        const int c = R.child[i];
        if (c >= 0) R.val[i] = 0.25f * (((R.val[c] + R.val[c + 1]) + R.val[c + 2]) + R.val[c + 3]);
    }
}

__global__ void __launch_bounds__(256) k_ref_out(CamParams C, RefArgs R, float *img) {
    const int pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= C.nx * C.ny) return;
    const int j = pix / C.nx, k = pix - j * C.nx;
    float acc = 0.f;
    for (int sub = 0; sub < R.ns2; ++sub) acc += R.val[pixel_node0(C, R.ns2, j, k, sub)];
    img[pix] = acc / (float)R.ns2;
}

__global__ void __launch_bounds__(256) k_ref_gw0(CamParams C, RefArgs R, const float *dimg) {
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= R.n0) return;
    int j, k, a, b;
    R.gw[i] = node0_pixel(C, R.ns2, i, j, k, a, b) ? dimg[j * C.nx + k] / (float)R.ns2 : 0.f;
}

__global__ void __launch_bounds__(256) k_ref_down(RefArgs R, int L) {
    const unsigned start = R.ctl[1 + L], end = R.ctl[2 + L];
    for (unsigned i = start + blockIdx.x * blockDim.x + threadIdx.x; i < end;
         i += gridDim.x * blockDim.x) {
        const int c = R.child[i];
        if (c >= 0) {
            const float g = 0.25f * R.gw[i];
            R.gw[c] = g;
            R.gw[c + 1] = g;
            R.gw[c + 2] = g;
            R.gw[c + 3] = g;
        }
    }
}

// the leaves' reverse sweeps; each block gathers a chunk of 256 leaves' cell
// gradients in its LDS table before the global atomics, as k_image_adj does
__global__ void __launch_bounds__(256) k_ref_adj(CamParams C, RefArgs R, double *g_alpha,
                                                 double *g_src) {
    extern __shared__ u64 adj_lds[];
    const unsigned H = 1u << C.hbits;
    u64 *ga = adj_lds, *gs = adj_lds + H;
    int *key = (int *)(adj_lds + 2 * H);
    const unsigned n = R.ctl[0] < R.cap ? R.ctl[0] : R.cap;
    uint32_t nseg = 0;
    for (unsigned base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
        for (unsigned t = threadIdx.x; t < H; t += blockDim.x) {
            key[t] = -1;
            ga[t] = 0ull;
            gs[t] = 0ull;
        }
        __syncthreads();
        const unsigned i = base + threadIdx.x;
        if (i < n && R.child[i] < 0) {
            const float g = R.gw[i];
            const float2 xy = R.xy[i];
            if (g != 0.f && xy.x < 0.5f * RT_REF_DEAD)
                adjoint_ray(C, xy.x, xy.y, g, key, ga, gs, g_alpha, g_src, nseg);
        }
        __syncthreads();
        adj_flush(key, ga, gs, H, g_alpha, g_src);
        __syncthreads();
    }
    count_crossings(C.icount ? C.icount + 1 : nullptr, nseg);
}

// point-source fluxes (erg/s/cm^2/Hz/sr over one pixel) passed by value
struct StarArg {
    float f[RT_MAX_STAR];
};

// Point stars: one block per star.  The star's pixel gets F e^-tau, tau along
// the ray from the star to the observer.  A star at the origin (pdspy's
// default) sees the grid along a radial ray: it stays in one (theta, phi) cell
// column and crosses every shell over its full width, so tau is a sum over the
// shells that the block's lanes share; any other star walks its ray cell by
// cell on one lane.  For the adjoint (img NULL) each cell on the ray gets
// -dimg F e^-tau l in g_alpha.
__global__ void k_stars(CamParams C, int nstar, const float *spos, StarArg sflux, float *img,
                        const float *dimg, double *g_alpha) {
    const int s = blockIdx.x;
    if (s >= nstar) return;
    const int lane = threadIdx.x;
    float qx = spos[3 * s], qy = spos[3 * s + 1], qz = spos[3 * s + 2];
    float xp = qx * C.ex[0] + qy * C.ex[1] + qz * C.ex[2];
    float yp = qx * C.ey[0] + qy * C.ey[1] + qz * C.ey[2];
    float X = xp * C.cpa - yp * C.spa;
    float Y = xp * C.spa + yp * C.cpa;
    int k = (int)floorf(X / C.pix + 0.5f) + C.nx / 2;
    int j = (int)floorf(Y / C.pix + 0.5f) + C.ny / 2;
    if (k < 0 || k >= C.nx || j < 0 || j >= C.ny) return;
    const DevGrid &g = C.g;
    if (qx == 0.f && qy == 0.f && qz == 0.f) {
        // radial ray: the (theta, phi) column of the direction
        Pos p;
        const float rm = 0.5f * (g.rw[0] + g.rw[1]);
        p.x = rm * C.n[0];
        p.y = rm * C.n[1];
        p.z = rm * C.n[2];
        p.dx = C.n[0];
        p.dy = C.n[1];
        p.dz = C.n[2];
        locate(g, p);
        const size_t col = (size_t)g.nr * (p.it + (size_t)g.nt * p.ip);
        float part = 0.f;
        for (int ir = lane; ir < g.nr; ir += blockDim.x)
            part += C.alpha[col + ir] * (g.rw[ir + 1] - g.rw[ir]);
        for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off);
        const float Is = sflux.f[s] * __expf(-part);
        if (img) {
            if (lane == 0) atomicAdd(&img[j * C.nx + k], Is);
            return;
        }
        const float gI = dimg[j * C.nx + k];
        if (gI == 0.f) return;
        const double qa = C.qs[0];
        for (int ir = lane; ir < g.nr; ir += blockDim.x)
            atomicAdd((u64 *)&g_alpha[col + ir],
                      (u64)__double2ll_rn((double)(-gI * Is * (g.rw[ir + 1] - g.rw[ir])) * qa));
        return;
    }
    if (lane != 0) return;
    Pos p;
    p.x = qx;
    p.y = qy;
    p.z = qz;
    p.dx = C.n[0];
    p.dy = C.n[1];
    p.dz = C.n[2];
    locate(g, p);
    float tau = 0.f;
    Pos p0 = p;
    uint32_t nzero = 0;
    for (int guard = 0; guard < (1 << 24) && p.ir < g.nr; ++guard) {
        int kind;
        float tw = next_wall(g, p, kind);
        const float rcell = g.rw[p.ir + 1];
        if (p.ir >= 0) tau += C.alpha[cell_index(g, p)] * tw;
        if (cross_wall(g, p, tw, kind)) break;
        unstick(g, p, tw, rcell, nzero);
    }
    float Is = sflux.f[s] * __expf(-tau);
    if (img) {
        atomicAdd(&img[j * C.nx + k], Is);
        return;
    }
    float gI = dimg[j * C.nx + k];
    if (gI == 0.f) return;
    p = p0;
    nzero = 0;
    for (int guard = 0; guard < (1 << 24) && p.ir < g.nr; ++guard) {
        int kind;
        float tw = next_wall(g, p, kind);
        const float rcell = g.rw[p.ir + 1];
        if (p.ir >= 0)
            atomicAdd((u64 *)&g_alpha[cell_index(g, p)],
                      (u64)__double2ll_rn((double)(-gI * Is * tw) * C.qs[0]));
        if (cross_wall(g, p, tw, kind)) break;
        unstick(g, p, tw, rcell, nzero);
    }
}

struct KapArg {
    float k[RT_MAX_SPEC];   // absorption opacity at the camera wavelength
    float ks[RT_MAX_SPEC];  // scattering opacity (0 without a scattering source)
};

// alpha (per au) and the source function per cell: B_nu(T), or with a
// scattering source J (isotropic) (rho kabs B + rho ksca J) / (rho kext)
__global__ void k_cam_prep(const float *rho, int S, int ncell, KapArg kap, float lenu, const float *T,
                           float pre, float hnu_k, int mode, const float *J, float *alpha,
                           float *src) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    float aa = 0.f, as = 0.f;
    for (int s = 0; s < S; ++s) {
        const float r = rho[(size_t)s * ncell + c];
        aa = fmaf(r, kap.k[s], aa);
        as = fmaf(r, kap.ks[s], as);
    }
    const float a = aa + as;
    alpha[c] = a * lenu;
    float S_ = 0.f;
    if (mode == 0) {
        const float B = planck_f(pre, hnu_k, T[c]);
        S_ = J ? (a > 0.f ? (aa * B + as * J[c]) / a : 0.f) : B;
    }
    src[c] = S_;
}

// chain rule to log rho_s and T.  With a scattering source (J held fixed):
// d alpha / d log rho_s = rho_s kext_s, dS / d log rho_s = rho_s (kabs_s B +
// ksca_s J - S kext_s) / (rho kext), dS / dT = (rho kabs / rho kext) dB/dT.
__global__ void k_cam_post(const double *g_alpha, const double *g_src, const double *qs,
                           const float *rho, int S, int ncell, KapArg kap, float lenu,
                           const float *T, float pre, float hnu_k, const float *J,
                           const float *src, float *g_logrho, float *g_T) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    // the fixed-point sums back to FP64
    const double qa = qs[0], qsrc = qs[1];
    const double ga = qa > 0.0 ? (double)((const long long *)g_alpha)[c] / qa : 0.0;
    const double gsv = qsrc > 0.0 ? (double)((const long long *)g_src)[c] / qsrc : 0.0;
    if (!J) {
        for (int s = 0; s < S; ++s)
            g_logrho[(size_t)s * ncell + c] =
                (float)(ga * (double)rho[(size_t)s * ncell + c] * (double)kap.k[s] * (double)lenu);
        g_T[c] = (float)(gsv * (double)dplanck_f(pre, hnu_k, T[c]));
        return;
    }
    const double gs = gsv;
    double aa = 0.0, a = 0.0;
    for (int s = 0; s < S; ++s) {
        const double r = rho[(size_t)s * ncell + c];
        aa += r * (double)kap.k[s];
        a += r * ((double)kap.k[s] + (double)kap.ks[s]);
    }
    const double B = planck_f(pre, hnu_k, T[c]), Jc = J[c], Sc = src[c];
    for (int s = 0; s < S; ++s) {
        const double r = rho[(size_t)s * ncell + c];
        const double ke = (double)kap.k[s] + (double)kap.ks[s];
        const double dS = a > 0.0 ? r * ((double)kap.k[s] * B + (double)kap.ks[s] * Jc - Sc * ke) / a
                                  : 0.0;
        g_logrho[(size_t)s * ncell + c] = (float)(ga * r * ke * (double)lenu + gs * dS);
    }
    g_T[c] = (float)(gs * (a > 0.0 ? aa / a : 0.0) * (double)dplanck_f(pre, hnu_k, T[c]));
}

__global__ void k_zero_f64(double *p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// Units of the adjoint's fixed-point sums, from bounds that hold whatever rays
// an image traces.  The cotangents of all rays sum to at most ||dimg||_1 (the
// rays of a pixel share its dL/dI, sub-pixels a quarter of their parent's).  A
// contribution to dL/dalpha_c is g dI/d(dtau) l with |dI/d(dtau)| <= S_max
// (the largest source: I and every emission term lie in [0, S_max]) and
// l <= 2 R_out; a star adds g F l with F its pixel flux.  A ray crosses a cell
// at most 4 times.  So |sum_c| <= 8 R_out max(S_max, F) (1 + nstar) ||dimg||_1
// for alpha and 4 ||dimg||_1 for S, and units of 2^-61 of those bounds keep
// every partial sum below 2^61.  One block; the reductions run in a fixed
// order, so the units (and with them the gradient) are bit-reproducible.
__global__ void __launch_bounds__(256) k_adj_scale(const float *dimg, int npix, const float *src,
                                                   int ncell, float fstar, int nstar,
                                                   float rout, double *qs) {
    __shared__ double s_l1[256];
    __shared__ float s_max[256];
    const int t = threadIdx.x;
    double l1 = 0.0;
    for (int i = t; i < npix; i += 256) l1 += fabs((double)dimg[i]);
    float smax = fstar;
    for (int i = t; i < ncell; i += 256) smax = fmaxf(smax, src[i]);
    s_l1[t] = l1;
    s_max[t] = smax;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h) {
            s_l1[t] += s_l1[t + h];
            s_max[t] = fmaxf(s_max[t], s_max[t + h]);
        }
        __syncthreads();
    }
    if (t == 0) {
        const double two61 = 2305843009213693952.0;
        const double ba = 8.0 * (double)rout * (double)s_max[0] * (1.0 + nstar) * s_l1[0];
        const double bs = 4.0 * s_l1[0];
        qs[0] = ba > 0.0 ? two61 / ba : 0.0;
        qs[1] = bs > 0.0 ? two61 / bs : 0.0;
    }
}

static inline unsigned nblk(size_t n, int b) { return (unsigned)((n + b - 1) / b); }

// opacity at the camera wavelength: log-log interpolation of the model's table
// (power-law extrapolation beyond the ends)
static double kappa_at(const rt_model *m, const std::vector<double> &tab, int s, double lam);

double kappa_at_lambda(const rt_model *m, const std::vector<double> &tab, int s, double lam) {
    return kappa_at(m, tab, s, lam);
}

static double kappa_at(const rt_model *m, const std::vector<double> &tab, int s, double lam) {
    const int nl = m->nlam;
    const double *k = &tab[(size_t)s * nl];
    if (nl == 1) return k[0];
    int i = 0;
    if (lam <= m->lam[0])
        i = 0;
    else if (lam >= m->lam[nl - 1])
        i = nl - 2;
    else
        while (i < nl - 2 && m->lam[i + 1] < lam) ++i;
    double l0 = std::log(m->lam[i]), l1 = std::log(m->lam[i + 1]);
    if (!(k[i] > 0) || !(k[i + 1] > 0)) {
        double w = (lam - m->lam[i]) / (m->lam[i + 1] - m->lam[i]);
        return k[i] + w * (k[i + 1] - k[i]);
    }
    double w = (std::log(lam) - l0) / (l1 - l0);
    return std::exp(std::log(k[i]) + w * (std::log(k[i + 1]) - std::log(k[i])));
}

// per-cell scratch of the camera kernels (allocated once per model) and the
// cells' sizes for the sub-pixeling: the least of dr, r_c dtheta and (with phi
// cells) r_c sin(theta_c) dphi at the cell centre, in au
int image_buffers(rt_model *m) {
    rt_ctx *ctx = m->ctx;
    size_t nc = (size_t)m->ncell;
    if (!m->d_alpha) {
        HIPCHK(ctx, hipMalloc(&m->d_alpha, nc * sizeof(float)));
        HIPCHK(ctx, hipMalloc(&m->d_src, nc * sizeof(float)));
        HIPCHK(ctx, hipMalloc(&m->d_galpha, nc * sizeof(double)));
        HIPCHK(ctx, hipMalloc(&m->d_gsrc, nc * sizeof(double)));
        HIPCHK(ctx, hipMalloc(&m->d_qscale, 2 * sizeof(double)));
    }
    if (!m->d_csize) {
        std::vector<float> cs(nc);
        for (int ip = 0; ip < m->np; ++ip)
            for (int it = 0; it < m->nt; ++it)
                for (int ir = 0; ir < m->nr; ++ir) {
                    const double rc = 0.5 * (m->rw[ir] + m->rw[ir + 1]);
                    double l = m->rw[ir + 1] - m->rw[ir];
                    l = std::min(l, rc * (m->tw[it + 1] - m->tw[it]));
                    if (m->np > 1)
                        l = std::min(l, rc * std::sin(0.5 * (m->tw[it] + m->tw[it + 1])) *
                                            (m->pw[ip + 1] - m->pw[ip]));
                    cs[(size_t)ir + (size_t)m->nr * (it + (size_t)m->nt * ip)] =
                        (float)(l / RT_LEN_UNIT);
                }
        HIPCHK(ctx, hipMalloc(&m->d_csize, nc * sizeof(float)));
        HIPCHK(ctx, hipMemcpy(m->d_csize, cs.data(), nc * sizeof(float), hipMemcpyHostToDevice));
    }
    return RT_OK;
}

int refine_reserve(rt_ctx *ctx, size_t n) {
    RefinePool &P = ctx->ref;
    if (n > ((size_t)1 << 30)) return set_err(ctx, RT_E_LIMIT, "sub-pixel pool of %zu nodes > 2^30", n);
    if (P.cap >= n && P.ctl) return RT_OK;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    void *old[] = {P.xy, P.val, P.child, P.gw};
    for (void *q : old)
        if (q) (void)hipFree(q);
    P.xy = nullptr;
    P.val = nullptr;
    P.child = nullptr;
    P.gw = nullptr;
    P.cap = 0;
    if (hipMalloc(&P.xy, n * sizeof(float2)) != hipSuccess ||
        hipMalloc(&P.val, n * sizeof(float)) != hipSuccess ||
        hipMalloc(&P.child, n * sizeof(int)) != hipSuccess ||
        hipMalloc(&P.gw, n * sizeof(float)) != hipSuccess)
        return set_err(ctx, RT_E_NOMEM, "sub-pixel pool of %zu nodes", n);
    if (!P.ctl) {
        HIPCHK(ctx, hipMalloc(&P.ctl, RT_REF_NCTL * sizeof(unsigned)));
        std::vector<unsigned> z(RT_REF_NCTL, 0u);
        HIPCHK(ctx, hipMemcpy(P.ctl, z.data(), RT_REF_NCTL * sizeof(unsigned), hipMemcpyHostToDevice));
    }
    P.cap = n;
    return RT_OK;
}

static int setup_camera(rt_model *m, const rt_camera *cam, CamParams &C, KapArg &kap, float &pre,
                        float &hnu_k) {
    rt_ctx *ctx = m->ctx;
    if (!cam || cam->nx < 1 || cam->ny < 1 || !(cam->pix_cm > 0) || !(cam->lam_um > 0))
        return set_err(ctx, RT_E_ARG, "invalid camera");
    if (cam->mode != 0 && cam->mode != 1) return set_err(ctx, RT_E_ARG, "camera mode must be 0 or 1");
    const double d2r = M_PI / 180.0;
    double i = cam->incl_deg * d2r, ph = cam->phi_deg * d2r, pa = cam->posang_deg * d2r;
    double n[3] = {std::sin(i) * std::sin(ph), -std::sin(i) * std::cos(ph), std::cos(i)};
    double ex[3] = {std::cos(ph), std::sin(ph), 0.0};
    double ey[3] = {n[1] * ex[2] - n[2] * ex[1], n[2] * ex[0] - n[0] * ex[2], n[0] * ex[1] - n[1] * ex[0]};
    C.g = dev_grid(m);
    C.icount = ctx->d_icount;
    for (int a = 0; a < 3; ++a) {
        C.n[a] = (float)n[a];
        C.ex[a] = (float)ex[a];
        C.ey[a] = (float)ey[a];
    }
    C.cpa = (float)std::cos(pa);
    C.spa = (float)std::sin(pa);
    C.pix = (float)(cam->pix_cm / RT_LEN_UNIT);
    C.nx = cam->nx;
    C.ny = cam->ny;
    C.nsub = cam->nsub > 0 ? cam->nsub : 1;
    C.mode = cam->mode;
    // ray tile: tile_w x tile_h pixels per block (default 64 x 1: one row run)
    C.tw = cam->tile_w > 0 ? cam->tile_w : 64;
    C.th = cam->tile_h > 0 ? cam->tile_h : 64 / C.tw;
    if (C.th < 1) C.th = 1;
    if (C.tw * C.th > 256 || (C.tw * C.th) % 64)
        return set_err(ctx, RT_E_ARG, "ray tile %d x %d: need a multiple of 64 pixels, at most 256",
                       C.tw, C.th);
    C.ntx = (C.nx + C.tw - 1) / C.tw;
    C.hbits = 8;
    while ((1u << C.hbits) < 4u * (unsigned)(C.tw * C.th)) ++C.hbits;
    C.tau_stop = (float)(cam->tau_stop > 0 ? cam->tau_stop : 40.0);
    for (int s = 0; s < RT_MAX_SPEC; ++s) {
        kap.k[s] = s < m->nspec ? (float)kappa_at(m, m->kabs, s, cam->lam_um) : 0.f;
        kap.ks[s] = (s < m->nspec && cam->d_jscat) ? (float)kappa_at(m, m->ksca, s, cam->lam_um)
                                                   : 0.f;
    }
    double nu = 2.99792458e10 / (cam->lam_um * 1e-4);
    pre = (float)(2.0 * 6.6260755e-27 * nu * nu * nu / (2.99792458e10 * 2.99792458e10));
    hnu_k = (float)(6.6260755e-27 * nu / 1.380658e-16);
    int rc = image_buffers(m);
    if (rc) return rc;
    C.alpha = m->d_alpha;
    C.src = m->d_src;
    C.csize = m->d_csize;
    return 0;
}

static bool stream_capturing(hipStream_t sm) {
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(sm, &st) == hipSuccess && st == hipStreamCaptureStatusActive;
}

// The sub-pixelled image (d_img) or, with d_dimg, the gradients w.r.t. the
// per-cell alpha and S accumulated into m->d_galpha / m->d_gsrc.  Outside a
// graph capture a pool too small for the refinement is grown and the image
// traced again; inside one the pool must have been reserved (rt_refine_reserve)
// and a sub-pixel without room stays a leaf (counted for rt_refine_stats).
static int refine_run(rt_model *m, const rt_camera *cam, CamParams C, float *d_img,
                      const float *d_dimg) {
    rt_ctx *ctx = m->ctx;
    hipStream_t sm = ctx->stream;
    const int nref = std::min(cam->nrefine, RT_MAX_REFINE);
    const int ns2 = C.nsub * C.nsub;
    const unsigned ntiles = (unsigned)(C.ntx * ((C.ny + C.th - 1) / C.th));
    const size_t n0 = (size_t)ntiles * (size_t)(C.tw * C.th) * (size_t)ns2;
    const bool capturing = stream_capturing(sm);
    // a fresh pool starts at 2^21 nodes (40 MB); a reserved one is used as
    // it is while it holds the pixels' rays twice
    size_t want = ctx->ref.cap ? std::max(ctx->ref.cap, 2 * n0)
                               : std::max((size_t)1 << 21, 8 * n0);
    if (capturing) {
        if (ctx->ref.cap < 2 * n0)
            return set_err(ctx, RT_E_ARG, "sub-pixel pool (%zu nodes) too small to capture an image "
                                          "of %zu rays: reserve it first", ctx->ref.cap, n0);
        want = ctx->ref.cap;
    }
    const unsigned gblk = (unsigned)(ctx->n_cu > 0 ? 8 * ctx->n_cu : 2048);
    for (;;) {
        int rc = refine_reserve(ctx, want);
        if (rc) return rc;
        RefinePool &P = ctx->ref;
        RefArgs R{P.xy, P.val, P.child, P.gw, P.ctl, (unsigned)P.cap,
                  (float)(cam->refine_crit > 0 ? cam->refine_crit : 1.0), nref, C.pix / C.nsub,
                  (unsigned)n0, ns2};
        hipLaunchKernelGGL(k_ref_init, dim3(nblk(n0, 256)), dim3(256), 0, sm, C, R);
        for (int L = 0; L <= nref; ++L) {
            hipLaunchKernelGGL(k_ref_trace, dim3(gblk), dim3(256), 0, sm, C, R, L);
            if (L < nref) hipLaunchKernelGGL(k_ref_close, dim3(1), dim3(64), 0, sm, R, L);
        }
        if (!capturing) {
            unsigned uns = 0;
            HIPCHK(ctx, hipMemcpyAsync(&uns, P.ctl + RT_REF_UNSPLIT, sizeof(uns),
                                       hipMemcpyDeviceToHost, sm));
            HIPCHK(ctx, hipStreamSynchronize(sm));
            if (uns && P.cap < ((size_t)1 << 30)) {
                unsigned z = 0;
                HIPCHK(ctx, hipMemcpy(P.ctl + RT_REF_UNSPLIT, &z, sizeof(z), hipMemcpyHostToDevice));
                want = std::min((size_t)1 << 30, 4 * P.cap);
                continue;  // trace again in a larger pool
            }
        }
        if (!d_dimg) {
            for (int L = nref - 1; L >= 0; --L)
                hipLaunchKernelGGL(k_ref_combine, dim3(gblk), dim3(256), 0, sm, R, L);
            hipLaunchKernelGGL(k_ref_out, dim3(nblk((size_t)C.nx * C.ny, 256)), dim3(256), 0, sm, C,
                               R, d_img);
        } else {
            hipLaunchKernelGGL(k_ref_gw0, dim3(nblk(n0, 256)), dim3(256), 0, sm, C, R, d_dimg);
            for (int L = 0; L < nref; ++L)
                hipLaunchKernelGGL(k_ref_down, dim3(gblk), dim3(256), 0, sm, R, L);
            CamParams Ca = C;
            Ca.hbits = 10;  // >= 4 entries per ray of a 256-ray chunk
            const size_t lds = (size_t)(1u << Ca.hbits) * (2 * sizeof(u64) + sizeof(int));
            hipLaunchKernelGGL(k_ref_adj, dim3(gblk), dim3(256), lds, sm, Ca, R, m->d_galpha,
                               m->d_gsrc);
        }
        HIPCHK(ctx, hipGetLastError());
        return RT_OK;
    }
}

static void star_fluxes(const rt_model *m, const rt_camera *cam, StarArg &fl) {
    double nu = 2.99792458e10 / (cam->lam_um * 1e-4);
    for (int s = 0; s < RT_MAX_STAR; ++s) fl.f[s] = 0.f;
    for (int s = 0; s < m->nstar && s < RT_MAX_STAR; ++s) {
        double R = m->stars[5 * s], T = m->stars[5 * s + 1];
        double x = 6.6260755e-27 * nu / (1.380658e-16 * T);
        double B = 2.0 * 6.6260755e-27 * nu * nu * nu / (2.99792458e10 * 2.99792458e10) / std::expm1(x);
        fl.f[s] = (float)(M_PI * R * R * B / (cam->pix_cm * cam->pix_cm));
    }
}

int image_run(rt_model *m, const rt_camera *cam, const float *d_T, const float *d_rho, float *d_img) {
    rt_ctx *ctx = m->ctx;
    CamParams C{};
    KapArg kap{};
    float pre, hnu_k;
    int rc = setup_camera(m, cam, C, kap, pre, hnu_k);
    if (rc) return rc;
    if (!d_T && cam->mode == 0) return set_err(ctx, RT_E_ARG, "temperature pointer is NULL");
    const float *rho = d_rho ? d_rho : m->d_rho;
    hipStream_t sm = ctx->stream;
    hipLaunchKernelGGL(k_cam_prep, dim3(nblk(m->ncell, 256)), dim3(256), 0, sm, rho, m->nspec,
                       m->ncell, kap, (float)RT_LEN_UNIT, d_T, pre, hnu_k, cam->mode,
                       cam->d_jscat, m->d_alpha, m->d_src);
    const unsigned ntiles = (unsigned)(C.ntx * ((C.ny + C.th - 1) / C.th));
    if (cam->nrefine > 0) {
        rc = refine_run(m, cam, C, d_img, nullptr);
        if (rc) return rc;
    } else {
        hipLaunchKernelGGL(k_image, dim3(ntiles), dim3(C.tw * C.th), 0, sm, C, d_img);
    }
    if (cam->incl_star && cam->mode == 0 && m->nstar > 0) {
        StarArg fl;
        star_fluxes(m, cam, fl);
        hipLaunchKernelGGL(k_stars, dim3(m->nstar), dim3(64), 0, sm, C, m->nstar, m->d_spos, fl, d_img,
                           (const float *)nullptr, (double *)nullptr);
    }
    HIPCHK(ctx, hipGetLastError());
    return RT_OK;
}

int image_adjoint_run(rt_model *m, const rt_camera *cam, const float *d_T, const float *d_rho,
                      const float *d_dimg, float *d_g_logrho, float *d_g_T) {
    rt_ctx *ctx = m->ctx;
    if (cam && cam->mode != 0) return set_err(ctx, RT_E_ARG, "adjoint needs mode 0 (intensity)");
    CamParams C{};
    KapArg kap{};
    float pre, hnu_k;
    int rc = setup_camera(m, cam, C, kap, pre, hnu_k);
    if (rc) return rc;
    if (!d_T || !d_dimg || !d_g_logrho || !d_g_T) return set_err(ctx, RT_E_ARG, "NULL pointer");
    const float *rho = d_rho ? d_rho : m->d_rho;
    hipStream_t sm = ctx->stream;
    size_t nc = (size_t)m->ncell;
    hipLaunchKernelGGL(k_cam_prep, dim3(nblk(nc, 256)), dim3(256), 0, sm, rho, m->nspec, m->ncell,
                       kap, (float)RT_LEN_UNIT, d_T, pre, hnu_k, 0, cam->d_jscat, m->d_alpha,
                       m->d_src);
    hipLaunchKernelGGL(k_zero_f64, dim3(nblk(nc, 256)), dim3(256), 0, sm, m->d_galpha, nc);
    hipLaunchKernelGGL(k_zero_f64, dim3(nblk(nc, 256)), dim3(256), 0, sm, m->d_gsrc, nc);
    const bool stars = cam->incl_star && m->nstar > 0;
    StarArg fl;
    star_fluxes(m, cam, fl);
    float fmax_star = 0.f;
    for (int s = 0; stars && s < m->nstar && s < RT_MAX_STAR; ++s) fmax_star = std::max(fmax_star, fl.f[s]);
    hipLaunchKernelGGL(k_adj_scale, dim3(1), dim3(256), 0, sm, d_dimg, C.nx * C.ny, m->d_src,
                       m->ncell, fmax_star, stars ? m->nstar : 0,
                       (float)(m->rw[m->nr] / RT_LEN_UNIT), m->d_qscale);
    C.qs = m->d_qscale;
    const unsigned ntiles = (unsigned)(C.ntx * ((C.ny + C.th - 1) / C.th));
    if (cam->nrefine > 0) {
        rc = refine_run(m, cam, C, nullptr, d_dimg);
        if (rc) return rc;
    } else {
        const size_t adj_lds = (size_t)(1u << C.hbits) * (2 * sizeof(u64) + sizeof(int));
        hipLaunchKernelGGL(k_image_adj, dim3(ntiles), dim3(C.tw * C.th), adj_lds, sm, C, d_dimg,
                           m->d_galpha, m->d_gsrc);
    }
    if (stars)
        hipLaunchKernelGGL(k_stars, dim3(m->nstar), dim3(64), 0, sm, C, m->nstar, m->d_spos, fl,
                           (float *)nullptr, d_dimg, m->d_galpha);
    hipLaunchKernelGGL(k_cam_post, dim3(nblk(nc, 256)), dim3(256), 0, sm, m->d_galpha, m->d_gsrc,
                       (const double *)m->d_qscale, rho, m->nspec, m->ncell, kap,
                       (float)RT_LEN_UNIT, d_T, pre, hnu_k, cam->d_jscat, m->d_src, d_g_logrho,
                       d_g_T);
    HIPCHK(ctx, hipGetLastError());
    return RT_OK;
}
