// MI355X (gfx950) FDTD time-stepping engine: fused Yee/UPML/constitutive/ADE/deposit kernels
// and the float64 flux transform, behind the C ABI of include/meep_fdtd.h.
//
// Arithmetic contract (reference docs/design/floating-point.md and stepping.py): every
// float32 operation rounds once in the reference's association order, no contraction
// (built with -ffp-contract=off), subnormals kept (-fno-gpu-flush-denormals-to-zero). The
// deposit and the off-diagonal row are formed in double and rounded once.
//
// Device layout: every volume x fastest, element (i,j,k) at (k*ny+j)*nx+i, each axis periodic
// (an absorber is a material, reference stepping.py:2984). One `SimDev` descriptor per
// simulation lives in device memory; kernels walk a list of 256-cell tiles, so a single run
// and a batch of uneven grids run the same persistent kernels. The tile walk is
// XCD-contiguous: workgroup b runs on XCD b%8, and each XCD sweeps its own contiguous eighth
// of the tile list so that the neighbour rows and planes a tile reads were fetched into that
// XCD's L2 by the tiles just before it.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/meep_fdtd.h"

namespace {

constexpr int kTile = 256;       // cells per tile = threads per workgroup
constexpr int kMaxPoles = FDTD_MAX_POLES;
constexpr int kRing = 32;        // flux samples buffered before one float64 transform pass

thread_local std::string g_error;

int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_error = buf;
    return code;
}

#define HIP_TRY(expr)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail(FDTD_EHIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),    \
                        __FILE__, __LINE__);                                                 \
    } while (0)

// One source box: deposit into D (field 0) or B (field 1) component `comp`.
struct SrcDev {
    int field, comp;
    int lo[3], cnt[3];
    int ncur;
    double dt;
    const double *amp_re, *amp_im;  // per box cell, x fastest
    const double *cur;              // (re, im) per step
};

struct SimDev {
    int n[3];
    int present[3];
    float dtdx;
    int pml;
    int offdiag;
    int npol;
    int nsrc;
    int step;       // completed steps (driver.py:3592)
    int hsep;       // MEEP's one-time H start done (stepping.py:2789)
    int ppar;       // which P buffer holds P^n
    int ncell;
    float *D[3], *B[3], *E[3], *H[3];
    float *fuB[3], *fuD[3], *fwE[3], *fwH[3];
    const float *inv[3];
    const float *u[3][3];
    const float4 *line[2][3];  // (kms, sinv, kps, member) per row, [half][axis]
    float pc[kMaxPoles][4];
    int pdriven[kMaxPoles][3];
    float *pbuf[kMaxPoles][3][2];
    const float *psig[kMaxPoles][3];
    float psig_s[kMaxPoles][3];
    const SrcDev *src;
};

// ------------------------------------------------------------------------------------------
// Tile walk
// ------------------------------------------------------------------------------------------
struct TileWalk {
    int next, stride, end;
};

__device__ inline TileWalk tile_walk(int ntiles, int xcd_remap) {
    const int G = gridDim.x, b = blockIdx.x;
    TileWalk w;
    if (xcd_remap && (G % 8) == 0 && G >= 8) {
        const int x = b % 8, s = b / 8, S = G / 8;
        const int lo = (int)(((long long)ntiles * x) / 8);
        const int hi = (int)(((long long)ntiles * (x + 1)) / 8);
        w.next = lo + s;
        w.stride = S;
        w.end = hi;
    } else {
        w.next = b;
        w.stride = G;
        w.end = ntiles;
    }
    return w;
}

struct Cell {
    int at, i, j, k;
    int xp, yp, zp;  // offsets to the +1 neighbour along x, y, z (periodic)
    int xm, ym, zm;  // offsets to the -1 neighbour
};

__device__ inline Cell locate(const SimDev &s, int at) {
    Cell c;
    const int nx = s.n[0], ny = s.n[1], nz = s.n[2];
    const int jk = at / nx;
    c.at = at;
    c.i = at - jk * nx;
    c.k = jk / ny;
    c.j = jk - c.k * ny;
    const int plane = nx * ny;
    c.xp = (c.i + 1 < nx) ? 1 : -(nx - 1);
    c.xm = (c.i > 0) ? -1 : (nx - 1);
    c.yp = (c.j + 1 < ny) ? nx : -(ny - 1) * nx;
    c.ym = (c.j > 0) ? -nx : (ny - 1) * nx;
    c.zp = (c.k + 1 < nz) ? plane : -(nz - 1) * plane;
    c.zm = (c.k > 0) ? -plane : (nz - 1) * plane;
    return c;
}

__device__ inline int coord(const Cell &c, int axis) { return axis == 0 ? c.i : (axis == 1 ? c.j : c.k); }
__device__ inline int plus(const Cell &c, int axis) { return axis == 0 ? c.xp : (axis == 1 ? c.yp : c.zp); }
__device__ inline int minus(const Cell &c, int axis) { return axis == 0 ? c.xm : (axis == 1 ? c.ym : c.zm); }

// Curl with MEEP's partner presence (stepping.py:2265): ((g1s-g1)+g2)-g2s, or the one-partner
// forms dtdx*(g1s-g1) and (-dtdx)*(g2s-g2).
__device__ inline float curl_of(float dtdx, int p1, int p2, float g1, float g1s, float g2,
                                float g2s) {
    if (p1 && p2) {
        float t = g1s - g1;
        t = t + g2;
        t = t - g2s;
        return dtdx * t;
    }
    if (p1) return dtdx * (g1s - g1);
    if (p2) return (-dtdx) * (g2s - g2);
    return 0.0f;
}

// The curl forms under an absorber (stepping.py:2629, :2662): the split-field recurrence on the
// dsigu layer's cells, the dsig-only form everywhere else (exactly f - curl where both are 1).
__device__ inline float curl_apply_pml(float f, float curl, float4 lsig, float4 lsigu, float *fu,
                                       int at) {
    if (lsigu.w != 0.0f) {
        const float prev = fu[at];
        float u = lsig.x * prev;
        u = u - curl;
        u = lsig.y * u;
        fu[at] = u;
        float v = lsigu.x * f;
        v = v + u;
        v = v - prev;
        return lsigu.y * v;
    }
    float v = lsig.x * f;
    v = v - curl;
    return lsig.y * v;
}

__device__ inline bool plus_zero(double v) { return __double_as_longlong(v) == 0; }

// Sources of one field/component deposited into value `f` at this cell (sources.py:1357).
__device__ inline float deposit(const SimDev &s, int field, int comp, const Cell &c, float f) {
    for (int q = 0; q < s.nsrc; ++q) {
        const SrcDev &d = s.src[q];
        if (d.field != field || d.comp != comp) continue;
        const int a = c.i - d.lo[0], b = c.j - d.lo[1], e = c.k - d.lo[2];
        if (a < 0 || b < 0 || e < 0 || a >= d.cnt[0] || b >= d.cnt[1] || e >= d.cnt[2]) continue;
        if (s.step >= d.ncur) continue;
        const double cre = d.cur[2 * s.step], cim = d.cur[2 * s.step + 1];
        if (plus_zero(cre) && plus_zero(cim)) continue;
        const int p = (e * d.cnt[1] + b) * d.cnt[0] + a;
        double r = d.amp_re[p] * cre;
        r = r - d.amp_im[p] * cim;
        r = r * d.dt;
        double w = (double)f;
        w = w - r;
        f = (float)w;
    }
    return f;
}

// ------------------------------------------------------------------------------------------
// H pair: step_B -> magnetic deposit -> update_H, one cell, component C.
// ------------------------------------------------------------------------------------------
template <int C, bool PML>
__device__ inline void hpair_component(const SimDev &s, const Cell &c, const float e[3],
                                       const float en[3][3]) {
    // en[comp][axis] = E[comp] at the +1 neighbour along axis
    constexpr int first = (C + 2) % 3, fa = (C + 1) % 3;
    constexpr int second = (C + 1) % 3, sa = (C + 2) % 3;
    constexpr int dsig = (C + 1) % 3, dsigu = (C + 2) % 3;
    const float curl = curl_of(s.dtdx, s.present[fa], s.present[sa], e[first], en[first][fa],
                               e[second], en[second][sa]);
    float *B = s.B[C];
    float b = B[c.at];
    if (PML) {
        const float4 lsig = s.line[1][dsig][coord(c, dsig)];
        const float4 lsigu = s.line[1][dsigu][coord(c, dsigu)];
        b = curl_apply_pml(b, curl, lsig, lsigu, s.fuB[C], c.at);
    } else {
        b = b - curl;
    }
    if (s.nsrc) b = deposit(s, 1, C, c, b);
    B[c.at] = b;
    if (PML) {
        const float4 lw = s.line[0][C][coord(c, C)];
        float *H = s.H[C];
        if (lw.w != 0.0f) {
            float *W = s.fwH[C];
            float h, prev;
            if (!s.hsep) {
                h = b;
                prev = b;
            } else {
                h = H[c.at];
                prev = W[c.at];
            }
            W[c.at] = b;
            const float a = lw.z * b;
            const float m = lw.x * prev;
            H[c.at] = h + (a - m);
        } else {
            H[c.at] = b;
        }
    }
}

template <bool PML>
__device__ inline void hpair_cell(const SimDev &s, const Cell &c) {
    float e[3], en[3][3];
#pragma unroll
    for (int q = 0; q < 3; ++q) e[q] = s.E[q][c.at];
    // neighbours each component's curl partners read (+1 along the partner's derivative axis)
    en[0][1] = s.E[0][c.at + c.yp];
    en[0][2] = s.E[0][c.at + c.zp];
    en[1][2] = s.E[1][c.at + c.zp];
    en[1][0] = s.E[1][c.at + c.xp];
    en[2][0] = s.E[2][c.at + c.xp];
    en[2][1] = s.E[2][c.at + c.yp];
    en[0][0] = en[1][1] = en[2][2] = 0.0f;
    hpair_component<0, PML>(s, c, e, en);
    hpair_component<1, PML>(s, c, e, en);
    hpair_component<2, PML>(s, c, e, en);
}

__global__ void __launch_bounds__(kTile) hpair_kernel(const SimDev *__restrict__ sims,
                                                      const int2 *__restrict__ tiles, int ntiles,
                                                      int xcd_remap) {
    const TileWalk w = tile_walk(ntiles, xcd_remap);
    for (int t = w.next; t < w.end; t += w.stride) {
        const int2 tile = tiles ? tiles[t] : make_int2(0, t * kTile);
        const SimDev &s = sims[tile.x];
        const int at = tile.y + (int)threadIdx.x;
        if (at >= s.ncell) continue;
        const Cell c = locate(s, at);
        if (s.pml)
            hpair_cell<true>(s, c);
        else
            hpair_cell<false>(s, c);
    }
}

// ------------------------------------------------------------------------------------------
// D pair: step_D -> electric deposit -> update_E -> update_P, one cell.
// ------------------------------------------------------------------------------------------
template <int C, bool PML>
__device__ inline float dcurl_component(const SimDev &s, const Cell &c, const float h[3],
                                        const float hn[3][3]) {
    constexpr int first = (C + 2) % 3, fa = (C + 1) % 3;
    constexpr int second = (C + 1) % 3, sa = (C + 2) % 3;
    constexpr int dsig = (C + 1) % 3, dsigu = (C + 2) % 3;
    const float curl = curl_of(s.dtdx, s.present[fa], s.present[sa], h[first], hn[first][fa],
                               h[second], hn[second][sa]);
    float d = s.D[C][c.at];
    if (PML) {
        const float4 lsig = s.line[0][dsig][coord(c, dsig)];
        const float4 lsigu = s.line[0][dsigu][coord(c, dsigu)];
        d = curl_apply_pml(d, curl, lsig, lsigu, s.fuD[C], c.at);
    } else {
        d = d - curl;
    }
    if (s.nsrc) d = deposit(s, 0, C, c, d);
    return d;
}

// E from the constitutive product `cst`, then every pole's P update for component C.
template <int C, bool PML>
__device__ inline void e_and_p(const SimDev &s, const Cell &c, float cst, const float p[kMaxPoles]) {
    float *E = s.E[C];
    float w;
    if (PML) {
        const float4 lw = s.line[1][C][coord(c, C)];
        if (lw.w != 0.0f) {
            float *W = s.fwE[C];
            const float prev = W[c.at];
            W[c.at] = cst;
            const float a = lw.z * cst;
            const float m = lw.x * prev;
            E[c.at] = E[c.at] + (a - m);
            w = cst;
        } else {
            E[c.at] = cst;
            w = cst;
        }
    } else {
        E[c.at] = cst;
        w = cst;
    }
    const int par = s.ppar;
    for (int q = 0; q < s.npol; ++q) {
        if (!s.pdriven[q][C]) continue;
        float *out = s.pbuf[q][C][par ^ 1];
        const float pp = out[c.at];
        const float sg = s.psig[q][C] ? s.psig[q][C][c.at] : s.psig_s[q][C];
        float t = s.pc[q][0] * p[q];
        t = t - s.pc[q][1] * pp;
        float dr = sg * w;
        dr = s.pc[q][2] * dr;
        t = t + dr;
        out[c.at] = s.pc[q][3] * t;
    }
}

template <int C, bool PML>
__device__ inline void dpair_component(const SimDev &s, const Cell &c, const float h[3],
                                       const float hn[3][3]) {
    const float d = dcurl_component<C, PML>(s, c, h, hn);
    s.D[C][c.at] = d;
    float p[kMaxPoles];
    float dmp = d;
    const int par = s.ppar;
    for (int q = 0; q < s.npol; ++q) {
        p[q] = 0.0f;
        if (!s.pdriven[q][C]) continue;
        p[q] = s.pbuf[q][C][par][c.at];
        dmp = dmp - p[q];
    }
    const float cst = s.inv[C][c.at] * dmp;
    e_and_p<C, PML>(s, c, cst, p);
}

template <bool PML>
__device__ inline void load_h(const SimDev &s, const Cell &c, float h[3], float hn[3][3]) {
    float *const *Hs = PML ? s.H : s.B;  // without an absorber H is B (fields.py:1380)
#pragma unroll
    for (int q = 0; q < 3; ++q) h[q] = Hs[q][c.at];
    // MEEP's negated strides: the D curl reads its partners one cell DOWN
    hn[0][1] = Hs[0][c.at + c.ym];
    hn[0][2] = Hs[0][c.at + c.zm];
    hn[1][2] = Hs[1][c.at + c.zm];
    hn[1][0] = Hs[1][c.at + c.xm];
    hn[2][0] = Hs[2][c.at + c.xm];
    hn[2][1] = Hs[2][c.at + c.ym];
    hn[0][0] = hn[1][1] = hn[2][2] = 0.0f;
}

template <bool PML>
__device__ inline void dpair_cell(const SimDev &s, const Cell &c) {
    float h[3], hn[3][3];
    load_h<PML>(s, c, h, hn);
    dpair_component<0, PML>(s, c, h, hn);
    dpair_component<1, PML>(s, c, h, hn);
    dpair_component<2, PML>(s, c, h, hn);
}

__global__ void __launch_bounds__(kTile) dpair_kernel(const SimDev *__restrict__ sims,
                                                      const int2 *__restrict__ tiles, int ntiles,
                                                      int xcd_remap) {
    const TileWalk w = tile_walk(ntiles, xcd_remap);
    for (int t = w.next; t < w.end; t += w.stride) {
        const int2 tile = tiles ? tiles[t] : make_int2(0, t * kTile);
        const SimDev &s = sims[tile.x];
        if (s.offdiag) continue;  // stepped by the two-kernel row path below
        const int at = tile.y + (int)threadIdx.x;
        if (at >= s.ncell) continue;
        const Cell c = locate(s, at);
        if (s.pml)
            dpair_cell<true>(s, c);
        else
            dpair_cell<false>(s, c);
    }
}

// Off-diagonal chi1inv: the row reads the partners' D - sum P at neighbouring cells, so the
// D curl (with its deposit) completes over the whole grid first, then E and P.
template <bool PML>
__device__ inline void dcurl_cell(const SimDev &s, const Cell &c) {
    float h[3], hn[3][3];
    load_h<PML>(s, c, h, hn);
    s.D[0][c.at] = dcurl_component<0, PML>(s, c, h, hn);
    s.D[1][c.at] = dcurl_component<1, PML>(s, c, h, hn);
    s.D[2][c.at] = dcurl_component<2, PML>(s, c, h, hn);
}

__global__ void __launch_bounds__(kTile) dcurl_kernel(const SimDev *__restrict__ sims,
                                                      const int2 *__restrict__ tiles, int ntiles,
                                                      int xcd_remap) {
    const TileWalk w = tile_walk(ntiles, xcd_remap);
    for (int t = w.next; t < w.end; t += w.stride) {
        const int2 tile = tiles ? tiles[t] : make_int2(0, t * kTile);
        const SimDev &s = sims[tile.x];
        if (!s.offdiag) continue;
        const int at = tile.y + (int)threadIdx.x;
        if (at >= s.ncell) continue;
        const Cell c = locate(s, at);
        if (s.pml)
            dcurl_cell<true>(s, c);
        else
            dcurl_cell<false>(s, c);
    }
}

__device__ inline float dmp_at(const SimDev &s, int comp, int at) {
    float v = s.D[comp][at];
    for (int q = 0; q < s.npol; ++q)
        if (s.pdriven[q][comp]) v = v - s.pbuf[q][comp][s.ppar][at];
    return v;
}

// The row of stepping._offdiagonal_terms (stepping.py:1594-1687), summed in double.
template <int C, bool PML>
__device__ inline void erow_component(const SimDev &s, const Cell &c) {
    float p[kMaxPoles];
    float dmp = s.D[C][c.at];
    for (int q = 0; q < s.npol; ++q) {
        p[q] = 0.0f;
        if (!s.pdriven[q][C]) continue;
        p[q] = s.pbuf[q][C][s.ppar][c.at];
        dmp = dmp - p[q];
    }
    float cst;
    const float *u1 = s.u[C][(C + 1) % 3], *u2 = s.u[C][(C + 2) % 3];
    if (!u1 && !u2) {
        cst = s.inv[C][c.at] * dmp;
    } else {
        double row = (double)(dmp * s.inv[C][c.at]);
#pragma unroll
        for (int o = 1; o <= 2; ++o) {
            const int pa = (C + o) % 3;
            const float *U = o == 1 ? u1 : u2;
            if (!U) continue;
            // u[q]*(g[q] + g[q - s_partner]) at q = here and q = here + s_own; a shift along
            // the own axis leaves the coordinate along the partner's axis unchanged
            const int up = c.at + plus(c, C);
            const int down = minus(c, pa);
            const float pair0 = dmp_at(s, pa, c.at) + dmp_at(s, pa, c.at + down);
            const float pair1 = dmp_at(s, pa, up) + dmp_at(s, pa, up + down);
            const float prod0 = U[c.at] * pair0;
            const float prod1 = U[up] * pair1;
            const float summed = prod0 + prod1;
            row = row + 0.25 * (double)summed;
        }
        cst = (float)row;
    }
    e_and_p<C, PML>(s, c, cst, p);
}

__global__ void __launch_bounds__(kTile) erow_kernel(const SimDev *__restrict__ sims,
                                                     const int2 *__restrict__ tiles, int ntiles,
                                                     int xcd_remap) {
    const TileWalk w = tile_walk(ntiles, xcd_remap);
    for (int t = w.next; t < w.end; t += w.stride) {
        const int2 tile = tiles ? tiles[t] : make_int2(0, t * kTile);
        const SimDev &s = sims[tile.x];
        if (!s.offdiag) continue;
        const int at = tile.y + (int)threadIdx.x;
        if (at >= s.ncell) continue;
        const Cell c = locate(s, at);
        if (s.pml) {
            erow_component<0, true>(s, c);
            erow_component<1, true>(s, c);
            erow_component<2, true>(s, c);
        } else {
            erow_component<0, false>(s, c);
            erow_component<1, false>(s, c);
            erow_component<2, false>(s, c);
        }
    }
}

// E = chi1inv * (D - sum P): the reference's on-demand get_E for runs that do not store E.
__global__ void prime_e_kernel(SimDev *sims, int ntiles) {
    const int t = blockIdx.x;
    if (t >= ntiles) return;
    const SimDev &s = sims[0];
    const int at = t * kTile + (int)threadIdx.x;
    if (at >= s.ncell) return;
    for (int q = 0; q < 3; ++q) s.E[q][at] = s.inv[q][at] * dmp_at(s, q, at);
}

// Advance each simulation's step counter, H-start flag and P parity after its D pair.
__global__ void advance_kernel(SimDev *sims, int count) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= count) return;
    sims[q].step += 1;
    sims[q].hsep = 1;
    if (sims[q].npol) sims[q].ppar ^= 1;
}

// ------------------------------------------------------------------------------------------
// Flux planes: cell-centred tangential samples into a ring, then a float64 transform pass.
// ------------------------------------------------------------------------------------------
struct PlaneDev {
    int normal;
    int lo[3], cnt[3];
    int cells;
    int nf;
    int dec;
    double scale;
    double dt;
    const double *omega;
    float *ring;     // [kRing][4][cells]
    double *times;   // [kRing] electric sample time t (magnetic is t - dt/2)
    double *phase;   // [kRing][2][nf] (re, im) scratch of one transform pass
    double *acc;     // [4][nf][cells] (re, im)
};

// MEEP yee2cent over the box (dft.py:352): average along every axis where the component's
// Yee shift is 0, x first, then y, then z, each 0.5f*(a+b) in float32.
__device__ inline float centred(const SimDev &s, const float *f, int comp_shift_mask, int i, int j,
                                int k) {
    const int nx = s.n[0], ny = s.n[1], nz = s.n[2];
    const int ex = !(comp_shift_mask & 1), ey = !(comp_shift_mask & 2), ez = !(comp_shift_mask & 4);
    const int i1 = (i + 1) % nx, j1 = (j + 1) % ny, k1 = (k + 1) % nz;
    float v[2][2][2];
    for (int a = 0; a <= ex; ++a)
        for (int b = 0; b <= ey; ++b)
            for (int c = 0; c <= ez; ++c) {
                const int ii = a ? i1 : i, jj = b ? j1 : j, kk = c ? k1 : k;
                v[a][b][c] = f[(kk * ny + jj) * nx + ii];
            }
    if (ex)
        for (int b = 0; b <= ey; ++b)
            for (int c = 0; c <= ez; ++c) v[0][b][c] = 0.5f * (v[0][b][c] + v[1][b][c]);
    if (ey)
        for (int c = 0; c <= ez; ++c) v[0][0][c] = 0.5f * (v[0][0][c] + v[0][1][c]);
    if (ez) v[0][0][0] = 0.5f * (v[0][0][0] + v[0][0][1]);
    return v[0][0][0];
}

// Yee shift bit masks (bit a set = shift 1 along axis a), fields.py:218-223.
__device__ __host__ inline int e_shift(int c) { return 1 << c; }
__device__ __host__ inline int h_shift(int c) { return 7 & ~(1 << c); }

__global__ void flux_sample_kernel(const SimDev *sims, PlaneDev *planes, int nplanes, int slot) {
    const PlaneDev &p = planes[blockIdx.y];
    const SimDev &s = sims[0];
    if (s.step % p.dec) return;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (blockIdx.x == 0 && threadIdx.x == 0) p.times[slot] = (double)s.step * p.dt;
    if (q >= p.cells) return;
    const int a = q % p.cnt[0], r = q / p.cnt[0];
    const int b = r % p.cnt[1], e = r / p.cnt[1];
    const int i = (p.lo[0] + a) % s.n[0], j = (p.lo[1] + b) % s.n[1], k = (p.lo[2] + e) % s.n[2];
    const int t1 = (p.normal + 1) % 3, t2 = (p.normal + 2) % 3;
    float *out = p.ring + (size_t)slot * 4 * p.cells;
    float *const *Hs = s.pml ? s.H : s.B;
    out[0 * p.cells + q] = centred(s, s.E[t1], e_shift(t1), i, j, k);
    out[1 * p.cells + q] = centred(s, s.E[t2], e_shift(t2), i, j, k);
    out[2 * p.cells + q] = centred(s, Hs[t1], h_shift(t1), i, j, k);
    out[3 * p.cells + q] = centred(s, Hs[t2], h_shift(t2), i, j, k);
}

// exp(i*omega*t)*scale in double for each buffered sample time, E at t and H at t - dt/2.
__global__ void flux_phase_kernel(PlaneDev *planes, int nslots) {
    PlaneDev &p = planes[blockIdx.y];
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nslots * p.nf) return;
    const int m = q / p.nf, f = q % p.nf;
    const double t = p.times[m];
    double sE, cE, sH, cH;
    sincos(p.omega[f] * t, &sE, &cE);
    sincos(p.omega[f] * (t - 0.5 * p.dt), &sH, &cH);
    double *ph = p.phase + (size_t)m * 4 * p.nf;
    ph[2 * f] = cE * p.scale;
    ph[2 * f + 1] = sE * p.scale;
    ph[2 * p.nf + 2 * f] = cH * p.scale;
    ph[2 * p.nf + 2 * f + 1] = sH * p.scale;
}

// acc[comp][f][cell] += sum_m phase[m][E|H][f] * ring[m][comp][cell], in float64.
__global__ void __launch_bounds__(256) flux_accumulate_kernel(PlaneDev *planes, int nslots) {
    PlaneDev &p = planes[blockIdx.z];
    const int comp = blockIdx.y / p.nf, f = blockIdx.y % p.nf;
    if (comp >= 4) return;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= p.cells) return;
    const int side = comp < 2 ? 0 : 1;
    double re = 0.0, im = 0.0;
    for (int m = 0; m < nslots; ++m) {
        const double v = (double)p.ring[((size_t)m * 4 + comp) * p.cells + q];
        const double *ph = p.phase + (size_t)m * 4 * p.nf + side * 2 * p.nf + 2 * f;
        re += ph[0] * v;
        im += ph[1] * v;
    }
    double *a = p.acc + (((size_t)comp * p.nf + f) * p.cells + q) * 2;
    a[0] += re;
    a[1] += im;
}

}  // namespace

// ==========================================================================================
// Host side
// ==========================================================================================
namespace {
int g_blocks_per_cu = 8;
int g_xcd_remap = 1;
int g_num_cu = 0;

int num_cu() {
    if (g_num_cu == 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) return 256;
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return 256;
        g_num_cu = prop.multiProcessorCount;
    }
    return g_num_cu;
}

int persistent_grid(int ntiles) {
    int g = num_cu() * g_blocks_per_cu;
    g = (g / 8) * 8;
    if (g < 8) g = 8;
    if (ntiles < g) {
        g = ((ntiles + 7) / 8) * 8;
        if (g < 8) g = 8;
    }
    return g;
}

struct PlaneHost {
    PlaneDev dev;
    int filled = 0;  // samples in the ring
};
}  // namespace

struct fdtd_sim {
    fdtd_config cfg;
    SimDev host;            // host copy of the descriptor (pointers are device pointers)
    SimDev *dev = nullptr;  // device copy
    hipStream_t stream = nullptr;
    std::vector<void *> allocs;
    std::vector<SrcDev> srcs;
    SrcDev *dev_src = nullptr;
    std::vector<PlaneHost> planes;
    PlaneDev *dev_planes = nullptr;
    uint64_t bytes = 0;
    int ntiles = 0;
    bool dirty = true;
    int host_step = 0;      // host mirror of the device counter

    int alloc(void **p, size_t bytes_) {
        hipError_t e = hipMalloc(p, bytes_);
        if (e != hipSuccess) return fail(FDTD_ENOMEM, "hipMalloc(%zu) failed: %s", bytes_, hipGetErrorString(e));
        allocs.push_back(*p);
        bytes += bytes_;
        return FDTD_OK;
    }
    int volume(float **p, float fill_one = 0.0f);
    int sync_desc() {
        if (!dirty) return FDTD_OK;
        HIP_TRY(hipMemcpyAsync(dev, &host, sizeof(SimDev), hipMemcpyHostToDevice, stream));
        dirty = false;
        return FDTD_OK;
    }
    int pull_desc() {
        // host edits not yet pushed win; otherwise the counters may have advanced on the
        // device, so refresh the host copy before editing it
        if (dirty) return sync_desc();
        HIP_TRY(hipMemcpyAsync(&host, dev, sizeof(SimDev), hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        return FDTD_OK;
    }
};

namespace {
__global__ void fill_kernel(float *p, size_t n, float v) {
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) p[q] = v;
}
}  // namespace

int fdtd_sim::volume(float **p, float fill) {
    const size_t n = (size_t)host.ncell;
    int rc = alloc((void **)p, n * sizeof(float));
    if (rc) return rc;
    if (fill == 0.0f) {
        HIP_TRY(hipMemsetAsync(*p, 0, n * sizeof(float), stream));
    } else {
        fill_kernel<<<(unsigned)((n + 255) / 256), 256, 0, stream>>>(*p, n, fill);
        HIP_TRY(hipGetLastError());
    }
    return FDTD_OK;
}

struct fdtd_batch {
    std::vector<fdtd_sim *> sims;
    SimDev *dev = nullptr;  // contiguous descriptors
    int2 *tiles = nullptr;
    int ntiles = 0;
    hipStream_t stream = nullptr;
};

extern "C" {

const char *fdtd_last_error(void) { return g_error.c_str(); }

int fdtd_device_info(char *text, size_t size) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, dev));
    snprintf(text, size, "%s gcn=%s cu=%d l2=%d", prop.name, prop.gcnArchName,
             prop.multiProcessorCount, prop.l2CacheSize);
    return FDTD_OK;
}

int fdtd_set_launch(int32_t blocks_per_cu, int32_t xcd_remap) {
    if (blocks_per_cu < 1 || blocks_per_cu > 64) return fail(FDTD_EINVAL, "blocks_per_cu %d", blocks_per_cu);
    g_blocks_per_cu = blocks_per_cu;
    g_xcd_remap = xcd_remap ? 1 : 0;
    return FDTD_OK;
}

int fdtd_sim_create(const fdtd_config *cfg, fdtd_sim **out) {
    if (!cfg || !out) return fail(FDTD_EINVAL, "null argument");
    int64_t ncell = 1;
    for (int a = 0; a < 3; ++a) {
        if (cfg->n[a] < 1) return fail(FDTD_EINVAL, "axis %d has %d cells", a, cfg->n[a]);
        if (cfg->invariant[a] && cfg->n[a] != 1)
            return fail(FDTD_EINVAL, "invariant axis %d must have one cell", a);
        ncell *= cfg->n[a];
    }
    if (ncell >= (int64_t)1 << 31) return fail(FDTD_EINVAL, "grid of %lld cells exceeds 2^31", (long long)ncell);
    fdtd_sim *s = new (std::nothrow) fdtd_sim();
    if (!s) return fail(FDTD_ENOMEM, "host allocation");
    s->cfg = *cfg;
    memset(&s->host, 0, sizeof(SimDev));
    SimDev &h = s->host;
    for (int a = 0; a < 3; ++a) {
        h.n[a] = cfg->n[a];
        h.present[a] = cfg->invariant[a] ? 0 : 1;
    }
    h.dtdx = cfg->dtdx;
    h.pml = cfg->pml ? 1 : 0;
    h.ncell = (int)ncell;
    s->ntiles = (int)((ncell + kTile - 1) / kTile);
    int rc = FDTD_OK;
    hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete s;
        return fail(FDTD_EHIP, "hipStreamCreate: %s", hipGetErrorString(e));
    }
#define TRY(x) do { rc = (x); if (rc) { fdtd_sim_destroy(s); return rc; } } while (0)
    TRY(s->alloc((void **)&s->dev, sizeof(SimDev)));
    for (int c = 0; c < 3; ++c) {
        TRY(s->volume(&h.D[c]));
        TRY(s->volume(&h.B[c]));
        TRY(s->volume(&h.E[c]));
        TRY(s->volume((float **)&h.inv[c], 1.0f));
        if (h.pml) {
            TRY(s->volume(&h.H[c]));
            TRY(s->volume(&h.fuB[c]));
            TRY(s->volume(&h.fuD[c]));
            TRY(s->volume(&h.fwE[c]));
            TRY(s->volume(&h.fwH[c]));
        }
    }
    // layer lines default to the identity (kms = sinv = kps = 1, not a member)
    for (int half = 0; half < 2; ++half)
        for (int a = 0; a < 3; ++a) {
            void *p;
            TRY(s->alloc(&p, sizeof(float4) * h.n[a]));
            std::vector<float4> ones(h.n[a], make_float4(1.f, 1.f, 1.f, 0.f));
            e = hipMemcpyAsync(p, ones.data(), sizeof(float4) * h.n[a], hipMemcpyHostToDevice, s->stream);
            if (e != hipSuccess) { fdtd_sim_destroy(s); return fail(FDTD_EHIP, "hipMemcpy: %s", hipGetErrorString(e)); }
            e = hipStreamSynchronize(s->stream);
            if (e != hipSuccess) { fdtd_sim_destroy(s); return fail(FDTD_EHIP, "sync: %s", hipGetErrorString(e)); }
            h.line[half][a] = (const float4 *)p;
        }
#undef TRY
    s->dirty = true;
    *out = s;
    return FDTD_OK;
}

void fdtd_sim_destroy(fdtd_sim *s) {
    if (!s) return;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (void *p : s->allocs) (void)hipFree(p);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

int fdtd_sim_bytes(const fdtd_sim *s, uint64_t *bytes) {
    if (!s || !bytes) return fail(FDTD_EINVAL, "null argument");
    *bytes = s->bytes;
    return FDTD_OK;
}

int fdtd_set_layer(fdtd_sim *s, int32_t half, int32_t axis, const float *kms, const float *sinv,
                   const float *kps, const uint8_t *member) {
    if (!s || half < 0 || half > 1 || axis < 0 || axis > 2 || !kms || !sinv || !kps || !member)
        return fail(FDTD_EINVAL, "fdtd_set_layer: bad argument");
    const int n = s->host.n[axis];
    std::vector<float4> v(n);
    for (int q = 0; q < n; ++q) v[q] = make_float4(kms[q], sinv[q], kps[q], member[q] ? 1.f : 0.f);
    HIP_TRY(hipMemcpyAsync((void *)s->host.line[half][axis], v.data(), sizeof(float4) * n,
                           hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return FDTD_OK;
}

static int upload_volume(fdtd_sim *s, float *dst, const float *src) {
    HIP_TRY(hipMemcpyAsync(dst, src, sizeof(float) * (size_t)s->host.ncell, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return FDTD_OK;
}

int fdtd_set_inverse_epsilon(fdtd_sim *s, int32_t c, const float *vol) {
    if (!s || c < 0 || c > 2 || !vol) return fail(FDTD_EINVAL, "fdtd_set_inverse_epsilon: bad argument");
    return upload_volume(s, (float *)s->host.inv[c], vol);
}

int fdtd_set_offdiagonal(fdtd_sim *s, int32_t row, int32_t partner, const float *vol) {
    if (!s || row < 0 || row > 2 || partner < 0 || partner > 2 || row == partner || !vol)
        return fail(FDTD_EINVAL, "fdtd_set_offdiagonal: bad argument");
    int rc = s->pull_desc();
    if (rc) return rc;
    if (!s->host.u[row][partner]) {
        float *p;
        rc = s->volume(&p);
        if (rc) return rc;
        s->host.u[row][partner] = p;
    }
    s->host.offdiag = 1;
    s->dirty = true;
    return upload_volume(s, (float *)s->host.u[row][partner], vol);
}

int fdtd_add_pole(fdtd_sim *s, const float consts[4], const float *const sigma[3],
                  const float sigma_scalar[3], const int32_t driven[3], int32_t *pole) {
    if (!s || !consts || !sigma || !sigma_scalar || !driven) return fail(FDTD_EINVAL, "fdtd_add_pole: null argument");
    int rc = s->pull_desc();
    if (rc) return rc;
    SimDev &h = s->host;
    if (h.npol >= kMaxPoles) return fail(FDTD_EINVAL, "at most %d poles", kMaxPoles);
    const int q = h.npol;
    for (int k = 0; k < 4; ++k) h.pc[q][k] = consts[k];
    for (int c = 0; c < 3; ++c) {
        h.pdriven[q][c] = driven[c] ? 1 : 0;
        h.psig_s[q][c] = sigma_scalar[c];
        h.psig[q][c] = nullptr;
        if (!driven[c]) continue;
        rc = s->volume(&h.pbuf[q][c][0]);
        if (rc) return rc;
        rc = s->volume(&h.pbuf[q][c][1]);
        if (rc) return rc;
        if (sigma[c]) {
            float *p;
            rc = s->volume(&p);
            if (rc) return rc;
            rc = upload_volume(s, p, sigma[c]);
            if (rc) return rc;
            h.psig[q][c] = p;
        }
    }
    h.npol = q + 1;
    s->dirty = true;
    if (pole) *pole = q;
    return FDTD_OK;
}

int fdtd_add_source(fdtd_sim *s, int32_t field, int32_t comp, const int32_t lo[3], const int32_t cnt[3],
                    const double *amp_re, const double *amp_im, const double *current, int32_t nsteps,
                    double dt) {
    if (!s || field < 0 || field > 1 || comp < 0 || comp > 2 || !lo || !cnt || !amp_re || !amp_im ||
        !current || nsteps < 0)
        return fail(FDTD_EINVAL, "fdtd_add_source: bad argument");
    int64_t npts = 1;
    for (int a = 0; a < 3; ++a) {
        if (cnt[a] < 1 || lo[a] < 0 || lo[a] + cnt[a] > s->host.n[a])
            return fail(FDTD_EINVAL, "source box axis %d [%d,+%d) outside [0,%d)", a, lo[a], cnt[a], s->host.n[a]);
        npts *= cnt[a];
    }
    int rc = s->pull_desc();
    if (rc) return rc;
    SrcDev d{};
    d.field = field;
    d.comp = comp;
    for (int a = 0; a < 3; ++a) { d.lo[a] = lo[a]; d.cnt[a] = cnt[a]; }
    d.ncur = nsteps;
    d.dt = dt;
    void *p;
    rc = s->alloc(&p, sizeof(double) * npts);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(p, amp_re, sizeof(double) * npts, hipMemcpyHostToDevice, s->stream));
    d.amp_re = (const double *)p;
    rc = s->alloc(&p, sizeof(double) * npts);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(p, amp_im, sizeof(double) * npts, hipMemcpyHostToDevice, s->stream));
    d.amp_im = (const double *)p;
    rc = s->alloc(&p, sizeof(double) * 2 * (size_t)(nsteps > 0 ? nsteps : 1));
    if (rc) return rc;
    if (nsteps) HIP_TRY(hipMemcpyAsync(p, current, sizeof(double) * 2 * nsteps, hipMemcpyHostToDevice, s->stream));
    d.cur = (const double *)p;
    s->srcs.push_back(d);
    rc = s->alloc(&p, sizeof(SrcDev) * s->srcs.size());
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(p, s->srcs.data(), sizeof(SrcDev) * s->srcs.size(), hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->host.src = (const SrcDev *)p;
    s->host.nsrc = (int)s->srcs.size();
    s->dirty = true;
    return FDTD_OK;
}

static float *volume_by_name(fdtd_sim *s, const char *name) {
    SimDev &h = s->host;
    auto comp = [](char ch) { return ch == 'x' ? 0 : ch == 'y' ? 1 : ch == 'z' ? 2 : -1; };
    const size_t L = strlen(name);
    if (L == 2) {
        const int c = comp(name[1]);
        if (c < 0) return nullptr;
        switch (name[0]) {
            case 'D': return h.D[c];
            case 'B': return h.B[c];
            case 'E': return h.E[c];
            case 'H': return h.pml ? h.H[c] : nullptr;
        }
        return nullptr;
    }
    if (L == 5 && !strncmp(name, "fu_", 3)) {
        const int c = comp(name[4]);
        if (c < 0 || !h.pml) return nullptr;
        return name[3] == 'B' ? h.fuB[c] : name[3] == 'D' ? h.fuD[c] : nullptr;
    }
    if (L == 6 && !strncmp(name, "f_w_", 4)) {
        const int c = comp(name[5]);
        if (c < 0 || !h.pml) return nullptr;
        return name[4] == 'E' ? h.fwE[c] : name[4] == 'H' ? h.fwH[c] : nullptr;
    }
    int k = -1;
    char e = 0, cc = 0;
    if (sscanf(name, "P[%d].%c%c", &k, &e, &cc) == 3 && !strncmp(name, "P[", 2)) {
        const int c = comp(cc);
        if (k < 0 || k >= h.npol || e != 'E' || c < 0 || !h.pdriven[k][c]) return nullptr;
        return h.pbuf[k][c][h.ppar];
    }
    if (sscanf(name, "P_prev[%d].%c%c", &k, &e, &cc) == 3) {
        const int c = comp(cc);
        if (k < 0 || k >= h.npol || e != 'E' || c < 0 || !h.pdriven[k][c]) return nullptr;
        return h.pbuf[k][c][h.ppar ^ 1];
    }
    return nullptr;
}

int fdtd_upload(fdtd_sim *s, const char *name, const float *host, size_t count) {
    if (!s || !name || !host) return fail(FDTD_EINVAL, "fdtd_upload: null argument");
    int rc = s->pull_desc();
    if (rc) return rc;
    if (count != (size_t)s->host.ncell) return fail(FDTD_EINVAL, "upload %s: %zu floats, grid has %d", name, count, s->host.ncell);
    float *dst = volume_by_name(s, name);
    if (!dst) return fail(FDTD_EINVAL, "no volume named '%s' in this simulation", name);
    return upload_volume(s, dst, host);
}

int fdtd_download(fdtd_sim *s, const char *name, float *host, size_t count) {
    if (!s || !name || !host) return fail(FDTD_EINVAL, "fdtd_download: null argument");
    int rc = s->pull_desc();
    if (rc) return rc;
    if (count != (size_t)s->host.ncell) return fail(FDTD_EINVAL, "download %s: %zu floats, grid has %d", name, count, s->host.ncell);
    float *src = volume_by_name(s, name);
    if (!src) return fail(FDTD_EINVAL, "no volume named '%s' in this simulation", name);
    HIP_TRY(hipMemcpyAsync(host, src, sizeof(float) * count, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return FDTD_OK;
}

int fdtd_get_step(fdtd_sim *s, int32_t *step) {
    if (!s || !step) return fail(FDTD_EINVAL, "null argument");
    int rc = s->pull_desc();
    if (rc) return rc;
    *step = s->host.step;
    return FDTD_OK;
}

int fdtd_set_step(fdtd_sim *s, int32_t step, int32_t h_separated) {
    if (!s || step < 0) return fail(FDTD_EINVAL, "bad argument");
    int rc = s->pull_desc();
    if (rc) return rc;
    s->host.step = step;
    s->host.hsep = h_separated ? 1 : 0;
    s->dirty = true;
    return s->sync_desc();
}

int fdtd_prime_e(fdtd_sim *s) {
    if (!s) return fail(FDTD_EINVAL, "null argument");
    int rc = s->sync_desc();
    if (rc) return rc;
    prime_e_kernel<<<s->ntiles, kTile, 0, s->stream>>>(s->dev, s->ntiles);
    HIP_TRY(hipGetLastError());
    return FDTD_OK;
}

static int launch_h(SimDev *dev, const int2 *tiles, int ntiles, hipStream_t st) {
    const int g = persistent_grid(ntiles);
    hpair_kernel<<<g, kTile, 0, st>>>(dev, tiles, ntiles, g_xcd_remap);
    HIP_TRY(hipGetLastError());
    return FDTD_OK;
}

static int launch_d(SimDev *dev, const int2 *tiles, int ntiles, int count, bool any_offdiag,
                    bool all_offdiag, hipStream_t st) {
    const int g = persistent_grid(ntiles);
    if (!all_offdiag) dpair_kernel<<<g, kTile, 0, st>>>(dev, tiles, ntiles, g_xcd_remap);
    if (any_offdiag) {
        dcurl_kernel<<<g, kTile, 0, st>>>(dev, tiles, ntiles, g_xcd_remap);
        erow_kernel<<<g, kTile, 0, st>>>(dev, tiles, ntiles, g_xcd_remap);
    }
    advance_kernel<<<(count + 63) / 64, 64, 0, st>>>(dev, count);
    HIP_TRY(hipGetLastError());
    return FDTD_OK;
}

int fdtd_step_h(fdtd_sim *s) {
    if (!s) return fail(FDTD_EINVAL, "null argument");
    int rc = s->sync_desc();
    if (rc) return rc;
    return launch_h(s->dev, nullptr, s->ntiles, s->stream);
}

static int flux_flush(fdtd_sim *s, PlaneHost &ph, int plane_index);

static int flux_after_step(fdtd_sim *s, int step_after) {
    if (s->planes.empty()) return FDTD_OK;
    // every plane shares the ring slot index when decimations agree; sample each on its own
    for (size_t q = 0; q < s->planes.size(); ++q) {
        PlaneHost &ph = s->planes[q];
        if (step_after % ph.dev.dec) continue;
        dim3 grid((ph.dev.cells + 255) / 256, 1);
        flux_sample_kernel<<<grid, 256, 0, s->stream>>>(s->dev, s->dev_planes + q, 1, ph.filled);
        HIP_TRY(hipGetLastError());
        ph.filled += 1;
        if (ph.filled == kRing) {
            int rc = flux_flush(s, ph, (int)q);
            if (rc) return rc;
        }
    }
    return FDTD_OK;
}

int fdtd_step_d(fdtd_sim *s) {
    if (!s) return fail(FDTD_EINVAL, "null argument");
    int rc = s->sync_desc();
    if (rc) return rc;
    rc = launch_d(s->dev, nullptr, s->ntiles, 1, s->host.offdiag, s->host.offdiag, s->stream);
    if (rc) return rc;
    s->host.step += 1;  // mirror (the device counter is authoritative)
    s->host.hsep = 1;
    if (s->host.npol) s->host.ppar ^= 1;
    return flux_after_step(s, s->host.step);
}

int fdtd_run(fdtd_sim *s, int32_t nsteps) {
    if (!s || nsteps < 0) return fail(FDTD_EINVAL, "bad argument");
    for (int q = 0; q < nsteps; ++q) {
        int rc = fdtd_step_h(s);
        if (rc) return rc;
        rc = fdtd_step_d(s);
        if (rc) return rc;
    }
    return FDTD_OK;
}

int fdtd_synchronize(fdtd_sim *s) {
    if (!s) return fail(FDTD_EINVAL, "null argument");
    HIP_TRY(hipStreamSynchronize(s->stream));
    return FDTD_OK;
}

int fdtd_add_flux_plane(fdtd_sim *s, int32_t normal, const int32_t lo[3], const int32_t cnt[3],
                        int32_t nf, const double *omega, double scale, int32_t dec, int32_t *plane) {
    if (!s || normal < 0 || normal > 2 || !lo || !cnt || nf < 1 || !omega || dec < 1)
        return fail(FDTD_EINVAL, "fdtd_add_flux_plane: bad argument");
    if (cnt[normal] != 1) return fail(FDTD_EINVAL, "a flux plane is one cell thick along its normal");
    int cells = 1;
    for (int a = 0; a < 3; ++a) {
        if (cnt[a] < 1 || lo[a] < 0 || lo[a] >= s->host.n[a] || cnt[a] > s->host.n[a])
            return fail(FDTD_EINVAL, "flux box axis %d [%d,+%d) outside the grid", a, lo[a], cnt[a]);
        cells *= cnt[a];
    }
    PlaneHost ph;
    PlaneDev &p = ph.dev;
    p.normal = normal;
    for (int a = 0; a < 3; ++a) { p.lo[a] = lo[a]; p.cnt[a] = cnt[a]; }
    p.cells = cells;
    p.nf = nf;
    p.dec = dec;
    p.scale = scale;
    p.dt = s->cfg.dt;
    int rc;
    void *m;
    rc = s->alloc(&m, sizeof(double) * nf);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(m, omega, sizeof(double) * nf, hipMemcpyHostToDevice, s->stream));
    p.omega = (const double *)m;
    rc = s->alloc(&m, sizeof(float) * (size_t)kRing * 4 * cells);
    if (rc) return rc;
    p.ring = (float *)m;
    rc = s->alloc(&m, sizeof(double) * kRing);
    if (rc) return rc;
    p.times = (double *)m;
    rc = s->alloc(&m, sizeof(double) * (size_t)kRing * 4 * nf);
    if (rc) return rc;
    p.phase = (double *)m;
    rc = s->alloc(&m, sizeof(double) * 2 * 4 * (size_t)nf * cells);
    if (rc) return rc;
    HIP_TRY(hipMemsetAsync(m, 0, sizeof(double) * 2 * 4 * (size_t)nf * cells, s->stream));
    p.acc = (double *)m;
    s->planes.push_back(ph);
    rc = s->alloc(&m, sizeof(PlaneDev) * s->planes.size());
    if (rc) return rc;
    std::vector<PlaneDev> devs;
    for (auto &x : s->planes) devs.push_back(x.dev);
    HIP_TRY(hipMemcpyAsync(m, devs.data(), sizeof(PlaneDev) * devs.size(), hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->dev_planes = (PlaneDev *)m;
    if (plane) *plane = (int)s->planes.size() - 1;
    return FDTD_OK;
}

static int flux_flush(fdtd_sim *s, PlaneHost &ph, int q) {
    if (ph.filled == 0) return FDTD_OK;
    PlaneDev &p = ph.dev;
    dim3 g1((ph.filled * p.nf + 255) / 256, 1);
    flux_phase_kernel<<<g1, 256, 0, s->stream>>>(s->dev_planes + q, ph.filled);
    dim3 g2((p.cells + 255) / 256, 4 * p.nf, 1);
    flux_accumulate_kernel<<<g2, 256, 0, s->stream>>>(s->dev_planes + q, ph.filled);
    HIP_TRY(hipGetLastError());
    ph.filled = 0;
    return FDTD_OK;
}

int fdtd_flux_read(fdtd_sim *s, int32_t plane, double *out, size_t count) {
    if (!s || plane < 0 || plane >= (int)s->planes.size() || !out) return fail(FDTD_EINVAL, "bad argument");
    PlaneHost &ph = s->planes[plane];
    const size_t need = (size_t)2 * 4 * ph.dev.nf * ph.dev.cells;
    if (count != need) return fail(FDTD_EINVAL, "flux read needs %zu doubles, got %zu", need, count);
    int rc = flux_flush(s, ph, plane);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(out, ph.dev.acc, sizeof(double) * need, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return FDTD_OK;
}

int fdtd_time_pairs(fdtd_sim *s, int32_t repeats, float out_ms[2]) {
    if (!s || repeats < 1 || !out_ms) return fail(FDTD_EINVAL, "bad argument");
    int rc = s->sync_desc();
    if (rc) return rc;
    hipEvent_t a, b, c;
    HIP_TRY(hipEventCreate(&a));
    HIP_TRY(hipEventCreate(&b));
    HIP_TRY(hipEventCreate(&c));
    float th = 0.f, td = 0.f;
    for (int r = 0; r < repeats; ++r) {
        HIP_TRY(hipEventRecord(a, s->stream));
        rc = launch_h(s->dev, nullptr, s->ntiles, s->stream);
        if (rc) return rc;
        HIP_TRY(hipEventRecord(b, s->stream));
        const int g = persistent_grid(s->ntiles);
        dpair_kernel<<<g, kTile, 0, s->stream>>>(s->dev, nullptr, s->ntiles, g_xcd_remap);
        HIP_TRY(hipEventRecord(c, s->stream));
        advance_kernel<<<1, 64, 0, s->stream>>>(s->dev, 1);
        HIP_TRY(hipEventSynchronize(c));
        float x = 0.f, y = 0.f;
        HIP_TRY(hipEventElapsedTime(&x, a, b));
        HIP_TRY(hipEventElapsedTime(&y, b, c));
        th += x;
        td += y;
    }
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->host.step += repeats;
    s->host.hsep = 1;
    if (s->host.npol && (repeats & 1)) s->host.ppar ^= 1;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    (void)hipEventDestroy(c);
    out_ms[0] = th / repeats;
    out_ms[1] = td / repeats;
    return FDTD_OK;
}

// ------------------------------------------------------------------------------------------
// Batch
// ------------------------------------------------------------------------------------------
int fdtd_batch_create(fdtd_sim *const *sims, int32_t count, fdtd_batch **out) {
    if (!sims || count < 1 || !out) return fail(FDTD_EINVAL, "fdtd_batch_create: bad argument");
    fdtd_batch *b = new (std::nothrow) fdtd_batch();
    if (!b) return fail(FDTD_ENOMEM, "host allocation");
    std::vector<SimDev> descs;
    std::vector<int2> tiles;
    for (int q = 0; q < count; ++q) {
        fdtd_sim *s = sims[q];
        if (!s) { delete b; return fail(FDTD_EINVAL, "null simulation %d", q); }
        if (!s->planes.empty()) { delete b; return fail(FDTD_EINVAL, "batched simulations carry no flux planes here"); }
        int rc = s->pull_desc();
        if (rc) { delete b; return rc; }
        b->sims.push_back(s);
        descs.push_back(s->host);
        for (int t = 0; t < s->ntiles; ++t) tiles.push_back(make_int2(q, t * kTile));
    }
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { delete b; return fail(FDTD_EHIP, "hipStreamCreate: %s", hipGetErrorString(e)); }
    if (hipMalloc((void **)&b->dev, sizeof(SimDev) * count) != hipSuccess ||
        hipMalloc((void **)&b->tiles, sizeof(int2) * tiles.size()) != hipSuccess) {
        fdtd_batch_destroy(b);
        return fail(FDTD_ENOMEM, "batch descriptors");
    }
    if (hipMemcpy(b->dev, descs.data(), sizeof(SimDev) * count, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->tiles, tiles.data(), sizeof(int2) * tiles.size(), hipMemcpyHostToDevice) != hipSuccess) {
        fdtd_batch_destroy(b);
        return fail(FDTD_EHIP, "batch descriptor upload");
    }
    b->ntiles = (int)tiles.size();
    *out = b;
    return FDTD_OK;
}

void fdtd_batch_destroy(fdtd_batch *b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    // hand the advanced counters back to each simulation's own descriptor
    if (b->dev) {
        std::vector<SimDev> descs(b->sims.size());
        if (hipMemcpy(descs.data(), b->dev, sizeof(SimDev) * descs.size(), hipMemcpyDeviceToHost) == hipSuccess)
            for (size_t q = 0; q < descs.size(); ++q) {
                b->sims[q]->host = descs[q];
                b->sims[q]->dirty = true;
            }
        (void)hipFree(b->dev);
    }
    if (b->tiles) (void)hipFree(b->tiles);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

int fdtd_batch_run(fdtd_batch *b, int32_t nsteps) {
    if (!b || nsteps < 0) return fail(FDTD_EINVAL, "bad argument");
    bool any_off = false, all_off = true;
    for (auto *s : b->sims) {
        any_off |= s->host.offdiag != 0;
        all_off &= s->host.offdiag != 0;
    }
    for (int q = 0; q < nsteps; ++q) {
        int rc = launch_h(b->dev, b->tiles, b->ntiles, b->stream);
        if (rc) return rc;
        rc = launch_d(b->dev, b->tiles, b->ntiles, (int)b->sims.size(), any_off, all_off, b->stream);
        if (rc) return rc;
    }
    return FDTD_OK;
}

int fdtd_batch_synchronize(fdtd_batch *b) {
    if (!b) return fail(FDTD_EINVAL, "null argument");
    HIP_TRY(hipStreamSynchronize(b->stream));
    return FDTD_OK;
}

}  // extern "C"
