// MI355X (gfx950) FDTD time-stepping engine: fused Yee/UPML/constitutive/ADE/deposit kernels
// and the float64 flux transform, behind the C ABI of include/meep_fdtd.h.
//
// Arithmetic contract (reference docs/design/floating-point.md and stepping.py): every
// float32 operation rounds once in the reference's association order, no contraction
// (built with -ffp-contract=off), subnormals kept (-fno-gpu-flush-denormals-to-zero). The
// deposit and the off-diagonal row are formed in double and rounded once.
//
// Device layout: every volume x fastest, element (i,j,k) at (k*ny+j)*pitch+i with the row
// pitch nx rounded up to a multiple of 4 floats, every axis periodic (an absorber is a
// material, reference stepping.py:2984). One `SimDev` descriptor per simulation lives in device
// memory. Kernels walk a list of work items of kThreads runs of V x-adjacent cells (each thread
// owns one run of one row), so a single run and a batch of uneven grids run the same kernels.
// A grid's runs split into its box -- the whole rows whose y and z are on no absorber layer
// line, where only the x line's forms can hold (off every line the layer forms are exactly the
// lossless ones) -- and a shell of the rows on a y or z line; where no simulation runs the
// general forms the two are stepped by two kernels, each with the registers its forms need. The launch is one workgroup per item or a persistent
// grid; with the XCD remap, workgroup b (which the dispatcher places on XCD b%8) takes items
// from the b%8-th contiguous eighth of the list, so the neighbour rows and planes an item reads
// were fetched into that XCD's own L2 by the items just before it. Step counters live on the device and one wave
// advances them after each D pair, so a step sequence carries no host state.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <xmmintrin.h>

// 1 in the build whose kernels flush float32 and float64 subnormals (denormal modes at flush)
#ifndef FDTD_FLUSH_SUBNORMALS
#define FDTD_FLUSH_SUBNORMALS 0
#endif
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <map>
#include <new>
#include <string>
#include <unordered_map>
#include <utility>
#include <numeric>
#include <vector>

#include "../../include/meep_fdtd.h"

namespace {

constexpr int kThreads = 256;   // threads per workgroup
constexpr int kPad = 4;         // row pitch multiple (floats)
constexpr int kMaxPoles = FDTD_MAX_POLES;
constexpr int kLosslessPoles = 4;  // poles the lossless periodic instantiations step; a
                                   // simulation with more runs the general ones
constexpr int kRing = 128;      // flux samples buffered before one float64 transform pass
constexpr int kMaxRuns = 2;     // runs of a layer membership line held in the descriptor
constexpr int kFoldItem = 1024;  // ghost-plane cells per fold_kernel workgroup
constexpr int kMaxPeriod = 1024;  // steps a decimated schedule's captured period may hold

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

// One source box: deposit into D (field 0) or B (field 1) component `comp`, or (field 2) an
// electric polarization source, whose dipole the constitutive update subtracts from D.
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
    int pitch;      // floats per stored row: nx rounded up to kPad
    int present[3];
    float dtdx;
    int pml;
    int offdiag;
    int npol;
    int nsrc;
    int npsrc;      // electric polarization (integrated) sources among them
    int step;       // completed steps (driver.py:3592)
    int hsep;       // MEEP's one-time H start done (stepping.py:2789)
    int ppar;       // which P buffer holds P^n
    int ncell;      // nx*ny*nz
    int npad;       // pitch*ny*nz, the allocated floats per volume
    double dt;
    float *D[3], *B[3], *E[3], *H[3];
    float *fuB[3], *fuD[3], *fwE[3], *fwH[3];
    const float *inv[3];
    // inverse epsilon as an index into a table of its distinct words, when it has few: the D
    // pair streams 1 or 2 bytes per cell instead of 4 (the words are the same)
    const uint8_t *inv8[3];
    const uint16_t *inv16[3];
    const float *invtab[3];
    int invn[3];  // words in invtab[c]
    const float *u[3][3];
    const float4 *line[2][3];  // (kms, sinv, kps, member) per row, [half][axis]
    int mr[2][3][2 * kMaxRuns];  // the membership lines as up to kMaxRuns [lo, hi) runs
    const int2 *prow;  // per row (j + k*ny): the x range [lo, hi) where a pole may hold or
                       // gain polarization (empty: lo == hi)
    float pc[kMaxPoles][4];
    int pdriven[kMaxPoles][3];
    float *pbuf[kMaxPoles][3][2];
    const float *psig[kMaxPoles][3];
    float psig_s[kMaxPoles][3];
    const SrcDev *src;
    int src_lo[3][3], src_hi[3][3];  // per field (D, B, polarization): union of the boxes, [lo, hi)
    // per field (0 = D, 1 = B) and component: a conductivity's condfac and condinv words and,
    // under an absorber, its f_cond history; null where the component is lossless
    const float *cf[2][3], *ci[2][3];
    float *fcond[2][3];
    int cond;      // any conductivity
    int metal[3];  // metallic (PEC) axes
    // mirror folds (grid.py:404-575): 0 none, 1 over metallic boundaries (no stored far wall:
    // a zero ghost past the top), 2 over periodic ones (one stored slot past the window); the
    // plane's phase (+1 even, -1 odd) and, for 2, the row the far ghost images
    int fold[3], fphase[3], frow[3];
    // complex storage (MEEP's two realnum parts, stepped alike with the same real
    // coefficients): which part this descriptor steps and the other part's descriptor in the
    // same launch; per axis the Bloch connection of the word read across the window edge (0
    // none, 1 copy, 2 negation, 3 a product by the phase ph, its parts rounded to float32)
    int part;
    const SimDev *other;
    int pkind[3];
    float ph[3][2];
    // The work split of every fused launch: the interior box [x0,x1) x [j0,j1) x [k0,k1) holds
    // cells on no layer line of either lattice (x0, x1 multiples of 4), where the layer forms
    // are exactly the lossless ones and every H read is B's word; every other run of V cells
    // with a real cell is listed (its padded index) in shell[V = 1, 2, 4 -> 0, 1, 2].
    int ibox[6];
    const int *shell[3];
};

// ------------------------------------------------------------------------------------------
// Tile walk and the step counters
// ------------------------------------------------------------------------------------------
struct TileWalk {
    int next, stride, end;
    int xcd, sb;  // super-block order: tiles in blocks of `sb`, block q on XCD q % 8
};

// Which tiles this workgroup takes, in order. `remap` 0: tile b, b+G, ...; 1: each XCD takes a
// contiguous eighth of the list; sb >= 2: the list is cut into super-blocks of sb tiles dealt
// round-robin to the 8 XCDs, so an XCD's next tile usually follows its last one (the j-1 rows
// it reads are in its own L2) while the eight XCDs still sweep neighbouring memory together.
// With sb = tiles per plane / 8 the k-1 plane is on the same XCD too.
__device__ inline TileWalk tile_walk(int ntiles, int remap) {
    const int G = gridDim.x, b = blockIdx.x;
    TileWalk w;
    w.xcd = 0;
    w.sb = 0;
    if (remap == 1 && (G % 8) == 0) {
        const int x = b % 8, s = b / 8, S = G / 8;
        w.next = (int)(((long long)ntiles * x) / 8) + s;
        w.stride = S;
        w.end = (int)(((long long)ntiles * (x + 1)) / 8);
    } else if (remap >= 2 && (G % 8) == 0) {
        w.xcd = b % 8;
        w.sb = remap;
        w.next = b / 8;
        w.stride = G / 8;
        w.end = ((ntiles + 8 * remap - 1) / (8 * remap)) * remap;  // slots per XCD
    } else {
        w.next = b;
        w.stride = G;
        w.end = ntiles;
    }
    return w;
}

// The tile of walk position g (may be >= ntiles in super-block order: skipped).
__device__ inline int tile_at(const TileWalk &w, int g) {
    return w.sb ? ((g / w.sb) * 8 + w.xcd) * w.sb + g % w.sb : g;
}

// Advance each simulation's step counter, H-start flag and P parity by `by` steps (one wave;
// kept out of the fused kernels so no workgroup of a launch sees a counter change). Eager steps
// advance by 1 after each D pair; a captured sequence of steps advances once, at its end, by
// its length, and its kernels read the counters through the step offset they were launched
// with (Clock).
__global__ void advance_kernel(SimDev *sims, int count, int by) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= count || by <= 0) return;
    sims[q].step += by;
    sims[q].hsep = 1;
    if (sims[q].npol) sims[q].ppar ^= by & 1;
}

// The counters a kernel launched `adv` steps after the descriptor's last advance works with:
// the step, whether H has started (update_H has run once), and the P buffer parity.
struct Clock {
    int step, hsep, ppar;
};

__device__ inline Clock clock_of(const SimDev &s, int adv) {
    Clock c;
    c.step = s.step + adv;
    c.hsep = (s.hsep || adv > 0) ? 1 : 0;
    c.ppar = s.npol ? (s.ppar ^ (adv & 1)) : s.ppar;
    return c;
}

// V x-adjacent cells (i0..i0+V-1) of one stored row; cells with i >= nx are row padding,
// computed on garbage and never read as a neighbour.
struct Run {
    int at;   // padded index of element 0
    int row;  // padded index of the row's cell 0
    int i0, j, k;
    int yp, ym, zp, zm;  // offsets to the +-1 rows along y and z (periodic)
};

__device__ inline Run locate(const SimDev &s, int at) {
    Run q;
    const int pitch = s.pitch, ny = s.n[1], nz = s.n[2];
    const int r = at / pitch;
    q.at = at;
    q.i0 = at - r * pitch;
    q.row = r * pitch;
    q.k = r / ny;
    q.j = r - q.k * ny;
    const int plane = pitch * ny;
    q.yp = (q.j + 1 < ny) ? pitch : -(ny - 1) * pitch;
    q.ym = (q.j > 0) ? -pitch : (ny - 1) * pitch;
    q.zp = (q.k + 1 < nz) ? plane : -(nz - 1) * plane;
    q.zm = (q.k > 0) ? -plane : (nz - 1) * plane;
    return q;
}

// A run given by its coordinates (i0 a multiple of V inside the stored row).
__device__ inline Run locate_ijk(const SimDev &s, int i0, int j, int k) {
    Run q;
    const int pitch = s.pitch, ny = s.n[1], nz = s.n[2];
    q.i0 = i0;
    q.j = j;
    q.k = k;
    q.row = (k * ny + j) * pitch;
    q.at = q.row + i0;
    const int plane = pitch * ny;
    q.yp = (j + 1 < ny) ? pitch : -(ny - 1) * pitch;
    q.ym = (j > 0) ? -pitch : (ny - 1) * pitch;
    q.zp = (k + 1 < nz) ? plane : -(nz - 1) * plane;
    q.zm = (k > 0) ? -plane : (nz - 1) * plane;
    return q;
}

template <int V>
struct VF {
    float v[V];
};

// Every field pointer the kernels follow is device memory: say so, so that the accesses are
// global (vector-memory only) rather than flat ones, which also count against the scalar
// memory counter and make each wait for a descriptor word wait for every field load in flight.
#define GMEM __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ GMEM T *gp(T *p) { return (GMEM T *)p; }
template <class T>
__device__ __forceinline__ const GMEM T *gp(const T *p) { return (const GMEM T *)p; }

// One work item of a fused launch: (simulation, kind, first, count) -- kind 0 the interior
// box's runs first .. first+count-1 (numbered x fastest, then y, then z), kind 1 the shell
// list's entries first .. first+count-1. Thread t of the workgroup takes entry first + t.
template <int V>
__device__ inline Run work_run(const SimDev &s, const int4 &w, int t) {
    if (w.y == 0) {
        const int q = w.z + t;
        const int nqx = (s.ibox[1] - s.ibox[0]) / V, nj = s.ibox[3] - s.ibox[2];
        const int r = q / nqx;
        const int kk = r / nj;
        return locate_ijk(s, s.ibox[0] + (q - r * nqx) * V, s.ibox[2] + (r - kk * nj),
                          s.ibox[4] + kk);
    }
    return locate(s, gp(s.shell[V == 4 ? 2 : V - 1])[w.z + t]);
}

typedef float f4v __attribute__((ext_vector_type(4)));
typedef float f2v __attribute__((ext_vector_type(2)));

template <int V>
__device__ inline VF<V> ldv(const float *p, int at) {
    VF<V> r;
    const GMEM float *g = gp(p) + at;
    if constexpr (V == 4) {
        const f4v t = *reinterpret_cast<const GMEM f4v *>(g);
        r.v[0] = t.x; r.v[1] = t.y; r.v[2] = t.z; r.v[3] = t.w;
    } else if constexpr (V == 2) {
        const f2v t = *reinterpret_cast<const GMEM f2v *>(g);
        r.v[0] = t.x; r.v[1] = t.y;
    } else {
        r.v[0] = *g;
    }
    return r;
}

// Load a run read exactly once in the launch (no neighbour reuses it): non-temporal, so it
// does not push out of the caches the neighbour rows and planes other loads come back for.
template <int V>
__device__ inline VF<V> ldv_once(const float *p, int at) {
    VF<V> r;
    const GMEM float *g = gp(p) + at;
    if constexpr (V == 4) {
        const f4v t = __builtin_nontemporal_load(reinterpret_cast<const GMEM f4v *>(g));
        r.v[0] = t.x; r.v[1] = t.y; r.v[2] = t.z; r.v[3] = t.w;
    } else if constexpr (V == 2) {
        const f2v t = __builtin_nontemporal_load(reinterpret_cast<const GMEM f2v *>(g));
        r.v[0] = t.x; r.v[1] = t.y;
    } else {
        r.v[0] = __builtin_nontemporal_load(g);
    }
    return r;
}

// Store a run; with NT the store is marked non-temporal (streamed past the caches, for the
// outputs no later read of the same launch wants).

template <int V, bool NT = false>
__device__ inline void stv(float *p, int at, const VF<V> &x) {
    GMEM float *g = gp(p) + at;
    if constexpr (V == 4) {
        const f4v v = {x.v[0], x.v[1], x.v[2], x.v[3]};
        if constexpr (NT) __builtin_nontemporal_store(v, reinterpret_cast<GMEM f4v *>(g));
        else *reinterpret_cast<GMEM f4v *>(g) = v;
    } else if constexpr (V == 2) {
        const f2v v = {x.v[0], x.v[1]};
        if constexpr (NT) __builtin_nontemporal_store(v, reinterpret_cast<GMEM f2v *>(g));
        else *reinterpret_cast<GMEM f2v *>(g) = v;
    } else {
        if constexpr (NT) __builtin_nontemporal_store(x.v[0], g);
        else *g = x.v[0];
    }
}

// The +1 (x-up) neighbours: v[m+1] inside the run, the next run's first cell after it, and
// the row's cell 0 past cell nx-1 (periodic wrap).
template <int V>
__device__ inline VF<V> x_up(const SimDev &s, const Run &q, const float *f, const VF<V> &v) {
    const int nx = s.n[0];
    VF<V> r;
#pragma unroll
    for (int m = 0; m + 1 < V; ++m) r.v[m] = v.v[m + 1];
    const GMEM float *gf = gp(f);
    r.v[V - 1] = (q.i0 + V < nx) ? gf[q.at + V] : gf[q.row];
    if (q.i0 + V > nx && q.i0 < nx) {
        const float w0 = gf[q.row];
        const int last = nx - 1 - q.i0;
#pragma unroll
        for (int m = 0; m < V; ++m)
            if (m == last) r.v[m] = w0;
    }
    return r;
}

// The -1 (x-down) neighbours: v[m-1], the previous run's last cell, cell nx-1 before cell 0.
template <int V>
__device__ inline VF<V> x_down(const SimDev &s, const Run &q, const float *f, const VF<V> &v) {
    VF<V> r;
#pragma unroll
    for (int m = 1; m < V; ++m) r.v[m] = v.v[m - 1];
    const GMEM float *gf = gp(f);
    r.v[0] = q.i0 > 0 ? gf[q.at - 1] : gf[q.row + s.n[0] - 1];
    return r;
}

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
        GMEM float *gfu = gp(fu);
        const float prev = gfu[at];
        float u = lsig.x * prev;
        u = u - curl;
        u = lsig.y * u;
        gfu[at] = u;
        float v = lsigu.x * f;
        v = v + u;
        v = v - prev;
        return lsigu.y * v;
    }
    float v = lsig.x * f;
    v = v - curl;
    return lsig.y * v;
}

// A component with a conductivity (stepping.py:2820-2841 without an absorber, :2844-2983 under
// one; MEEP step_generic.cpp:84-210): X = curl, cf = condfac, ci = condinv. On the dsig layer's
// cells the f_cond stage fcnd' = ((cf*fcnd) - X)*ci runs first and its increment fcnd' - fcnd,
// formed before it is added, replaces -X in the layer's recurrences; on the dsigu layer's cells
// alone the split field is fu' = ((cf*fu) - X)*ci; elsewhere f' = ((cf*f) - X)*ci.
__device__ inline float curl_apply_cond(float f, float curl, float cfv, float civ, bool in_s,
                                        bool in_u, float4 ls, float4 lu, float *fu, float *fc,
                                        int at) {
    if (in_s) {
        GMEM float *gfc = gp(fc);
        const float old = gfc[at];
        float n = cfv * old;
        n = n - curl;
        n = n * civ;
        gfc[at] = n;
        const float inc = n - old;
        if (in_u) {
            GMEM float *gfu = gp(fu);
            const float prev = gfu[at];
            float u = ls.x * prev;
            u = u + inc;
            u = u * ls.y;
            gfu[at] = u;
            float v = lu.x * f;
            v = v + u;
            v = v - prev;
            return v * lu.y;
        }
        float v = ls.x * f;
        v = v + inc;
        return v * ls.y;
    }
    if (in_u) {
        GMEM float *gfu = gp(fu);
        const float prev = gfu[at];
        float u = cfv * prev;
        u = u - curl;
        u = u * civ;
        gfu[at] = u;
        float v = lu.x * f;
        v = v + u;
        v = v - prev;
        return v * lu.y;
    }
    float v = cfv * f;
    v = v - curl;
    return v * civ;
}

__device__ inline bool plus_zero(double v) { return __double_as_longlong(v) == 0; }

// Sources of one field/component deposited into the run's values (sources.py:1357,
// host_writes.py:140): f = float(double(f) - ((amp_re*J_re - amp_im*J_im)*dt)), in list order.
template <int V, bool GEN>
__device__ inline void deposit_run(const SimDev &s, int step, int field, int comp, const Run &q,
                                   VF<V> &x) {
    // GEN: a conductive component scales the current by its condinv word, in the same double
    // expression (MEEP step.cpp:299-307, sources.py:1357-1400)
    const float *ci = GEN ? s.ci[field][comp] : nullptr;
    // one test against the union of this field's boxes before any per-source work
    if (q.j < s.src_lo[field][1] || q.j >= s.src_hi[field][1] || q.k < s.src_lo[field][2] ||
        q.k >= s.src_hi[field][2] || q.i0 + V <= s.src_lo[field][0] || q.i0 >= s.src_hi[field][0])
        return;
    for (int n = 0; n < s.nsrc; ++n) {
        const GMEM SrcDev &d = gp(s.src)[n];
        if (d.field != field || d.comp != comp) continue;
        const int b = q.j - d.lo[1], e = q.k - d.lo[2];
        if (b < 0 || e < 0 || b >= d.cnt[1] || e >= d.cnt[2]) continue;
        if (q.i0 + V <= d.lo[0] || q.i0 >= d.lo[0] + d.cnt[0]) continue;
        if (step >= d.ncur) continue;
        const double cre = gp(d.cur)[2 * step], cim = gp(d.cur)[2 * step + 1];
        if (plus_zero(cre) && plus_zero(cim)) continue;
#pragma unroll
        for (int m = 0; m < V; ++m) {
            const int a = q.i0 + m - d.lo[0];
            if (a < 0 || a >= d.cnt[0]) continue;
            const int p = (e * d.cnt[1] + b) * d.cnt[0] + a;
            double r = gp(d.amp_re)[p] * cre;
            r = r - gp(d.amp_im)[p] * cim;
            r = r * d.dt;
            if constexpr (GEN) {
                if (ci) {  // a 0 word: the cell's chunk keeps no conductivity (MEEP's NULL)
                    const float cw = gp(ci)[q.at + m];
                    if (cw != 0.0f) r = r * (double)cw;
                }
            }
            double w = (double)x.v[m];
            w = w - r;
            x.v[m] = (float)w;
        }
    }
}

// The electric polarization (integrated) sources of component `comp` at the run: MEEP's
// f_minus_p (update_eh.cpp:66-140, sources.py:1697-1747). The dipole P, the step's table word
// (at (n+1)*dt), is subtracted from the copy of D that the constitutive update reads:
// x = float(double(x) - (amp_re*P_re - amp_im*P_im)), once per source in list order, with no
// dt and no conductivity; D itself keeps its word.
template <int V>
__device__ inline void polarize_run(const SimDev &s, int step, int comp, const Run &q, VF<V> &x) {
    if (q.j < s.src_lo[2][1] || q.j >= s.src_hi[2][1] || q.k < s.src_lo[2][2] ||
        q.k >= s.src_hi[2][2] || q.i0 + V <= s.src_lo[2][0] || q.i0 >= s.src_hi[2][0])
        return;
    for (int n = 0; n < s.nsrc; ++n) {
        const GMEM SrcDev &d = gp(s.src)[n];
        if (d.field != 2 || d.comp != comp) continue;
        const int b = q.j - d.lo[1], e = q.k - d.lo[2];
        if (b < 0 || e < 0 || b >= d.cnt[1] || e >= d.cnt[2]) continue;
        if (q.i0 + V <= d.lo[0] || q.i0 >= d.lo[0] + d.cnt[0]) continue;
        if (step >= d.ncur) continue;
        const double pre = gp(d.cur)[2 * step], pim = gp(d.cur)[2 * step + 1];
        if (plus_zero(pre) && plus_zero(pim)) continue;
#pragma unroll
        for (int m = 0; m < V; ++m) {
            const int a = q.i0 + m - d.lo[0];
            if (a < 0 || a >= d.cnt[0]) continue;
            const int p = (e * d.cnt[1] + b) * d.cnt[0] + a;
            double r = gp(d.amp_re)[p] * pre;
            r = r - gp(d.amp_im)[p] * pim;
            double w = (double)x.v[m];
            w = w - r;
            x.v[m] = (float)w;
        }
    }
}

__device__ inline int coord_of(const Run &q, int axis, int m) {
    return axis == 0 ? q.i0 + m : (axis == 1 ? q.j : q.k);
}

// Does the one-rank chunk owning row c of `axis` (sub-lattice `half`) carry that axis's layer?
// Pure comparisons against the runs in the descriptor: no memory access. Off the layer every
// coefficient is exactly 1 (the region plan's precondition, chunk_map.py:1031), so the forms
// there need no coefficient words at all.
__device__ inline bool member(const SimDev &s, int half, int axis, int c) {
    const int *r = s.mr[half][axis];
    bool in = false;
#pragma unroll
    for (int t = 0; t < kMaxRuns; ++t) in |= (c >= r[2 * t]) & (c < r[2 * t + 1]);
    return in;
}

__device__ inline float4 coeff(const SimDev &s, int half, int axis, int c, bool in) {
    if (!in) return make_float4(1.f, 1.f, 1.f, 0.f);
    const GMEM float *l = reinterpret_cast<const GMEM float *>(gp(s.line[half][axis]) + c);
    return make_float4(l[0], l[1], l[2], l[3]);
}

// Metallic walls (stepping.py:140-146, :2429-2530, :2589-2626, :3094-3183). Past a perfect
// conductor there is no field, so a curl operand beyond the wall reads zero; the samples ON a
// wall -- stored cell 0 along the axis of a component whose Yee shift there is 0 (B_c along c,
// D_c along the other two) -- are not stepped (their curl is dropped) and are held at zero after
// the sub-step's sources.
__device__ inline bool on_wall(const SimDev &s, const Run &q, int axis, int m) {
    return s.metal[axis] && coord_of(q, axis, m) == 0;
}

// The ghosts of a mirror-folded axis (stepping.py:2589-2626): stored cell 0 of a component
// whose Yee shift on the axis is 0 and, over periodic boundaries, the last slot of a shift-1
// one. Their curl is dropped and the fold kernel writes them from their images. family 0: D
// (shift 1 on its own axis), 1: B (shift 0 on its own axis).
__device__ inline bool ghost_at(const SimDev &s, const Run &q, int family, int comp, int m) {
    bool g = false;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (!s.fold[a]) continue;
        const bool shift0 = family ? (a == comp) : (a != comp);
        const int c = coord_of(q, a, m);
        g |= shift0 && c == 0;
        g |= s.fold[a] == 2 && !shift0 && c == s.n[a] - 1;
    }
    return g;
}

// MEEP's connection of the word a stencil reads at the SAME stored index across a Bloch-phased
// axis's window edge (plane 0; the engine stores MEEP's owned window there): below the window
// conj(eikna), above it eikna (stepping._bloch_ghost_operand :2545-2586). A phase within 1e-13
// of +1 is a copy, of -1 a negation (boundaries.cpp:359-366); otherwise complex<realnum>(phase)
// times the word, (a*c) - (b*d) for the real part and (a*d) + (b*c) for the imaginary one,
// each operation rounded on its own (step.cpp:183-192): that reads the other part's word.
__device__ inline float connect_value(const SimDev &s, int axis, bool below, float mine,
                                      float other) {
    const int k = s.pkind[axis];
    if (k == 2) return -mine;
    if (k != 3) return mine;
    const float a = s.ph[axis][0], b = below ? -s.ph[axis][1] : s.ph[axis][1];
    const float x = a * mine, y = b * other;
    return s.part == 0 ? x - y : x + y;
}

__device__ inline float connect_word(const SimDev &s, int axis, bool below, float mine,
                                     const float *other, int at) {
    return connect_value(s, axis, below, mine, s.pkind[axis] == 3 ? gp(other)[at] : 0.0f);
}

// Past the top of a metallic axis, walled or folded, a +1 curl operand reads zero.
__device__ inline bool zero_top(const SimDev &s, int axis) {
    return s.metal[axis] || s.fold[axis] == 1;
}

template <int V>
__device__ inline void zero_vf(VF<V> &x) {
#pragma unroll
    for (int m = 0; m < V; ++m) x.v[m] = 0.0f;
}

// The H pair's +1 operands past the high wall.
template <int V>
__device__ inline void metal_up(const SimDev &s, const Run &q, VF<V> (&en)[3][3]) {
    if (zero_top(s, 1) && q.j == s.n[1] - 1) { zero_vf(en[0][1]); zero_vf(en[2][1]); }
    if (zero_top(s, 2) && q.k == s.n[2] - 1) { zero_vf(en[0][2]); zero_vf(en[1][2]); }
    if (zero_top(s, 0)) {
#pragma unroll
        for (int m = 0; m < V; ++m)
            if (q.i0 + m == s.n[0] - 1) en[1][0].v[m] = en[2][0].v[m] = 0.0f;
    }
}

// The D pair's -1 operands past the low wall.
template <int V>
__device__ inline void metal_down(const SimDev &s, const Run &q, VF<V> (&hn)[3][3]) {
    if (s.metal[1] && q.j == 0) { zero_vf(hn[0][1]); zero_vf(hn[2][1]); }
    if (s.metal[2] && q.k == 0) { zero_vf(hn[0][2]); zero_vf(hn[1][2]); }
    if (s.metal[0]) {
#pragma unroll
        for (int m = 0; m < V; ++m)
            if (q.i0 + m == 0) hn[1][0].v[m] = hn[2][0].v[m] = 0.0f;
    }
}

// ------------------------------------------------------------------------------------------
// H pair: step_B -> magnetic deposit -> update_H.
// en[comp][axis]: E[comp] at the +1 neighbour along axis (only the pairs a curl reads).
// ------------------------------------------------------------------------------------------
template <int V, int C, bool PML, bool NT, bool GEN>
__device__ inline void hpair_component(const SimDev &s, const Clock &ck, const Run &q,
                                       const VF<V> (&e)[3], const VF<V> (&en)[3][3], VF<V> b) {
    constexpr int first = (C + 2) % 3, fa = (C + 1) % 3;
    constexpr int second = (C + 1) % 3, sa = (C + 2) % 3;
    constexpr int dsig = (C + 1) % 3, dsigu = (C + 2) % 3;
#pragma unroll
    for (int m = 0; m < V; ++m) {
        float g1 = e[first].v[m], g2 = e[second].v[m];
        if constexpr (GEN) {  // the unshifted partners below a phased axis's window edge
            if (s.pkind[fa] && coord_of(q, fa, m) == 0)
                g1 = connect_word(s, fa, true, g1, s.other ? s.other->E[first] : nullptr, q.at + m);
            if (s.pkind[sa] && coord_of(q, sa, m) == 0)
                g2 = connect_word(s, sa, true, g2, s.other ? s.other->E[second] : nullptr, q.at + m);
        }
        float curl = curl_of(s.dtdx, s.present[fa], s.present[sa], g1, en[first][fa].v[m], g2,
                             en[second][sa].v[m]);
        if constexpr (GEN) {
            if (on_wall(s, q, C, m) || ghost_at(s, q, 1, C, m)) curl = 0.0f;
            // a condinv word of 0 marks a cell whose chunk keeps no conductivity array: it
            // steps the lossless forms (structure.cpp:883-900)
            const float civ = s.cf[1][C] ? gp(s.ci[1][C])[q.at + m] : 0.0f;
            if (civ != 0.0f) {
                const int cs = coord_of(q, dsig, m), cu = coord_of(q, dsigu, m);
                const bool in_s = PML && member(s, 1, dsig, cs), in_u = PML && member(s, 1, dsigu, cu);
                b.v[m] = curl_apply_cond(b.v[m], curl, gp(s.cf[1][C])[q.at + m], civ, in_s, in_u,
                                         coeff(s, 1, dsig, cs, in_s), coeff(s, 1, dsigu, cu, in_u),
                                         s.fuB[C], s.fcond[1][C], q.at + m);
                continue;
            }
        }
        if (PML) {
            const int cs = coord_of(q, dsig, m), cu = coord_of(q, dsigu, m);
            const bool in_s = member(s, 1, dsig, cs), in_u = member(s, 1, dsigu, cu);
            if (in_s || in_u)
                b.v[m] = curl_apply_pml(b.v[m], curl, coeff(s, 1, dsig, cs, in_s),
                                        coeff(s, 1, dsigu, cu, in_u), s.fuB[C], q.at + m);
            else
                b.v[m] = b.v[m] - curl;
        } else {
            b.v[m] = b.v[m] - curl;
        }
    }
    if (s.nsrc) deposit_run<V, GEN>(s, ck.step, 1, C, q, b);
    if constexpr (GEN) {
#pragma unroll
        for (int m = 0; m < V; ++m)
            if (on_wall(s, q, C, m)) b.v[m] = 0.0f;
    }
    stv<V, NT>(s.B[C], q.at, b);
    if (PML) {
        // update_H: H accumulates on the own-axis layer's rows; everywhere else H is B's word
        // (update_eh.cpp:162-170), so nothing is stored there and readers take B.
#pragma unroll
        for (int m = 0; m < V; ++m) {
            const int cw = coord_of(q, C, m);
            // a fold's ghost takes its B from its image first: the fold kernel updates its H
            if (member(s, 0, C, cw) && !(GEN && ghost_at(s, q, 1, C, m))) {
                const float4 lw = coeff(s, 0, C, cw, true);
                float hold = b.v[m], prev = b.v[m];
                if (ck.hsep) {
                    hold = gp(s.H[C])[q.at + m];
                    prev = gp(s.fwH[C])[q.at + m];
                }
                gp(s.fwH[C])[q.at + m] = b.v[m];
                const float a = lw.z * b.v[m];
                const float mm = lw.x * prev;
                gp(s.H[C])[q.at + m] = hold + (a - mm);
            }
        }
    }
}

template <int V, bool PML, bool NT, bool GEN>
__device__ inline void hpair_run(const SimDev &s, const Clock &ck, const Run &q) {
    VF<V> e[3], en[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c) e[c] = ldv<V>(s.E[c], q.at);
    en[0][1] = ldv<V>(s.E[0], q.at + q.yp);
    en[0][2] = ldv<V>(s.E[0], q.at + q.zp);
    en[1][2] = ldv<V>(s.E[1], q.at + q.zp);
    en[1][0] = x_up<V>(s, q, s.E[1], e[1]);
    en[2][0] = x_up<V>(s, q, s.E[2], e[2]);
    en[2][1] = ldv<V>(s.E[2], q.at + q.yp);
    if constexpr (GEN) metal_up<V>(s, q, en);
    // every stream's load issued before the first store: the stores may alias as far as the
    // compiler knows, so loads left behind them would wait a memory round trip each
    VF<V> b[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) b[c] = NT ? ldv_once<V>(s.B[c], q.at) : ldv<V>(s.B[c], q.at);
    hpair_component<V, 0, PML, NT, GEN>(s, ck, q, e, en, b[0]);
    hpair_component<V, 1, PML, NT, GEN>(s, ck, q, e, en, b[1]);
    hpair_component<V, 2, PML, NT, GEN>(s, ck, q, e, en, b[2]);
}

template <int V, bool NT, bool GEN>
__global__ void __launch_bounds__(kThreads) hpair_kernel(const SimDev *__restrict__ sims,
                                                         const int4 *__restrict__ tiles,
                                                         int ntiles, int xcd_remap, int adv) {
    const TileWalk w = tile_walk(ntiles, xcd_remap);
    for (int g = w.next; g < w.end; g += w.stride) {
        const int t = tile_at(w, g);
        if (t >= ntiles) continue;
        const int4 tile = tiles[t];
        if ((int)threadIdx.x >= tile.w) continue;
        const SimDev &s = sims[tile.x];
        const Run q = work_run<V>(s, tile, threadIdx.x);
        const Clock ck = clock_of(s, adv);
        if (s.pml)
            hpair_run<V, true, NT, GEN>(s, ck, q);
        else
            hpair_run<V, false, NT, GEN>(s, ck, q);
    }
}

// ------------------------------------------------------------------------------------------
// D pair: step_D -> electric deposit -> update_E -> update_P.
// ------------------------------------------------------------------------------------------
// The array the D curl reads component c of H from at lane m: H on c's own-axis layer rows, B
// elsewhere (as ld_hb chooses).
__device__ inline const float *hb_of(const SimDev &s, const Run &q, int c, int m) {
    return member(s, 0, c, coord_of(q, c, m)) ? s.H[c] : s.B[c];
}

template <int V, int C, bool PML, bool GEN>
__device__ inline VF<V> dcurl_component(const SimDev &s, const Clock &ck, const Run &q,
                                        const VF<V> (&h)[3], const VF<V> (&hn)[3][3], VF<V> d) {
    constexpr int first = (C + 2) % 3, fa = (C + 1) % 3;
    constexpr int second = (C + 1) % 3, sa = (C + 2) % 3;
    constexpr int dsig = (C + 1) % 3, dsigu = (C + 2) % 3;
#pragma unroll
    for (int m = 0; m < V; ++m) {
        float g1 = h[first].v[m], g2 = h[second].v[m];
        if constexpr (GEN) {  // the unshifted partners above a phased axis's window edge
            if (s.pkind[fa] && coord_of(q, fa, m) == 0)
                g1 = connect_word(s, fa, false, g1, s.other ? hb_of(*s.other, q, first, m) : nullptr,
                                  q.at + m);
            if (s.pkind[sa] && coord_of(q, sa, m) == 0)
                g2 = connect_word(s, sa, false, g2, s.other ? hb_of(*s.other, q, second, m) : nullptr,
                                  q.at + m);
        }
        float curl = curl_of(s.dtdx, s.present[fa], s.present[sa], g1, hn[first][fa].v[m], g2,
                             hn[second][sa].v[m]);
        if constexpr (GEN) {
            if (on_wall(s, q, (C + 1) % 3, m) || on_wall(s, q, (C + 2) % 3, m) ||
                ghost_at(s, q, 0, C, m))
                curl = 0.0f;
            const float civ = s.cf[0][C] ? gp(s.ci[0][C])[q.at + m] : 0.0f;
            if (civ != 0.0f) {  // 0: the cell's chunk keeps no conductivity (lossless forms)
                const int cs = coord_of(q, dsig, m), cu = coord_of(q, dsigu, m);
                const bool in_s = PML && member(s, 0, dsig, cs), in_u = PML && member(s, 0, dsigu, cu);
                d.v[m] = curl_apply_cond(d.v[m], curl, gp(s.cf[0][C])[q.at + m], civ, in_s, in_u,
                                         coeff(s, 0, dsig, cs, in_s), coeff(s, 0, dsigu, cu, in_u),
                                         s.fuD[C], s.fcond[0][C], q.at + m);
                continue;
            }
        }
        if (PML) {
            const int cs = coord_of(q, dsig, m), cu = coord_of(q, dsigu, m);
            const bool in_s = member(s, 0, dsig, cs), in_u = member(s, 0, dsigu, cu);
            // the dsigu recurrence reads the dsig row's words wherever it runs: 1 off the line,
            // except on a metallic axis's wall row, which no chunk owns but whose words are
            // graded from the wall (the region plan broadcasts the whole vector there)
            bool read_s = in_s;
            if constexpr (GEN) read_s |= cs == 0 && s.metal[dsig];
            if (in_s || in_u)
                d.v[m] = curl_apply_pml(d.v[m], curl, coeff(s, 0, dsig, cs, read_s),
                                        coeff(s, 0, dsigu, cu, in_u), s.fuD[C], q.at + m);
            else
                d.v[m] = d.v[m] - curl;
        } else {
            d.v[m] = d.v[m] - curl;
        }
    }
    if (s.nsrc) deposit_run<V, GEN>(s, ck.step, 0, C, q, d);
    if constexpr (GEN) {
#pragma unroll
        for (int m = 0; m < V; ++m)
            if (on_wall(s, q, (C + 1) % 3, m) || on_wall(s, q, (C + 2) % 3, m)) d.v[m] = 0.0f;
    }
    return d;
}

// E from the constitutive product `cst`, then every pole's P update for component C. The
// drive W is f_w on the own-axis layer's cells and E elsewhere: the constitutive product both.
template <int V, int C, bool PML, bool NT, int NP>
__device__ inline void e_and_p(const SimDev &s, const Clock &ck, const Run &q, const VF<V> &cst,
                               const VF<V> (&p)[NP], bool active) {
    VF<V> e;
    if (PML) {
#pragma unroll
        for (int m = 0; m < V; ++m) {
            const int cw = coord_of(q, C, m);
            float ev = cst.v[m];
            if (member(s, 1, C, cw)) {
                const float4 lw = coeff(s, 1, C, cw, true);
                const float prev = gp(s.fwE[C])[q.at + m];
                gp(s.fwE[C])[q.at + m] = cst.v[m];
                const float a = lw.z * cst.v[m];
                const float mm = lw.x * prev;
                ev = gp(s.E[C])[q.at + m] + (a - mm);
            }
            e.v[m] = ev;
        }
    } else {
        e = cst;
    }
    stv<V, NT>(s.E[C], q.at, e);
    if (!active) return;  // no polarization on this row, now or later: P stays +0
    const int par = ck.ppar;
#pragma unroll
    for (int pq = 0; pq < NP; ++pq) {
        if (pq >= s.npol) break;
        if (!s.pdriven[pq][C]) continue;
        float *out = s.pbuf[pq][C][par ^ 1];
        const VF<V> pp = ldv<V>(out, q.at);
        const float *sv = s.psig[pq][C];
        VF<V> sg;
        if (sv) {
            sg = ldv<V>(sv, q.at);
        } else {
#pragma unroll
            for (int m = 0; m < V; ++m) sg.v[m] = s.psig_s[pq][C];
        }
        VF<V> r;
#pragma unroll
        for (int m = 0; m < V; ++m) {
            float t = s.pc[pq][0] * p[pq].v[m];
            t = t - s.pc[pq][1] * pp.v[m];
            float dr = sg.v[m] * cst.v[m];
            dr = s.pc[pq][2] * dr;
            t = t + dr;
            r.v[m] = s.pc[pq][3] * t;
        }
        stv<V>(out, q.at, r);
    }
}

// D - sum P for component C from the D values in `d`, loading each driving pole's P^n.
__device__ inline bool pole_run_active(const SimDev &s, const Run &q, int V) {
    if (!s.npol) return false;
    if (!s.prow) return true;
    const GMEM int *r = reinterpret_cast<const GMEM int *>(gp(s.prow) + q.j + q.k * s.n[1]);
    return q.i0 < r[1] && q.i0 + V > r[0];
}

template <int V, int C, int NP>
__device__ inline VF<V> minus_poles(const SimDev &s, const Clock &ck, const Run &q,
                                    const VF<V> &d, VF<V> (&p)[NP], bool active) {
    VF<V> dmp = d;
    const int par = ck.ppar;
#pragma unroll
    for (int pq = 0; pq < NP; ++pq) {
#pragma unroll
        for (int m = 0; m < V; ++m) p[pq].v[m] = 0.0f;
        if (pq >= s.npol || !active) continue;
        if (!s.pdriven[pq][C]) continue;
        p[pq] = ldv<V>(s.pbuf[pq][C][par], q.at);
#pragma unroll
        for (int m = 0; m < V; ++m) dmp.v[m] = dmp.v[m] - p[pq].v[m];
    }
    return dmp;
}

// Inverse epsilon in two parts: the index words (or nothing, for a float volume) are loaded
// first of everything the D pair reads, so the table look-ups that depend on them go out while
// the field loads are still in flight instead of after all of them.
struct InvWords {
    uint32_t w[2];
};

// INV: the storage every simulation of the launch uses for all three components -- 0 float,
// 1 byte indices, 2 16-bit indices -- or 3, mixed (chosen per component at run time).
template <int V, int INV, bool NT>
__device__ inline InvWords load_inv_words(const SimDev &s, int c, int at) {
    InvWords r;
    r.w[0] = r.w[1] = 0u;
    if (INV == 0) return r;
    if ((INV == 1) || (INV == 3 && s.inv8[c])) {
        const GMEM uint8_t *g = gp(s.inv8[c]) + at;
        if (V == 1) r.w[0] = NT ? __builtin_nontemporal_load(g) : *g;
        else if (V == 2) {
            const GMEM uint16_t *h = reinterpret_cast<const GMEM uint16_t *>(g);
            r.w[0] = NT ? __builtin_nontemporal_load(h) : *h;
        } else {
            const GMEM uint32_t *h = reinterpret_cast<const GMEM uint32_t *>(g);
            r.w[0] = NT ? __builtin_nontemporal_load(h) : *h;
        }
    } else if ((INV == 2) || (INV == 3 && s.inv16[c])) {
        const GMEM uint16_t *g = gp(s.inv16[c]) + at;
        if (V == 1) r.w[0] = *g;
        else if (V == 2) r.w[0] = *reinterpret_cast<const GMEM uint32_t *>(g);
        else {
            typedef uint32_t u2v __attribute__((ext_vector_type(2)));
            const u2v x = *reinterpret_cast<const GMEM u2v *>(g);
            r.w[0] = x.x;
            r.w[1] = x.y;
        }
    }
    return r;
}

template <int V, int INV>
__device__ inline VF<V> inv_values(const SimDev &s, int c, int at, const InvWords &iw) {
    VF<V> r;
    if (INV == 0) return ldv<V>(s.inv[c], at);
    if ((INV == 1) || (INV == 3 && s.inv8[c])) {
#pragma unroll
        for (int m = 0; m < V; ++m) r.v[m] = gp(s.invtab[c])[(iw.w[0] >> (8 * m)) & 0xffu];
        return r;
    }
    if ((INV == 2) || (INV == 3 && s.inv16[c])) {
#pragma unroll
        for (int m = 0; m < V; ++m) r.v[m] = gp(s.invtab[c])[(iw.w[m / 2] >> (16 * (m % 2))) & 0xffffu];
        return r;
    }
    return ldv<V>(s.inv[c], at);
}

template <int V, int C, bool PML, bool NT, bool GEN>
__device__ inline void dpair_component(const SimDev &s, const Clock &ck, const Run &q,
                                       const VF<V> (&h)[3], const VF<V> (&hn)[3][3],
                                       const VF<V> &d0, const VF<V> &iv) {
    VF<V> d = dcurl_component<V, C, PML, GEN>(s, ck, q, h, hn, d0);
    stv<V, NT>(s.D[C], q.at, d);
    // the constitutive update reads D less the polarization sources' dipole (general forms)
    if constexpr (GEN) {
        if (s.npsrc) polarize_run<V>(s, ck.step, C, q, d);
    }
    constexpr int NP = GEN ? kMaxPoles : kLosslessPoles;
    VF<V> p[NP];
    // a cell where every pole's sigma and stored P are zero keeps P = +0 for good (its update
    // is g1inv*((A*0 - g1*0) + od*(0*w)) = +0), and D - (+0) is D: runs with no such cell
    // skip the pole work (a run straddling the range updates its +0 cells to +0)
    const bool active = pole_run_active(s, q, V);
    const VF<V> dmp = minus_poles<V, C, NP>(s, ck, q, d, p, active);
    VF<V> cst;
#pragma unroll
    for (int m = 0; m < V; ++m) cst.v[m] = iv.v[m] * dmp.v[m];
    e_and_p<V, C, PML, NT, NP>(s, ck, q, cst, p, active);
}

// H component c at the run shifted by `off` (a shift across c's own axis): the stored H on
// c's own-axis layer rows, B elsewhere. Only Hx's rows vary along a run.
template <int V>
__device__ inline VF<V> ld_hb(const SimDev &s, const Run &q, int c, int off) {
    if (c != 0 || V == 1) {
        const bool in = member(s, 0, c, coord_of(q, c, 0));
        return ldv<V>(in ? s.H[c] : s.B[c], q.at + off);
    }
    VF<V> r = ldv<V>(s.B[0], q.at + off);
#pragma unroll
    for (int m = 0; m < V; ++m)
        if (member(s, 0, 0, q.i0 + m)) r.v[m] = gp(s.H[0])[q.at + off + m];
    return r;
}

template <int V, bool PML>
__device__ inline void load_h(const SimDev &s, const Run &q, VF<V> (&h)[3], VF<V> (&hn)[3][3]) {
    if (PML) {
#pragma unroll
        for (int c = 0; c < 3; ++c) h[c] = ld_hb<V>(s, q, c, 0);
        // MEEP's negated strides: the D curl reads its partners one cell DOWN
        hn[0][1] = ld_hb<V>(s, q, 0, q.ym);
        hn[0][2] = ld_hb<V>(s, q, 0, q.zm);
        hn[1][2] = ld_hb<V>(s, q, 1, q.zm);
        hn[2][1] = ld_hb<V>(s, q, 2, q.ym);
        const float *hy = member(s, 0, 1, q.j) ? s.H[1] : s.B[1];
        const float *hz = member(s, 0, 2, q.k) ? s.H[2] : s.B[2];
        hn[1][0] = x_down<V>(s, q, hy, h[1]);
        hn[2][0] = x_down<V>(s, q, hz, h[2]);
        return;
    }
    float *const *Hs = s.B;  // without an absorber H is B (fields.py:1380)
#pragma unroll
    for (int c = 0; c < 3; ++c) h[c] = ldv<V>(Hs[c], q.at);
    hn[0][1] = ldv<V>(Hs[0], q.at + q.ym);
    hn[0][2] = ldv<V>(Hs[0], q.at + q.zm);
    hn[1][2] = ldv<V>(Hs[1], q.at + q.zm);
    hn[1][0] = x_down<V>(s, q, Hs[1], h[1]);
    hn[2][0] = x_down<V>(s, q, Hs[2], h[2]);
    hn[2][1] = ldv<V>(Hs[2], q.at + q.ym);
}

template <int V, bool PML, bool NT, int INV, bool GEN>
__device__ inline void dpair_run(const SimDev &s, const Clock &ck, const Run &q) {
    InvWords iw[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) iw[c] = load_inv_words<V, INV, NT>(s, c, q.at);
    VF<V> h[3], hn[3][3];
    load_h<V, PML>(s, q, h, hn);
    if constexpr (GEN) metal_down<V>(s, q, hn);
    // every stream's load issued before the first store (see hpair_run)
    VF<V> d[3], iv[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) d[c] = NT ? ldv_once<V>(s.D[c], q.at) : ldv<V>(s.D[c], q.at);
#pragma unroll
    for (int c = 0; c < 3; ++c) iv[c] = inv_values<V, INV>(s, c, q.at, iw[c]);
    dpair_component<V, 0, PML, NT, GEN>(s, ck, q, h, hn, d[0], iv[0]);
    dpair_component<V, 1, PML, NT, GEN>(s, ck, q, h, hn, d[1], iv[1]);
    dpair_component<V, 2, PML, NT, GEN>(s, ck, q, h, hn, d[2], iv[2]);
}

// ------------------------------------------------------------------------------------------
// The shell's layer forms, loads first (lossless simulations with at most one pole). Each
// component's three layer roles -- the dsig and dsigu lines of its curl and its own axis's line
// of the constitutive update -- lie on the three different axes, so exactly one of them varies
// along the run (x) and the other two hold for all V cells. Every operand of the three
// components (the layer words each form uses, the split fields, the f_w histories and old E or
// H words, the pole's state) is loaded before the first store: on CDNA one in-order counter
// covers loads and stores, and loads issued after a store would each wait for it. The
// arithmetic is curl_apply_pml's and e_and_p's, in the same order, so the words are the same.
// ------------------------------------------------------------------------------------------
template <int V, int AXIS>
struct Role {  // a y or z line: one membership and one word pair for the whole run
    bool in_, any;
    float2 kw_;
    __device__ bool in(int) const { return in_; }
    __device__ float2 kw(int) const { return kw_; }
};

template <int V>
struct Role<V, 0> {  // the x line: per cell
    bool in_[V], any;
    float2 kw_[V];
    __device__ bool in(int m) const { return in_[m]; }
    __device__ float2 kw(int m) const { return kw_[m]; }
};

// Line (half, AXIS)'s role over the run: the words its form multiplies by, (kms, sinv), or with
// `use_kps` (kms, kps); the identity's off the line. XO: the run is in the box, whose y and z
// are on no line, so only the x line can hold.
template <int V, int AXIS, bool XO>
__device__ inline Role<V, AXIS> load_role(const SimDev &s, const Run &q, int half, bool use_kps) {
    Role<V, AXIS> r;
    if constexpr (XO && AXIS != 0) {
        r.in_ = false;
        r.any = false;
        r.kw_ = make_float2(1.0f, 1.0f);
    } else if constexpr (AXIS == 0) {
        r.any = false;
#pragma unroll
        for (int m = 0; m < V; ++m) {
            r.in_[m] = member(s, half, 0, q.i0 + m);
            const float4 c = coeff(s, half, 0, q.i0 + m, r.in_[m]);
            r.kw_[m] = make_float2(c.x, use_kps ? c.z : c.y);
            r.any |= r.in_[m];
        }
    } else {
        const int cc = AXIS == 1 ? q.j : q.k;
        r.in_ = member(s, half, AXIS, cc);
        const float4 c = coeff(s, half, AXIS, cc, r.in_);
        r.kw_ = make_float2(c.x, use_kps ? c.z : c.y);
        r.any = r.in_;
    }
    return r;
}

// One D-pair component's layer operands.
template <int V, int C>
struct DLayer {
    Role<V, (C + 1) % 3> rs;  // dsig of the curl (integer lattice)
    Role<V, (C + 2) % 3> ru;  // dsigu of the curl
    Role<V, C> rw;            // own axis of update_E (half-integer lattice)
    VF<V> fu, fw, e;          // split field, f_w_E and E where a cell of the run needs them
    VF<V> p, pp, sg;          // the pole's P^n, P^{n-1} and sigma (POLE)
};

template <int V, int C, bool POLE, bool XO>
__device__ inline void load_dlayer(const SimDev &s, const Clock &ck, const Run &q, bool active,
                                   DLayer<V, C> &o) {
    o.rs = load_role<V, (C + 1) % 3, XO>(s, q, 0, false);
    o.ru = load_role<V, (C + 2) % 3, XO>(s, q, 0, false);
    o.rw = load_role<V, C, XO>(s, q, 1, true);
    if (o.ru.any) o.fu = ldv<V>(s.fuD[C], q.at);
    if (o.rw.any) {
        o.fw = ldv<V>(s.fwE[C], q.at);
        o.e = ldv<V>(s.E[C], q.at);
    }
    if constexpr (POLE) {
        if (active && s.pdriven[0][C]) {
            o.p = ldv<V>(s.pbuf[0][C][ck.ppar], q.at);
            o.pp = ldv<V>(s.pbuf[0][C][ck.ppar ^ 1], q.at);
            if (s.psig[0][C]) o.sg = ldv<V>(s.psig[0][C], q.at);
            else {
#pragma unroll
                for (int m = 0; m < V; ++m) o.sg.v[m] = s.psig_s[0][C];
            }
        }
    }
}

// The component's new words, then its stores (after every component's loads).
template <int V, int C, bool NT, bool POLE>
__device__ inline void dlayer_component(const SimDev &s, const Clock &ck, const Run &q,
                                        const VF<V> (&h)[3], const VF<V> (&hn)[3][3],
                                        const VF<V> &d0, const VF<V> &iv, bool active,
                                        const DLayer<V, C> &o) {
    constexpr int first = (C + 2) % 3, fa = (C + 1) % 3;
    constexpr int second = (C + 1) % 3, sa = (C + 2) % 3;
    VF<V> d = d0, fu = o.fu;
#pragma unroll
    for (int m = 0; m < V; ++m) {
        const float curl = curl_of(s.dtdx, s.present[fa], s.present[sa], h[first].v[m],
                                   hn[first][fa].v[m], h[second].v[m], hn[second][sa].v[m]);
        const float2 ls = o.rs.kw(m), lu = o.ru.kw(m);
        if (o.ru.in(m)) {  // curl_apply_pml's split-field recurrence
            const float prev = o.fu.v[m];
            float u = ls.x * prev;
            u = u - curl;
            u = ls.y * u;
            fu.v[m] = u;
            float v = lu.x * d.v[m];
            v = v + u;
            v = v - prev;
            d.v[m] = lu.y * v;
        } else if (o.rs.in(m)) {  // the dsig-only form
            float v = ls.x * d.v[m];
            v = v - curl;
            d.v[m] = ls.y * v;
        } else {
            d.v[m] = d.v[m] - curl;
        }
    }
    if (s.nsrc) deposit_run<V, false>(s, ck.step, 0, C, q, d);
    const bool pole = POLE && active && s.pdriven[0][C];
    VF<V> dmp = d;
    if (pole) {
#pragma unroll
        for (int m = 0; m < V; ++m) dmp.v[m] = dmp.v[m] - o.p.v[m];
    }
    VF<V> cst, e, fw = o.fw;
#pragma unroll
    for (int m = 0; m < V; ++m) cst.v[m] = iv.v[m] * dmp.v[m];
#pragma unroll
    for (int m = 0; m < V; ++m) {
        float ev = cst.v[m];
        if (o.rw.in(m)) {  // update_E on the own-axis layer (e_and_p)
            const float2 lw = o.rw.kw(m);  // (kms, kps)
            const float prev = o.fw.v[m];
            fw.v[m] = cst.v[m];
            const float a = lw.y * cst.v[m];
            const float mm = lw.x * prev;
            ev = o.e.v[m] + (a - mm);
        }
        e.v[m] = ev;
    }
    stv<V, NT>(s.D[C], q.at, d);
    if (o.ru.any) stv<V>(s.fuD[C], q.at, fu);  // cells off the line keep their word
    if (o.rw.any) stv<V>(s.fwE[C], q.at, fw);
    stv<V, NT>(s.E[C], q.at, e);
    if (pole) {
        VF<V> r;
#pragma unroll
        for (int m = 0; m < V; ++m) {
            float t = s.pc[0][0] * o.p.v[m];
            t = t - s.pc[0][1] * o.pp.v[m];
            float dr = o.sg.v[m] * cst.v[m];
            dr = s.pc[0][2] * dr;
            t = t + dr;
            r.v[m] = s.pc[0][3] * t;
        }
        stv<V>(s.pbuf[0][C][ck.ppar ^ 1], q.at, r);
    }
}

// The box's H operands: H is stored only on its own-axis layer rows, and in the box y and z
// are on no line, so Hy and Hz are B's words and only Hx may be H's (at the x layer's cells).
template <int V>
__device__ inline void load_h_box(const SimDev &s, const Run &q, VF<V> (&h)[3], VF<V> (&hn)[3][3]) {
    h[0] = ld_hb<V>(s, q, 0, 0);
    h[1] = ldv<V>(s.B[1], q.at);
    h[2] = ldv<V>(s.B[2], q.at);
    hn[0][1] = ld_hb<V>(s, q, 0, q.ym);
    hn[0][2] = ld_hb<V>(s, q, 0, q.zm);
    hn[1][2] = ldv<V>(s.B[1], q.at + q.zm);
    hn[1][0] = x_down<V>(s, q, s.B[1], h[1]);
    hn[2][0] = x_down<V>(s, q, s.B[2], h[2]);
    hn[2][1] = ldv<V>(s.B[2], q.at + q.ym);
}

template <int V, bool NT, int INV, bool POLE, bool XO>
__device__ inline void dshell_run(const SimDev &s, const Clock &ck, const Run &q) {
    InvWords iw[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) iw[c] = load_inv_words<V, INV, NT>(s, c, q.at);
    VF<V> h[3], hn[3][3];
    if constexpr (XO) load_h_box<V>(s, q, h, hn);
    else load_h<V, true>(s, q, h, hn);
    VF<V> d[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) d[c] = NT ? ldv_once<V>(s.D[c], q.at) : ldv<V>(s.D[c], q.at);
    const bool active = POLE && pole_run_active(s, q, V);
    DLayer<V, 0> l0;
    DLayer<V, 1> l1;
    DLayer<V, 2> l2;
    load_dlayer<V, 0, POLE, XO>(s, ck, q, active, l0);
    load_dlayer<V, 1, POLE, XO>(s, ck, q, active, l1);
    load_dlayer<V, 2, POLE, XO>(s, ck, q, active, l2);
    VF<V> iv[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) iv[c] = inv_values<V, INV>(s, c, q.at, iw[c]);
    dlayer_component<V, 0, NT, POLE>(s, ck, q, h, hn, d[0], iv[0], active, l0);
    dlayer_component<V, 1, NT, POLE>(s, ck, q, h, hn, d[1], iv[1], active, l1);
    dlayer_component<V, 2, NT, POLE>(s, ck, q, h, hn, d[2], iv[2], active, l2);
}

// One H-pair component's layer operands: dsig and dsigu of the curl on the half-integer lattice,
// the own-axis integer-lattice line of update_H (H and f_w_H where a cell needs them).
template <int V, int C>
struct HLayer {
    Role<V, (C + 1) % 3> rs;
    Role<V, (C + 2) % 3> ru;
    Role<V, C> rw;
    VF<V> fu, hw, fw;
};

template <int V, int C, bool XO>
__device__ inline void load_hlayer(const SimDev &s, const Clock &ck, const Run &q, HLayer<V, C> &o) {
    o.rs = load_role<V, (C + 1) % 3, XO>(s, q, 1, false);
    o.ru = load_role<V, (C + 2) % 3, XO>(s, q, 1, false);
    o.rw = load_role<V, C, XO>(s, q, 0, true);
    if (o.ru.any) o.fu = ldv<V>(s.fuB[C], q.at);
    if (o.rw.any && ck.hsep) {
        o.hw = ldv<V>(s.H[C], q.at);
        o.fw = ldv<V>(s.fwH[C], q.at);
    }
}

template <int V, int C, bool NT>
__device__ inline void hlayer_component(const SimDev &s, const Clock &ck, const Run &q,
                                        const VF<V> (&e)[3], const VF<V> (&en)[3][3],
                                        const VF<V> &b0, const HLayer<V, C> &o) {
    constexpr int first = (C + 2) % 3, fa = (C + 1) % 3;
    constexpr int second = (C + 1) % 3, sa = (C + 2) % 3;
    VF<V> b = b0, fu = o.fu;
#pragma unroll
    for (int m = 0; m < V; ++m) {
        const float curl = curl_of(s.dtdx, s.present[fa], s.present[sa], e[first].v[m],
                                   en[first][fa].v[m], e[second].v[m], en[second][sa].v[m]);
        const float2 ls = o.rs.kw(m), lu = o.ru.kw(m);
        if (o.ru.in(m)) {
            const float prev = o.fu.v[m];
            float u = ls.x * prev;
            u = u - curl;
            u = ls.y * u;
            fu.v[m] = u;
            float v = lu.x * b.v[m];
            v = v + u;
            v = v - prev;
            b.v[m] = lu.y * v;
        } else if (o.rs.in(m)) {
            float v = ls.x * b.v[m];
            v = v - curl;
            b.v[m] = ls.y * v;
        } else {
            b.v[m] = b.v[m] - curl;
        }
    }
    if (s.nsrc) deposit_run<V, false>(s, ck.step, 1, C, q, b);
    stv<V, NT>(s.B[C], q.at, b);
    if (o.ru.any) stv<V>(s.fuB[C], q.at, fu);
    if (!o.rw.any) return;
    // update_H on the own-axis layer: H accumulates there; elsewhere H is B's word and H is
    // not stored, so the words go out per cell
    GMEM float *gh = gp(s.H[C]), *gf = gp(s.fwH[C]);
#pragma unroll
    for (int m = 0; m < V; ++m) {
        if (!o.rw.in(m)) continue;
        const float2 lw = o.rw.kw(m);  // (kms, kps)
        const float hold = ck.hsep ? o.hw.v[m] : b.v[m];
        const float prev = ck.hsep ? o.fw.v[m] : b.v[m];
        gf[q.at + m] = b.v[m];
        const float a = lw.y * b.v[m];
        const float mm = lw.x * prev;
        gh[q.at + m] = hold + (a - mm);
    }
}

template <int V, bool NT, bool XO>
__device__ inline void hshell_run(const SimDev &s, const Clock &ck, const Run &q) {
    VF<V> e[3], en[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c) e[c] = ldv<V>(s.E[c], q.at);
    en[0][1] = ldv<V>(s.E[0], q.at + q.yp);
    en[0][2] = ldv<V>(s.E[0], q.at + q.zp);
    en[1][2] = ldv<V>(s.E[1], q.at + q.zp);
    en[1][0] = x_up<V>(s, q, s.E[1], e[1]);
    en[2][0] = x_up<V>(s, q, s.E[2], e[2]);
    en[2][1] = ldv<V>(s.E[2], q.at + q.yp);
    VF<V> b[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) b[c] = NT ? ldv_once<V>(s.B[c], q.at) : ldv<V>(s.B[c], q.at);
    HLayer<V, 0> l0;
    HLayer<V, 1> l1;
    HLayer<V, 2> l2;
    load_hlayer<V, 0, XO>(s, ck, q, l0);
    load_hlayer<V, 1, XO>(s, ck, q, l1);
    load_hlayer<V, 2, XO>(s, ck, q, l2);
    hlayer_component<V, 0, NT>(s, ck, q, e, en, b[0], l0);
    hlayer_component<V, 1, NT>(s, ck, q, e, en, b[1], l1);
    hlayer_component<V, 2, NT>(s, ck, q, e, en, b[2], l2);
}

// The lossless groups' launches: the box's rows and the shell's rows as two kernels, each
// with the registers its own forms need -- the box's with the x line's layer forms only (or,
// where no member has an absorber, the lossless forms alone: dint, hint), the shell's with every
// role.
template <int V, bool NT, int INV>
__global__ void __launch_bounds__(kThreads) dint_kernel(const SimDev *__restrict__ sims,
                                                        const int4 *__restrict__ items,
                                                        int nitems, int xcd_remap, int adv) {
    const TileWalk w = tile_walk(nitems, xcd_remap);
    for (int g = w.next; g < w.end; g += w.stride) {
        const int t = tile_at(w, g);
        if (t >= nitems) continue;
        const int4 it = items[t];
        if ((int)threadIdx.x >= it.w) continue;
        const SimDev &s = sims[it.x];
        dpair_run<V, false, NT, INV, false>(s, clock_of(s, adv), work_run<V>(s, it, threadIdx.x));
    }
}

template <int V, bool NT, int INV, bool POLE, bool XO>
__global__ void __launch_bounds__(kThreads) dshell_kernel(const SimDev *__restrict__ sims,
                                                          const int4 *__restrict__ items,
                                                          int nitems, int xcd_remap, int adv) {
    const TileWalk w = tile_walk(nitems, xcd_remap);
    for (int g = w.next; g < w.end; g += w.stride) {
        const int t = tile_at(w, g);
        if (t >= nitems) continue;
        const int4 it = items[t];
        if ((int)threadIdx.x >= it.w) continue;
        const SimDev &s = sims[it.x];
        dshell_run<V, NT, INV, POLE, XO>(s, clock_of(s, adv), work_run<V>(s, it, threadIdx.x));
    }
}

// The box's D pair over LDS-staged halo tiles: a workgroup takes a tile of TX x TY cells of
// one plane (one cell a thread, TX = kThreads / TY) and first stages the H words the tile's
// curls read within the plane -- Hx and Hz with the row below the tile, Hy and Hz with the
// column left of it -- so each word is loaded from memory once per tile, and the j-1 and i-1
// neighbours come from LDS; the k-1 plane is read from memory (the caches hold it). The words
// are those load_h_box reads, so the arithmetic (dshell_run's, with XO) is the same.
template <int TY>
struct DTileStage {
    static constexpr int TX = kThreads / TY;
    float hx[TY + 1][TX];      // row -1 .. TY-1
    float hz[TY + 1][TX + 1];  // row -1 .. TY-1, column -1 .. TX-1
    float hy[TY][TX + 1];      // column -1 .. TX-1
};

// Work item (simulation, 2, tile, 0): tiles numbered x fastest, then y, then z over the box.
template <int TY>
__device__ inline bool dtile_cell(const SimDev &s, int tile, int tx, int ty, int &i, int &j, int &k,
                                  int &x0, int &y0) {
    constexpr int TX = kThreads / TY;
    const int bxn = (s.ibox[1] - s.ibox[0] + TX - 1) / TX;
    const int byn = (s.ibox[3] - s.ibox[2] + TY - 1) / TY;
    const int bx = tile % bxn, r = tile / bxn;
    const int by = r % byn, bz = r / byn;
    x0 = s.ibox[0] + bx * TX;
    y0 = s.ibox[2] + by * TY;
    i = x0 + tx;
    j = y0 + ty;
    k = s.ibox[4] + bz;
    return i < s.ibox[1] && j < s.ibox[3];
}

template <int TY, bool NT, int INV, bool POLE>
__global__ void __launch_bounds__(kThreads) dtile_kernel(const SimDev *__restrict__ sims,
                                                         const int4 *__restrict__ items,
                                                         int nitems, int xcd_remap, int adv) {
    constexpr int TX = kThreads / TY;
    __shared__ DTileStage<TY> st;
    const int tx = (int)threadIdx.x % TX, ty = (int)threadIdx.x / TX;
    const TileWalk w = tile_walk(nitems, xcd_remap);
    for (int g = w.next; g < w.end; g += w.stride) {
        const int t = tile_at(w, g);
        if (t >= nitems) continue;  // the same for every thread of the workgroup
        const int4 it = items[t];
        const SimDev &s = sims[it.x];
        int i, j, k, x0, y0;
        const bool live = dtile_cell<TY>(s, it.z, tx, ty, i, j, k, x0, y0);
        // every memory operand is requested before the stage's barrier, so the barrier waits
        // while they are in flight
        Run q;
        Clock ck;
        InvWords iw[3];
        VF<1> h[3], hn[3][3], d[3];
        DLayer<1, 0> l0;
        DLayer<1, 1> l1;
        DLayer<1, 2> l2;
        bool active = false;
        float hx = 0.f, hy = 0.f, hz = 0.f, hxm = 0.f, hzm = 0.f, hyl = 0.f, hzl = 0.f;
        if (live) {
            q = locate_ijk(s, i, j, k);
            ck = clock_of(s, adv);
#pragma unroll
            for (int c = 0; c < 3; ++c) iw[c] = load_inv_words<1, INV, NT>(s, c, q.at);
            hx = ld_hb<1>(s, q, 0, 0).v[0];
            hy = gp(s.B[1])[q.at];
            hz = gp(s.B[2])[q.at];
            if (ty == 0) {  // the row below the tile (periodic)
                const Run qm = locate_ijk(s, i, y0 > 0 ? y0 - 1 : s.n[1] - 1, k);
                hxm = ld_hb<1>(s, qm, 0, 0).v[0];
                hzm = gp(s.B[2])[qm.at];
            }
            if (tx == 0) {  // the column left of the tile (the wrap within the stored row)
                const int at = q.row + (x0 > 0 ? x0 - 1 : s.n[0] - 1);
                hyl = gp(s.B[1])[at];
                hzl = gp(s.B[2])[at];
            }
            hn[0][2] = ld_hb<1>(s, q, 0, q.zm);
            hn[1][2] = ldv<1>(s.B[1], q.at + q.zm);
#pragma unroll
            for (int c = 0; c < 3; ++c) d[c] = NT ? ldv_once<1>(s.D[c], q.at) : ldv<1>(s.D[c], q.at);
            active = POLE && pole_run_active(s, q, 1);
            load_dlayer<1, 0, POLE, true>(s, ck, q, active, l0);
            load_dlayer<1, 1, POLE, true>(s, ck, q, active, l1);
            load_dlayer<1, 2, POLE, true>(s, ck, q, active, l2);
        }
        __syncthreads();  // the previous tile's reads of the stage are done
        if (live) {
            st.hx[ty + 1][tx] = hx;
            st.hy[ty][tx + 1] = hy;
            st.hz[ty + 1][tx + 1] = hz;
            if (ty == 0) {
                st.hx[0][tx] = hxm;
                st.hz[0][tx + 1] = hzm;
            }
            if (tx == 0) {
                st.hy[ty][0] = hyl;
                st.hz[ty + 1][0] = hzl;
            }
        }
        __syncthreads();
        if (!live) continue;
        h[0].v[0] = hx;
        h[1].v[0] = hy;
        h[2].v[0] = hz;
        hn[0][1].v[0] = st.hx[ty][tx];
        hn[2][1].v[0] = st.hz[ty][tx + 1];
        hn[1][0].v[0] = st.hy[ty][tx];
        hn[2][0].v[0] = st.hz[ty + 1][tx];
        VF<1> iv[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) iv[c] = inv_values<1, INV>(s, c, q.at, iw[c]);
        dlayer_component<1, 0, NT, POLE>(s, ck, q, h, hn, d[0], iv[0], active, l0);
        dlayer_component<1, 1, NT, POLE>(s, ck, q, h, hn, d[1], iv[1], active, l1);
        dlayer_component<1, 2, NT, POLE>(s, ck, q, h, hn, d[2], iv[2], active, l2);
    }
}

template <int V, bool NT>
__global__ void __launch_bounds__(kThreads) hint_kernel(const SimDev *__restrict__ sims,
                                                        const int4 *__restrict__ items,
                                                        int nitems, int xcd_remap, int adv) {
    const TileWalk w = tile_walk(nitems, xcd_remap);
    for (int g = w.next; g < w.end; g += w.stride) {
        const int t = tile_at(w, g);
        if (t >= nitems) continue;
        const int4 it = items[t];
        if ((int)threadIdx.x >= it.w) continue;
        const SimDev &s = sims[it.x];
        hpair_run<V, false, NT, false>(s, clock_of(s, adv), work_run<V>(s, it, threadIdx.x));
    }
}

template <int V, bool NT, bool XO>
__global__ void __launch_bounds__(kThreads) hshell_kernel(const SimDev *__restrict__ sims,
                                                          const int4 *__restrict__ items,
                                                          int nitems, int xcd_remap, int adv) {
    const TileWalk w = tile_walk(nitems, xcd_remap);
    for (int g = w.next; g < w.end; g += w.stride) {
        const int t = tile_at(w, g);
        if (t >= nitems) continue;
        const int4 it = items[t];
        if ((int)threadIdx.x >= it.w) continue;
        const SimDev &s = sims[it.x];
        hshell_run<V, NT, XO>(s, clock_of(s, adv), work_run<V>(s, it, threadIdx.x));
    }
}

template <int V, bool NT, int INV, bool GEN>
__global__ void __launch_bounds__(kThreads) dpair_kernel(const SimDev *__restrict__ sims,
                                                         const int4 *__restrict__ tiles,
                                                         int ntiles, int xcd_remap, int adv) {
    const TileWalk w = tile_walk(ntiles, xcd_remap);
    for (int g = w.next; g < w.end; g += w.stride) {
        const int t = tile_at(w, g);
        if (t >= ntiles) continue;
        const int4 tile = tiles[t];
        if ((int)threadIdx.x >= tile.w) continue;
        const SimDev &s = sims[tile.x];
        if (s.offdiag) continue;  // stepped by the two-kernel row path below
        const Run q = work_run<V>(s, tile, threadIdx.x);
        const Clock ck = clock_of(s, adv);
        if (s.pml)
            dpair_run<V, true, NT, INV, GEN>(s, ck, q);
        else
            dpair_run<V, false, NT, INV, GEN>(s, ck, q);
    }
}

// Off-diagonal chi1inv: the row reads the partners' D - sum P at neighbouring cells, so the
// D curl (with its deposit) completes over the whole grid first, then E and P.
template <int V, bool PML>
__device__ inline void dcurl_run(const SimDev &s, const Clock &ck, const Run &q) {
    VF<V> h[3], hn[3][3];
    load_h<V, PML>(s, q, h, hn);
    metal_down<V>(s, q, hn);  // the general forms: zero below a metallic wall
    VF<V> d[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) d[c] = ldv<V>(s.D[c], q.at);
    stv<V>(s.D[0], q.at, dcurl_component<V, 0, PML, true>(s, ck, q, h, hn, d[0]));
    stv<V>(s.D[1], q.at, dcurl_component<V, 1, PML, true>(s, ck, q, h, hn, d[1]));
    stv<V>(s.D[2], q.at, dcurl_component<V, 2, PML, true>(s, ck, q, h, hn, d[2]));
}

template <int V>
__global__ void __launch_bounds__(kThreads) dcurl_kernel(const SimDev *__restrict__ sims,
                                                         const int4 *__restrict__ tiles,
                                                         int ntiles, int xcd_remap, int adv) {
    const TileWalk w = tile_walk(ntiles, xcd_remap);
    for (int g = w.next; g < w.end; g += w.stride) {
        const int t = tile_at(w, g);
        if (t >= ntiles) continue;
        const int4 tile = tiles[t];
        if ((int)threadIdx.x >= tile.w) continue;
        const SimDev &s = sims[tile.x];
        if (!s.offdiag) continue;
        const Run q = work_run<V>(s, tile, threadIdx.x);
        const Clock ck = clock_of(s, adv);
        if (s.pml)
            dcurl_run<V, true>(s, ck, q);
        else
            dcurl_run<V, false>(s, ck, q);
    }
}

__device__ inline float dmp_at(const SimDev &s, int ppar, int comp, int at) {
    float v = gp(s.D[comp])[at];
#pragma unroll
    for (int pq = 0; pq < kMaxPoles; ++pq) {
        if (pq >= s.npol) break;
        if (s.pdriven[pq][comp]) v = v - gp(s.pbuf[pq][comp][ppar])[at];
    }
    return v;
}

// The electric polarization sources of component `comp` subtracted from the word x at padded
// index `at`, as polarize_run does for a run (the same products and roundings, source by source
// in list order): MEEP's f_minus_p at a cell the off-diagonal rows read beside their own.
__device__ inline float polarize_at(const SimDev &s, int step, int comp, int at, float x) {
    const int i = at % s.pitch, r = at / s.pitch, j = r % s.n[1], k = r / s.n[1];
    if (i < s.src_lo[2][0] || i >= s.src_hi[2][0] || j < s.src_lo[2][1] || j >= s.src_hi[2][1] ||
        k < s.src_lo[2][2] || k >= s.src_hi[2][2])
        return x;
    for (int n = 0; n < s.nsrc; ++n) {
        const GMEM SrcDev &d = gp(s.src)[n];
        if (d.field != 2 || d.comp != comp) continue;
        const int a = i - d.lo[0], b = j - d.lo[1], e = k - d.lo[2];
        if (a < 0 || b < 0 || e < 0 || a >= d.cnt[0] || b >= d.cnt[1] || e >= d.cnt[2]) continue;
        if (step >= d.ncur) continue;
        const double pre = gp(d.cur)[2 * step], pim = gp(d.cur)[2 * step + 1];
        if (plus_zero(pre) && plus_zero(pim)) continue;
        const int p = (e * d.cnt[1] + b) * d.cnt[0] + a;
        double rr = gp(d.amp_re)[p] * pre;
        rr = rr - gp(d.amp_im)[p] * pim;
        double w = (double)x;
        w = w - rr;
        x = (float)w;
    }
    return x;
}

// f_minus_p less the poles' P at padded index `at`: (D - dipole) - P, the order of the
// constitutive update's own read.
__device__ inline float fmp_at(const SimDev &s, int step, int ppar, int comp, int at) {
    float v = gp(s.D[comp])[at];
    if (s.npsrc) v = polarize_at(s, step, comp, at, v);
#pragma unroll
    for (int pq = 0; pq < kMaxPoles; ++pq) {
        if (pq >= s.npol) break;
        if (s.pdriven[pq][comp]) v = v - gp(s.pbuf[pq][comp][ppar])[at];
    }
    return v;
}

// Padded-index offset of the +1 / -1 neighbour of cell (i, j, k) along `axis` (periodic).
__device__ inline int step_off(const SimDev &s, const Run &q, int i, int axis, int dir) {
    if (axis == 0) {
        if (dir > 0) return (i + 1 < s.n[0]) ? 1 : -(s.n[0] - 1);
        return (i > 0) ? -1 : (s.n[0] - 1);
    }
    if (axis == 1) return dir > 0 ? q.yp : q.ym;
    return dir > 0 ? q.zp : q.zm;
}

// The off-diagonal row's partner pair of one part at cell c3 (padded index `at`): D - P of
// partner `pa` there, read across a phased partner axis's window edge as the word above it,
// plus its lower neighbour along `pa` -- zero past a metallic wall, the mirror ghost on a
// folded axis (the parity of D_pa about its own plane, -phase, times stored row 2), else the
// periodic neighbour (stepping._offdiagonal_terms :1594-1687 with _shift_down :2485).
__device__ inline float offdiag_pair(const SimDev &s, const SimDev *o, int step, int ppar, int pa,
                                     int at, const int c3[3], int down) {
    float v = fmp_at(s, step, ppar, pa, at);
    if (s.pkind[pa] && c3[pa] == 0)
        v = connect_value(s, pa, false, v,
                          (s.pkind[pa] == 3 && o) ? fmp_at(*o, step, ppar, pa, at) : 0.0f);
    float lower;
    if (s.metal[pa] && c3[pa] == 0) {
        lower = 0.0f;
    } else if (s.fold[pa] && c3[pa] == 0) {
        const int stride = pa == 0 ? 1 : (pa == 1 ? s.pitch : s.pitch * s.n[1]);
        const float m = fmp_at(s, step, ppar, pa, at + 2 * stride);
        lower = s.fphase[pa] > 0 ? -m : m;
    } else {
        lower = fmp_at(s, step, ppar, pa, at + down);
    }
    return v + lower;
}

// The row of stepping._offdiagonal_terms (stepping.py:1594-1687), summed in double. On a folded
// or metallic own axis the product above the top row reads zero (_shift_up's zero face); at
// plane 0 of a phased own axis the product here is the word below the window, which for a
// product connection is formed from both parts' products.
template <int V, int C, bool PML>
__device__ inline void erow_component(const SimDev &s, const Clock &ck, const Run &q) {
    VF<V> p[kMaxPoles];
    const bool active = pole_run_active(s, q, V);
    VF<V> d = ldv<V>(s.D[C], q.at);
    // the polarization sources' dipole (on a folded grid, whose ghost rows carry their images'
    // boxes; a real off-diagonal row refuses such sources)
    if (s.npsrc) polarize_run<V>(s, ck.step, C, q, d);
    const VF<V> dmp = minus_poles<V, C, kMaxPoles>(s, ck, q, d, p, active);
    const VF<V> iv = ldv<V>(s.inv[C], q.at);
    VF<V> cst;
    const float *u1 = s.u[C][(C + 1) % 3], *u2 = s.u[C][(C + 2) % 3];
#pragma unroll
    for (int m = 0; m < V; ++m) {
        const int at = q.at + m, i = q.i0 + m;
        if (i >= s.n[0] || (!u1 && !u2)) {
            cst.v[m] = iv.v[m] * dmp.v[m];
            continue;
        }
        double row = (double)(dmp.v[m] * iv.v[m]);
        const int c3[3] = {i, q.j, q.k};
#pragma unroll
        for (int o = 1; o <= 2; ++o) {
            const int pa = (C + o) % 3;
            const float *U = o == 1 ? u1 : u2;
            if (!U) continue;
            // u[r]*(g[r] + g[r - s_partner]) at r = here and r = here + s_own; a shift along
            // the own axis leaves the coordinate along the partner's axis unchanged
            const bool top_c = (s.metal[C] || s.fold[C]) && c3[C] == s.n[C] - 1;
            const int up = at + step_off(s, q, i, C, +1);
            const int iu = (C == 0) ? ((i + 1 < s.n[0]) ? i + 1 : 0) : i;
            int u3[3] = {iu, q.j, q.k};
            if (C == 1) u3[1] = (q.j + 1 < s.n[1]) ? q.j + 1 : 0;
            if (C == 2) u3[2] = (q.k + 1 < s.n[2]) ? q.k + 1 : 0;
            const int down0 = step_off(s, q, i, pa, -1);
            const int down1 = step_off(s, q, iu, pa, -1);
            float prod0 = gp(U)[at] * offdiag_pair(s, s.other, ck.step, ck.ppar, pa, at, c3, down0);
            if (s.pkind[C] && c3[C] == 0) {
                float other = 0.0f;
                if (s.pkind[C] == 3 && s.other)
                    other = gp(U)[at] * offdiag_pair(*s.other, &s, ck.step, ck.ppar, pa, at, c3, down0);
                prod0 = connect_value(s, C, true, prod0, other);
            }
            float prod1 = 0.0f;
            if (!top_c) prod1 = gp(U)[up] * offdiag_pair(s, s.other, ck.step, ck.ppar, pa, up, u3, down1);
            float summed = prod0 + prod1;
            if ((s.metal[(C + 1) % 3] && c3[(C + 1) % 3] == 0) ||
                (s.metal[(C + 2) % 3] && c3[(C + 2) % 3] == 0))
                summed = 0.0f;
            row = row + 0.25 * (double)summed;
        }
        cst.v[m] = (float)row;
    }
    e_and_p<V, C, PML, false, kMaxPoles>(s, ck, q, cst, p, active);
}

template <int V>
__global__ void __launch_bounds__(kThreads) erow_kernel(const SimDev *__restrict__ sims,
                                                        const int4 *__restrict__ tiles,
                                                        int ntiles, int xcd_remap, int adv) {
    const TileWalk w = tile_walk(ntiles, xcd_remap);
    for (int g = w.next; g < w.end; g += w.stride) {
        const int t = tile_at(w, g);
        if (t >= ntiles) continue;
        const int4 tile = tiles[t];
        if ((int)threadIdx.x >= tile.w) continue;
        const SimDev &s = sims[tile.x];
        if (!s.offdiag) continue;
        const Run q = work_run<V>(s, tile, threadIdx.x);
        const Clock ck = clock_of(s, adv);
        if (s.pml) {
            erow_component<V, 0, true>(s, ck, q);
            erow_component<V, 1, true>(s, ck, q);
            erow_component<V, 2, true>(s, ck, q);
        } else {
            erow_component<V, 0, false>(s, ck, q);
            erow_component<V, 1, false>(s, ck, q);
            erow_component<V, 2, false>(s, ck, q);
        }
    }
}

// ------------------------------------------------------------------------------------------
// Mirror folds: the ghost planes from their images, after the sub-step's sources
// (fill_symmetry_bc_* stepping.py:2004-2063, zero_metal_* :3094-3183, then
// fill_folded_far_ghosts_* :2083-2145). A ghost's image is the stored cell that every ghost
// coordinate maps to -- row 2 for the near plane, frow for the far one -- which no fold writes
// and which the pair left final (its sources deposited). The reference's plane-by-plane passes
// leave the image times the near planes' parities, then +0 if the cell lies on a wall (the
// wall is cleared after the near fill), then times the far planes' parities (the far fill
// comes last): the word written here, signs of zero included. An item is (simulation, plane
// 2*axis + far, first cell, cells) of one folded plane; a cell that a component has as a ghost
// on two planes is written from the first of them (x before y before z, near before far).
// Family 1 (B) then updates H at the ghosts on the own-axis layer's rows, as the H pair does
// everywhere else; family 0 (D) leaves E to the row kernel that follows it.
// ------------------------------------------------------------------------------------------
__device__ inline float mirror_parity(int family, int comp, int axis, int phase) {
    // MEEP's phase_shift (reference fields.py:121-160): a vector (D) is odd about the plane
    // normal to its own axis and even otherwise, a pseudovector (B) the opposite; both times
    // the plane's declared phase
    const bool own = comp == axis;
    const int sign = family ? (own ? 1 : -1) : (own ? -1 : 1);
    return (float)(sign * phase);
}

__global__ void __launch_bounds__(256) fold_kernel(const SimDev *__restrict__ sims,
                                                   const int4 *__restrict__ items, int family,
                                                   int adv) {
    const int4 it = items[blockIdx.x];
    const SimDev &s = sims[it.x];
    const int a = it.y >> 1, far = it.y & 1;
    const int b1 = (a + 1) % 3, b2 = (a + 2) % 3;
    const Clock ck = clock_of(s, adv);
    for (int t = threadIdx.x; t < it.w; t += blockDim.x) {
        const int p = it.z + t;
        int c3[3];
        c3[a] = far ? s.n[a] - 1 : 0;
        c3[b1] = p % s.n[b1];
        c3[b2] = p / s.n[b1];
        for (int comp = 0; comp < 3; ++comp) {
            int src[3] = {c3[0], c3[1], c3[2]};
            float near_par = 1.0f, far_par = 1.0f;
            int first = -1;  // the first plane (2*axis + far) on which the cell is a ghost
            bool wall = false;  // on a wall of a (walled, unfolded) metallic axis
            for (int x = 0; x < 3; ++x) {
                const bool shift0 = family ? (x == comp) : (x != comp);
                wall |= s.metal[x] && shift0 && c3[x] == 0;
                if (!s.fold[x]) continue;
                int code = -1;
                if (shift0 && c3[x] == 0) {
                    src[x] = 2;
                    code = 2 * x;
                    near_par *= mirror_parity(family, comp, x, s.fphase[x]);
                } else if (s.fold[x] == 2 && !shift0 && c3[x] == s.n[x] - 1) {
                    src[x] = s.frow[x];
                    code = 2 * x + 1;
                    far_par *= mirror_parity(family, comp, x, s.fphase[x]);
                }
                if (code >= 0 && first < 0) first = code;
            }
            if (first != it.y) continue;
            const int at = (c3[2] * s.n[1] + c3[1]) * s.pitch + c3[0];
            const int sat = (src[2] * s.n[1] + src[1]) * s.pitch + src[0];
            float *F = family ? s.B[comp] : s.D[comp];
            float w = 0.0f;
            if (!wall) {
                const float v = gp(F)[sat];
                w = near_par < 0.0f ? -v : v;
            }
            if (far_par < 0.0f) w = -w;
            gp(F)[at] = w;
            if (family && s.pml && member(s, 0, comp, c3[comp])) {
                // update_H at the ghost: the H pair's own-axis layer form
                const float4 lw = coeff(s, 0, comp, c3[comp], true);
                float hold = w, prev = w;
                if (ck.hsep) {
                    hold = gp(s.H[comp])[at];
                    prev = gp(s.fwH[comp])[at];
                }
                gp(s.fwH[comp])[at] = w;
                const float aa = lw.z * w;
                const float mm = lw.x * prev;
                gp(s.H[comp])[at] = hold + (aa - mm);
            }
        }
    }
}

// E = chi1inv * (D - sum P): the reference's on-demand get_E for runs that do not store E.
__global__ void prime_e_kernel(SimDev *sims, int count) {
    const SimDev &s = sims[0];
    const int at = blockIdx.x * blockDim.x + threadIdx.x;
    if (at >= count) return;
    for (int c = 0; c < 3; ++c) s.E[c][at] = s.inv[c][at] * dmp_at(s, s.ppar, c, at);
}

// ------------------------------------------------------------------------------------------
// Flux planes: cell-centred tangential samples into a ring, then a float64 transform pass.
// ------------------------------------------------------------------------------------------
// A DFT region: the cells [lo, lo+cnt) of a simulation, with nslot (1..4) sampled components.
// A flux plane is the region one cell thick along its normal with the four tangential
// components E1, E2, H1, H2, cell-centred.
struct PlaneDev {
    const SimDev *sim;  // the descriptor the plane's simulation is stepped through
    int nslot;
    int comp[4];        // per slot: 0..2 = Ex..Ez, 3..5 = Hx..Hz
    int yee;            // 1: the component's own Yee site; 0: the cell centre (yee2cent)
    int hmask;          // bit r: slot r is magnetic (sampled at t - dt/2)
    int lo[3], cnt[3];
    int cells;
    int row;            // ring row stride: cells rounded up to kCB
    int nf;
    int dec;
    int first_step;     // sampling counts steps after this one
    int base;           // ring cursor: sample n (0-based since first_step) goes to slot
                        // (n - base) mod kRing, so that the transform pass after an early
                        // read (which moves base by the samples it took) starts at slot 0
    double scale;
    const double *omega;
    float *ring;        // [kRing][nslot][row]
    double *times;      // [kRing] electric sample time t (magnetic is t - dt/2)
    double *phase;      // [kRing][2][nf] (re, im)
    double *acc;        // [nslot][nf][cells] (re, im)
};

// MEEP yee2cent (dft.py:352): average along every axis where the component's Yee shift is 0,
// x first, then y, then z, each 0.5f*(a+b) in float32. `mask` bit a = shift 1 on axis a.
__device__ inline float centred(const SimDev &s, const float *f, int mask, int i, int j, int k) {
    const int ex = !(mask & 1), ey = !(mask & 2), ez = !(mask & 4);
    const int i1 = (i + 1 < s.n[0]) ? i + 1 : 0;
    const int j1 = (j + 1 < s.n[1]) ? j + 1 : 0;
    const int k1 = (k + 1 < s.n[2]) ? k + 1 : 0;
    const int py = s.pitch, pz = s.pitch * s.n[1];
    const GMEM float *gf = gp(f);
    auto at = [&](int a, int b, int c) { return gf[c * pz + b * py + a]; };
    float v000 = at(i, j, k), v100 = 0.f, v010 = 0.f, v110 = 0.f;
    float v001 = 0.f, v101 = 0.f, v011 = 0.f, v111 = 0.f;
    if (ex) v100 = at(i1, j, k);
    if (ey) v010 = at(i, j1, k);
    if (ex && ey) v110 = at(i1, j1, k);
    if (ez) v001 = at(i, j, k1);
    if (ex && ez) v101 = at(i1, j, k1);
    if (ey && ez) v011 = at(i, j1, k1);
    if (ex && ey && ez) v111 = at(i1, j1, k1);
    if (ex) {
        v000 = 0.5f * (v000 + v100);
        v010 = 0.5f * (v010 + v110);
        v001 = 0.5f * (v001 + v101);
        v011 = 0.5f * (v011 + v111);
    }
    if (ey) {
        v000 = 0.5f * (v000 + v010);
        v001 = 0.5f * (v001 + v011);
    }
    if (ez) v000 = 0.5f * (v000 + v001);
    return v000;
}

// The H component c at cell (i, j, k): stored H on c's own-axis layer rows, B elsewhere.
__device__ inline float h_at(const SimDev &s, int c, int i, int j, int k) {
    const int idx = (k * s.n[1] + j) * s.pitch + i;
    if (!s.pml) return gp(s.B[c])[idx];
    const int cc = c == 0 ? i : (c == 1 ? j : k);
    return member(s, 0, c, cc) ? gp(s.H[c])[idx] : gp(s.B[c])[idx];
}

// yee2cent for an H component: its Yee shift is 0 only along its own axis c, so the average
// is of the cell and its +1 neighbour along c (which may sit on the other side of the layer).
__device__ inline float centred_h(const SimDev &s, int c, int i, int j, int k) {
    int i1 = i, j1 = j, k1 = k;
    if (c == 0) i1 = (i + 1 < s.n[0]) ? i + 1 : 0;
    else if (c == 1) j1 = (j + 1 < s.n[1]) ? j + 1 : 0;
    else k1 = (k + 1 < s.n[2]) ? k + 1 : 0;
    return 0.5f * (h_at(s, c, i, j, k) + h_at(s, c, i1, j1, k1));
}

// Component c of a region at cell (i, j, k): at the component's own Yee site, or centred.
__device__ inline float sample_of(const SimDev &s, int c, int yee, int i, int j, int k) {
    if (c < 3) {
        if (yee) return gp(s.E[c])[(k * s.n[1] + j) * s.pitch + i];
        return centred(s, s.E[c], 1 << c, i, j, k);  // E_c: Yee shift 1 along its own axis
    }
    if (yee) return h_at(s, c - 3, i, j, k);
    return centred_h(s, c - 3, i, j, k);
}

// One launch samples every region of a simulation (or of a batch) that accumulates this step.
// `adv`: steps since the descriptors' last counter advance (the sample follows the step's
// advance, so a captured step t samples at the descriptor's step + t + 1).
__global__ void flux_sample_kernel(const PlaneDev *planes, int nplanes, int adv) {
    const int pi = blockIdx.y;
    if (pi >= nplanes) return;
    const PlaneDev &p = planes[pi];
    const SimDev &s = *p.sim;
    const int step = s.step + adv;
    const int since = step - p.first_step;
    if (since <= 0 || since % p.dec) return;
    int slot = (since / p.dec - 1 - p.base) % kRing;
    if (slot < 0) slot += kRing;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (blockIdx.x == 0 && threadIdx.x == 0) gp(p.times)[slot] = (double)step * s.dt;
    if (q >= p.cells) return;
    const int a = q % p.cnt[0], r = q / p.cnt[0];
    const int b = r % p.cnt[1], e = r / p.cnt[1];
    const int i = (p.lo[0] + a) % s.n[0], j = (p.lo[1] + b) % s.n[1], k = (p.lo[2] + e) % s.n[2];
    GMEM float *out = gp(p.ring) + (size_t)slot * p.nslot * p.row;
    for (int r = 0; r < p.nslot; ++r) out[r * p.row + q] = sample_of(s, p.comp[r], p.yee, i, j, k);
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
    sincos(p.omega[f] * (t - 0.5 * p.sim->dt), &sH, &cH);
    double *ph = p.phase + (size_t)m * 4 * p.nf;
    ph[2 * f] = cE * p.scale;
    ph[2 * f + 1] = sE * p.scale;
    ph[2 * p.nf + 2 * f] = cH * p.scale;
    ph[2 * p.nf + 2 * f + 1] = sH * p.scale;
}

// acc[comp][f][cell] += sum_m phase[m][E|H][f] * ring[m][comp][cell], in float64: per plane and
// component a (2*nf x nslots) by (nslots x cells) product. Each thread owns kCB consecutive
// cells and kFB frequencies of accumulators; the workgroup stages the phase words of its
// frequencies in LDS (every lane reads them as broadcasts) and the frequency range is split
// over blockIdx.z so that small planes still fill the machine.
constexpr int kCB = 4;   // cells per thread
constexpr int kFB = 8;   // frequencies per register block
constexpr int kFZ = 16;  // frequencies per workgroup

__global__ void __launch_bounds__(256) flux_accumulate_kernel(
    const float *__restrict__ ring, const double *__restrict__ phase, double *__restrict__ acc,
    int cells, int row, int nf, int nslots, int nslot, int hmask) {
    __shared__ double2 ph_s[kRing][kFZ];
    const int comp = blockIdx.y;
    const int side = (hmask >> comp) & 1;
    const int fz0 = blockIdx.z * kFZ;
    const int fzn = nf - fz0 < kFZ ? nf - fz0 : kFZ;
    for (int x = threadIdx.x; x < kRing * kFZ; x += blockDim.x) {
        const int m = x / kFZ, fi = x % kFZ;
        double2 v = make_double2(0.0, 0.0);
        if (m < nslots && fi < fzn) {
            const double *ph = phase + (size_t)m * 4 * nf + side * 2 * nf + 2 * (fz0 + fi);
            v = make_double2(ph[0], ph[1]);
        }
        ph_s[m][fi] = v;
    }
    __syncthreads();
    const int q0 = (blockIdx.x * blockDim.x + threadIdx.x) * kCB;
    if (q0 >= cells) return;
    double *a = acc + (size_t)comp * nf * cells * 2;
    for (int f0 = 0; f0 < fzn; f0 += kFB) {
        double re[kFB][kCB], im[kFB][kCB];
#pragma unroll
        for (int fi = 0; fi < kFB; ++fi)
#pragma unroll
            for (int c = 0; c < kCB; ++c) re[fi][c] = im[fi][c] = 0.0;
        for (int m = 0; m < nslots; ++m) {
            const float4 sv = *reinterpret_cast<const float4 *>(ring + ((size_t)m * nslot + comp) * row + q0);
            const double sm[kCB] = {(double)sv.x, (double)sv.y, (double)sv.z, (double)sv.w};
#pragma unroll
            for (int fi = 0; fi < kFB; ++fi) {
                const double2 w = ph_s[m][f0 + fi];
#pragma unroll
                for (int c = 0; c < kCB; ++c) {
                    re[fi][c] = fma(w.x, sm[c], re[fi][c]);  // float64 by design: fused is closer
                    im[fi][c] = fma(w.y, sm[c], im[fi][c]);
                }
            }
        }
#pragma unroll
        for (int fi = 0; fi < kFB; ++fi) {
            if (f0 + fi >= fzn) break;
            double2 *cell = reinterpret_cast<double2 *>(a + ((size_t)(fz0 + f0 + fi) * cells + q0) * 2);
#pragma unroll
            for (int c = 0; c < kCB; ++c) {
                if (q0 + c >= cells) break;
                double2 old = cell[c];
                old.x += re[fi][c];
                old.y += im[fi][c];
                cell[c] = old;
            }
        }
    }
}

// The same product on the matrix cores: per component, C[n][cell] += sum_m P[n][m] R[m][cell]
// with n = 2f + (re, im), P the phase words and R the ring's samples widened to double, in
// 16x16x4 float64 MFMA steps. V_MFMA_F64_16X16X4F64 on a wave: lane l supplies A[l%16][l/16]
// and B[l/16][l%16] and holds D[l/16 + 4r][l%16] in accumulator r. Each wave owns CT tiles of
// 16 cells and kMT tiles of 16 n, so each phase word it loads feeds CT MFMAs and each ring word
// kMT; tiles wholly past the frequencies or the cells are skipped (with 100 frequencies, 3 of
// the 16 tiles a wave range covers). The ring rows are padded to 16 cells.
typedef double d4v __attribute__((ext_vector_type(4)));
constexpr int kMT = 4;  // 16-row n tiles per wave

// CT: 16-cell tiles per wave (each operand load feeds kMT or CT products). Measured on the
// 192^3 sphere's planes: one tile per wave, 129 us a pass; two, 183 us; four, 296 us (fewer
// waves resident: 8, 3 and 1 per SIMD).
template <int CT>
__global__ void __launch_bounds__(256) flux_accumulate_mfma_kernel(
    const float *__restrict__ ring, const double *__restrict__ phase, double *__restrict__ acc,
    int cells, int row, int nf, int nslots, int nslot, int hmask) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int comp = blockIdx.y;
    const int side = (hmask >> comp) & 1;
    const int cell0 = (blockIdx.x * 4 + wave) * 16 * CT;
    if (cell0 >= cells) return;
    const int nn = 2 * nf;
    const int n0 = blockIdx.z * 16 * kMT;
    const int kk = lane >> 4;      // this lane's k (slot) within a 4-slot step
    const int ni = lane & 15;      // this lane's n within a tile (A row) and cell (B column)
    // the tiles this wave multiplies: wave-uniform counts, so every MFMA runs on a full wave
    int mt = (nn - n0 + 15) / 16, ct = (cells - cell0 + 15) / 16;
    mt = mt < kMT ? mt : kMT;
    ct = ct < CT ? ct : CT;
    d4v c[kMT][CT];
#pragma unroll
    for (int t = 0; t < kMT; ++t)
#pragma unroll
        for (int u = 0; u < CT; ++u) c[t][u] = (d4v){0.0, 0.0, 0.0, 0.0};
    const GMEM float *rg = gp(ring) + (size_t)comp * row;
    const GMEM double *ph = gp(phase) + (size_t)side * nn;
    for (int m0 = 0; m0 < nslots; m0 += 4) {
        const int m = m0 + kk;
        const bool live = m < nslots;
        double b[CT], av[kMT];
#pragma unroll
        for (int u = 0; u < CT; ++u) {
            const int cell = cell0 + u * 16 + ni;  // < row (a multiple of 16) for u < ct
            b[u] = (live && u < ct) ? (double)rg[(size_t)m * nslot * row + cell] : 0.0;
        }
#pragma unroll
        for (int t = 0; t < kMT; ++t) {
            const int n = n0 + t * 16 + ni;
            av[t] = (live && n < nn) ? ph[(size_t)m * 4 * nf + n] : 0.0;
        }
#pragma unroll
        for (int t = 0; t < kMT; ++t) {
            if (t >= mt) break;
#pragma unroll
            for (int u = 0; u < CT; ++u) {
                if (u >= ct) break;
                c[t][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[t], b[u], c[t][u], 0, 0, 0);
            }
        }
    }
    GMEM double *out = gp(acc) + (size_t)comp * nf * cells * 2;
#pragma unroll
    for (int t = 0; t < kMT; ++t)
#pragma unroll
        for (int u = 0; u < CT; ++u) {
            const int cell = cell0 + u * 16 + ni;
            if (t >= mt || u >= ct || cell >= cells) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int n = n0 + t * 16 + kk + 4 * r;
                if (n >= nn) continue;
                GMEM double *w = out + ((size_t)(n >> 1) * cells + cell) * 2 + (n & 1);
                *w += c[t][u][r];
            }
        }
}

// acc *= (re + i im) over n complex float64 words (dft_chunk::scale_dft as reference
// dft.py:3274-3286 restates it: every stored element, E and H sides alike).
__global__ void flux_scale_kernel(double *acc, size_t n, double re, double im) {
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    GMEM double *a = gp(acc) + 2 * q;
    const double x = a[0], y = a[1];
    double r = x * re;
    r = r - y * im;
    double i = x * im;
    i = i + y * re;
    a[0] = r;
    a[1] = i;
}

// The flux spectrum of one plane reduced on the device (dft.py:3107-3130 as the host restates
// it): per frequency f, sum over the plane's cells of w * Re(E1 conj(H2) - E2 conj(H1)) from the
// float64 accumulators [E1, E2, H1, H2][f][cell], w per cell or 1. Stage 1: workgroup (b, f)
// sums the cells b, b + G, ... (G = gridDim.x workgroups of 256 threads per frequency) into
// partial[f][b]: each lane adds its cells in order, the wave folds its 64 lanes with cross-lane
// shuffles and the four waves meet in LDS. Stage 2 sums partial[f][0..G) in order. Both are
// float64; the host's own sum associates differently and agrees to rounding.
constexpr int kSpecBlocks = 128;  // stage-1 workgroups per frequency (at most)

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__global__ void __launch_bounds__(256) flux_spectrum_partial_kernel(
    const double *__restrict__ acc, const double *__restrict__ w, int cells, int nf,
    double *__restrict__ partial) {
    __shared__ double part[4];
    const int f = blockIdx.y, G = gridDim.x;
    const size_t comp = (size_t)nf * cells * 2;  // doubles per component
    const GMEM double *e1 = gp(acc) + (size_t)f * cells * 2;
    const GMEM double *e2 = e1 + comp;
    const GMEM double *h1 = e1 + 2 * comp;
    const GMEM double *h2 = e1 + 3 * comp;
    double s = 0.0;
    for (int c = blockIdx.x * 256 + (int)threadIdx.x; c < cells; c += G * 256) {
        double a = e1[2 * c] * h2[2 * c];
        a = a + e1[2 * c + 1] * h2[2 * c + 1];
        double b = e2[2 * c] * h1[2 * c];
        b = b + e2[2 * c + 1] * h1[2 * c + 1];
        double t = a - b;
        if (w) t = t * gp(w)[c];
        s = s + t;
    }
    s = wave_sum(s);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) part[wave] = s;
    __syncthreads();
    if (threadIdx.x == 0)
        gp(partial)[(size_t)f * G + blockIdx.x] = ((part[0] + part[1]) + part[2]) + part[3];
}

__global__ void flux_spectrum_final_kernel(const double *__restrict__ partial, int G, int nf,
                                           double *__restrict__ out) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    double s = 0.0;
    for (int b = 0; b < G; ++b) s = s + gp(partial)[(size_t)f * G + b];
    gp(out)[f] = s;
}

// One int of a device structure set in stream order (a plane's ring cursor after an early
// transform pass), with the value as a kernel argument: no host buffer has to outlive the call.
__global__ void set_int_kernel(int *p, int v) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *p = v;
}

__global__ void fill_kernel(float *p, size_t n, float v) {
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) p[q] = v;
}

__global__ void copy_probe_kernel(const float4 *__restrict__ src, float4 *__restrict__ dst,
                                  size_t n) {
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
         q += (size_t)gridDim.x * blockDim.x)
        dst[q] = src[q];
}

// The fused pairs' access pattern without the arithmetic: `nr` read streams and `nw` write
// streams of n floats, every stream touched once at the same index by one thread.
struct StreamSet {
    float *p[16];
};

__global__ void stream_probe_kernel(StreamSet set, int nr, int nw, int n) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float acc = 0.0f;
#pragma unroll
    for (int r = 0; r < 16; ++r)
        if (r < nr) acc += set.p[r][i];
#pragma unroll
    for (int w = 0; w < 16; ++w)
        if (w < nw) set.p[nr + w][i] = acc + (float)w;
}

// Read-only and write-only streams (the two halves of a copy, measured apart).
__global__ void read_probe_kernel(const float4 *__restrict__ src, float *__restrict__ out, size_t n) {
    float acc = 0.f;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
         q += (size_t)gridDim.x * blockDim.x) {
        const float4 v = src[q];
        acc += (v.x + v.y) + (v.z + v.w);
    }
    if (acc == 12345.678f) out[blockIdx.x] = acc;  // keeps the loads; never true for zeros
}

__global__ void write_probe_kernel(float4 *__restrict__ dst, size_t n) {
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
         q += (size_t)gridDim.x * blockDim.x)
        dst[q] = make_float4(0.f, 1.f, 2.f, 3.f);
}

// Four independent loads in flight per thread before the stores; `nt` marks the stores
// non-temporal (streamed past the caches).
template <bool NT>
__global__ void copy_probe4_kernel(const float4 *__restrict__ src, float4 *__restrict__ dst,
                                   size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += 4 * stride) {
        float4 v[4];
#pragma unroll
        for (int r = 0; r < 4; ++r)
            if (q + r * stride < n) v[r] = src[q + r * stride];
#pragma unroll
        for (int r = 0; r < 4; ++r)
            if (q + r * stride < n) {
                if (NT) {
                    __builtin_nontemporal_store(v[r].x, &dst[q + r * stride].x);
                    __builtin_nontemporal_store(v[r].y, &dst[q + r * stride].y);
                    __builtin_nontemporal_store(v[r].z, &dst[q + r * stride].z);
                    __builtin_nontemporal_store(v[r].w, &dst[q + r * stride].w);
                } else {
                    dst[q + r * stride] = v[r];
                }
            }
    }
}

// A one-pass copy with U float4 loads in flight per thread, all issued before the first store:
// workgroup b copies the U consecutive 256-float4 slabs starting at slab b*U (each slab one
// coalesced dwordx4 sweep of the 256 threads). NTL / NTS: non-temporal loads / stores. The grid
// covers n exactly (n a multiple of 256*U), so every access is in bounds.
template <int U, bool NTL, bool NTS>
__global__ void __launch_bounds__(256) copy_wide_kernel(const f4v *__restrict__ src,
                                                        f4v *__restrict__ dst) {
    const size_t base = (size_t)blockIdx.x * U * 256 + threadIdx.x;
    const GMEM f4v *s = gp(src);
    GMEM f4v *d = gp(dst);
    f4v v[U];
#pragma unroll
    for (int u = 0; u < U; ++u)
        v[u] = NTL ? __builtin_nontemporal_load(s + base + (size_t)u * 256) : s[base + (size_t)u * 256];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        if (NTS) __builtin_nontemporal_store(v[u], d + base + (size_t)u * 256);
        else d[base + (size_t)u * 256] = v[u];
    }
}

}  // namespace

// The device scratch of fdtd_flux_spectrum_noalloc: the stage-1 partial sums of one spectrum,
// [nf][kSpecBlocks] doubles for up to kScratchFreqs frequencies (1 MiB). Each GPU holds its own
// copy for the life of the process; every call on a GPU shares it.
constexpr int kScratchFreqs = 1024;
static __device__ __align__(256) double scratch[kSpecBlocks * kScratchFreqs];

// ==========================================================================================
// Host side
// ==========================================================================================
namespace {
struct LaunchShape {
    int vec = 1;            // cells per thread: 1, 2 or 4
    int blocks_per_cu = 8;  // 0: one workgroup per tile; else a persistent grid
    int xcd_remap = 0;
    int nt = 0;             // non-temporal stores of the fused pairs' outputs
    int svec = 1;           // cells per thread of the shell's lean kernel (its layer forms hold
                            // more operands, so fewer cells a thread keep more waves resident)
    int tile = 0;           // the D pair's box rows over LDS-staged halo tiles of this many rows
                            // (2, 4 or 8; one cell a thread), or 0: runs of vec cells
};
// How one simulation's (or batch's) launches run: each fused pair's launch shape ([0] the H
// pair's launches, [1] the D pair's and its row path), whether fdtd_run / fdtd_batch_run replay
// captured graphs, and the flux transform kernel (1, 2 or 4: MFMA tiles per wave; 0: vector).
// The words of a configuration: per pair (vec, blocks_per_cu, xcd_remap, nt, svec, tile),
// then graphs and the flux transform.
constexpr int kShapeWords = 6;
constexpr int kConfigWords = 2 * kShapeWords + 2;
struct LaunchConfig {
    LaunchShape shape[2];
    int graphs = 1;
    int flux_mfma = 1;
};
// The process defaults (the process-wide setters): every object without a configuration of
// its own (fdtd_sim_set_config, fdtd_batch_set_config) follows them, as they change.
LaunchConfig g_defaults;
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

int grid_for(const LaunchShape &sh, int ntiles) {
    int g = sh.blocks_per_cu > 0 ? num_cu() * sh.blocks_per_cu : ntiles;
    if (g > ntiles) g = ntiles;
    g = ((g + 7) / 8) * 8;
    return g < 8 ? 8 : g;
}

bool same_shape(const LaunchShape &a, const LaunchShape &b) {
    return a.vec == b.vec && a.blocks_per_cu == b.blocks_per_cu && a.xcd_remap == b.xcd_remap &&
           a.nt == b.nt && a.svec == b.svec && a.tile == b.tile;
}

struct PlaneHost {
    PlaneDev dev;
    int filled = 0;  // samples in the ring since the last transform pass
    // the device flux-spectrum reduction's buffers (fdtd_flux_spectrum), made on first use:
    // cell weights, stage-1 partial sums [nf][kSpecBlocks], the nf results
    double *spec_w = nullptr, *spec_partial = nullptr, *spec_out = nullptr;
};

// A set of simulations stepped by the same launches: one, or a batch.
struct Group {
    SimDev *dev = nullptr;          // descriptors, contiguous
    int nsims = 0;
    std::map<int, const int4 *> tiles;  // work items per vector width (interior and shell)
    std::map<int, int> ntiles;
    std::map<int, const int4 *> shell_tiles, int_tiles;  // the shell's items, the box's items
    std::map<int, int> nshell_tiles, nint_tiles;
    std::map<int, const int4 *> lds_tiles;  // the box's LDS halo tiles, per tile row count
    std::map<int, int> nlds_tiles;
    int max_npol = 0;     // the most poles a member has
    bool any_pml = false;  // some member has an absorber (its box rows hold x-layer cells)
    bool any_offdiag = false, all_offdiag = false;
    bool general = false;  // some member has a conductivity, a metallic wall or more than
                           // kLosslessPoles poles: the pairs' general instantiation (the others
                           // keep the lossless periodic code)
    int inv_mode = 3;  // the dpair_kernel INV instantiation that serves every member
    const LaunchConfig *cfg = &g_defaults;  // the owner's launch configuration
    const int4 *fold_items = nullptr;  // the folded members' ghost planes (fold_kernel)
    int nfold = 0;
};

// Does a descriptor need the pairs' general instantiation: a conductivity, a metallic wall,
// more than kLosslessPoles poles or a polarization source?
bool needs_general(const SimDev &h) {
    return h.cond != 0 || h.metal[0] || h.metal[1] || h.metal[2] || h.npol > kLosslessPoles ||
           h.npsrc > 0 || h.fold[0] || h.fold[1] || h.fold[2] || h.pkind[0] || h.pkind[1] ||
           h.pkind[2];
}

// The one inverse-epsilon storage all three components of a descriptor use, or 3 (mixed).
int inv_mode_of(const SimDev &h) {
    int m = -1;
    for (int c = 0; c < 3; ++c) {
        const int mc = h.inv8[c] ? 1 : h.inv16[c] ? 2 : 0;
        if (m < 0) m = mc;
        else if (m != mc) return 3;
    }
    return m;
}

template <int V, bool NT, bool GEN>
void launch_dpair_c(const Group &g, int G, const int4 *tl, int nt, int rm, int adv, hipStream_t st) {
    switch (g.inv_mode) {
        case 0: dpair_kernel<V, NT, 0, GEN><<<G, kThreads, 0, st>>>(g.dev, tl, nt, rm, adv); break;
        case 1: dpair_kernel<V, NT, 1, GEN><<<G, kThreads, 0, st>>>(g.dev, tl, nt, rm, adv); break;
        case 2: dpair_kernel<V, NT, 2, GEN><<<G, kThreads, 0, st>>>(g.dev, tl, nt, rm, adv); break;
        default: dpair_kernel<V, NT, 3, GEN><<<G, kThreads, 0, st>>>(g.dev, tl, nt, rm, adv); break;
    }
}

// The general instantiations run only where a member has a conductivity, a metallic wall or
// more than kLosslessPoles poles: the lossless periodic pairs' code stays what it is.
template <int V, bool NT>
void launch_dpair(const Group &g, int G, const int4 *tl, int nt, int rm, int adv, hipStream_t st) {
    if (g.general) launch_dpair_c<V, NT, true>(g, G, tl, nt, rm, adv, st);
    else launch_dpair_c<V, NT, false>(g, G, tl, nt, rm, adv, st);
}

template <int V, bool NT>
void launch_hpair(const Group &g, int G, const int4 *tl, int nt, int rm, int adv, hipStream_t st) {
    if (g.general) hpair_kernel<V, NT, true><<<G, kThreads, 0, st>>>(g.dev, tl, nt, rm, adv);
    else hpair_kernel<V, NT, false><<<G, kThreads, 0, st>>>(g.dev, tl, nt, rm, adv);
}

// The two-kernel form: where no member runs the general forms, the box's rows and the shell's
// are two lean kernels (the D pair's also needs no off-diagonal row, float or byte-indexed
// inverse epsilon and at most one pole).
bool lean_d(const Group &g) {
    return !g.general && !g.any_offdiag && (g.inv_mode == 0 || g.inv_mode == 1) && g.max_npol <= 1;
}

// The fused pairs' launches `adv` steps after the descriptors' last counter advance (0 for an
// eager step; a captured step's offset within its sequence).
template <int V, bool NT>
void launch_hbox(const Group &g, int G, const int4 *l, int n, int rm, int adv, hipStream_t st) {
    if (g.any_pml) hshell_kernel<V, NT, true><<<G, kThreads, 0, st>>>(g.dev, l, n, rm, adv);
    else hint_kernel<V, NT><<<G, kThreads, 0, st>>>(g.dev, l, n, rm, adv);
}

int launch_h(const Group &g, hipStream_t st, int adv = 0) {
    const LaunchShape &sh = g.cfg->shape[0];
    const int V = sh.vec, SV = sh.svec;
    const int rm = sh.xcd_remap;
    if (!g.general) {  // the box's rows, then the shell's: two lean kernels
        const int ni = g.nint_tiles.at(V);
        const int4 *il = g.int_tiles.at(V);
        const int GI = grid_for(sh, ni);
        if (ni) {
            if (V == 4) { if (sh.nt) launch_hbox<4, true>(g, GI, il, ni, rm, adv, st); else launch_hbox<4, false>(g, GI, il, ni, rm, adv, st); }
            else if (V == 2) { if (sh.nt) launch_hbox<2, true>(g, GI, il, ni, rm, adv, st); else launch_hbox<2, false>(g, GI, il, ni, rm, adv, st); }
            else { if (sh.nt) launch_hbox<1, true>(g, GI, il, ni, rm, adv, st); else launch_hbox<1, false>(g, GI, il, ni, rm, adv, st); }
        }
        const int ns = g.nshell_tiles.at(SV);
        const int4 *sl = g.shell_tiles.at(SV);
        const int GS = grid_for(sh, ns);
        if (ns) {
            if (SV == 4) { if (sh.nt) hshell_kernel<4, true, false><<<GS, kThreads, 0, st>>>(g.dev, sl, ns, rm, adv); else hshell_kernel<4, false, false><<<GS, kThreads, 0, st>>>(g.dev, sl, ns, rm, adv); }
            else if (SV == 2) { if (sh.nt) hshell_kernel<2, true, false><<<GS, kThreads, 0, st>>>(g.dev, sl, ns, rm, adv); else hshell_kernel<2, false, false><<<GS, kThreads, 0, st>>>(g.dev, sl, ns, rm, adv); }
            else { if (sh.nt) hshell_kernel<1, true, false><<<GS, kThreads, 0, st>>>(g.dev, sl, ns, rm, adv); else hshell_kernel<1, false, false><<<GS, kThreads, 0, st>>>(g.dev, sl, ns, rm, adv); }
        }
        HIP_TRY(hipGetLastError());
        return FDTD_OK;
    }
    const int nt = g.ntiles.at(V);
    const int4 *tl = g.tiles.at(V);
    const int G = grid_for(sh, nt);
    if (sh.nt) {
        if (V == 4) launch_hpair<4, true>(g, G, tl, nt, rm, adv, st);
        else if (V == 2) launch_hpair<2, true>(g, G, tl, nt, rm, adv, st);
        else launch_hpair<1, true>(g, G, tl, nt, rm, adv, st);
    } else {
        if (V == 4) launch_hpair<4, false>(g, G, tl, nt, rm, adv, st);
        else if (V == 2) launch_hpair<2, false>(g, G, tl, nt, rm, adv, st);
        else launch_hpair<1, false>(g, G, tl, nt, rm, adv, st);
    }
    if (g.nfold) fold_kernel<<<g.nfold, 256, 0, st>>>(g.dev, g.fold_items, 1, adv);
    HIP_TRY(hipGetLastError());
    return FDTD_OK;
}

// The lean D kernels over a list: the box's rows (BOX: the x layer's forms where a member has
// an absorber, else the lossless forms) or the shell's (every layer role).
template <int V, bool NT, int INV, bool BOX>
void launch_dlean_inv(const Group &g, int G, const int4 *l, int n, int rm, int adv, hipStream_t st) {
    if (BOX && !g.any_pml) dint_kernel<V, NT, INV><<<G, kThreads, 0, st>>>(g.dev, l, n, rm, adv);
    else if (g.max_npol == 1) dshell_kernel<V, NT, INV, true, BOX><<<G, kThreads, 0, st>>>(g.dev, l, n, rm, adv);
    else dshell_kernel<V, NT, INV, false, BOX><<<G, kThreads, 0, st>>>(g.dev, l, n, rm, adv);
}

template <int V, bool NT, bool BOX>
void launch_dlean(const Group &g, int G, const int4 *l, int n, int rm, int adv, hipStream_t st) {
    if (g.inv_mode == 1) launch_dlean_inv<V, NT, 1, BOX>(g, G, l, n, rm, adv, st);
    else launch_dlean_inv<V, NT, 0, BOX>(g, G, l, n, rm, adv, st);
}

template <int TY>
void launch_dtile(const Group &g, int G, const int4 *l, int n, int rm, int adv, hipStream_t st,
                  bool nt) {
    const bool pole = g.max_npol == 1;
    if (g.inv_mode == 1) {
        if (nt) { if (pole) dtile_kernel<TY, true, 1, true><<<G, kThreads, 0, st>>>(g.dev, l, n, rm, adv); else dtile_kernel<TY, true, 1, false><<<G, kThreads, 0, st>>>(g.dev, l, n, rm, adv); }
        else { if (pole) dtile_kernel<TY, false, 1, true><<<G, kThreads, 0, st>>>(g.dev, l, n, rm, adv); else dtile_kernel<TY, false, 1, false><<<G, kThreads, 0, st>>>(g.dev, l, n, rm, adv); }
    } else {
        if (nt) { if (pole) dtile_kernel<TY, true, 0, true><<<G, kThreads, 0, st>>>(g.dev, l, n, rm, adv); else dtile_kernel<TY, true, 0, false><<<G, kThreads, 0, st>>>(g.dev, l, n, rm, adv); }
        else { if (pole) dtile_kernel<TY, false, 0, true><<<G, kThreads, 0, st>>>(g.dev, l, n, rm, adv); else dtile_kernel<TY, false, 0, false><<<G, kThreads, 0, st>>>(g.dev, l, n, rm, adv); }
    }
}

// ... then the counter advance by `advance_by` steps (0 inside a captured sequence, which
// advances once at its end).
int launch_d(const Group &g, hipStream_t st, int adv = 0, int advance_by = 1) {
    const LaunchShape &sh = g.cfg->shape[1];
    const int V = sh.vec, SV = sh.svec;
    const int rm = sh.xcd_remap;
    const int nt = g.ntiles.at(V);
    const int4 *tl = g.tiles.at(V);
    const int G = grid_for(sh, nt);
    if (lean_d(g)) {  // the box's rows, then the shell's
        const int ni = g.nint_tiles.at(V);
        const int4 *il = g.int_tiles.at(V);
        const int GI = grid_for(sh, ni);
        const int nl = sh.tile && V == 1 && g.any_pml ? g.nlds_tiles.at(sh.tile) : 0;
        if (nl) {  // the box over LDS-staged halo tiles
            const int4 *ll = g.lds_tiles.at(sh.tile);
            const int GL = grid_for(sh, nl);
            if (sh.tile == 2) launch_dtile<2>(g, GL, ll, nl, rm, adv, st, sh.nt);
            else if (sh.tile == 4) launch_dtile<4>(g, GL, ll, nl, rm, adv, st, sh.nt);
            else launch_dtile<8>(g, GL, ll, nl, rm, adv, st, sh.nt);
        } else if (ni) {
            if (V == 4) { if (sh.nt) launch_dlean<4, true, true>(g, GI, il, ni, rm, adv, st); else launch_dlean<4, false, true>(g, GI, il, ni, rm, adv, st); }
            else if (V == 2) { if (sh.nt) launch_dlean<2, true, true>(g, GI, il, ni, rm, adv, st); else launch_dlean<2, false, true>(g, GI, il, ni, rm, adv, st); }
            else { if (sh.nt) launch_dlean<1, true, true>(g, GI, il, ni, rm, adv, st); else launch_dlean<1, false, true>(g, GI, il, ni, rm, adv, st); }
        }
        const int ns = g.nshell_tiles.at(SV);
        const int4 *sl = g.shell_tiles.at(SV);
        const int GS = grid_for(sh, ns);
        if (ns) {
            if (SV == 4) { if (sh.nt) launch_dlean<4, true, false>(g, GS, sl, ns, rm, adv, st); else launch_dlean<4, false, false>(g, GS, sl, ns, rm, adv, st); }
            else if (SV == 2) { if (sh.nt) launch_dlean<2, true, false>(g, GS, sl, ns, rm, adv, st); else launch_dlean<2, false, false>(g, GS, sl, ns, rm, adv, st); }
            else { if (sh.nt) launch_dlean<1, true, false>(g, GS, sl, ns, rm, adv, st); else launch_dlean<1, false, false>(g, GS, sl, ns, rm, adv, st); }
        }
    } else if (!g.all_offdiag) {
        if (sh.nt) {
            if (V == 4) launch_dpair<4, true>(g, G, tl, nt, rm, adv, st);
            else if (V == 2) launch_dpair<2, true>(g, G, tl, nt, rm, adv, st);
            else launch_dpair<1, true>(g, G, tl, nt, rm, adv, st);
        } else {
            if (V == 4) launch_dpair<4, false>(g, G, tl, nt, rm, adv, st);
            else if (V == 2) launch_dpair<2, false>(g, G, tl, nt, rm, adv, st);
            else launch_dpair<1, false>(g, G, tl, nt, rm, adv, st);
        }
    }
    if (g.any_offdiag) {
        // (a folded member steps this path too: its ghost planes are filled between the two)
        if (V == 4) dcurl_kernel<4><<<G, kThreads, 0, st>>>(g.dev, tl, nt, rm, adv);
        else if (V == 2) dcurl_kernel<2><<<G, kThreads, 0, st>>>(g.dev, tl, nt, rm, adv);
        else dcurl_kernel<1><<<G, kThreads, 0, st>>>(g.dev, tl, nt, rm, adv);
        if (g.nfold) fold_kernel<<<g.nfold, 256, 0, st>>>(g.dev, g.fold_items, 0, adv);
        if (V == 4) erow_kernel<4><<<G, kThreads, 0, st>>>(g.dev, tl, nt, rm, adv);
        else if (V == 2) erow_kernel<2><<<G, kThreads, 0, st>>>(g.dev, tl, nt, rm, adv);
        else erow_kernel<1><<<G, kThreads, 0, st>>>(g.dev, tl, nt, rm, adv);
    }
    if (advance_by > 0) advance_kernel<<<(g.nsims + 63) / 64, 64, 0, st>>>(g.dev, g.nsims, advance_by);
    HIP_TRY(hipGetLastError());
    return FDTD_OK;
}
// A captured sequence of kRing steps (fused pairs, flux sampling and the transform passes that
// fall due, each launched with its step's offset; one counter advance by kRing at the end),
// replayed while the launch shape and every plane's ring
// phase are what they were at capture. Plane rings fill once per kRing steps when every plane
// samples each step, so one capture serves every later period.
// (g_defaults.graphs: replay captured step sequences where eligible; g_defaults.flux_mfma: the
// flux transform on the matrix cores with this many 16-cell tiles per
                      // wave (1, 2 or 4), or 0: the vector FMA kernel

// The kernel instantiations a group's launches select (its general/lossless forms, inverse
// epsilon storage and off-diagonal path): a captured sequence is valid only for the same ones.
int group_signature(const Group &g) {
    return (g.general ? 1 : 0) | (g.inv_mode << 1) | (g.any_offdiag ? 8 : 0) |
           (g.all_offdiag ? 16 : 0);
}

struct StepGraph {
    hipGraphExec_t exec = nullptr;
    hipStream_t st = nullptr;  // the stream its replays are queued on
    std::vector<int> phase;
    LaunchShape shape[2];
    int flux_mode = -1;  // the transform kernel the sequence was captured with
    int signature = -1;  // group_signature at capture
    int64_t captures = 0, replays = 0;  // over the owner's life (fdtd_graph_stats)
    // Replays of the exec may still be queued on its stream: they finish before it is destroyed
    // (and before any buffer its launches read is freed, which callers do after a reset).
    void reset() {
        if (exec) {
            if (st) (void)hipStreamSynchronize(st);
            (void)hipGraphExecDestroy(exec);
        }
        exec = nullptr;
        st = nullptr;
        phase.clear();
    }
    bool matches(const std::vector<int> &ph, const Group &g) const {
        return exec && ph == phase && same_shape(shape[0], g.cfg->shape[0]) &&
               same_shape(shape[1], g.cfg->shape[1]) && flux_mode == g.cfg->flux_mfma &&
               signature == group_signature(g);
    }
    ~StepGraph() { reset(); }
};
}  // namespace

struct fdtd_sim {
    fdtd_config cfg;
    SimDev host;            // host copy of the descriptor (pointers are device pointers)
    SimDev *dev = nullptr;  // device copy
    hipStream_t stream = nullptr;
    std::vector<void *> allocs;
    std::vector<SrcDev> srcs;
    std::vector<PlaneHost> planes;
    PlaneDev *dev_planes = nullptr;
    int max_plane_cells = 0;
    std::vector<uint8_t> hmember[3];  // integer-lattice layer membership: where H is stored
    std::vector<int2> prow;           // per row (j + k*ny): x range of pole work; see mark_rows
    uint64_t bytes = 0;
    bool dirty = true;
    bool in_batch = false;
    bool destroy_pending = false;  // fdtd_sim_destroy was called while in a batch
    StepGraph graph;
    std::vector<hipEvent_t> timed_ev;  // fdtd_run_timed: around each fused pair of a chunk
    // the work split (see SimDev::ibox): per V (1, 2, 4 -> 0, 1, 2) the shell list's length,
    // the interior's run count and the lone simulation's work items
    int nshell[3] = {0, 0, 0}, ninterior[3] = {0, 0, 0};
    const int4 *work[3] = {nullptr, nullptr, nullptr};
    int nwork[3] = {0, 0, 0};
    const int4 *shellw[3] = {nullptr, nullptr, nullptr};  // the shell's items alone
    int nshellw[3] = {0, 0, 0};
    const int4 *ldsw[3] = {nullptr, nullptr, nullptr};    // the box's halo tiles of 2, 4, 8 rows
    int nldsw[3] = {0, 0, 0};
    const int4 *intw[3] = {nullptr, nullptr, nullptr};    // the interior box's items alone
    int nintw[3] = {0, 0, 0};
    LaunchConfig own_cfg;  // the simulation's own launch configuration, when has_cfg
    bool has_cfg = false;
    const LaunchConfig *config() const { return has_cfg ? &own_cfg : &g_defaults; }
    fdtd_sim *partner = nullptr;  // the other part of complex storage (fdtd_link_parts)
    const int4 *foldw = nullptr;  // the ghost planes of the folded axes (fold_kernel items)
    int nfoldw = 0;
    std::vector<int4> fold_host;      // the same items, for a batch's concatenation
    std::vector<void *> work_allocs;  // the device lists above (replaced when the lines change)
    bool work_dirty = true;           // the layer lines changed since the split was built

    int alloc(void **p, size_t bytes_) {
        hipError_t e = hipMalloc(p, bytes_);
        if (e != hipSuccess) return fail(FDTD_ENOMEM, "hipMalloc(%zu) failed: %s", bytes_, hipGetErrorString(e));
        allocs.push_back(*p);
        bytes += bytes_;
        return FDTD_OK;
    }
    int work_alloc(void **p, size_t bytes_) {
        hipError_t e = hipMalloc(p, bytes_);
        if (e != hipSuccess) return fail(FDTD_ENOMEM, "hipMalloc(%zu) failed: %s", bytes_, hipGetErrorString(e));
        work_allocs.push_back(*p);
        work_bytes += bytes_;
        return FDTD_OK;
    }
    uint64_t work_bytes = 0;
    int volume(float **p, float fill = 0.0f) {
        const size_t n = (size_t)host.npad;
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
    int prepare_work();
    // before a launch: the work split is current, then the descriptor is on the device
    int ready() {
        if (work_dirty) {
            int rc = prepare_work();
            if (rc) return rc;
        }
        return sync_desc();
    }
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
    Group group() const {
        Group g;
        g.dev = dev;
        g.nsims = 1;
        for (int vi = 0; vi < 3; ++vi) {
            g.tiles[1 << vi] = work[vi];
            g.ntiles[1 << vi] = nwork[vi];
            g.shell_tiles[1 << vi] = shellw[vi];
            g.nshell_tiles[1 << vi] = nshellw[vi];
            g.int_tiles[1 << vi] = intw[vi];
            g.nint_tiles[1 << vi] = nintw[vi];
            g.lds_tiles[2 << vi] = ldsw[vi];
            g.nlds_tiles[2 << vi] = nldsw[vi];
        }
        g.max_npol = host.npol;
        g.any_pml = host.pml != 0;
        g.any_offdiag = g.all_offdiag = host.offdiag != 0;
        g.general = needs_general(host);
        g.inv_mode = inv_mode_of(host);
        g.fold_items = foldw;
        g.nfold = nfoldw;
        g.cfg = config();
        return g;
    }
    int push_planes() {
        if (planes.empty()) return FDTD_OK;
        std::vector<PlaneDev> devs;
        for (auto &x : planes) devs.push_back(x.dev);
        HIP_TRY(hipMemcpyAsync(dev_planes, devs.data(), sizeof(PlaneDev) * devs.size(),
                               hipMemcpyHostToDevice, stream));
        return FDTD_OK;
    }
};

// The box of a simulation's grid: its whole rows [0, pitch) x [j0, j1) x [k0, k1) where j and
// k are on no layer line of either lattice (per axis the longest range no membership run
// covers), so that only the x line's forms can hold in it. Without an absorber, the whole
// padded grid.
static void interior_box(const SimDev &h, int box[6]) {
    if (!h.pml) {
        const int whole[6] = {0, h.pitch, 0, h.n[1], 0, h.n[2]};
        for (int t = 0; t < 6; ++t) box[t] = whole[t];
        return;
    }
    box[0] = 0;
    box[1] = h.pitch;
    for (int a = 1; a < 3; ++a) {
        const int n = h.n[a];
        std::vector<uint8_t> covered(n, 0);
        for (int half = 0; half < 2; ++half)
            for (int r = 0; r < kMaxRuns; ++r)
                for (int c = std::max(0, h.mr[half][a][2 * r]); c < std::min(n, h.mr[half][a][2 * r + 1]); ++c)
                    covered[c] = 1;
        int best_lo = 0, best_hi = 0;
        for (int c = 0; c < n;) {
            if (covered[c]) { ++c; continue; }
            int e = c;
            while (e < n && !covered[e]) ++e;
            if (e - c > best_hi - best_lo) { best_lo = c; best_hi = e; }
            c = e;
        }
        if (best_hi < best_lo) best_hi = best_lo;
        box[2 * a] = best_lo;
        box[2 * a + 1] = best_hi;
    }
}

// Build the work split: the interior box, the shell lists for V = 1, 2, 4 (every run of V
// cells holding a real cell that is not wholly inside the box, in address order) and the lone
// simulation's work items (kThreads runs each), interior and shell items merged by address.
int fdtd_sim::prepare_work() {
    graph.reset();  // a captured sequence launched the old lists (its replays finish first)
    if (stream) HIP_TRY(hipStreamSynchronize(stream));
    for (void *p : work_allocs) (void)hipFree(p);
    work_allocs.clear();
    work_bytes = 0;
    SimDev &h = host;
    int box[6];
    interior_box(h, box);
    for (int t = 0; t < 6; ++t) h.ibox[t] = box[t];
    const int nx = h.n[0], ny = h.n[1], nz = h.n[2];
    const bool empty = box[0] >= box[1] || box[2] >= box[3] || box[4] >= box[5];
    for (int vi = 0; vi < 3; ++vi) {
        const int V = 1 << vi;
        std::vector<int> list;
        if (h.pml) {
            for (int k = 0; k < nz; ++k)
                for (int j = 0; j < ny; ++j) {
                    const bool row_in = !empty && j >= box[2] && j < box[3] && k >= box[4] && k < box[5];
                    for (int x = 0; x < nx; x += V) {
                        if (row_in && x >= box[0] && x + V <= box[1]) continue;
                        list.push_back((k * ny + j) * h.pitch + x);
                    }
                }
        }
        const int64_t nint = empty ? 0 : (int64_t)((box[1] - box[0]) / V) * (box[3] - box[2]) * (box[5] - box[4]);
        if (nint + (int64_t)list.size() >= ((int64_t)1 << 31))
            return fail(FDTD_EINVAL, "grid too large for the work lists");
        ninterior[vi] = (int)nint;
        nshell[vi] = (int)list.size();
        // interior runs numbered x fastest: the address of run q
        const int nqx = empty ? 1 : (box[1] - box[0]) / V, nj = empty ? 1 : box[3] - box[2];
        auto interior_at = [&](int q) {
            const int r = q / nqx, kk = r / nj;
            return ((box[4] + kk) * ny + box[2] + (r - kk * nj)) * h.pitch + box[0] + (q - r * nqx) * V;
        };
        std::vector<int4> items;
        int qi = 0, qs = 0;
        while (qi < nint || qs < (int)list.size()) {
            const bool take_interior =
                qs >= (int)list.size() || (qi < nint && interior_at(qi) < list[qs]);
            if (take_interior) {
                const int c = (int)std::min<int64_t>(kThreads, nint - qi);
                items.push_back(make_int4(0, 0, qi, c));
                qi += c;
            } else {
                const int c = std::min(kThreads, (int)list.size() - qs);
                items.push_back(make_int4(0, 1, qs, c));
                qs += c;
            }
        }
        void *p = nullptr;
        int rc = work_alloc(&p, sizeof(int) * std::max<size_t>(1, list.size()));
        if (rc) return rc;
        if (!list.empty())
            HIP_TRY(hipMemcpy(p, list.data(), sizeof(int) * list.size(), hipMemcpyHostToDevice));
        h.shell[vi] = (const int *)p;
        rc = work_alloc(&p, sizeof(int4) * std::max<size_t>(1, items.size()));
        if (rc) return rc;
        if (!items.empty())
            HIP_TRY(hipMemcpy(p, items.data(), sizeof(int4) * items.size(), hipMemcpyHostToDevice));
        work[vi] = (const int4 *)p;
        nwork[vi] = (int)items.size();
        // the shell's items alone and the box's alone, for the two-kernel form
        std::vector<int4> shell_items, int_items;
        for (const int4 &w : items) (w.y == 1 ? shell_items : int_items).push_back(w);
        for (int which = 0; which < 2; ++which) {
            const std::vector<int4> &v = which ? int_items : shell_items;
            rc = work_alloc(&p, sizeof(int4) * std::max<size_t>(1, v.size()));
            if (rc) return rc;
            if (!v.empty()) HIP_TRY(hipMemcpy(p, v.data(), sizeof(int4) * v.size(), hipMemcpyHostToDevice));
            if (which) { intw[vi] = (const int4 *)p; nintw[vi] = (int)v.size(); }
            else { shellw[vi] = (const int4 *)p; nshellw[vi] = (int)v.size(); }
        }
    }
    // the box's LDS halo tiles of TY = 2, 4, 8 rows (TX = kThreads / TY cells along x)
    for (int ti = 0; ti < 3; ++ti) {
        const int TY = 2 << ti, TX = kThreads / TY;
        int64_t n = 0;
        if (h.pml && !empty)
            n = (int64_t)((box[1] - box[0] + TX - 1) / TX) * ((box[3] - box[2] + TY - 1) / TY) *
                (box[5] - box[4]);
        if (n >= ((int64_t)1 << 31)) return fail(FDTD_EINVAL, "grid too large for the work lists");
        std::vector<int4> items((size_t)n);
        for (int64_t t = 0; t < n; ++t) items[(size_t)t] = make_int4(0, 2, (int)t, 0);
        void *p = nullptr;
        int rc = work_alloc(&p, sizeof(int4) * std::max<size_t>(1, items.size()));
        if (rc) return rc;
        if (!items.empty())
            HIP_TRY(hipMemcpy(p, items.data(), sizeof(int4) * items.size(), hipMemcpyHostToDevice));
        ldsw[ti] = (const int4 *)p;
        nldsw[ti] = (int)n;
    }
    // the folded axes' ghost planes, in pieces of up to kFoldItem cells
    fold_host.clear();
    for (int a = 0; a < 3; ++a) {
        if (!h.fold[a]) continue;
        const int cells = h.n[(a + 1) % 3] * h.n[(a + 2) % 3];
        for (int far = 0; far < (h.fold[a] == 2 ? 2 : 1); ++far)
            for (int c0 = 0; c0 < cells; c0 += kFoldItem)
                fold_host.push_back(make_int4(0, 2 * a + far, c0, std::min(kFoldItem, cells - c0)));
    }
    foldw = nullptr;
    nfoldw = (int)fold_host.size();
    if (nfoldw) {
        void *p = nullptr;
        int rc = work_alloc(&p, sizeof(int4) * fold_host.size());
        if (rc) return rc;
        HIP_TRY(hipMemcpy(p, fold_host.data(), sizeof(int4) * fold_host.size(), hipMemcpyHostToDevice));
        foldw = (const int4 *)p;
    }
    work_dirty = false;
    dirty = true;
    return FDTD_OK;
}

struct fdtd_batch {
    std::vector<fdtd_sim *> sims;
    Group group;
    std::vector<void *> allocs;
    PlaneDev *planes = nullptr;     // every simulation's planes, pointing at group.dev
    int nplanes = 0;
    int max_plane_cells = 0;
    hipStream_t stream = nullptr;
    StepGraph graph;
    LaunchConfig own_cfg;  // the batch's own launch configuration, when has_cfg
    bool has_cfg = false;
    const LaunchConfig *config() const { return has_cfg ? &own_cfg : &g_defaults; }
};

namespace {
int flux_flush_plane(PlaneDev *dev_plane, PlaneHost &ph, hipStream_t st, int mfma) {
    if (ph.filled == 0) return FDTD_OK;
    const PlaneDev &p = ph.dev;
    flux_phase_kernel<<<dim3((ph.filled * p.nf + 255) / 256, 1), 256, 0, st>>>(dev_plane, ph.filled);
    if (mfma) {
        const int ct = mfma;  // 16-cell tiles per wave: 1, 2 or 4
        const int per_block = 4 * 16 * ct;
        const dim3 grid((p.cells + per_block - 1) / per_block, p.nslot,
                        (2 * p.nf + 16 * kMT - 1) / (16 * kMT));
        if (ct == 4)
            flux_accumulate_mfma_kernel<4><<<grid, 256, 0, st>>>(p.ring, p.phase, p.acc, p.cells, p.row,
                                                                 p.nf, ph.filled, p.nslot, p.hmask);
        else if (ct == 2)
            flux_accumulate_mfma_kernel<2><<<grid, 256, 0, st>>>(p.ring, p.phase, p.acc, p.cells, p.row,
                                                                 p.nf, ph.filled, p.nslot, p.hmask);
        else
            flux_accumulate_mfma_kernel<1><<<grid, 256, 0, st>>>(p.ring, p.phase, p.acc, p.cells, p.row,
                                                                 p.nf, ph.filled, p.nslot, p.hmask);
    } else {
        const int per_block = 256 * kCB;
        flux_accumulate_kernel<<<dim3((p.cells + per_block - 1) / per_block, p.nslot, (p.nf + kFZ - 1) / kFZ),
                                 256, 0, st>>>(p.ring, p.phase, p.acc, p.cells, p.row, p.nf, ph.filled,
                                               p.nslot, p.hmask);
    }
    // the next sample lands in slot 0: move the cursor by what this pass took (a full ring
    // leaves it where it was, so a captured period's passes never write it)
    const int base = (p.base + ph.filled) % kRing;
    if (base != p.base) {
        ph.dev.base = base;
        set_int_kernel<<<1, 64, 0, st>>>(&dev_plane->base, base);
    }
    HIP_TRY(hipGetLastError());
    ph.filled = 0;
    return FDTD_OK;
}

bool samples_now(const PlaneHost &ph, int step_after) {
    const int since = step_after - ph.dev.first_step;
    return since > 0 && since % ph.dev.dec == 0;
}

void mirror_advance(fdtd_sim *s) {
    s->host.step += 1;  // mirror of the device counter the advance launch moved
    s->host.hsep = 1;
    if (s->host.npol) s->host.ppar ^= 1;
}

// The simulations one launch sequence steps (one, or a batch) with their flux planes, in
// descriptor order: simulation q's planes follow simulation q-1's in `planes`.
// HIP events around every fused-pair launch of the steps a Stepper runs: ev[3t..3t+2] around
// step t of a chunk.
struct Timing {
    hipEvent_t *ev = nullptr;
    double h_ms = 0.0, d_ms = 0.0;
};

struct Stepper {
    Group g;
    fdtd_sim *const *sims = nullptr;
    int nsims = 0;
    PlaneDev *planes = nullptr;
    int nplanes = 0;
    int max_plane_cells = 0;
    hipStream_t st = nullptr;
    StepGraph *graph = nullptr;
    Timing *tm = nullptr;
};

// Add the pair durations of `nsteps` recorded steps (waits for the last one).
int add_timing(Timing *tm, int nsteps) {
    HIP_TRY(hipEventSynchronize(tm->ev[3 * (nsteps - 1) + 2]));
    for (int t = 0; t < nsteps; ++t) {
        float x = 0.f, y = 0.f;
        HIP_TRY(hipEventElapsedTime(&x, tm->ev[3 * t], tm->ev[3 * t + 1]));
        HIP_TRY(hipEventElapsedTime(&y, tm->ev[3 * t + 1], tm->ev[3 * t + 2]));
        tm->h_ms += x;
        tm->d_ms += y;
    }
    return FDTD_OK;
}

// Flux bookkeeping after a step: one sampling launch for every plane that samples now, then a
// transform pass for each ring that is full. With launch == false only the host mirrors move
// (the step itself runs inside a replayed graph). With transforms == false no pass is due: the
// caller has left every ring room for the samples ahead (a decimated period's graph, whose
// passes run between its replays).
int flux_after_step(const Stepper &k, bool launch, int adv = 0, bool transforms = true) {
    if (!k.nplanes) return FDTD_OK;
    bool any = false;
    for (int q = 0; q < k.nsims; ++q)
        for (auto &ph : k.sims[q]->planes) any |= samples_now(ph, k.sims[q]->host.step);
    if (!any) return FDTD_OK;
    if (launch) {
        flux_sample_kernel<<<dim3((k.max_plane_cells + 255) / 256, k.nplanes), 256, 0, k.st>>>(
            k.planes, k.nplanes, adv);
        HIP_TRY(hipGetLastError());
    }
    int pi = 0;
    for (int q = 0; q < k.nsims; ++q)
        for (size_t r = 0; r < k.sims[q]->planes.size(); ++r, ++pi) {
            PlaneHost &ph = k.sims[q]->planes[r];
            if (!samples_now(ph, k.sims[q]->host.step)) continue;
            ph.filled += 1;
            if (ph.filled < kRing || !transforms) continue;
            if (!launch) {
                ph.filled = 0;
                continue;
            }
            int rc = flux_flush_plane(k.planes + pi, ph, k.st, k.g.cfg->flux_mfma);
            if (rc) return rc;
        }
    return FDTD_OK;
}

// One whole step. `t` < 0: a lone step (the counters advance after its D pair); else step t of a
// chunk (eager or being captured) whose launches read the counters at offset t and whose
// counters advance once, by the chunk's length, at its end (chunk_end).
int step_once(const Stepper &k, bool launch, int t = -1, bool transforms = true) {
    const int adv = t < 0 ? 0 : t;
    if (launch) {
        hipEvent_t *ev = (k.tm && t >= 0) ? k.tm->ev + 3 * t : nullptr;
        if (ev) HIP_TRY(hipEventRecord(ev[0], k.st));
        int rc = launch_h(k.g, k.st, adv);
        if (rc) return rc;
        if (ev) HIP_TRY(hipEventRecord(ev[1], k.st));
        rc = launch_d(k.g, k.st, adv, t < 0 ? 1 : 0);
        if (rc) return rc;
        if (ev) HIP_TRY(hipEventRecord(ev[2], k.st));
    }
    for (int q = 0; q < k.nsims; ++q) mirror_advance(k.sims[q]);
    return flux_after_step(k, launch, t < 0 ? 0 : t + 1, transforms);
}

// The counter advance that closes a chunk of n steps.
int chunk_end(const Stepper &k, int n) {
    advance_kernel<<<(k.g.nsims + 63) / 64, 64, 0, k.st>>>(k.g.dev, k.g.nsims, n);
    HIP_TRY(hipGetLastError());
    return FDTD_OK;
}

// n <= kRing steps launched eagerly as one chunk (then their pair durations, when timed).
int eager_chunk(const Stepper &k, int n) {
    for (int t = 0; t < n; ++t) {
        int rc = step_once(k, true, t);
        if (rc) return rc;
    }
    int rc = chunk_end(k, n);
    if (rc == FDTD_OK && k.tm) rc = add_timing(k.tm, n);
    return rc;
}

struct Mirror {
    std::vector<SimDev> descs;
    std::vector<int> fills;
};

Mirror save_mirror(const Stepper &k) {
    Mirror m;
    for (int q = 0; q < k.nsims; ++q) {
        m.descs.push_back(k.sims[q]->host);
        for (auto &ph : k.sims[q]->planes) m.fills.push_back(ph.filled);
    }
    return m;
}

void restore_mirror(const Stepper &k, const Mirror &m) {
    size_t f = 0;
    for (int q = 0; q < k.nsims; ++q) {
        k.sims[q]->host = m.descs[q];
        for (auto &ph : k.sims[q]->planes) ph.filled = m.fills[f++];
    }
}

// Capture `steps` steps into k.graph and launch it once (the host mirrors advance as the
// capture walks the steps). transforms: the ring passes that fall due are in the sequence (a
// kRing period of planes that sample every step), or none are (a decimated period). On failure
// nothing has run and the mirrors are put back.
int capture_period(const Stepper &k, const std::vector<int> &phase, int steps = kRing,
                   bool transforms = true) {
    k.graph->reset();
    (void)num_cu();  // device query outside the capture
    const Mirror before = save_mirror(k);
    HIP_TRY(hipStreamBeginCapture(k.st, hipStreamCaptureModeThreadLocal));
    int rc = FDTD_OK;
    for (int t = 0; t < steps && rc == FDTD_OK; ++t) rc = step_once(k, true, t, transforms);
    if (rc == FDTD_OK) rc = chunk_end(k, steps);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(k.st, &graph);
    hipGraphExec_t exec = nullptr;
    if (rc == FDTD_OK && e != hipSuccess)
        rc = fail(FDTD_EHIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    if (rc == FDTD_OK) {
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (e != hipSuccess) rc = fail(FDTD_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    if (graph) (void)hipGraphDestroy(graph);
    if (rc != FDTD_OK) {
        restore_mirror(k, before);
        return rc;
    }
    k.graph->exec = exec;
    k.graph->st = k.st;
    k.graph->phase = phase;
    k.graph->shape[0] = k.g.cfg->shape[0];
    k.graph->shape[1] = k.g.cfg->shape[1];
    k.graph->flux_mode = k.g.cfg->flux_mfma;
    k.graph->signature = group_signature(k.g);
    k.graph->captures += 1;
    HIP_TRY(hipGraphLaunch(exec, k.st));
    return FDTD_OK;
}

// The period of a decimated schedule: the least common multiple of the planes' decimations,
// taken as many times as fits in kRing steps, or 0 when no such period holds (it would exceed
// kMaxPeriod steps, or a plane would take kRing - 1 or more samples in it).
int decimated_period(const Stepper &k) {
    long long l = 1;
    int dmin = 1 << 30;
    for (int q = 0; q < k.nsims; ++q)
        for (auto &ph : k.sims[q]->planes) {
            const long long d = ph.dev.dec;
            l = l / std::gcd(l, d) * d;
            if (l > kMaxPeriod) return 0;
            dmin = std::min(dmin, ph.dev.dec);
        }
    const int p = l <= kRing ? (int)(l * (kRing / l)) : (int)l;
    return p / dmin < kRing - 1 ? p : 0;
}

// Transform now every ring that the next `period` steps would fill (at most kRing - 1 samples
// may stand in a ring when a decimated period's replay starts, so none fills inside it).
int flush_before_period(const Stepper &k, int period) {
    int pi = 0;
    for (int q = 0; q < k.nsims; ++q)
        for (size_t r = 0; r < k.sims[q]->planes.size(); ++r, ++pi) {
            PlaneHost &ph = k.sims[q]->planes[r];
            if (ph.filled + period / ph.dev.dec <= kRing - 1) continue;
            int rc = flux_flush_plane(k.planes + pi, ph, k.st, k.g.cfg->flux_mfma);
            if (rc) return rc;
        }
    return FDTD_OK;
}

// nsteps whole steps: replayed graphs -- kRing-step periods with their transform passes where
// every plane samples each step, or periods of the decimations' least common multiple with the
// passes between replays where some plane decimates -- and eager launches for the rest.
int run_steps(const Stepper &k, int nsteps) {
    int done = 0;
    bool every = true;
    for (int q = 0; q < k.nsims; ++q)
        for (auto &ph : k.sims[q]->planes) every &= ph.dev.dec == 1;
    const int period = every ? kRing : decimated_period(k);
    bool eligible = k.g.cfg->graphs && k.graph && !k.tm && period > 0 && nsteps >= period;
    if (eligible && !every) {
        // the graph holds the pairs and the sampling launches of one period; it replays while
        // each plane's decimation phase is the one it was captured at
        std::vector<int> phase{1, period};
        for (int q = 0; q < k.nsims; ++q)
            for (auto &ph : k.sims[q]->planes)
                phase.push_back(((k.sims[q]->host.step - ph.dev.first_step) % ph.dev.dec +
                                 ph.dev.dec) % ph.dev.dec);
        if (!k.graph->matches(phase, k.g)) {
            int rc = flush_before_period(k, period);
            if (rc == FDTD_OK) rc = capture_period(k, phase, period, false);
            if (rc) return rc;
            done = period;
        }
        for (; done + period <= nsteps; done += period) {
            int rc = flush_before_period(k, period);
            if (rc) return rc;
            for (int t = 0; t < period; ++t) {
                rc = step_once(k, false, -1, false);
                if (rc) return rc;
            }
            HIP_TRY(hipGraphLaunch(k.graph->exec, k.st));
            k.graph->replays += 1;
        }
        eligible = false;
    }
    std::vector<int> phase{0};
    for (int q = 0; q < k.nsims && eligible; ++q)
        for (auto &ph : k.sims[q]->planes) phase.push_back(ph.filled);
    if (eligible) {
        if (!k.graph->matches(phase, k.g)) {
            int rc = capture_period(k, phase);
            if (rc) return rc;
            done = kRing;
        }
        for (; done + kRing <= nsteps; done += kRing) {
            for (int t = 0; t < kRing; ++t) {
                int rc = step_once(k, false);
                if (rc) return rc;
            }
            HIP_TRY(hipGraphLaunch(k.graph->exec, k.st));
            k.graph->replays += 1;
        }
    }
    while (done < nsteps) {
        const int n = nsteps - done < kRing ? nsteps - done : kRing;
        int rc = eager_chunk(k, n);
        if (rc) return rc;
        done += n;
    }
    return FDTD_OK;
}

Stepper lone(fdtd_sim *&s) {
    Stepper k;
    k.g = s->group();
    k.sims = &s;
    k.nsims = 1;
    k.planes = s->dev_planes;
    k.nplanes = (int)s->planes.size();
    k.max_plane_cells = s->max_plane_cells;
    k.st = s->stream;
    k.graph = &s->graph;
    return k;
}

int sim_step(fdtd_sim *s) {
    int rc = s->ready();
    if (rc) return rc;
    return step_once(lone(s), true);
}
}  // namespace

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

int fdtd_device_memory(uint64_t *free_bytes, uint64_t *total_bytes) {
    if (!free_bytes || !total_bytes) return fail(FDTD_EINVAL, "null argument");
    size_t f = 0, t = 0;
    HIP_TRY(hipMemGetInfo(&f, &t));
    *free_bytes = f;
    *total_bytes = t;
    return FDTD_OK;
}

int fdtd_set_launch(int32_t vec, int32_t blocks_per_cu, int32_t xcd_remap, int32_t nt_stores) {
    if (vec != 1 && vec != 2 && vec != 4) return fail(FDTD_EINVAL, "vec must be 1, 2 or 4, got %d", vec);
    if (blocks_per_cu < 0 || blocks_per_cu > 64) return fail(FDTD_EINVAL, "blocks_per_cu %d", blocks_per_cu);
    for (int pair = 0; pair < 2; ++pair) {
        int rc = fdtd_set_launch_pair(pair, vec, blocks_per_cu, xcd_remap, nt_stores);
        if (rc) return rc;
    }
    return FDTD_OK;
}

int fdtd_set_launch_pair(int32_t pair, int32_t vec, int32_t blocks_per_cu, int32_t xcd_remap,
                         int32_t nt_stores) {
    if (pair != 0 && pair != 1) return fail(FDTD_EINVAL, "pair must be 0 (H) or 1 (D), got %d", pair);
    if (vec != 1 && vec != 2 && vec != 4) return fail(FDTD_EINVAL, "vec must be 1, 2 or 4, got %d", vec);
    if (blocks_per_cu < 0 || blocks_per_cu > 64) return fail(FDTD_EINVAL, "blocks_per_cu %d", blocks_per_cu);
    if (xcd_remap < 0 || xcd_remap > 4096) return fail(FDTD_EINVAL, "xcd_remap %d", xcd_remap);
    LaunchShape &sh = g_defaults.shape[pair];
    sh.vec = vec;
    sh.blocks_per_cu = blocks_per_cu;
    sh.xcd_remap = xcd_remap;
    sh.nt = nt_stores ? 1 : 0;
    return FDTD_OK;
}

int fdtd_set_launch_shell(int32_t pair, int32_t vec) {
    if (pair != 0 && pair != 1) return fail(FDTD_EINVAL, "pair must be 0 (H) or 1 (D), got %d", pair);
    if (vec != 1 && vec != 2 && vec != 4) return fail(FDTD_EINVAL, "vec must be 1, 2 or 4, got %d", vec);
    g_defaults.shape[pair].svec = vec;
    return FDTD_OK;
}

int fdtd_get_launch_shell(int32_t pair, int32_t *vec) {
    if ((pair != 0 && pair != 1) || !vec) return fail(FDTD_EINVAL, "bad argument");
    *vec = g_defaults.shape[pair].svec;
    return FDTD_OK;
}

int fdtd_set_launch_tile(int32_t rows) {
    if (rows != 0 && rows != 2 && rows != 4 && rows != 8)
        return fail(FDTD_EINVAL, "halo tile rows must be 0, 2, 4 or 8, got %d", rows);
    g_defaults.shape[1].tile = rows;
    return FDTD_OK;
}

int fdtd_get_launch_tile(int32_t *rows) {
    if (!rows) return fail(FDTD_EINVAL, "null argument");
    *rows = g_defaults.shape[1].tile;
    return FDTD_OK;
}

int fdtd_get_launch_pair(int32_t pair, int32_t out[4]) {
    if (!out || (pair != 0 && pair != 1)) return fail(FDTD_EINVAL, "bad argument");
    out[0] = g_defaults.shape[pair].vec;
    out[1] = g_defaults.shape[pair].blocks_per_cu;
    out[2] = g_defaults.shape[pair].xcd_remap;
    out[3] = g_defaults.shape[pair].nt;
    return FDTD_OK;
}

int fdtd_get_launch(int32_t out[4]) {
    if (!out) return fail(FDTD_EINVAL, "null argument");
    return fdtd_get_launch_pair(0, out);
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
    const int64_t pitch = ((cfg->n[0] + kPad - 1) / kPad) * kPad;
    const int64_t npad = pitch * cfg->n[1] * cfg->n[2];
    if (npad >= ((int64_t)1 << 31) - 4 * kThreads * 4)
        return fail(FDTD_EINVAL, "grid of %lld padded cells exceeds the 2^31 index range", (long long)npad);
    if (cfg->n[2] >= 32768) return fail(FDTD_EINVAL, "at most 32767 cells along z, got %d", cfg->n[2]);
    if (!(cfg->dt > 0.0)) return fail(FDTD_EINVAL, "dt must be positive");
    fdtd_sim *s = new (std::nothrow) fdtd_sim();
    if (!s) return fail(FDTD_ENOMEM, "host allocation");
    s->cfg = *cfg;
    memset(&s->host, 0, sizeof(SimDev));
    SimDev &h = s->host;
    for (int a = 0; a < 3; ++a) {
        h.n[a] = cfg->n[a];
        h.present[a] = cfg->invariant[a] ? 0 : 1;
    }
    h.pitch = (int)pitch;
    h.dtdx = cfg->dtdx;
    h.dt = cfg->dt;
    h.pml = cfg->pml ? 1 : 0;
    h.ncell = (int)ncell;
    h.npad = (int)npad;
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
            const int len = a == 0 ? h.pitch : h.n[a];  // the x line covers the row padding
            TRY(s->alloc(&p, sizeof(float4) * len));
            std::vector<float4> ones(len, make_float4(1.f, 1.f, 1.f, 0.f));
            e = hipMemcpyAsync(p, ones.data(), sizeof(float4) * len, hipMemcpyHostToDevice, s->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
            if (e != hipSuccess) { fdtd_sim_destroy(s); return fail(FDTD_EHIP, "line upload: %s", hipGetErrorString(e)); }
            h.line[half][a] = (const float4 *)p;
        }
#undef TRY
    s->dirty = true;
    *out = s;
    return FDTD_OK;
}

void fdtd_sim_destroy(fdtd_sim *s) {
    if (!s) return;
    if (s->in_batch) {
        // the batch's descriptors and plane copies point into this simulation's memory: its
        // release waits for fdtd_batch_destroy, which hands the counters back and then frees it
        s->destroy_pending = true;
        (void)fail(FDTD_ESTATE, "simulation is stepped by a batch: released when the batch is destroyed");
        return;
    }
    if (s->partner) s->partner->partner = nullptr;  // the other part is stepped alone no more
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    s->graph.reset();  // before the memory its launches read is freed
    for (hipEvent_t e : s->timed_ev) (void)hipEventDestroy(e);
    for (void *p : s->allocs) (void)hipFree(p);
    for (void *p : s->work_allocs) (void)hipFree(p);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

int fdtd_sim_bytes(const fdtd_sim *s, uint64_t *bytes) {
    if (!s || !bytes) return fail(FDTD_EINVAL, "null argument");
    *bytes = s->bytes + s->work_bytes;
    return FDTD_OK;
}

int fdtd_set_layer(fdtd_sim *s, int32_t half, int32_t axis, const float *kms, const float *sinv,
                   const float *kps, const uint8_t *member) {
    if (!s || half < 0 || half > 1 || axis < 0 || axis > 2 || !kms || !sinv || !kps || !member)
        return fail(FDTD_EINVAL, "fdtd_set_layer: bad argument");
    if (s->in_batch) return fail(FDTD_ESTATE, "the simulation is stepped by a batch");
    const int n = s->host.n[axis];
    std::vector<float4> v(n);
    for (int q = 0; q < n; ++q) v[q] = make_float4(kms[q], sinv[q], kps[q], member[q] ? 1.f : 0.f);
    if (half == 0) s->hmember[axis].assign(member, member + n);
    int runs[2 * kMaxRuns];
    for (int t = 0; t < 2 * kMaxRuns; ++t) runs[t] = 0;  // empty runs [0, 0)
    int nr = 0;
    for (int q = 0; q < n;) {
        if (!member[q]) { ++q; continue; }
        int e = q;
        while (e < n && member[e]) ++e;
        if (nr == kMaxRuns) return fail(FDTD_EINVAL, "layer membership of axis %d has more than %d runs", axis, kMaxRuns);
        runs[2 * nr] = q;
        runs[2 * nr + 1] = e;
        ++nr;
        q = e;
    }
    int rc = s->pull_desc();
    if (rc) return rc;
    for (int t = 0; t < 2 * kMaxRuns; ++t) s->host.mr[half][axis][t] = runs[t];
    s->dirty = true;
    s->work_dirty = true;  // the interior box and shell lists follow the lines
    HIP_TRY(hipMemcpyAsync((void *)s->host.line[half][axis], v.data(), sizeof(float4) * n,
                           hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return FDTD_OK;
}

// Host <-> device volume transfers. The device rows are padded to the pitch; the rows are
// re-pitched on the host and moved as ONE contiguous copy (a pitched 2-D copy from pageable
// memory goes row by row and costs ~10 us per row).
// Host transfers go through one pinned staging buffer: the runtime copies pageable memory with
// blit kernels that wait behind a batch occupying the device, which held an admission (uploads
// of a new solve) to the batch's pace; from pinned memory the copy engines run beside it. The
// buffer is two fixed halves of kPinPiece bytes, allocated once per process, and a transfer of
// any size streams through them piece by piece: an upload's memcpy of one piece into one half
// overlaps the copy engine's read of the previous piece from the other. Every transfer has
// completed when the call returns.
static constexpr size_t kPinPiece = (size_t)16 << 20;
static std::mutex g_pin_mu;
static char *g_pin = nullptr;  // 2 * kPinPiece bytes

static int copy_pinned(fdtd_sim *s, void *dev, const void *host, size_t bytes, bool to_device) {
    std::lock_guard<std::mutex> lock(g_pin_mu);
    if (!g_pin) HIP_TRY(hipHostMalloc((void **)&g_pin, 2 * kPinPiece, hipHostMallocDefault));
    char *d = (char *)dev;
    const char *hin = (const char *)host;
    char *hout = (char *)host;
    size_t k = 0;
    for (size_t off = 0; off < bytes; off += kPinPiece, ++k) {
        const size_t piece = std::min(kPinPiece, bytes - off);
        char *half = g_pin + (k & 1) * kPinPiece;
        if (to_device) {
            // this half was last read by piece k - 2, which the previous iteration's wait ended
            std::memcpy(half, hin + off, piece);
            HIP_TRY(hipStreamSynchronize(s->stream));  // piece k - 1 leaves the other half
            HIP_TRY(hipMemcpyAsync(d + off, half, piece, hipMemcpyHostToDevice, s->stream));
        } else {
            HIP_TRY(hipMemcpyAsync(half, d + off, piece, hipMemcpyDeviceToHost, s->stream));
            HIP_TRY(hipStreamSynchronize(s->stream));
            std::memcpy(hout + off, half, piece);
        }
    }
    HIP_TRY(hipStreamSynchronize(s->stream));
    return FDTD_OK;
}

static int upload_volume(fdtd_sim *s, float *dst, const float *src) {
    const SimDev &h = s->host;
    const size_t rows = (size_t)h.n[1] * h.n[2];
    if (h.pitch == h.n[0]) return copy_pinned(s, dst, src, sizeof(float) * rows * h.n[0], true);
    std::vector<float> stage(rows * h.pitch, 0.0f);
    for (size_t r = 0; r < rows; ++r)
        std::memcpy(stage.data() + r * h.pitch, src + r * h.n[0], sizeof(float) * h.n[0]);
    return copy_pinned(s, dst, stage.data(), sizeof(float) * stage.size(), true);
}

static int download_volume(fdtd_sim *s, float *host, const float *src) {
    const SimDev &h = s->host;
    const size_t rows = (size_t)h.n[1] * h.n[2];
    if (h.pitch == h.n[0]) return copy_pinned(s, (void *)src, host, sizeof(float) * rows * h.n[0], false);
    std::vector<float> stage(rows * h.pitch);
    int rc = copy_pinned(s, (void *)src, stage.data(), sizeof(float) * stage.size(), false);
    if (rc) return rc;
    for (size_t r = 0; r < rows; ++r)
        std::memcpy(host + r * h.n[0], stage.data() + r * h.pitch, sizeof(float) * h.n[0]);
    return FDTD_OK;
}

int fdtd_set_boundary(fdtd_sim *s, int32_t axis, int32_t kind) {
    if (!s || axis < 0 || axis > 2 || (kind != 0 && kind != 1))
        return fail(FDTD_EINVAL, "fdtd_set_boundary: bad argument");
    if (s->in_batch) return fail(FDTD_ESTATE, "the simulation is stepped by a batch");
    int rc = s->pull_desc();
    if (rc) return rc;
    SimDev &h = s->host;
    if (kind == 1) {
        if (!h.present[axis]) return fail(FDTD_EINVAL, "axis %d is invariant: it has no wall", axis);
        if (h.fold[axis])
            return fail(FDTD_EINVAL, "axis %d is folded: its metallic boundary is the fold's kind", axis);
    }
    h.metal[axis] = kind;
    s->graph.reset();
    s->dirty = true;
    return FDTD_OK;
}

// Whether a descriptor takes the two-kernel D path (dcurl, then the row kernel): an
// off-diagonal chi1inv row, or a mirror fold (whose ghost planes are filled between the two).
static int split_d(const SimDev &h) {
    for (int r = 0; r < 3; ++r)
        for (int p = 0; p < 3; ++p)
            if (h.u[r][p]) return 1;
    return (h.fold[0] || h.fold[1] || h.fold[2]) ? 1 : 0;
}

int fdtd_set_fold(fdtd_sim *s, int32_t axis, int32_t kind, int32_t phase, int32_t far_row) {
    if (!s || axis < 0 || axis > 2 || kind < 0 || kind > 2)
        return fail(FDTD_EINVAL, "fdtd_set_fold: bad argument");
    if (s->in_batch) return fail(FDTD_ESTATE, "the simulation is stepped by a batch");
    int rc = s->pull_desc();
    if (rc) return rc;
    SimDev &h = s->host;
    if (kind) {
        if (!h.present[axis]) return fail(FDTD_EINVAL, "axis %d is invariant: it has no mirror plane", axis);
        if (phase != 1 && phase != -1)
            return fail(FDTD_EINVAL, "a mirror's phase is +1 or -1, got %d", (int)phase);
        if (h.n[axis] < 4) return fail(FDTD_EINVAL, "a folded axis needs 4 stored cells, axis %d has %d", axis, h.n[axis]);
        if (kind == 2 && (far_row < 3 || far_row >= h.n[axis] - 1))
            return fail(FDTD_EINVAL, "far ghost row %d is not an interior row of axis %d", (int)far_row, axis);
        if (!s->planes.empty())
            return fail(FDTD_EINVAL, "fdtd_set_fold: flux and DFT regions on a folded grid are "
                                     "not supported (fold before adding none)");
    }
    h.fold[axis] = kind;
    h.fphase[axis] = kind ? phase : 0;
    h.frow[axis] = kind == 2 ? far_row : 0;
    if (kind) h.metal[axis] = 0;  // a fold's ghost row is no wall
    h.offdiag = split_d(h);
    s->graph.reset();
    s->work_dirty = true;  // the ghost planes are work items
    s->dirty = true;
    return FDTD_OK;
}

int fdtd_set_conductivity(fdtd_sim *s, int32_t field, int32_t c, const float *sigma) {
    if (!s || (field != 0 && field != 1) || c < 0 || c > 2)
        return fail(FDTD_EINVAL, "fdtd_set_conductivity: bad argument");
    if (s->in_batch) return fail(FDTD_ESTATE, "the simulation is stepped by a batch");
    int prc = s->pull_desc();
    if (prc) return prc;
    SimDev &h = s->host;
    const size_t ncell = (size_t)h.ncell;
    bool any = false;
    for (size_t q = 0; sigma && q < ncell; ++q) {
        if (!std::isfinite(sigma[q])) return fail(FDTD_EINVAL, "a conductivity must be finite");
        any |= sigma[q] != 0.0f;
    }
    s->graph.reset();
    if (!any) {  // lossless: the plain forms, word for word (the volumes stay allocated)
        h.cf[field][c] = h.ci[field][c] = nullptr;
        h.fcond[field][c] = nullptr;
    } else {
        // condfac = 1 - sigma*dt2 in float32, dt2 = float32(dt)*0.5 (step_generic.cpp:88);
        // condinv = 1/(1 + (sigma*dt)*0.5) in double from the float32 word, rounded once
        // (structure.cpp:699) -- fields.set_d_conductivity, reference fields.py:836-864
        const float dt2 = (float)h.dt * 0.5f;
        std::vector<float> cf(ncell), ci(ncell);
        for (size_t q = 0; q < ncell; ++q) {
            const float prod = sigma[q] * dt2;
            cf[q] = 1.0f - prod;
            const double wide = (double)sigma[q] * h.dt;
            ci[q] = (float)(1.0 / (1.0 + wide * 0.5));
        }
        float *pf = const_cast<float *>(h.cf[field][c]), *pi = const_cast<float *>(h.ci[field][c]);
        int rc = FDTD_OK;
        if (!pf) rc = s->volume(&pf);
        if (rc == FDTD_OK && !pi) rc = s->volume(&pi);
        if (rc) return rc;
        rc = upload_volume(s, pf, cf.data());
        if (rc == FDTD_OK) rc = upload_volume(s, pi, ci.data());
        if (rc) return rc;
        h.cf[field][c] = pf;
        h.ci[field][c] = pi;
        if (h.pml && !h.fcond[field][c]) {  // f_cond history (fields.py:754-773)
            float *fc = nullptr;
            rc = s->volume(&fc);
            if (rc) return rc;
            h.fcond[field][c] = fc;
        }
    }
    h.cond = 0;
    for (int f = 0; f < 2; ++f)
        for (int a = 0; a < 3; ++a) h.cond |= h.cf[f][a] != nullptr;
    s->dirty = true;
    return FDTD_OK;
}

int fdtd_subnormal_mode(void) { return FDTD_FLUSH_SUBNORMALS; }

int fdtd_host_flush(int32_t on) {
    const unsigned bits = 0x8040u;  // MXCSR FTZ (bit 15) | DAZ (bit 6)
    const unsigned csr = _mm_getcsr();
    _mm_setcsr(on ? (csr | bits) : (csr & ~bits));
    return FDTD_OK;
}

int fdtd_set_conductive_cells(fdtd_sim *s, int32_t field, int32_t c, const uint8_t *cells) {
    if (!s || (field != 0 && field != 1) || c < 0 || c > 2 || !cells)
        return fail(FDTD_EINVAL, "fdtd_set_conductive_cells: bad argument");
    if (s->in_batch) return fail(FDTD_ESTATE, "the simulation is stepped by a batch");
    int rc = s->pull_desc();
    if (rc) return rc;
    SimDev &h = s->host;
    if (!h.ci[field][c])
        return fail(FDTD_ESTATE, "%c%c has no conductivity to confine", field ? 'B' : 'D', 'x' + c);
    // a condinv word of 0 (no finite sigma gives it) marks a cell stepped by the lossless forms
    std::vector<float> ci((size_t)h.ncell);
    rc = download_volume(s, ci.data(), h.ci[field][c]);
    if (rc) return rc;
    for (size_t q = 0; q < ci.size(); ++q)
        if (!cells[q]) ci[q] = 0.0f;
    rc = upload_volume(s, (float *)h.ci[field][c], ci.data());
    if (rc) return rc;
    s->graph.reset();
    s->dirty = true;
    return FDTD_OK;
}

int fdtd_set_inverse_epsilon(fdtd_sim *s, int32_t c, const float *vol) {
    if (!s || c < 0 || c > 2 || !vol) return fail(FDTD_EINVAL, "fdtd_set_inverse_epsilon: bad argument");
    if (s->in_batch) return fail(FDTD_ESTATE, "the simulation is stepped by a batch");
    int rc = upload_volume(s, (float *)s->host.inv[c], vol);
    if (rc) return rc;
    // the distinct words, by bit pattern (so -0 and NaN payloads stay themselves)
    SimDev &h = s->host;
    // (neighbouring cells mostly share a word: the last one found skips the hash)
    std::unordered_map<uint32_t, uint32_t> index;
    std::vector<float> table;
    const size_t ncell = (size_t)h.ncell;
    uint32_t last = 0;
    bool have_last = false;
    for (size_t q = 0; q < ncell && table.size() <= 65536; ++q) {
        uint32_t w;
        std::memcpy(&w, vol + q, 4);
        if (have_last && w == last) continue;
        if (index.emplace(w, (uint32_t)table.size()).second) table.push_back(vol[q]);
        last = w;
        have_last = true;
    }
    rc = s->pull_desc();
    if (rc) return rc;
    h.inv8[c] = nullptr;
    h.inv16[c] = nullptr;
    h.invtab[c] = nullptr;
    h.invn[c] = 0;
    s->dirty = true;
    s->graph.reset();  // a captured sequence launched the kernels of the old storage
    if (table.size() > 65536) return FDTD_OK;  // the float volume serves
    const bool narrow = table.size() <= 256;
    const size_t rows = (size_t)h.n[1] * h.n[2], esize = narrow ? 1 : 2;
    std::vector<uint8_t> idx((size_t)h.npad * esize, 0);
    uint32_t last_k = 0;
    have_last = false;
    for (size_t r = 0; r < rows; ++r)
        for (int i = 0; i < h.n[0]; ++i) {
            uint32_t w;
            std::memcpy(&w, vol + r * h.n[0] + i, 4);
            if (!have_last || w != last) {
                last = w;
                last_k = index[w];
                have_last = true;
            }
            const uint32_t k = last_k;
            const size_t at = r * h.pitch + i;
            if (narrow) idx[at] = (uint8_t)k;
            else std::memcpy(&idx[at * 2], &k, 2);
        }
    void *dev_idx = nullptr, *dev_tab = nullptr;
    rc = s->alloc(&dev_idx, idx.size());
    if (rc) return rc;
    rc = s->alloc(&dev_tab, sizeof(float) * table.size());
    if (rc) return rc;
    rc = copy_pinned(s, dev_idx, idx.data(), idx.size(), true);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(dev_tab, table.data(), sizeof(float) * table.size(), hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (narrow) h.inv8[c] = (const uint8_t *)dev_idx;
    else h.inv16[c] = (const uint16_t *)dev_idx;
    h.invtab[c] = (const float *)dev_tab;
    h.invn[c] = (int)table.size();
    return FDTD_OK;
}

int fdtd_work_split(fdtd_sim *s, int32_t vec, int32_t out[8]) {
    if (!s || !out || (vec != 1 && vec != 2 && vec != 4)) return fail(FDTD_EINVAL, "fdtd_work_split: bad argument");
    if (s->work_dirty) {
        int rc = s->prepare_work();
        if (rc) return rc;
    }
    const int vi = vec == 4 ? 2 : vec - 1;
    for (int t = 0; t < 6; ++t) out[t] = s->host.ibox[t];
    out[6] = s->ninterior[vi];
    out[7] = s->nshell[vi];
    return FDTD_OK;
}

int fdtd_pole_rows(fdtd_sim *s, int64_t *active, int64_t *rows) {
    if (!s || !active || !rows) return fail(FDTD_EINVAL, "null argument");
    *rows = (int64_t)s->host.n[1] * s->host.n[2];
    int64_t a = 0;
    for (const int2 &v : s->prow) a += v.y > v.x ? v.y - v.x : 0;
    *active = s->host.npol ? a : 0;  // cells, in units of rows
    *rows = *rows * s->host.n[0];
    return FDTD_OK;
}

int fdtd_inverse_epsilon_bytes(fdtd_sim *s, int32_t out[3]) {
    if (!s || !out) return fail(FDTD_EINVAL, "null argument");
    for (int c = 0; c < 3; ++c)
        out[c] = s->host.inv8[c] ? 1 : s->host.inv16[c] ? 2 : 4;
    return FDTD_OK;
}

int fdtd_set_offdiagonal(fdtd_sim *s, int32_t row, int32_t partner, const float *vol) {
    if (!s || row < 0 || row > 2 || partner < 0 || partner > 2 || row == partner || !vol)
        return fail(FDTD_EINVAL, "fdtd_set_offdiagonal: bad argument");
    if (s->in_batch) return fail(FDTD_ESTATE, "the simulation is stepped by a batch");
    s->graph.reset();  // the off-diagonal path launches other kernels
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

// Per row (j + k*ny), the x range [lo, hi) where pole work must run: it covers every nonzero
// word (bitwise, so -0 counts) of `vol` (host layout, x fastest, unpadded), or the whole row
// when `all`. Ranges only grow. Outside them every pole's P stays +0 -- exact only while the
// pole constants A, gamma1 and 1/gamma1 are non-negative, which the caller checks.
static int mark_rows(fdtd_sim *s, const float *vol, bool all) {
    SimDev &h = s->host;
    const size_t rows = (size_t)h.n[1] * h.n[2];
    if (s->prow.empty()) s->prow.assign(rows, make_int2(0, 0));
    for (size_t r = 0; r < rows; ++r) {
        int lo = h.n[0], hi = 0;
        if (all) {
            lo = 0;
            hi = h.n[0];
        } else if (vol) {
            for (int i = 0; i < h.n[0]; ++i) {
                uint32_t w;
                std::memcpy(&w, vol + r * h.n[0] + i, 4);
                if (w) {
                    lo = std::min(lo, i);
                    hi = i + 1;
                }
            }
        }
        if (lo >= hi) continue;
        int2 &cur = s->prow[r];
        if (cur.x >= cur.y) cur = make_int2(lo, hi);
        else cur = make_int2(std::min(cur.x, lo), std::max(cur.y, hi));
    }
    if (!h.prow) {
        void *p = nullptr;
        int rc = s->alloc(&p, rows * sizeof(int2));
        if (rc) return rc;
        h.prow = (const int2 *)p;
        s->dirty = true;
    }
    HIP_TRY(hipMemcpyAsync((void *)h.prow, s->prow.data(), rows * sizeof(int2), hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return FDTD_OK;
}

static bool nonnegative(float v) {
    uint32_t w;
    std::memcpy(&w, &v, 4);
    return !(w >> 31) && std::isfinite(v);
}

int fdtd_add_pole(fdtd_sim *s, const float consts[4], const float *const sigma[3],
                  const float sigma_scalar[3], const int32_t driven[3], int32_t *pole) {
    if (!s || !consts || !sigma || !sigma_scalar || !driven) return fail(FDTD_EINVAL, "fdtd_add_pole: null argument");
    if (s->in_batch) return fail(FDTD_ESTATE, "the simulation is stepped by a batch");
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
    s->graph.reset();  // a captured sequence steps the poles (and instantiation) it had
    const bool exact_skip = nonnegative(consts[0]) && nonnegative(consts[1]) && nonnegative(consts[3]);
    for (int c = 0; c < 3; ++c) {
        if (!driven[c]) continue;
        // a scalar sigma drives every row; a volume the rows where it is nonzero
        const bool all = !exact_skip || !sigma[c];
        rc = mark_rows(s, sigma[c], all);  // the volume is not read when `all`
        if (rc) return rc;
    }
    if (pole) *pole = q;
    return FDTD_OK;
}

int fdtd_add_source(fdtd_sim *s, int32_t field, int32_t comp, const int32_t lo[3], const int32_t cnt[3],
                    const double *amp_re, const double *amp_im, const double *current, int32_t nsteps,
                    double dt) {
    if (!s || field < 0 || field > 3 || comp < 0 || comp > 2 || !lo || !cnt || !amp_re || !amp_im ||
        !current || nsteps < 0)
        return fail(FDTD_EINVAL, "fdtd_add_source: bad argument");
    if (field == 3)
        return fail(FDTD_EINVAL, "fdtd_add_source: a magnetic polarization (integrated) source is "
                                 "not supported: H is kept apart from B only on the layer rows, "
                                 "and such a source makes H differ from B at its points");
    if (s->in_batch) return fail(FDTD_ESTATE, "the simulation is stepped by a batch");
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
    s->host.npsrc += field == 2;
    for (int f = 0; f < 3; ++f)
        for (int a = 0; a < 3; ++a) {
            s->host.src_lo[f][a] = 1 << 30;
            s->host.src_hi[f][a] = -1;
        }
    for (const SrcDev &x : s->srcs)
        for (int a = 0; a < 3; ++a) {
            s->host.src_lo[x.field][a] = std::min(s->host.src_lo[x.field][a], x.lo[a]);
            s->host.src_hi[x.field][a] = std::max(s->host.src_hi[x.field][a], x.lo[a] + x.cnt[a]);
        }
    s->dirty = true;
    return FDTD_OK;
}

// cfg[0..4] the H pair's shape (vec, blocks_per_cu, xcd_remap, nt, svec), [5..9] the D pair's,
// [10] graphs, [11] flux transform; checked as the process-wide setters check them.
static int parse_config(const int32_t *cfg, LaunchConfig *out) {
    for (int pair = 0; pair < 2; ++pair) {
        const int32_t *c = cfg + kShapeWords * pair;
        if (c[0] != 1 && c[0] != 2 && c[0] != 4) return fail(FDTD_EINVAL, "vec must be 1, 2 or 4, got %d", c[0]);
        if (c[1] < 0 || c[1] > 64) return fail(FDTD_EINVAL, "blocks_per_cu %d", c[1]);
        if (c[2] < 0 || c[2] > 4096) return fail(FDTD_EINVAL, "xcd_remap %d", c[2]);
        if (c[4] != 1 && c[4] != 2 && c[4] != 4) return fail(FDTD_EINVAL, "shell vec must be 1, 2 or 4, got %d", c[4]);
        LaunchShape &sh = out->shape[pair];
        sh.vec = c[0];
        sh.blocks_per_cu = c[1];
        sh.xcd_remap = c[2];
        sh.nt = c[3] ? 1 : 0;
        sh.svec = c[4];
        if (c[5] != 0 && c[5] != 2 && c[5] != 4 && c[5] != 8)
            return fail(FDTD_EINVAL, "halo tile rows must be 0, 2, 4 or 8, got %d", c[5]);
        if (c[5] && (pair != 1 || c[0] != 1))
            return fail(FDTD_EINVAL, "halo tiles serve the D pair at one cell a thread");
        sh.tile = c[5];
    }
    const int32_t *tail = cfg + 2 * kShapeWords;
    if (tail[1] != 0 && tail[1] != 1 && tail[1] != 2 && tail[1] != 4)
        return fail(FDTD_EINVAL, "flux transform must be 0, 1, 2 or 4, got %d", tail[1]);
    out->graphs = tail[0] ? 1 : 0;
    out->flux_mfma = tail[1];
    return FDTD_OK;
}

static void write_config(const LaunchConfig &c, int32_t out[kConfigWords]) {
    for (int pair = 0; pair < 2; ++pair) {
        const LaunchShape &sh = c.shape[pair];
        int32_t *o = out + kShapeWords * pair;
        o[0] = sh.vec;
        o[1] = sh.blocks_per_cu;
        o[2] = sh.xcd_remap;
        o[3] = sh.nt;
        o[4] = sh.svec;
        o[5] = sh.tile;
    }
    out[2 * kShapeWords] = c.graphs;
    out[2 * kShapeWords + 1] = c.flux_mfma;
}

int fdtd_sim_set_config(fdtd_sim *s, const int32_t *cfg) {
    if (!s) return fail(FDTD_EINVAL, "null simulation");
    if (s->in_batch) return fail(FDTD_ESTATE, "the simulation is stepped by a batch");
    LaunchConfig c;
    if (cfg) {
        int rc = parse_config(cfg, &c);
        if (rc) return rc;
    }
    s->graph.reset();  // a captured sequence launched with the old shapes
    s->own_cfg = c;
    s->has_cfg = cfg != nullptr;
    return FDTD_OK;
}

int fdtd_sim_get_config(fdtd_sim *s, int32_t out[kConfigWords]) {
    if (!s || !out) return fail(FDTD_EINVAL, "null argument");
    write_config(*s->config(), out);
    return FDTD_OK;
}

int fdtd_batch_set_config(fdtd_batch *b, const int32_t *cfg) {
    if (!b) return fail(FDTD_EINVAL, "null batch");
    LaunchConfig c;
    if (cfg) {
        int rc = parse_config(cfg, &c);
        if (rc) return rc;
    }
    b->graph.reset();
    b->own_cfg = c;
    b->has_cfg = cfg != nullptr;
    return FDTD_OK;
}

int fdtd_batch_get_config(fdtd_batch *b, int32_t out[kConfigWords]) {
    if (!b || !out) return fail(FDTD_EINVAL, "null argument");
    write_config(*b->config(), out);
    return FDTD_OK;
}

int fdtd_link_parts(fdtd_sim *re, fdtd_sim *im, const int32_t kinds[3], const float phase[6]) {
    if (!re || !im || re == im || !kinds || !phase) return fail(FDTD_EINVAL, "fdtd_link_parts: bad argument");
    if (re->in_batch || im->in_batch) return fail(FDTD_ESTATE, "a simulation is stepped by a batch");
    if (re->partner || im->partner) return fail(FDTD_ESTATE, "a simulation is already linked");
    for (int a = 0; a < 3; ++a) {
        if (kinds[a] < 0 || kinds[a] > 3) return fail(FDTD_EINVAL, "connection kind %d on axis %d", kinds[a], a);
        if (re->host.n[a] != im->host.n[a]) return fail(FDTD_EINVAL, "the parts' grids differ");
        if (kinds[a] && (re->host.fold[a] || re->host.metal[a] || !re->host.present[a]))
            return fail(FDTD_EINVAL, "axis %d carries a Bloch connection but does not wrap", a);
    }
    if (!re->planes.empty() || !im->planes.empty())
        return fail(FDTD_EINVAL, "fdtd_link_parts: flux and DFT regions on complex storage are not "
                                 "supported");
    fdtd_sim *parts[2] = {re, im};
    for (int p = 0; p < 2; ++p) {
        int rc = parts[p]->pull_desc();
        if (rc) return rc;
        SimDev &h = parts[p]->host;
        h.part = p;
        h.other = nullptr;
        for (int a = 0; a < 3; ++a) {
            h.pkind[a] = kinds[a];
            h.ph[a][0] = phase[2 * a];
            h.ph[a][1] = phase[2 * a + 1];
        }
        parts[p]->graph.reset();
        parts[p]->dirty = true;
    }
    re->partner = im;
    im->partner = re;
    return FDTD_OK;
}

int fdtd_graph_stats(fdtd_sim *s, int64_t out[2]) {
    if (!s || !out) return fail(FDTD_EINVAL, "null argument");
    out[0] = s->graph.captures;
    out[1] = s->graph.replays;
    return FDTD_OK;
}

int fdtd_batch_graph_stats(fdtd_batch *b, int64_t out[2]) {
    if (!b || !out) return fail(FDTD_EINVAL, "null argument");
    out[0] = b->graph.captures;
    out[1] = b->graph.replays;
    return FDTD_OK;
}

int fdtd_set_source_currents(fdtd_sim *s, int32_t index, const double *current, int32_t nsteps) {
    if (!s || !current || nsteps < 0) return fail(FDTD_EINVAL, "fdtd_set_source_currents: bad argument");
    if (s->in_batch) return fail(FDTD_ESTATE, "the simulation is stepped by a batch");
    if (index < 0 || index >= (int)s->srcs.size())
        return fail(FDTD_EINVAL, "source %d of %d", (int)index, (int)s->srcs.size());
    int rc = s->pull_desc();
    if (rc) return rc;
    // a new table (the old one stays allocated until the simulation is destroyed: a replay
    // queued before this call may still read it), then the descriptors that point at it
    void *p;
    rc = s->alloc(&p, sizeof(double) * 2 * (size_t)(nsteps > 0 ? nsteps : 1));
    if (rc) return rc;
    if (nsteps) HIP_TRY(hipMemcpyAsync(p, current, sizeof(double) * 2 * nsteps, hipMemcpyHostToDevice, s->stream));
    s->srcs[index].cur = (const double *)p;
    s->srcs[index].ncur = nsteps;
    rc = s->alloc(&p, sizeof(SrcDev) * s->srcs.size());
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(p, s->srcs.data(), sizeof(SrcDev) * s->srcs.size(), hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->host.src = (const SrcDev *)p;
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
    if (L == 9 && !strncmp(name, "f_cond_", 7)) {
        const int c = comp(name[8]);
        const int f = name[7] == 'D' ? 0 : name[7] == 'B' ? 1 : -1;
        if (c < 0 || f < 0) return nullptr;
        return h.fcond[f][c];
    }
    int k = -1;
    char e = 0, cc = 0;
    if (!strncmp(name, "P[", 2) && sscanf(name, "P[%d].%c%c", &k, &e, &cc) == 3) {
        const int c = comp(cc);
        if (k < 0 || k >= h.npol || e != 'E' || c < 0 || !h.pdriven[k][c]) return nullptr;
        return h.pbuf[k][c][h.ppar];
    }
    if (!strncmp(name, "P_prev[", 7) && sscanf(name, "P_prev[%d].%c%c", &k, &e, &cc) == 3) {
        const int c = comp(cc);
        if (k < 0 || k >= h.npol || e != 'E' || c < 0 || !h.pdriven[k][c]) return nullptr;
        return h.pbuf[k][c][h.ppar ^ 1];
    }
    return nullptr;
}

int fdtd_upload(fdtd_sim *s, const char *name, const float *host, size_t count) {
    if (!s || !name || !host) return fail(FDTD_EINVAL, "fdtd_upload: null argument");
    if (s->in_batch) return fail(FDTD_ESTATE, "the simulation is stepped by a batch");
    int rc = s->pull_desc();
    if (rc) return rc;
    if (count != (size_t)s->host.ncell) return fail(FDTD_EINVAL, "upload %s: %zu floats, grid has %d", name, count, s->host.ncell);
    float *dst = volume_by_name(s, name);
    if (!dst) return fail(FDTD_EINVAL, "no volume named '%s' in this simulation", name);
    if (name[0] == 'P') {  // a stored polarization: its nonzero rows need the pole work
        rc = mark_rows(s, host, false);
        if (rc) return rc;
    }
    return upload_volume(s, dst, host);
}

int fdtd_download(fdtd_sim *s, const char *name, float *host, size_t count) {
    if (!s || !name || !host) return fail(FDTD_EINVAL, "fdtd_download: null argument");
    if (s->in_batch) return fail(FDTD_ESTATE, "the simulation is stepped by a batch");
    int rc = s->pull_desc();
    if (rc) return rc;
    if (count != (size_t)s->host.ncell) return fail(FDTD_EINVAL, "download %s: %zu floats, grid has %d", name, count, s->host.ncell);
    float *src = volume_by_name(s, name);
    if (!src) return fail(FDTD_EINVAL, "no volume named '%s' in this simulation", name);
    const SimDev &h = s->host;
    rc = download_volume(s, host, src);
    if (rc) return rc;
    if (name[0] == 'H' && name[2] == 0 && h.pml && h.hsep) {
        // after the first update_H the stored H is current only on its own-axis layer rows;
        // everywhere else H is B's word
        const int c = name[1] - 'x';
        std::vector<float> b((size_t)h.ncell);
        rc = download_volume(s, b.data(), h.B[c]);
        if (rc) return rc;
        const std::vector<uint8_t> &line = s->hmember[c];
        for (int k = 0; k < h.n[2]; ++k)
            for (int j = 0; j < h.n[1]; ++j)
                for (int i = 0; i < h.n[0]; ++i) {
                    const int cc = c == 0 ? i : (c == 1 ? j : k);
                    const bool member = !line.empty() && line[cc];
                    const size_t idx = ((size_t)k * h.n[1] + j) * h.n[0] + i;
                    if (!member) host[idx] = b[idx];
                }
    }
    return FDTD_OK;
}

int fdtd_get_step(fdtd_sim *s, int32_t *step) {
    if (!s || !step) return fail(FDTD_EINVAL, "null argument");
    if (s->in_batch) {
        *step = s->host.step;  // the batch mirrors every step it runs
        return FDTD_OK;
    }
    int rc = s->pull_desc();
    if (rc) return rc;
    *step = s->host.step;
    return FDTD_OK;
}

int fdtd_set_step(fdtd_sim *s, int32_t step, int32_t h_separated) {
    if (!s || step < 0) return fail(FDTD_EINVAL, "bad argument");
    if (s->in_batch) return fail(FDTD_ESTATE, "the simulation is stepped by a batch");
    if (!s->planes.empty()) return fail(FDTD_ESTATE, "the step count is fixed once a flux plane samples");
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
    prime_e_kernel<<<(s->host.npad + 255) / 256, 256, 0, s->stream>>>(s->dev, s->host.npad);
    HIP_TRY(hipGetLastError());
    return FDTD_OK;
}

int fdtd_step_h(fdtd_sim *s) {
    if (s && s->partner)
        return fail(FDTD_ESTATE, "a part of complex storage steps only beside its other part, in a batch");
    if (!s) return fail(FDTD_EINVAL, "null argument");
    if (s->in_batch) return fail(FDTD_ESTATE, "the simulation is stepped by a batch");
    int rc = s->ready();
    if (rc) return rc;
    return launch_h(s->group(), s->stream);
}

int fdtd_step_d(fdtd_sim *s) {
    if (s && s->partner)
        return fail(FDTD_ESTATE, "a part of complex storage steps only beside its other part, in a batch");
    if (!s) return fail(FDTD_EINVAL, "null argument");
    if (s->in_batch) return fail(FDTD_ESTATE, "the simulation is stepped by a batch");
    int rc = s->ready();
    if (rc) return rc;
    rc = launch_d(s->group(), s->stream);
    if (rc) return rc;
    mirror_advance(s);
    return flux_after_step(lone(s), true);
}

// The transform pass over every plane's buffered samples (its ring then starts empty).
static int flush_all(fdtd_sim *s) {
    for (size_t r = 0; r < s->planes.size(); ++r) {
        int rc = flux_flush_plane(s->dev_planes + r, s->planes[r], s->stream, s->config()->flux_mfma);
        if (rc) return rc;
    }
    return FDTD_OK;
}

int fdtd_flush(fdtd_sim *s) {
    if (!s) return fail(FDTD_EINVAL, "null argument");
    if (s->in_batch) return fail(FDTD_ESTATE, "the simulation is stepped by a batch");
    return flush_all(s);
}

int fdtd_run(fdtd_sim *s, int32_t nsteps) {
    if (s && s->partner)
        return fail(FDTD_ESTATE, "a part of complex storage steps only beside its other part, in a batch");
    if (!s || nsteps < 0) return fail(FDTD_EINVAL, "bad argument");
    if (s->in_batch) return fail(FDTD_ESTATE, "the simulation is stepped by a batch");
    int rc = s->ready();
    if (rc) return rc;
    return run_steps(lone(s), nsteps);
}

int fdtd_set_flux_transform(int32_t mfma) {
    if (mfma != 0 && mfma != 1 && mfma != 2 && mfma != 4)
        return fail(FDTD_EINVAL, "flux transform: 0 (vector), or 1, 2, 4 (matrix cores), got %d", mfma);
    g_defaults.flux_mfma = mfma;
    return FDTD_OK;
}

int fdtd_set_graphs(int32_t enable) {
    g_defaults.graphs = enable ? 1 : 0;
    return FDTD_OK;
}

int fdtd_synchronize(fdtd_sim *s) {
    if (!s) return fail(FDTD_EINVAL, "null argument");
    HIP_TRY(hipStreamSynchronize(s->stream));
    return FDTD_OK;
}

// A DFT region's device state: the ring, the sample times, the phase table and the float64
// accumulators, registered after the simulation's other regions.
static int add_region(fdtd_sim *s, const int32_t lo[3], const int32_t cnt[3], int nslot,
                      const int32_t *comps, int yee, int32_t nf, const double *omega,
                      double scale, int32_t dec, int32_t *plane) {
    s->graph.reset();  // a captured sequence samples the regions it was captured with
    int cells = 1;
    for (int a = 0; a < 3; ++a) {
        if (cnt[a] < 1 || lo[a] < 0 || lo[a] >= s->host.n[a] || cnt[a] > s->host.n[a])
            return fail(FDTD_EINVAL, "region box axis %d [%d,+%d) outside the grid", a, lo[a], cnt[a]);
        if ((int64_t)cells * cnt[a] > (1 << 30)) return fail(FDTD_EINVAL, "region too large");
        cells *= cnt[a];
    }
    int rc = s->pull_desc();
    if (rc) return rc;
    PlaneHost ph;
    PlaneDev &p = ph.dev;
    p.sim = s->dev;
    p.nslot = nslot;
    p.hmask = 0;
    for (int r = 0; r < 4; ++r) p.comp[r] = r < nslot ? comps[r] : 0;
    for (int r = 0; r < nslot; ++r) p.hmask |= (comps[r] >= 3 ? 1 : 0) << r;
    p.yee = yee ? 1 : 0;
    for (int a = 0; a < 3; ++a) { p.lo[a] = lo[a]; p.cnt[a] = cnt[a]; }
    p.cells = cells;
    p.row = ((cells + 15) / 16) * 16;  // whole 16-cell tiles for the matrix-core transform
    p.nf = nf;
    p.dec = dec;
    p.first_step = s->host.step;
    p.base = 0;
    p.scale = scale;
    void *m;
    rc = s->alloc(&m, sizeof(double) * nf);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(m, omega, sizeof(double) * nf, hipMemcpyHostToDevice, s->stream));
    p.omega = (const double *)m;
    rc = s->alloc(&m, sizeof(float) * (size_t)kRing * nslot * p.row);
    if (rc) return rc;
    HIP_TRY(hipMemsetAsync(m, 0, sizeof(float) * (size_t)kRing * nslot * p.row, s->stream));
    p.ring = (float *)m;
    rc = s->alloc(&m, sizeof(double) * kRing);
    if (rc) return rc;
    p.times = (double *)m;
    rc = s->alloc(&m, sizeof(double) * (size_t)kRing * 4 * nf);
    if (rc) return rc;
    p.phase = (double *)m;
    rc = s->alloc(&m, sizeof(double) * 2 * (size_t)nslot * nf * cells);
    if (rc) return rc;
    HIP_TRY(hipMemsetAsync(m, 0, sizeof(double) * 2 * (size_t)nslot * nf * cells, s->stream));
    p.acc = (double *)m;
    s->planes.push_back(ph);
    rc = s->alloc(&m, sizeof(PlaneDev) * s->planes.size());
    if (rc) return rc;
    s->dev_planes = (PlaneDev *)m;
    if (cells > s->max_plane_cells) s->max_plane_cells = cells;
    rc = s->push_planes();
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (plane) *plane = (int)s->planes.size() - 1;
    return FDTD_OK;
}

int fdtd_add_flux_plane(fdtd_sim *s, int32_t normal, const int32_t lo[3], const int32_t cnt[3],
                        int32_t nf, const double *omega, double scale, int32_t dec, int32_t *plane) {
    if (!s || normal < 0 || normal > 2 || !lo || !cnt || nf < 1 || !omega || dec < 1)
        return fail(FDTD_EINVAL, "fdtd_add_flux_plane: bad argument");
    if (cnt[normal] != 1) return fail(FDTD_EINVAL, "a flux plane is one cell thick along its normal");
    if (s->in_batch) return fail(FDTD_ESTATE, "the simulation is stepped by a batch");
    const int t1 = (normal + 1) % 3, t2 = (normal + 2) % 3;
    const int32_t comps[4] = {t1, t2, 3 + t1, 3 + t2};
    return add_region(s, lo, cnt, 4, comps, 0, nf, omega, scale, dec, plane);
}

int fdtd_add_dft_region(fdtd_sim *s, const int32_t lo[3], const int32_t cnt[3], int32_t ncomp,
                        const int32_t *comps, int32_t yee_grid, int32_t nf, const double *omega,
                        double scale, int32_t dec, int32_t *region) {
    if (!s || !lo || !cnt || ncomp < 1 || ncomp > 4 || !comps || nf < 1 || !omega || dec < 1)
        return fail(FDTD_EINVAL, "fdtd_add_dft_region: bad argument");
    for (int r = 0; r < ncomp; ++r)
        if (comps[r] < 0 || comps[r] > 5) return fail(FDTD_EINVAL, "component code %d (0..5: Ex..Hz)", comps[r]);
    if (s->in_batch) return fail(FDTD_ESTATE, "the simulation is stepped by a batch");
    return add_region(s, lo, cnt, ncomp, comps, yee_grid, nf, omega, scale, dec, region);
}

int fdtd_flux_read(fdtd_sim *s, int32_t plane, double *out, size_t count) {
    if (!s || plane < 0 || plane >= (int)s->planes.size() || !out) return fail(FDTD_EINVAL, "bad argument");
    if (s->in_batch) return fail(FDTD_ESTATE, "the simulation is stepped by a batch");
    PlaneHost &ph = s->planes[plane];
    const size_t need = (size_t)2 * ph.dev.nslot * ph.dev.nf * ph.dev.cells;
    if (count != need) return fail(FDTD_EINVAL, "flux read needs %zu doubles, got %zu", need, count);
    int rc = flux_flush_plane(s->dev_planes + plane, ph, s->stream, s->config()->flux_mfma);
    if (rc) return rc;
    return copy_pinned(s, ph.dev.acc, out, sizeof(double) * need, false);
}

// Bring a plane's accumulators up to date (the ring's pending samples transformed in) before
// they are replaced or scaled: the pending samples belong to the run before the edit.
static int flux_plane_for_edit(fdtd_sim *s, int32_t plane, size_t count, bool check_count) {
    if (!s || plane < 0 || plane >= (int)s->planes.size()) return fail(FDTD_EINVAL, "bad argument");
    if (s->in_batch) return fail(FDTD_ESTATE, "the simulation is stepped by a batch");
    PlaneHost &ph = s->planes[plane];
    const size_t need = (size_t)2 * ph.dev.nslot * ph.dev.nf * ph.dev.cells;
    if (check_count && count != need)
        return fail(FDTD_EINVAL, "flux write needs %zu doubles, got %zu", need, count);
    return flux_flush_plane(s->dev_planes + plane, ph, s->stream, s->config()->flux_mfma);
}

int fdtd_flux_write(fdtd_sim *s, int32_t plane, const double *in, size_t count) {
    if (!in) return fail(FDTD_EINVAL, "null argument");
    int rc = flux_plane_for_edit(s, plane, count, true);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(s->planes[plane].dev.acc, in, sizeof(double) * count,
                           hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return FDTD_OK;
}

int fdtd_flux_scale(fdtd_sim *s, int32_t plane, double re, double im) {
    int rc = flux_plane_for_edit(s, plane, 0, false);
    if (rc) return rc;
    const PlaneDev &p = s->planes[plane].dev;
    const size_t n = (size_t)p.nslot * p.nf * p.cells;
    flux_scale_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s->stream>>>(p.acc, n, re, im);
    HIP_TRY(hipGetLastError());
    return FDTD_OK;
}

// A flux plane (the four tangential components, cell-centred), as opposed to a DFT region.
static bool is_flux_plane(const PlaneDev &p) {
    return p.nslot == 4 && !p.yee && p.comp[0] < 3 && p.comp[1] == (p.comp[0] + 1) % 3 &&
           p.comp[2] == p.comp[0] + 3 && p.comp[3] == p.comp[1] + 3;
}

// The two reduction launches of a plane's spectrum into out_dev[nf] (see the kernels).
static int launch_spectrum(const PlaneDev &p, const double *w_dev, double *partial,
                           double *out_dev, hipStream_t st) {
    const int G = std::max(1, std::min(kSpecBlocks, (p.cells + 255) / 256));
    flux_spectrum_partial_kernel<<<dim3(G, p.nf), 256, 0, st>>>(p.acc, w_dev, p.cells, p.nf,
                                                                 partial);
    flux_spectrum_final_kernel<<<(p.nf + 63) / 64, 64, 0, st>>>(partial, G, p.nf, out_dev);
    HIP_TRY(hipGetLastError());
    return FDTD_OK;
}

int fdtd_flux_spectrum(fdtd_sim *s, int32_t plane, const double *weights, double *out,
                       size_t count) {
    if (!s || plane < 0 || plane >= (int)s->planes.size() || !out)
        return fail(FDTD_EINVAL, "fdtd_flux_spectrum: bad argument");
    if (s->in_batch) return fail(FDTD_ESTATE, "the simulation is stepped by a batch");
    PlaneHost &ph = s->planes[plane];
    const PlaneDev &p = ph.dev;
    if (!is_flux_plane(p)) return fail(FDTD_EINVAL, "region %d is not a flux plane", plane);
    if (count != (size_t)p.nf) return fail(FDTD_EINVAL, "the plane has %d frequencies, got %zu", p.nf, count);
    int rc = flux_flush_plane(s->dev_planes + plane, ph, s->stream, s->config()->flux_mfma);
    if (rc) return rc;
    if (!ph.spec_partial) {
        rc = s->alloc((void **)&ph.spec_partial, sizeof(double) * (size_t)p.nf * kSpecBlocks);
        if (rc == FDTD_OK) rc = s->alloc((void **)&ph.spec_out, sizeof(double) * (size_t)p.nf);
        if (rc) return rc;
    }
    if (weights && !ph.spec_w) {
        rc = s->alloc((void **)&ph.spec_w, sizeof(double) * (size_t)p.cells);
        if (rc) return rc;
    }
    if (weights)
        HIP_TRY(hipMemcpyAsync(ph.spec_w, weights, sizeof(double) * (size_t)p.cells,
                               hipMemcpyHostToDevice, s->stream));
    rc = launch_spectrum(p, weights ? ph.spec_w : nullptr, ph.spec_partial, ph.spec_out, s->stream);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(out, ph.spec_out, sizeof(double) * (size_t)p.nf, hipMemcpyDeviceToHost,
                           s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return FDTD_OK;
}

// The device address of `scratch` on the current GPU, looked up once per device with that
// device current (host code never passes the array's own name, which is a host-side stand-in).
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[64] = {};

static int scratch_address(double **out) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return fail(FDTD_EINVAL, "device ordinal %d", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(FDTD_EHIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        g_scratch_ptr[dev] = (double *)p;
    }
    *out = g_scratch_ptr[dev];
    return FDTD_OK;
}

int fdtd_flux_spectrum_noalloc(fdtd_sim *s, int32_t plane, const double *weights, double *out,
                               size_t count) {
    if (!s || plane < 0 || plane >= (int)s->planes.size() || !out)
        return fail(FDTD_EINVAL, "fdtd_flux_spectrum_noalloc: bad argument");
    if (s->in_batch) return fail(FDTD_ESTATE, "the simulation is stepped by a batch");
    PlaneHost &ph = s->planes[plane];
    const PlaneDev &p = ph.dev;
    if (!is_flux_plane(p)) return fail(FDTD_EINVAL, "region %d is not a flux plane", plane);
    if (count != (size_t)p.nf) return fail(FDTD_EINVAL, "the plane has %d frequencies, got %zu", p.nf, count);
    if (p.nf > kScratchFreqs)
        return fail(FDTD_EINVAL, "fdtd_flux_spectrum_noalloc takes at most %d frequencies, the plane has %d",
                    kScratchFreqs, p.nf);
    double *partial = nullptr;
    int rc = scratch_address(&partial);
    if (rc) return rc;
    rc = flux_flush_plane(s->dev_planes + plane, ph, s->stream, s->config()->flux_mfma);
    if (rc) return rc;
    // stage 1 writes every partial[f][b] stage 2 reads (b < the launch's G): nothing to clear
    return launch_spectrum(p, weights, partial, out, s->stream);
}

int fdtd_time_pairs(fdtd_sim *s, int32_t repeats, float out_ms[2]) {
    if (s && s->partner)
        return fail(FDTD_ESTATE, "a part of complex storage steps only beside its other part, in a batch");
    if (!s || repeats < 1 || !out_ms) return fail(FDTD_EINVAL, "bad argument");
    if (s->in_batch) return fail(FDTD_ESTATE, "the simulation is stepped by a batch");
    if (!s->planes.empty()) return fail(FDTD_ESTATE, "fdtd_time_pairs steps a simulation without flux planes");
    int rc = s->ready();
    if (rc) return rc;
    const Group g = s->group();
    hipEvent_t a, b, c;
    HIP_TRY(hipEventCreate(&a));
    HIP_TRY(hipEventCreate(&b));
    HIP_TRY(hipEventCreate(&c));
    float th = 0.f, td = 0.f;
    for (int r = 0; r < repeats; ++r) {
        HIP_TRY(hipEventRecord(a, s->stream));
        rc = launch_h(g, s->stream);
        if (rc) return rc;
        HIP_TRY(hipEventRecord(b, s->stream));
        rc = launch_d(g, s->stream);
        if (rc) return rc;
        HIP_TRY(hipEventRecord(c, s->stream));
        HIP_TRY(hipEventSynchronize(c));
        float x = 0.f, y = 0.f;
        HIP_TRY(hipEventElapsedTime(&x, a, b));
        HIP_TRY(hipEventElapsedTime(&y, b, c));
        th += x;
        td += y;
        mirror_advance(s);
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    (void)hipEventDestroy(c);
    out_ms[0] = th / repeats;
    out_ms[1] = td / repeats;
    return FDTD_OK;
}

int fdtd_run_timed(fdtd_sim *s, int32_t nsteps, float out_ms[3]) {
    if (s && s->partner)
        return fail(FDTD_ESTATE, "a part of complex storage steps only beside its other part, in a batch");
    if (!s || nsteps < 0 || !out_ms) return fail(FDTD_EINVAL, "bad argument");
    if (s->in_batch) return fail(FDTD_ESTATE, "the simulation is stepped by a batch");
    int rc = s->ready();
    if (rc) return rc;
    if (s->timed_ev.empty()) {
        s->timed_ev.resize(3 * kRing + 2);
        for (auto &e : s->timed_ev) HIP_TRY(hipEventCreate(&e));
    }
    // eager chunks of kRing steps (launched as fdtd_run launches the steps it does not replay)
    // with events around every fused pair, read after each chunk
    Timing tm;
    tm.ev = s->timed_ev.data();
    fdtd_sim *self = s;
    Stepper k = lone(self);
    k.tm = &tm;
    hipEvent_t t0 = s->timed_ev[3 * kRing], t1 = s->timed_ev[3 * kRing + 1];
    HIP_TRY(hipEventRecord(t0, s->stream));
    rc = run_steps(k, nsteps);
    if (rc) return rc;
    // every sample taken in the window is transformed in it: the passes still due run here
    rc = flush_all(s);
    if (rc) return rc;
    HIP_TRY(hipEventRecord(t1, s->stream));
    HIP_TRY(hipEventSynchronize(t1));
    float total = 0.f;
    HIP_TRY(hipEventElapsedTime(&total, t0, t1));
    out_ms[0] = total;
    out_ms[1] = nsteps ? (float)(tm.h_ms / nsteps) : 0.f;
    out_ms[2] = nsteps ? (float)(tm.d_ms / nsteps) : 0.f;
    return FDTD_OK;
}

int fdtd_copy_probe(uint64_t bytes, int32_t repeats, int32_t mode, float *ms) {
    if (bytes < 1024 || repeats < 1 || !ms || mode < 0 || mode > 8) return fail(FDTD_EINVAL, "bad argument");
    // the one-pass wide copies (modes 5-8) take whole 256*8-float4 groups
    const size_t n = mode >= 5 ? (size_t)(bytes / (16 * 256 * 8)) * 256 * 8 : (size_t)(bytes / 16);
    if (n == 0) return fail(FDTD_EINVAL, "copy probe: %llu bytes is too small", (unsigned long long)bytes);
    float4 *a = nullptr, *b = nullptr;
    HIP_TRY(hipMalloc((void **)&a, n * 16));
    if (hipMalloc((void **)&b, n * 16) != hipSuccess) { (void)hipFree(a); return fail(FDTD_ENOMEM, "copy probe"); }
    HIP_TRY(hipMemset(a, 0, n * 16));
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    const int grid = num_cu() * 8;
    auto go = [&]() {
        if (mode == 0) copy_probe_kernel<<<grid, 256>>>(a, b, n);
        else if (mode == 1) copy_probe4_kernel<false><<<grid, 256>>>(a, b, n);
        else if (mode == 2) copy_probe4_kernel<true><<<grid, 256>>>(a, b, n);
        else if (mode == 3) read_probe_kernel<<<grid, 256>>>(a, (float *)b, n);
        else if (mode == 4) write_probe_kernel<<<grid, 256>>>(b, n);
        else if (mode == 5) copy_wide_kernel<8, true, true><<<(unsigned)(n / 2048), 256>>>((const f4v *)a, (f4v *)b);
        else if (mode == 6) copy_wide_kernel<8, false, false><<<(unsigned)(n / 2048), 256>>>((const f4v *)a, (f4v *)b);
        else if (mode == 7) copy_wide_kernel<4, true, true><<<(unsigned)(n / 1024), 256>>>((const f4v *)a, (f4v *)b);
        else copy_wide_kernel<4, false, true><<<(unsigned)(n / 1024), 256>>>((const f4v *)a, (f4v *)b);
    };
    go();
    HIP_TRY(hipEventRecord(e0, 0));
    for (int r = 0; r < repeats; ++r) go();
    HIP_TRY(hipEventRecord(e1, 0));
    HIP_TRY(hipEventSynchronize(e1));
    float t = 0.f;
    HIP_TRY(hipEventElapsedTime(&t, e0, e1));
    *ms = t / repeats;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    return FDTD_OK;
}

int fdtd_stream_probe(int32_t nread, int32_t nwrite, int32_t n, uint64_t stagger,
                      int32_t repeats, float *ms) {
    if (nread < 0 || nwrite < 1 || nread + nwrite > 16 || n < 256 || repeats < 1 || !ms ||
        stagger > (1u << 20))
        return fail(FDTD_EINVAL, "fdtd_stream_probe: bad argument");
    const int q = nread + nwrite;
    const size_t span = (((size_t)n * 4 + (2u << 20) - 1) / (2u << 20)) * (2u << 20);
    const size_t total = q * span + (size_t)q * stagger + 4096;
    char *arena = nullptr;
    HIP_TRY(hipMalloc((void **)&arena, total));
    HIP_TRY(hipMemset(arena, 0, total));
    StreamSet set;
    for (int r = 0; r < 16; ++r) set.p[r] = nullptr;
    for (int r = 0; r < q; ++r) set.p[r] = (float *)(arena + r * span + (size_t)r * stagger);
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    const int grid = (n + 255) / 256;
    stream_probe_kernel<<<grid, 256>>>(set, nread, nwrite, n);
    HIP_TRY(hipEventRecord(e0, 0));
    for (int r = 0; r < repeats; ++r) stream_probe_kernel<<<grid, 256>>>(set, nread, nwrite, n);
    HIP_TRY(hipEventRecord(e1, 0));
    HIP_TRY(hipEventSynchronize(e1));
    float t = 0.f;
    HIP_TRY(hipEventElapsedTime(&t, e0, e1));
    *ms = t / repeats;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(arena);
    return FDTD_OK;
}

// ------------------------------------------------------------------------------------------
// Batch
// ------------------------------------------------------------------------------------------
// The batch's work items for vector width V: each member's own items (interior and shell,
// merged by address), re-numbered to the member's descriptor.
// One of every member's device lists for width index vi -- which 0: all items, 1: the
// shell's, 2: the box's -- concatenated.
static int batch_list(fdtd_batch *b, int vi, int which, const int4 **out, int *nout) {
    std::vector<int4> items;
    for (int q = 0; q < (int)b->sims.size(); ++q) {
        fdtd_sim *s = b->sims[q];
        const int4 *list = which == 0 ? s->work[vi] : which == 1 ? s->shellw[vi]
                           : which == 2 ? s->intw[vi] : s->ldsw[vi];
        const int n = which == 0 ? s->nwork[vi] : which == 1 ? s->nshellw[vi]
                      : which == 2 ? s->nintw[vi] : s->nldsw[vi];
        std::vector<int4> own(std::max(1, n));
        if (n) HIP_TRY(hipMemcpy(own.data(), list, sizeof(int4) * n, hipMemcpyDeviceToHost));
        for (int t = 0; t < n; ++t) {
            int4 w = own[t];
            w.x = q;
            items.push_back(w);
        }
    }
    int4 *d = nullptr;
    HIP_TRY(hipMalloc((void **)&d, sizeof(int4) * std::max<size_t>(1, items.size())));
    b->allocs.push_back(d);
    if (!items.empty())
        HIP_TRY(hipMemcpy(d, items.data(), sizeof(int4) * items.size(), hipMemcpyHostToDevice));
    *out = d;
    *nout = (int)items.size();
    return FDTD_OK;
}

// The batch's work items for vector width V: each member's own lists (all items, the shell's,
// the box's), re-numbered to the member's descriptor.
static int batch_tiles(fdtd_batch *b, int V) {
    if (b->group.tiles.count(V)) return FDTD_OK;
    const int vi = V == 4 ? 2 : V - 1;
    Group &g = b->group;
    const int4 *p = nullptr;
    int n = 0;
    int rc = batch_list(b, vi, 0, &p, &n);
    if (rc) return rc;
    g.tiles[V] = p;
    g.ntiles[V] = n;
    rc = batch_list(b, vi, 1, &p, &n);
    if (rc) return rc;
    g.shell_tiles[V] = p;
    g.nshell_tiles[V] = n;
    rc = batch_list(b, vi, 2, &p, &n);
    if (rc) return rc;
    g.int_tiles[V] = p;
    g.nint_tiles[V] = n;
    rc = batch_list(b, vi, 3, &p, &n);  // (the halo tiles of 2 << vi rows)
    if (rc) return rc;
    g.lds_tiles[2 << vi] = p;
    g.nlds_tiles[2 << vi] = n;
    return FDTD_OK;
}

int fdtd_batch_create(fdtd_sim *const *sims, int32_t count, fdtd_batch **out) {
    if (!sims || count < 1 || !out) return fail(FDTD_EINVAL, "fdtd_batch_create: bad argument");
    for (int q = 0; q < count; ++q) {
        if (!sims[q]) return fail(FDTD_EINVAL, "null simulation %d", q);
        if (sims[q]->in_batch) return fail(FDTD_ESTATE, "simulation %d is already in a batch", q);
        for (int r = 0; r < q; ++r)
            if (sims[r] == sims[q]) return fail(FDTD_EINVAL, "simulation %d appears twice", q);
    }
    fdtd_batch *b = new (std::nothrow) fdtd_batch();
    if (!b) return fail(FDTD_ENOMEM, "host allocation");
    std::vector<SimDev> descs;
    for (int q = 0; q < count; ++q) {
        fdtd_sim *s = sims[q];
        int rc = s->ready();  // the member's work split is what the batch's lists copy
        if (rc == FDTD_OK) rc = s->pull_desc();
        if (rc) { delete b; return rc; }
        HIP_TRY(hipStreamSynchronize(s->stream));
        b->sims.push_back(s);
        descs.push_back(s->host);
        b->group.any_offdiag |= s->host.offdiag != 0;
        b->group.general |= needs_general(s->host);
        b->group.max_npol = std::max(b->group.max_npol, s->host.npol);
        b->group.any_pml |= s->host.pml != 0;
    }
    b->group.all_offdiag = true;
    for (auto *s : b->sims) b->group.all_offdiag &= s->host.offdiag != 0;
    b->group.inv_mode = inv_mode_of(b->sims[0]->host);
    for (auto *s : b->sims)
        if (inv_mode_of(s->host) != b->group.inv_mode) b->group.inv_mode = 3;
    // the parts of complex storage read each other's words: point each at its partner's copy
    for (int q = 0; q < count; ++q) {
        descs[q].other = nullptr;
        fdtd_sim *mate = sims[q]->partner;
        if (!mate) continue;
        int r = 0;
        while (r < count && sims[r] != mate) ++r;
        if (r == count) {
            delete b;
            return fail(FDTD_EINVAL, "simulation %d is a part of complex storage whose other part "
                                     "is not in the batch", q);
        }
        descs[q].other = (const SimDev *)(uintptr_t)r;  // an index until the copy exists
    }
    b->group.nsims = count;
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { delete b; return fail(FDTD_EHIP, "hipStreamCreate: %s", hipGetErrorString(e)); }
    auto cleanup = [&](int code, const char *what) {
        fdtd_batch_destroy(b);
        return fail(code, "%s", what);
    };
    if (hipMalloc((void **)&b->group.dev, sizeof(SimDev) * count) != hipSuccess) return cleanup(FDTD_ENOMEM, "batch descriptors");
    b->allocs.push_back(b->group.dev);
    for (int q = 0; q < count; ++q)
        if (sims[q]->partner) descs[q].other = b->group.dev + (uintptr_t)descs[q].other;
    if (hipMemcpy(b->group.dev, descs.data(), sizeof(SimDev) * count, hipMemcpyHostToDevice) != hipSuccess)
        return cleanup(FDTD_EHIP, "batch descriptor upload");
    // every simulation's planes, re-pointed at the batch's descriptors
    std::vector<PlaneDev> planes;
    for (int q = 0; q < count; ++q)
        for (size_t r = 0; r < b->sims[q]->planes.size(); ++r) {
            PlaneDev p = b->sims[q]->planes[r].dev;
            p.sim = b->group.dev + q;
            planes.push_back(p);
            if (p.cells > b->max_plane_cells) b->max_plane_cells = p.cells;
        }
    b->nplanes = (int)planes.size();
    if (!planes.empty()) {
        if (hipMalloc((void **)&b->planes, sizeof(PlaneDev) * planes.size()) != hipSuccess) return cleanup(FDTD_ENOMEM, "batch planes");
        b->allocs.push_back(b->planes);
        if (hipMemcpy(b->planes, planes.data(), sizeof(PlaneDev) * planes.size(), hipMemcpyHostToDevice) != hipSuccess)
            return cleanup(FDTD_EHIP, "batch plane upload");
    }
    int rc = FDTD_OK;
    for (const LaunchShape &sh : b->config()->shape) {  // the widths the launches take
        if (rc == FDTD_OK) rc = batch_tiles(b, sh.vec);
        if (rc == FDTD_OK) rc = batch_tiles(b, sh.svec);
        if (rc == FDTD_OK && sh.tile) rc = batch_tiles(b, sh.tile / 2);  // its halo tiles
    }
    if (rc) { fdtd_batch_destroy(b); return rc; }
    // the folded members' ghost planes, re-numbered to their descriptors
    std::vector<int4> folds;
    for (int q = 0; q < count; ++q)
        for (int4 w : b->sims[q]->fold_host) {
            w.x = q;
            folds.push_back(w);
        }
    if (!folds.empty()) {
        int4 *d = nullptr;
        if (hipMalloc((void **)&d, sizeof(int4) * folds.size()) != hipSuccess) return cleanup(FDTD_ENOMEM, "batch fold items");
        b->allocs.push_back(d);
        if (hipMemcpy(d, folds.data(), sizeof(int4) * folds.size(), hipMemcpyHostToDevice) != hipSuccess)
            return cleanup(FDTD_EHIP, "batch fold item upload");
        b->group.fold_items = d;
        b->group.nfold = (int)folds.size();
    }
    for (auto *s : b->sims) s->in_batch = true;
    *out = b;
    return FDTD_OK;
}

void fdtd_batch_destroy(fdtd_batch *b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    b->graph.reset();  // before the tile lists and plane copies its launches read are freed
    // hand the advanced counters back to each simulation's own descriptor, and the planes'
    // ring cursors (moved by transform passes over the batch's copies) to its own plane copies
    std::vector<fdtd_sim *> release;
    if (b->group.dev && !b->sims.empty()) {
        std::vector<SimDev> descs(b->sims.size());
        if (hipMemcpy(descs.data(), b->group.dev, sizeof(SimDev) * descs.size(), hipMemcpyDeviceToHost) == hipSuccess)
            for (size_t q = 0; q < descs.size(); ++q) {
                b->sims[q]->host.step = descs[q].step;
                b->sims[q]->host.hsep = descs[q].hsep;
                b->sims[q]->host.ppar = descs[q].ppar;
                b->sims[q]->dirty = true;
            }
        for (auto *s : b->sims) {
            s->in_batch = false;
            if (s->destroy_pending) {
                release.push_back(s);
                continue;
            }
            if (s->push_planes() == FDTD_OK) (void)hipStreamSynchronize(s->stream);
        }
    }
    for (void *p : b->allocs) (void)hipFree(p);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    for (auto *s : release) fdtd_sim_destroy(s);
}

// The batch's stepper, with the tile lists of the widths its configuration launches.
static int batch_stepper(fdtd_batch *b, Stepper *k) {
    int rc = FDTD_OK;
    for (const LaunchShape &sh : b->config()->shape) {  // the widths the launches take
        if (rc == FDTD_OK) rc = batch_tiles(b, sh.vec);
        if (rc == FDTD_OK) rc = batch_tiles(b, sh.svec);
        if (rc == FDTD_OK && sh.tile) rc = batch_tiles(b, sh.tile / 2);  // its halo tiles
    }
    if (rc) return rc;
    k->g = b->group;
    k->g.cfg = b->config();
    k->sims = b->sims.data();
    k->nsims = (int)b->sims.size();
    k->planes = b->planes;
    k->nplanes = b->nplanes;
    k->max_plane_cells = b->max_plane_cells;
    k->st = b->stream;
    k->graph = &b->graph;
    return FDTD_OK;
}

int fdtd_batch_run(fdtd_batch *b, int32_t nsteps) {
    if (!b || nsteps < 0) return fail(FDTD_EINVAL, "bad argument");
    Stepper k;
    int rc = batch_stepper(b, &k);
    if (rc) return rc;
    return run_steps(k, nsteps);
}

int fdtd_batch_step_h(fdtd_batch *b) {
    if (!b) return fail(FDTD_EINVAL, "null argument");
    Stepper k;
    int rc = batch_stepper(b, &k);
    if (rc) return rc;
    return launch_h(k.g, k.st);
}

int fdtd_batch_step_d(fdtd_batch *b) {
    if (!b) return fail(FDTD_EINVAL, "null argument");
    Stepper k;
    int rc = batch_stepper(b, &k);
    if (rc) return rc;
    rc = launch_d(k.g, k.st);
    if (rc) return rc;
    for (int q = 0; q < k.nsims; ++q) mirror_advance(k.sims[q]);
    return flux_after_step(k, true);
}

int fdtd_batch_flush(fdtd_batch *b) {
    if (!b) return fail(FDTD_EINVAL, "null argument");
    int pi = 0;
    for (auto *s : b->sims)
        for (size_t r = 0; r < s->planes.size(); ++r, ++pi) {
            int rc = flux_flush_plane(b->planes + pi, s->planes[r], b->stream, b->config()->flux_mfma);
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
