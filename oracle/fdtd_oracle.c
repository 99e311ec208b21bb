/*
 * CPU restatement of the reference array path's per-step arithmetic (the ORACLE).
 *
 * Test infrastructure only: the parity tests, smoke() and bench.py's cpu_baseline leg call
 * it as the checker / the timed CPU baseline. The product never links or loads it.
 *
 * Layout is the reference's: every volume is float32 C-order (nx, ny, nz), z fastest, and
 * every axis wraps periodically (the fixture cases and the benchmark grids are periodic;
 * an absorber is a material, not a boundary, reference stepping.py:2984-3040).
 *
 * Each function restates one reference routine, operation for operation (one IEEE
 * rounding per operation, no contraction -- build with -ffp-contract=off):
 *
 *   oracle_step_curl  <- stepping.step_B :333 / step_D :463, _curl_operands :2199,
 *                        _curl_from_operands :2265 (association ((g1s-g1)+g2)-g2s and the
 *                        absent-partner forms), _apply_curl_regions :2705 with
 *                        _apply_pml_update :2629 and _apply_dsig_only_update :2662.
 *                        B: forward differences, half-integer coefficients;
 *                        D: MEEP's negated strides (backward differences), integer ones.
 *                        A component with a conductivity takes MEEP's conductive forms:
 *                        _apply_conductive_update :2820-2841 and _apply_conductive_pml_update
 *                        :2844-2983 (condfac/condinv from fields.set_d_conductivity
 *                        fields.py:775-870; the f_cond stage on the dsig layer's cells).
 *   oracle_update_h   <- stepping.update_H :1216, separate_h_from_b :2789,
 *                        _apply_constitutive_regions :2749, _apply_constitutive_pml :2991.
 *   oracle_update_e   <- stepping.update_E :1243, fields.displacement_minus_polarization
 *                        fields.py:1173 (D - P0 - P1 ...), _offdiagonal_terms :1594 and
 *                        _rounded_row :1690 (row summed in double, rounded once).
 *   oracle_update_p   <- dispersion.PolarizationState.update dispersion.py:689-742 with
 *                        fields.drive_field fields.py:1234 (w = f_w on the layer's cells).
 *   oracle_deposit    <- sources._deposit sources.py:1357 + host_writes.subtract_double
 *                        host_writes.py:140 (double product, one rounding to float32).
 *   oracle_dft_plane  <- dft.FluxMonitor.update dft.py:3031 with _sliced_component :442 and
 *                        _apply_yee_interpolation :352 (centre average, then phase*sample).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>
#include <xmmintrin.h>

#define IDX(i, j, k) (((int64_t)(i) * ny + (j)) * nz + (k))

typedef struct {
    int n[3];              /* nx, ny, nz */
    int present[3];        /* 1 when the axis is resolved (not translationally invariant) */
    float dtdx;
    int pml;               /* 0: no absorber anywhere (plain forms, no auxiliaries) */
    /* UPML 1-D vectors per [half][axis]: half 0 = integer lattice, 1 = half-integer */
    const float *kms[2][3];
    const float *sinv[2][3];
    const float *kps[2][3];
    const uint8_t *mem[2][3];  /* one-rank chunk membership lines */
    int metal[3];          /* 1: a metallic (PEC) axis -- stepping.py:140-146, :2429-2530 */
    /* a mirror-folded axis (grid.py:404-575, stepping.py:2004-2160, :2423-2626): 1 over
     * metallic boundaries (no stored far wall, a zero ghost past the top), 2 over periodic
     * ones (one stored slot past the window, the far ghost) */
    int fold[3];
    int fphase[3];         /* a fold's declared phase, +1 or -1 */
    /* Bloch connection per axis (boundaries.cpp:359-366, step.cpp:171-221): 0 none (a plain
     * periodic wrap), 1 a copy (phase within 1e-13 of +1), 2 a negation (-1), 3 a product by
     * the phase (ph: its two parts rounded to float32); on a phased axis the engine stores
     * MEEP's owned window, and the read of a neighbour's SAME stored index across the window
     * edge (plane 0) is the connected word (stepping._bloch_ghost_operand :2545-2586) */
    int pkind[3];
    float ph[3][2];
    /* complex storage as two real parts stepped alike: which part this call steps (0 the real,
     * 1 the imaginary); a product connection reads the other part's word */
    int part;
} ogrid;

/* MEEP's connection of one word across a phased axis's window edge: `mine` this part's word,
 * `other` the other part's; conj for the word BELOW the window (conj(eikna)), else above.
 * A product: libc++'s (a*c) - (b*d), (a*d) + (b*c), each operation rounded on its own
 * (part_products.connect_parts). */
static inline float conn(const ogrid *g, int axis, int conj, float mine, float other) {
    switch (g->pkind[axis]) {
        case 2: return -mine;
        case 3: {
            const float a = g->ph[axis][0], b = conj ? -g->ph[axis][1] : g->ph[axis][1];
            if (g->part == 0) {
                const float x = a * mine, y = b * other;
                return x - y;
            }
            const float x = a * mine, y = b * other;
            return x + y;
        }
        default: return mine;
    }
}

static inline int64_t nb(const ogrid *g, int64_t i, int64_t j, int64_t k, int axis, int step) {
    int64_t c[3] = {i, j, k};
    int n = g->n[axis];
    c[axis] = (c[axis] + step + n) % n;
    return ((c[0] * g->n[1]) + c[1]) * g->n[2] + c[2];
}

/* A curl operand one cell along `axis` (step = +1 forward, -1 backward): the wrapped neighbour,
 * or, past the wall of a metallic axis, zero -- no field beyond a perfect conductor
 * (_shift_up / _shift_down, stepping.py:2429-2530). */
static inline float rd(const ogrid *g, const float *G, int64_t i, int64_t j, int64_t k, int axis,
                       int step) {
    const int64_t c[3] = {i, j, k};
    const int64_t m = c[axis] + step;
    if (g->metal[axis] && (m < 0 || m >= g->n[axis])) return 0.0f;
    /* past the top of a folded metallic axis: the unstored wall plane, zero (_shift_up). The
     * other reads across a folded axis's faces feed a curl that the ghost mask drops. */
    if (g->fold[axis] == 1 && m >= g->n[axis]) return 0.0f;
    return G[nb(g, i, j, k, axis, step)];
}

/* One curl sub-step. family 0 = B (src = E), 1 = D (src = H). cf/ci/fc: per component, the
 * conductivity's condfac, condinv and (under an absorber) f_cond volumes, or NULL (lossless). */
void oracle_step_curl(const ogrid *g, int family, float *const f[3], float *const src[3],
                      float *const fu[3], const float *const cf[3], const float *const ci[3],
                      float *const fc[3], float *const src_other[3]) {
    const int nx = g->n[0], ny = g->n[1], nz = g->n[2];
    const int step = family == 0 ? +1 : -1;
    const int half = family == 0 ? 1 : 0;
    const float dtdx = g->dtdx;
    for (int c = 0; c < 3; ++c) {
        const int first = (c + 2) % 3, first_axis = (c + 1) % 3;
        const int second = (c + 1) % 3, second_axis = (c + 2) % 3;
        const int dsig = (c + 1) % 3, dsigu = (c + 2) % 3;
        const int p1 = g->present[first_axis], p2 = g->present[second_axis];
        float *restrict F = f[c];
        const float *G1 = src[first], *G2 = src[second];
        const float *O1 = src_other ? src_other[first] : 0, *O2 = src_other ? src_other[second] : 0;
        /* B's forward differences read the unshifted partner half a cell BELOW the target
         * (conj(eikna) at plane 0), D's backward ones half a cell ABOVE (eikna) */
        const int below = family == 0;
        float *U = fu ? fu[c] : 0;
        const float *CF = cf ? cf[c] : 0, *CI = ci ? ci[c] : 0;
        float *FC = fc ? fc[c] : 0;
#pragma omp parallel for schedule(static)
        for (int i = 0; i < nx; ++i) {
            for (int j = 0; j < ny; ++j) {
                for (int k = 0; k < nz; ++k) {
                    const int64_t at = IDX(i, j, k);
                    const int ic[3] = {i, j, k};
                    float curl;
                    float g1 = G1[at], g2 = G2[at];
                    if (g->pkind[first_axis] && ic[first_axis] == 0)
                        g1 = conn(g, first_axis, below, g1, O1 ? O1[at] : 0.0f);
                    if (g->pkind[second_axis] && ic[second_axis] == 0)
                        g2 = conn(g, second_axis, below, g2, O2 ? O2[at] : 0.0f);
                    if (p1 && p2) {
                        float g1s = rd(g, G1, i, j, k, first_axis, step);
                        float g2s = rd(g, G2, i, j, k, second_axis, step);
                        float t = g1s - g1;
                        t = t + g2;
                        t = t - g2s;
                        curl = dtdx * t;
                    } else if (p1) {
                        float t = rd(g, G1, i, j, k, first_axis, step) - g1;
                        curl = dtdx * t;
                    } else if (p2) {
                        float t = rd(g, G2, i, j, k, second_axis, step) - g2;
                        curl = (-dtdx) * t;
                    } else {
                        curl = 0.0f;
                    }
                    /* a component sitting ON a metallic wall (Yee shift 0 on that axis: B_c
                     * along c, D_c along the other two) at stored cell 0 is not stepped: its
                     * curl is dropped (_mask_non_owned_cells, stepping.py:2589-2626) */
                    for (int a = 0; a < 3; ++a) {
                        const int shift0 = family == 0 ? (a == c) : (a != c);
                        if (g->metal[a] && shift0 && ic[a] == 0) curl = 0.0f;
                        /* the ghosts of a folded axis are not owned: cell 0 of a shift-0
                         * component and, over periodic boundaries, the last slot of a shift-1
                         * one (_mask_non_owned_cells :2589-2626); the fills write them */
                        if (g->fold[a] && shift0 && ic[a] == 0) curl = 0.0f;
                        if (g->fold[a] == 2 && !shift0 && ic[a] == g->n[a] - 1) curl = 0.0f;
                    }
                    /* a condinv word of 0 (no finite sigma gives it) marks a cell whose chunk
                     * keeps no conductivity array: it steps the lossless forms below
                     * (structure.cpp:883-900) */
                    if (CF && CI[at] != 0.0f) {
                        /* X = curl, cf, ci: MEEP's eight step_curl forms reduce, per cell,
                         * to the four below (step_generic.cpp:84-210 as the reference
                         * restates them); the increment fcnd' - fcnd is formed first */
                        const float cfv = CF[at], civ = CI[at];
                        const int in_s = g->pml && g->mem[half][dsig][ic[dsig]];
                        const int in_u = g->pml && g->mem[half][dsigu][ic[dsigu]];
                        const float kms = g->pml ? g->kms[half][dsig][ic[dsig]] : 1.0f;
                        const float sinv = g->pml ? g->sinv[half][dsig][ic[dsig]] : 1.0f;
                        const float kms_u = g->pml ? g->kms[half][dsigu][ic[dsigu]] : 1.0f;
                        const float sinv_u = g->pml ? g->sinv[half][dsigu][ic[dsigu]] : 1.0f;
                        if (in_s) {
                            const float old = FC[at];
                            float n = cfv * old;
                            n = n - curl;
                            n = n * civ;
                            FC[at] = n;
                            const float inc = n - old;
                            if (in_u) {
                                const float prev = U[at];
                                float u = kms * prev;
                                u = u + inc;
                                u = u * sinv;
                                U[at] = u;
                                float v = kms_u * F[at];
                                v = v + u;
                                v = v - prev;
                                F[at] = v * sinv_u;
                            } else {
                                float v = kms * F[at];
                                v = v + inc;
                                F[at] = v * sinv;
                            }
                        } else if (in_u) {
                            const float prev = U[at];
                            float u = cfv * prev;
                            u = u - curl;
                            u = u * civ;
                            U[at] = u;
                            float v = kms_u * F[at];
                            v = v + u;
                            v = v - prev;
                            F[at] = v * sinv_u;
                        } else {
                            float v = cfv * F[at];
                            v = v - curl;
                            F[at] = v * civ;
                        }
                        continue;
                    }
                    if (!g->pml) {
                        F[at] = F[at] - curl;
                        continue;
                    }
                    const float kms = g->kms[half][dsig][ic[dsig]];
                    const float sinv = g->sinv[half][dsig][ic[dsig]];
                    if (g->mem[half][dsigu][ic[dsigu]]) {
                        const float kms_u = g->kms[half][dsigu][ic[dsigu]];
                        const float sinv_u = g->sinv[half][dsigu][ic[dsigu]];
                        float prev = U[at];
                        float u = kms * prev;
                        u = u - curl;
                        u = sinv * u;
                        U[at] = u;
                        float v = kms_u * F[at];
                        v = v + u;
                        v = v - prev;
                        F[at] = sinv_u * v;
                    } else {
                        float v = kms * F[at];
                        v = v - curl;
                        F[at] = sinv * v;
                    }
                }
            }
        }
    }
}

/* H from B: f_w accumulation on the dsigw (own-axis) layer's cells, H = B elsewhere.
 * `separated` 0 runs MEEP's first-call start (H and f_w copied from B on those cells). */
void oracle_update_h(const ogrid *g, float *const h[3], float *const b[3], float *const fw[3],
                     int separated) {
    const int nx = g->n[0], ny = g->n[1], nz = g->n[2];
    for (int c = 0; c < 3; ++c) {
        float *H = h[c], *W = fw[c];
        const float *B = b[c];
#pragma omp parallel for schedule(static)
        for (int i = 0; i < nx; ++i) {
            for (int j = 0; j < ny; ++j) {
                for (int k = 0; k < nz; ++k) {
                    const int64_t at = IDX(i, j, k);
                    const int ic[3] = {i, j, k};
                    if (g->mem[0][c][ic[c]]) {
                        if (!separated) { H[at] = B[at]; W[at] = B[at]; }
                        const float kps = g->kps[0][c][ic[c]], kms = g->kms[0][c][ic[c]];
                        float prev = W[at];
                        W[at] = B[at];
                        float a = kps * W[at];
                        float m = kms * prev;
                        H[at] = H[at] + (a - m);
                    } else {
                        H[at] = B[at];
                    }
                }
            }
        }
    }
}

#define ORACLE_MAX_POLES 8

typedef struct {
    int npol;                   /* at most ORACLE_MAX_POLES */
    const float *p[ORACLE_MAX_POLES][3];  /* P[k][c], NULL when pole k does not drive c */
    const float *u[3][3];       /* off-diagonal chi1inv[row][partner], NULL when absent */
    int offdiag;                /* any off-diagonal row present */
} oelectric;

static inline float dmp(const oelectric *x, float *const d[3], int c, int64_t at) {
    float v = d[c][at];
    for (int k = 0; k < x->npol; ++k)
        if (x->p[k][c]) v = v - x->p[k][c][at];
    return v;
}

/* A connection for an explicit part (the off-diagonal row forms the other part's product). */
static inline float conn_part(const ogrid *g, int axis, int conj, int part, float mine, float other) {
    if (g->pkind[axis] != 3) return g->pkind[axis] == 2 ? -mine : mine;
    const float a = g->ph[axis][0], b = conj ? -g->ph[axis][1] : g->ph[axis][1];
    const float x = a * mine, y = b * other;
    return part == 0 ? x - y : x + y;
}

/* The off-diagonal row's partner pair at cell (i, j, k) for one part: (D - P) of partner `pa`
 * there, read across a phased partner axis's window edge as the word above it, plus its lower
 * neighbour along `pa` -- zero past a metallic wall, the mirror ghost on a folded axis (parity
 * of D_pa about its own plane, -phase, times stored row 2), else the periodic neighbour
 * (stepping._offdiagonal_terms :1594-1687 with _shift_down :2485). `xm`/`dm` are this part's
 * polarizations and D, `xt`/`dt_` the other part's (NULL for real storage). */
static float offdiag_pair(const ogrid *g, int part, const oelectric *xm, float *const dm[3],
                          const oelectric *xt, float *const dt_[3], int pa, int i, int j, int k) {
    const int ic[3] = {i, j, k};
    const int64_t at = ((int64_t)i * g->n[1] + j) * g->n[2] + k;
    float v = dmp(xm, dm, pa, at);
    if (g->pkind[pa] && ic[pa] == 0) v = conn_part(g, pa, 0, part, v, xt ? dmp(xt, dt_, pa, at) : 0.0f);
    float lower;
    if (g->metal[pa] && ic[pa] == 0) {
        lower = 0.0f;
    } else if (g->fold[pa] && ic[pa] == 0) {
        int r[3] = {i, j, k};
        r[pa] = 2;
        const float m = dmp(xm, dm, pa, ((int64_t)r[0] * g->n[1] + r[1]) * g->n[2] + r[2]);
        lower = g->fphase[pa] > 0 ? -m : m;
    } else {
        lower = dmp(xm, dm, pa, nb(g, i, j, k, pa, -1));
    }
    return v + lower;
}

/* E from D: constitutive product (diagonal, or the off-diagonal row in double), then the
 * f_w accumulation on the own-axis layer's cells under an absorber. For complex storage `xo`
 * and `d_other` are the other part's polarizations and D (NULL for real storage). */
void oracle_update_e(const ogrid *g, const oelectric *x, float *const e[3], float *const d[3],
                     const float *const inv[3], float *const fw[3], const oelectric *xo,
                     float *const d_other[3]) {
    const int nx = g->n[0], ny = g->n[1], nz = g->n[2];
    for (int c = 0; c < 3; ++c) {
        float *E = e[c];
        float *W = fw ? fw[c] : 0;
#pragma omp parallel for schedule(static)
        for (int i = 0; i < nx; ++i) {
            for (int j = 0; j < ny; ++j) {
                for (int k = 0; k < nz; ++k) {
                    const int64_t at = IDX(i, j, k);
                    const int ic[3] = {i, j, k};
                    float cst;
                    int has_row = 0;
                    if (x->offdiag)
                        for (int o = 1; o <= 2; ++o) has_row |= x->u[c][(c + o) % 3] != 0;
                    if (!has_row) {
                        cst = inv[c][at] * dmp(x, d, c, at);
                    } else {
                        double row = (double)(dmp(x, d, c, at) * inv[c][at]);
                        for (int o = 1; o <= 2; ++o) {
                            const int pa = (c + o) % 3;
                            const float *U = x->u[c][pa];
                            if (!U) continue;
                            /* product(q) = u[q] * (g[q] + g[q - s_partner]); past a metallic
                             * wall the partner's lower neighbour and the product above read
                             * the shift helpers' zero ghost, and the summed term is zero on a
                             * wall plane of a metallic axis where this component's Yee shift
                             * is 0 (stepping.py:1594-1687, _mask_metallic_wall_coupling
                             * :1718-1745) */
                            /* the product above reads zero past the top of a metallic or a
                             * folded own axis (_shift_up's zero face); at plane 0 of a phased
                             * own axis the product here is the word below the window */
                            const int top_c = (g->metal[c] || g->fold[c]) && ic[c] == g->n[c] - 1;
                            const int64_t up = nb(g, i, j, k, c, +1);
                            int iu[3] = {i, j, k};
                            iu[c] = (iu[c] + 1) % g->n[c];
                            float prod0 = U[at] * offdiag_pair(g, g->part, x, d, xo, d_other, pa, i, j, k);
                            if (g->pkind[c] && ic[c] == 0) {
                                float other = 0.0f;
                                if (xo)
                                    other = U[at] * offdiag_pair(g, 1 - g->part, xo, d_other, x, d, pa, i, j, k);
                                prod0 = conn_part(g, c, 1, g->part, prod0, other);
                            }
                            float prod1 = 0.0f;
                            if (!top_c)
                                prod1 = U[up] * offdiag_pair(g, g->part, x, d, xo, d_other, pa, iu[0], iu[1], iu[2]);
                            float summed = prod0 + prod1;
                            for (int a = 0; a < 3; ++a)
                                if (a != c && g->metal[a] && ic[a] == 0) summed = 0.0f;
                            row = row + 0.25 * (double)summed;
                        }
                        cst = (float)row;
                    }
                    if (g->pml && g->mem[1][c][ic[c]]) {
                        const float kps = g->kps[1][c][ic[c]], kms = g->kms[1][c][ic[c]];
                        float prev = W[at];
                        W[at] = cst;
                        float a = kps * cst;
                        float m = kms * prev;
                        E[at] = E[at] + (a - m);
                    } else {
                        E[at] = cst;
                    }
                }
            }
        }
    }
}

/* One pole's P update for one component: out = gamma1inv*(((A*p)-(g1*pp))+(od*(s*w))).
 * `w` is f_w on the own-axis layer's cells under an absorber, E elsewhere. `sigma` NULL
 * means the scalar `sigma_scalar`. `out` may alias `pp` (the rotation's retired buffer). */
void oracle_update_p(const ogrid *g, int c, const float consts[4], const float *p,
                     const float *pp, const float *sigma, float sigma_scalar,
                     const float *e, const float *fw, float *out) {
    const int nx = g->n[0], ny = g->n[1], nz = g->n[2];
    const float A = consts[0], g1 = consts[1], od = consts[2], g1inv = consts[3];
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i) {
        for (int j = 0; j < ny; ++j) {
            for (int k = 0; k < nz; ++k) {
                const int64_t at = IDX(i, j, k);
                const int ic[3] = {i, j, k};
                const float w = (g->pml && fw && g->mem[1][c][ic[c]]) ? fw[at] : e[at];
                const float s = sigma ? sigma[at] : sigma_scalar;
                float t = A * p[at];
                t = t - g1 * pp[at];
                float dr = s * w;
                dr = od * dr;
                t = t + dr;
                out[at] = g1inv * t;
            }
        }
    }
}

/* f[idx] = float(double(f[idx]) - ((amp_re*c_re - amp_im*c_im) * dt) [* double(condinv)])
 * at each point; condinv (the written component's conductivity word, sources.py:1357-1400,
 * step.cpp:299-307) is null where that component is lossless, and 0 at a cell whose chunk
 * keeps no conductivity array (MEEP's NULL there). */
void oracle_deposit(float *f, const int64_t *idx, const double *amp_re, const double *amp_im,
                    int n, double c_re, double c_im, double dt, const float *condinv) {
    for (int q = 0; q < n; ++q) {
        double r = amp_re[q] * c_re;
        r = r - amp_im[q] * c_im;
        r = r * dt;
        if (condinv && condinv[idx[q]] != 0.0f) r = r * (double)condinv[idx[q]];
        double w = (double)f[idx[q]];
        w = w - r;
        f[idx[q]] = (float)w;
    }
}

/* Cell-centred sample of one component over a box [lo, lo+cnt) (MEEP's yee2cent): in each
 * axis where the component's Yee shift is 0 the field is averaged 0.5*(f[q]+f[q+1]),
 * x first, then y, then z, each in float32. Periodic wrap past the last row. */
void oracle_centre_sample(const ogrid *g, const float *f, const int shift[3], const int lo[3],
                          const int cnt[3], float *out) {
    const int ny = g->n[1], nz = g->n[2];
    (void)ny; (void)nz;
    for (int a = 0; a < cnt[0]; ++a)
        for (int b = 0; b < cnt[1]; ++b)
            for (int c = 0; c < cnt[2]; ++c) {
                /* gather the 2x2x2 (or smaller) neighbourhood in the reference's order */
                float v[2][2][2];
                const int ex = shift[0] == 0, ey = shift[1] == 0, ez = shift[2] == 0;
                for (int da = 0; da <= ex; ++da)
                    for (int db = 0; db <= ey; ++db)
                        for (int dc = 0; dc <= ez; ++dc) {
                            int64_t i = (lo[0] + a + da) % g->n[0];
                            int64_t j = (lo[1] + b + db) % g->n[1];
                            int64_t k = (lo[2] + c + dc) % g->n[2];
                            v[da][db][dc] = f[(i * g->n[1] + j) * g->n[2] + k];
                        }
                if (ex)
                    for (int db = 0; db <= ey; ++db)
                        for (int dc = 0; dc <= ez; ++dc)
                            v[0][db][dc] = 0.5f * (v[0][db][dc] + v[1][db][dc]);
                if (ey)
                    for (int dc = 0; dc <= ez; ++dc)
                        v[0][0][dc] = 0.5f * (v[0][0][dc] + v[0][1][dc]);
                if (ez) v[0][0][0] = 0.5f * (v[0][0][0] + v[0][0][1]);
                out[((int64_t)a * cnt[1] + b) * cnt[2] + c] = v[0][0][0];
            }
}

/* Reference-precision accumulation: acc(complex64) += phase(complex64) * sample(float32),
 * per part: re += pr*s, im += pi*s (each one float32 rounding). */
void oracle_dft_accumulate_c64(float *acc, const float *phase, const float *sample, int nf,
                               int64_t cells) {
    for (int f = 0; f < nf; ++f) {
        const float pr = phase[2 * f], pi = phase[2 * f + 1];
        float *a = acc + 2 * (int64_t)f * cells;
        for (int64_t q = 0; q < cells; ++q) {
            float s = sample[q];
            a[2 * q] = a[2 * q] + pr * s;
            a[2 * q + 1] = a[2 * q + 1] + pi * s;
        }
    }
}

/* Float64 accumulation against a float64 phase (the oracle for the engine's spectra). */
void oracle_dft_accumulate_c128(double *acc, const double *phase, const float *sample, int nf,
                                int64_t cells) {
#pragma omp parallel for schedule(static)
    for (int f = 0; f < nf; ++f) {
        const double pr = phase[2 * f], pi = phase[2 * f + 1];
        double *a = acc + 2 * (int64_t)f * cells;
        for (int64_t q = 0; q < cells; ++q) {
            double s = (double)sample[q];
            a[2 * q] += pr * s;
            a[2 * q + 1] += pi * s;
        }
    }
}

/* The subnormal policy of every thread the restatement computes on: on != 0 sets MXCSR's FTZ
 * (bit 15, a subnormal result is written as zero) and DAZ (bit 6, a subnormal operand is read
 * as zero), the x86 flush under which the reference mints its flush identity set
 * (meep_identity.py:43-52); 0 clears both. MXCSR is per thread, so it is written in the calling
 * thread and in every thread of an OpenMP team. */
void oracle_set_flush(int on) {
    const unsigned bits = 0x8040u;
#pragma omp parallel
    {
        const unsigned csr = _mm_getcsr();
        _mm_setcsr(on ? (csr | bits) : (csr & ~bits));
    }
    const unsigned csr = _mm_getcsr();
    _mm_setcsr(on ? (csr | bits) : (csr & ~bits));
}

int oracle_flush_state(void) { return (_mm_getcsr() & 0x8040u) == 0x8040u; }
