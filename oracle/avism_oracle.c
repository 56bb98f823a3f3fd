/*
 * avism_oracle.c -- CPU oracle (test infrastructure only) for the AVISM
 * grid-and-void path.  Every function restates one reference routine and
 * cites it.  Build: see oracle/Makefile (gcc -O2 -ffp-contract=off).
 *
 * This code is a checker.  The product path (avism_amd/csrc) never calls it.
 */
#include "avism_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define UM_CONST 9.1717E18f   /* commondata.f90:17 */
#define UV_CONST 299792.458f  /* commondata.f90:16 */
#define UL_CONST 10.98f       /* commondata.f90:15 */
#define NCELL_MIN 10          /* commondata.f90:10 */
#define NSEC 1                /* voidfinding.f90:101 */
#define NBUFF 1               /* voidfinding.f90:797 */

static const double PI_D = 3.141592653589793238462643383279502884;

/* =================================================================== grid */
/* MALLA (grid.f90:3-76) and the PBC extension (avism.f90:780-865). */
orc_grid* orc_grid_new(int32_t n, float L0, int32_t periodic) {
    int32_t nplus = 0, low1 = 1, low2 = n;
    float lado0plus = L0;
    if (periodic) {
        float sizing = 50.0f;                           /* avism.f90:790 */
        float dxp = L0 / (float)n;                      /* avism.f90:791 */
        if (L0 / 2.0f < sizing) sizing = L0 / 2.0f;     /* avism.f90:793 */
        int32_t a = n / 8, b = (int32_t)(sizing / dxp); /* avism.f90:795 */
        nplus = a > b ? a : b;
        low1 = -nplus + 1;
        low2 = n + nplus;
        lado0plus = L0 + (2.0f * (float)nplus) * dxp;   /* avism.f90:799 */
    }
    int32_t rlo = low1 < 0 ? low1 : 0;
    int32_t rhi = low2 > n + 1 ? low2 : n + 1;
    size_t cnt = (size_t)(rhi - rlo + 1);
    orc_grid* g = (orc_grid*)calloc(1, sizeof(orc_grid) + cnt * sizeof(float));
    g->n = n; g->periodic = periodic; g->nplus = nplus;
    g->low1 = low1; g->low2 = low2; g->next = low2 - low1 + 1;
    g->L0 = L0; g->lado0plus = lado0plus; g->rlo = rlo; g->rhi = rhi;
    /* MALLA with LADO = LADO0 - LADO0/NXX (avism.f90:817) */
    float lado = L0 - (L0 / (float)n);
    float A = -lado / 2.0f, B = lado / 2.0f;
    float C = (B - A) / (float)(n - 1);
    float* R = g->rad - rlo;  /* R[i] = RADX(i) */
    R[1] = A;
    for (int32_t i = 2; i <= n; i++) R[i] = R[1] + (float)(i - 1) * C;
    R[0] = R[1] - C;
    R[n + 1] = R[n] + C;
    float dx = R[2] - R[1];
    g->dx = dx; g->rx1 = R[1];
    if (periodic) {  /* ghost centres, avism.f90:837-861 */
        for (int32_t i = low1; i <= low2; i++) {
            if (i < 1) R[i] = R[1] - (float)(1 - i) * dx;
            else if (i > n) R[i] = R[n] + (float)(i - n) * dx;
        }
    }
    return g;
}

void orc_grid_free(orc_grid* g) { free(g); }

float orc_grid_rad(const orc_grid* g, int32_t i) { return g->rad[i - g->rlo]; }

void orc_grid_info(const orc_grid* g, int32_t* ints6, float* floats4, float* rad_ext) {
    ints6[0] = g->n; ints6[1] = g->periodic; ints6[2] = g->nplus;
    ints6[3] = g->low1; ints6[4] = g->low2; ints6[5] = g->next;
    floats4[0] = g->L0; floats4[1] = g->lado0plus; floats4[2] = g->dx; floats4[3] = g->rx1;
    if (rad_ext)
        for (int32_t i = g->low1; i <= g->low2; i++) rad_ext[i - g->low1] = orc_grid_rad(g, i);
}

/* ============================================================== cosmology */
/* avism.f90:469-479 and 599-606 */
void orc_cosmo(float ache, float omega0, float zeta, float* rote, float* rete, float* hte) {
    float ache0 = (float)((double)ache * 3.66e-3);
    float rodo = (float)(((double)omega0 * 3.0) * (double)(ache0 * ache0));
    float re0 = 1.0f / 10.98f;
    if (zeta < 0.0f) zeta = 0.0f;
    float t = 1.0f + zeta;
    *rote = rodo * ((t * t) * t);
    *rete = re0 / t;
    float r = re0 / *rete;
    float h = omega0 * (((r * r) * r) - 1.0f) + 1.0f;
    *hte = ache * sqrtf(h);
}

/* ==================================================================== kNN */
/* Exact k-nearest neighbours with the kd-tree's metric (kdtree.f90:1441-1465):
 * squared FP32 distance accumulated x, y, z in order, min-image under PBC.
 * The reference admits strictly closer points in traversal order
 * (kdtree.f90:658); here ties are broken by the lower particle index so the
 * result does not depend on a traversal order.  Search uses a uniform bin grid
 * scanned in Chebyshev shells until the k-th distance is below the distance to
 * the unscanned region. */
struct orc_knn {
    int64_t np;
    const float *x, *y, *z;
    float L0, lo, bsz;
    int32_t periodic, nb;
    int64_t* start;  /* nb^3 + 1 */
    int64_t* pid;    /* particle ids sorted by bin */
};

static inline int32_t bin_of(const orc_knn* t, float v) {
    int32_t b = (int32_t)floorf((v - t->lo) / t->bsz);
    if (b < 0) b = 0;
    if (b >= t->nb) b = t->nb - 1;
    return b;
}

orc_knn* orc_knn_new(int64_t np, const float* x, const float* y, const float* z,
                     float L0, int32_t periodic) {
    orc_knn* t = (orc_knn*)calloc(1, sizeof(orc_knn));
    t->np = np; t->x = x; t->y = y; t->z = z; t->L0 = L0; t->periodic = periodic;
    int32_t nb = (int32_t)cbrt((double)np / 2.0);
    if (nb < 1) nb = 1;
    if (nb > 256) nb = 256;
    t->nb = nb; t->lo = -L0 / 2.0f; t->bsz = L0 / (float)nb;
    size_t nbins = (size_t)nb * nb * nb;
    t->start = (int64_t*)calloc(nbins + 1, sizeof(int64_t));
    t->pid = (int64_t*)malloc(sizeof(int64_t) * (size_t)(np > 0 ? np : 1));
    int64_t* fill = (int64_t*)calloc(nbins, sizeof(int64_t));
    int64_t* key = (int64_t*)malloc(sizeof(int64_t) * (size_t)(np > 0 ? np : 1));
    for (int64_t p = 0; p < np; p++) {
        int64_t b = bin_of(t, x[p]) + (int64_t)nb * (bin_of(t, y[p]) + (int64_t)nb * bin_of(t, z[p]));
        key[p] = b;
        t->start[b + 1]++;
    }
    for (size_t b = 0; b < nbins; b++) t->start[b + 1] += t->start[b];
    for (int64_t p = 0; p < np; p++) t->pid[t->start[key[p]] + fill[key[p]]++] = p;
    free(fill); free(key);
    return t;
}

void orc_knn_free(orc_knn* t) {
    if (!t) return;
    free(t->start); free(t->pid); free(t);
}

static inline float axis_d(const orc_knn* t, float p, float q) {
    float d = p - q;
    if (t->periodic) {
        float a = fabsf(d), b = t->L0 - fabsf(d);
        d = a < b ? a : b;
    }
    return d;
}

static inline float dist2(const orc_knn* t, int64_t p, float qx, float qy, float qz) {
    float dx = axis_d(t, t->x[p], qx), dy = axis_d(t, t->y[p], qy), dz = axis_d(t, t->z[p], qz);
    float d = 0.0f;
    d = d + dx * dx;
    d = d + dy * dy;
    d = d + dz * dz;
    return d;
}

/* max-heap on (d2, idx) */
static inline int key_gt(float da, int64_t ia, float db, int64_t ib) {
    return da > db || (da == db && ia > ib);
}
static void heap_push(float* hd, int64_t* hi, int32_t* cnt, int32_t k, float d, int64_t i) {
    if (*cnt < k) {
        int32_t c = (*cnt)++;
        hd[c] = d; hi[c] = i;
        while (c > 0) {
            int32_t p = (c - 1) / 2;
            if (key_gt(hd[c], hi[c], hd[p], hi[p])) {
                float td = hd[c]; hd[c] = hd[p]; hd[p] = td;
                int64_t ti = hi[c]; hi[c] = hi[p]; hi[p] = ti;
                c = p;
            } else break;
        }
        return;
    }
    if (!key_gt(hd[0], hi[0], d, i)) return;  /* not closer than the worst */
    hd[0] = d; hi[0] = i;
    int32_t c = 0;
    for (;;) {
        int32_t l = 2 * c + 1, r = l + 1, m = c;
        if (l < k && key_gt(hd[l], hi[l], hd[m], hi[m])) m = l;
        if (r < k && key_gt(hd[r], hi[r], hd[m], hi[m])) m = r;
        if (m == c) break;
        float td = hd[c]; hd[c] = hd[m]; hd[m] = td;
        int64_t ti = hi[c]; hi[c] = hi[m]; hi[m] = ti;
        c = m;
    }
}

typedef struct { float d; int64_t i; } di_pair;
static int cmp_di(const void* a, const void* b) {
    const di_pair* x = (const di_pair*)a; const di_pair* y = (const di_pair*)b;
    if (x->d < y->d) return -1;
    if (x->d > y->d) return 1;
    return (x->i > y->i) - (x->i < y->i);
}

/* visit the bins of Chebyshev shell s around (bx,by,bz); returns 0 if the
 * shell lies entirely outside the (non-periodic) domain */
typedef void (*bin_visit)(void* ctx, int64_t bin);
static int visit_shell(const orc_knn* t, int32_t bx, int32_t by, int32_t bz, int32_t s,
                       bin_visit f, void* ctx) {
    int any = 0, nb = t->nb;
    for (int32_t dz = -s; dz <= s; dz++) {
        for (int32_t dy = -s; dy <= s; dy++) {
            int edge = (dz == -s || dz == s || dy == -s || dy == s);
            for (int32_t dx = -s; dx <= s; dx += (edge || s == 0) ? 1 : 2 * s) {
                int32_t cx = bx + dx, cy = by + dy, cz = bz + dz;
                if (t->periodic) {
                    cx = ((cx % nb) + nb) % nb; cy = ((cy % nb) + nb) % nb; cz = ((cz % nb) + nb) % nb;
                } else if (cx < 0 || cy < 0 || cz < 0 || cx >= nb || cy >= nb || cz >= nb) {
                    continue;
                }
                any = 1;
                f(ctx, cx + (int64_t)nb * (cy + (int64_t)nb * cz));
                if (s == 0) break;
            }
        }
    }
    return any;
}

typedef struct {
    const orc_knn* t; float qx, qy, qz; int32_t k, cnt; float* hd; int64_t* hi;
} knn_ctx;
static void knn_visit(void* c, int64_t b) {
    knn_ctx* q = (knn_ctx*)c;
    const orc_knn* t = q->t;
    for (int64_t j = t->start[b]; j < t->start[b + 1]; j++) {
        int64_t p = t->pid[j];
        heap_push(q->hd, q->hi, &q->cnt, q->k, dist2(t, p, q->qx, q->qy, q->qz), p);
    }
}

/* lower bound on the distance from q to any point outside shells 0..s */
static float outside_bound(const orc_knn* t, float q, int32_t b, int32_t s) {
    float lo_face = t->lo + (float)(b - s) * t->bsz;
    float hi_face = t->lo + (float)(b + s + 1) * t->bsz;
    float dlo = q - lo_face, dhi = hi_face - q;
    if (!t->periodic) {
        if (b - s <= 0) dlo = INFINITY;
        if (b + s + 1 >= t->nb) dhi = INFINITY;
    }
    return dlo < dhi ? dlo : dhi;
}

int32_t orc_knn_query(orc_knn* t, float qx, float qy, float qz, int32_t k,
                      int64_t* idx, float* d2) {
    if (k <= 0 || t->np <= 0) return 0;
    if (k > t->np) k = (int32_t)t->np;
    float* hd = (float*)malloc(sizeof(float) * k);
    int64_t* hi = (int64_t*)malloc(sizeof(int64_t) * k);
    knn_ctx c = {t, qx, qy, qz, k, 0, hd, hi};
    int32_t bx = bin_of(t, qx), by = bin_of(t, qy), bz = bin_of(t, qz);
    int32_t smax = t->nb;  /* enough to cover every bin */
    for (int32_t s = 0; s <= smax; s++) {
        if (t->periodic && 2 * s + 1 > t->nb) {
            /* remaining bins would repeat: finish with a full scan of the rest */
            c.cnt = 0;
            for (int64_t p = 0; p < t->np; p++)
                heap_push(hd, hi, &c.cnt, k, dist2(t, p, qx, qy, qz), p);
            break;
        }
        int any = visit_shell(t, bx, by, bz, s, knn_visit, &c);
        if (!t->periodic && !any && s > 0) {
            /* shell fully outside: continue only if nothing else remains */
            int32_t ext = bx > by ? bx : by; ext = ext > bz ? ext : bz;
            int32_t m1 = t->nb - 1 - bx, m2 = t->nb - 1 - by, m3 = t->nb - 1 - bz;
            int32_t far = ext; if (m1 > far) far = m1; if (m2 > far) far = m2; if (m3 > far) far = m3;
            if (s > far) break;
        }
        if (c.cnt == k) {
            float bd = outside_bound(t, qx, bx, s);
            float b2 = outside_bound(t, qy, by, s); if (b2 < bd) bd = b2;
            float b3 = outside_bound(t, qz, bz, s); if (b3 < bd) bd = b3;
            /* strict, with a relative margin for rounding of the face position */
            if (bd == INFINITY || hd[0] < bd * bd * 0.999f) break;
        }
    }
    int32_t cnt = c.cnt;
    di_pair* pr = (di_pair*)malloc(sizeof(di_pair) * (cnt > 0 ? cnt : 1));
    for (int32_t i = 0; i < cnt; i++) { pr[i].d = hd[i]; pr[i].i = hi[i]; }
    qsort(pr, cnt, sizeof(di_pair), cmp_di);
    for (int32_t i = 0; i < cnt; i++) { idx[i] = pr[i].i; d2[i] = pr[i].d; }
    free(pr); free(hd); free(hi);
    return cnt;
}

typedef struct {
    const orc_knn* t; float qx, qy, qz, r2; int64_t cnt, cap; di_pair* out;
} ball_ctx;
static void ball_visit(void* c, int64_t b) {
    ball_ctx* q = (ball_ctx*)c;
    const orc_knn* t = q->t;
    for (int64_t j = t->start[b]; j < t->start[b + 1]; j++) {
        int64_t p = t->pid[j];
        float d = dist2(t, p, q->qx, q->qy, q->qz);
        if (d <= q->r2) {
            if (q->cnt >= q->cap) {
                q->cap = q->cap * 2 + 16;
                q->out = (di_pair*)realloc(q->out, sizeof(di_pair) * q->cap);
            }
            q->out[q->cnt].d = d; q->out[q->cnt].i = p; q->cnt++;
        }
    }
}

/* ball_search (kdtree.f90:1015-1072): d2 <= radius**2 (inclusive, 1126) */
int64_t orc_ball_query(orc_knn* t, float qx, float qy, float qz, float r,
                       int64_t cap, int64_t* idx, float* d2) {
    ball_ctx c = {t, qx, qy, qz, r * r, 0, 64, NULL};
    c.out = (di_pair*)malloc(sizeof(di_pair) * c.cap);
    int32_t bx = bin_of(t, qx), by = bin_of(t, qy), bz = bin_of(t, qz);
    int32_t s_need = (int32_t)ceilf(r / t->bsz) + 1;
    if (t->periodic && 2 * s_need + 1 > t->nb) {
        for (int64_t p = 0; p < t->np; p++) {
            float d = dist2(t, p, qx, qy, qz);
            if (d <= c.r2) {
                if (c.cnt >= c.cap) { c.cap = c.cap * 2 + 16; c.out = (di_pair*)realloc(c.out, sizeof(di_pair) * c.cap); }
                c.out[c.cnt].d = d; c.out[c.cnt].i = p; c.cnt++;
            }
        }
    } else {
        for (int32_t s = 0; s <= s_need; s++) visit_shell(t, bx, by, bz, s, ball_visit, &c);
    }
    qsort(c.out, (size_t)c.cnt, sizeof(di_pair), cmp_di);
    int64_t n = c.cnt < cap ? c.cnt : cap;
    for (int64_t i = 0; i < n; i++) { idx[i] = c.out[i].i; d2[i] = c.out[i].d; }
    free(c.out);
    return c.cnt;
}

/* ============================================================ field stage */
void orc_convert_units(int64_t np, float* vx, float* vy, float* vz, float* m) {
    for (int64_t i = 0; i < np; i++) {  /* reader.f90:457-461 */
        m[i] = m[i] / UM_CONST;
        vx[i] = vx[i] / UV_CONST; vy[i] = vy[i] / UV_CONST; vz[i] = vz[i] / UV_CONST;
    }
}

/* M4 cubic spline of KERNEL_FUNC (particles.f90:943-950) / DDENS (525-533) */
static inline float m4(float q) {
    if (q <= 1.0f) return 1.0f - (1.5f * (q * q)) * (1.0f - 0.5f * q);
    if (q <= 2.0f) { float t = 2.0f - q; return 0.25f * ((t * t) * t); }
    return 0.0f;
}

/* The per-query loops below run under OpenMP (one query per iteration, each
 * with its own buffers).  Every result is a function of its own query, and
 * HPART is an order-free maximum, so the output does not depend on the
 * thread count. */
void orc_ppart_dens(orc_knn* t, int64_t np, const float* x, const float* y,
                    const float* z, const float* m, int32_t k,
                    float* part_dens, float* hk_out) {
    #pragma omp parallel
    {
        int64_t* idx = (int64_t*)malloc(sizeof(int64_t) * k);
        float* d2 = (float*)malloc(sizeof(float) * k);
        #pragma omp for schedule(dynamic, 256)
        for (int64_t p = 0; p < np; p++) {
            int32_t c = orc_knn_query(t, x[p], y[p], z[p], k, idx, d2);
            float hk = sqrtf(d2[c - 1]);               /* kdtree.f90:539 */
            double bas8 = 0.0;
            for (int32_t j = 0; j < c; j++) bas8 += (double)m[idx[j]];
            double vol = ((4.0 / 3.0) * PI_D) * (double)((hk * hk) * hk);
            part_dens[p] = (float)(bas8 / vol);          /* particles.f90:794 */
            if (hk_out) hk_out[p] = hk;
        }
        free(idx); free(d2);
    }
}

/* HPART = max(HPART, v) (particles.f90:882-885): non-negative floats order as
 * their bits, so an integer compare-and-swap loop gives the maximum in any
 * order */
static inline void atomic_max_f(float* a, float v) {
    uint32_t nv, cur;
    memcpy(&nv, &v, 4);
    cur = __atomic_load_n((uint32_t*)a, __ATOMIC_RELAXED);
    while (nv > cur && !__atomic_compare_exchange_n((uint32_t*)a, &cur, nv, 0, __ATOMIC_RELAXED,
                                                    __ATOMIC_RELAXED)) {
    }
}

/* common cell query of VVEL_INTERP_SPH_VW / H_DISTANCE (particles.f90:858-885):
 * k neighbours, or the ball of radius DX when d_k <= DX. Returns the count. */
static int64_t cell_neighbours(orc_knn* t, const orc_grid* g, int32_t k, float qx, float qy,
                               float qz, int64_t** idx, float** d, int64_t* cap) {
    int32_t c = orc_knn_query(t, qx, qy, qz, k, *idx, *d);
    float dk = sqrtf((*d)[c - 1]);
    if (dk > g->dx) {
        for (int32_t j = 0; j < c; j++) (*d)[j] = sqrtf((*d)[j]);
        return c;
    }
    int64_t n = orc_ball_query(t, qx, qy, qz, g->dx, *cap, *idx, *d);
    if (n > *cap) {
        *cap = n;
        *idx = (int64_t*)realloc(*idx, sizeof(int64_t) * n);
        *d = (float*)realloc(*d, sizeof(float) * n);
        n = orc_ball_query(t, qx, qy, qz, g->dx, *cap, *idx, *d);
    }
    for (int64_t j = 0; j < n; j++) (*d)[j] = sqrtf((*d)[j]);
    return n;
}

/* one cell (0-based linear inner index c) of VVEL_INTERP_SPH_VW
 * (particles.f90:856-925); hpart may be NULL */
static void vvel_cell(orc_knn* t, const orc_grid* g, int32_t k, const float* m, const float* part_dens,
                      const float* vx, const float* vy, const float* vz, float* hpart, size_t c,
                      float* u, float* v, float* w, int64_t** idx, float** d, int64_t* cap) {
    int32_t n = g->n;
    int32_t ix = (int32_t)(c % n) + 1, jy = (int32_t)((c / n) % n) + 1, kz = (int32_t)(c / ((size_t)n * n)) + 1;
    float qx = orc_grid_rad(g, ix), qy = orc_grid_rad(g, jy), qz = orc_grid_rad(g, kz);
    int64_t cnt = cell_neighbours(t, g, k, qx, qy, qz, idx, d, cap);
    if (hpart)
        for (int64_t j = 0; j < cnt; j++) atomic_max_f(&hpart[(*idx)[j]], (*d)[j]);
    float hk = (*d)[cnt - 1];
    double b = 0, bx = 0, by = 0, bz = 0;
    for (int64_t j = 0; j < cnt; j++) {
        int64_t q8 = (*idx)[j];
        float q = (*d)[j] / (hk / 2.0f);           /* particles.f90:941 */
        float wgt = m4(q);
        wgt = wgt * (m[q8] / part_dens[q8]);       /* 897-900 */
        double w8 = (double)wgt;
        b += w8;
        bx += w8 * (double)vx[q8];
        by += w8 * (double)vy[q8];
        bz += w8 * (double)vz[q8];
    }
    *u = (float)(bx / b); *v = (float)(by / b); *w = (float)(bz / b);
}

void orc_vvel_vw(orc_knn* t, const orc_grid* g, int32_t k, const float* m,
                 const float* part_dens, const float* vx, const float* vy,
                 const float* vz, const int8_t* smask, float* hpart,
                 float* u, float* v, float* w) {
    int64_t ncell = (int64_t)g->n * g->n * g->n;
    #pragma omp parallel
    {
        int64_t cap = k > 64 ? k : 64;
        int64_t* idx = (int64_t*)malloc(sizeof(int64_t) * cap);
        float* d = (float*)malloc(sizeof(float) * cap);
        #pragma omp for schedule(dynamic, 256)
        for (int64_t c = 0; c < ncell; c++) {
            if (smask && smask[c] == 0) continue;
            vvel_cell(t, g, k, m, part_dens, vx, vy, vz, hpart, (size_t)c, &u[c], &v[c], &w[c], &idx, &d, &cap);
        }
        free(idx); free(d);
    }
}

/* VVEL_INTERP_SPH_VW's velocities at the listed cells only (0-based linear
 * inner indices), for sampled comparisons on large grids: part_dens needs to
 * hold the neighbours' values */
void orc_vvel_cells(orc_knn* t, const orc_grid* g, int32_t k, const float* m,
                    const float* part_dens, const float* vx, const float* vy,
                    const float* vz, int64_t ncells, const int64_t* cells,
                    float* u, float* v, float* w) {
    #pragma omp parallel
    {
        int64_t cap = k > 64 ? k : 64;
        int64_t* idx = (int64_t*)malloc(sizeof(int64_t) * cap);
        float* d = (float*)malloc(sizeof(float) * cap);
        #pragma omp for schedule(dynamic, 64)
        for (int64_t i = 0; i < ncells; i++)
            vvel_cell(t, g, k, m, part_dens, vx, vy, vz, NULL, (size_t)cells[i], &u[i], &v[i], &w[i], &idx, &d,
                      &cap);
        free(idx); free(d);
    }
}

void orc_h_distance(orc_knn* t, const orc_grid* g, int32_t k, const int8_t* smask,
                    float* hpart) {
    int32_t n = g->n;
    int64_t ncell = (int64_t)n * n * n;
    #pragma omp parallel
    {
        int64_t cap = k > 64 ? k : 64;
        int64_t* idx = (int64_t*)malloc(sizeof(int64_t) * cap);
        float* d = (float*)malloc(sizeof(float) * cap);
        #pragma omp for schedule(dynamic, 256)
        for (int64_t c = 0; c < ncell; c++) {
            if (smask && smask[c] == 0) continue;
            int32_t ix = (int32_t)(c % n) + 1, jy = (int32_t)((c / n) % n) + 1, kz = (int32_t)(c / ((int64_t)n * n)) + 1;
            int64_t cnt = cell_neighbours(t, g, k, orc_grid_rad(g, ix), orc_grid_rad(g, jy),
                                          orc_grid_rad(g, kz), &idx, &d, &cap);
            for (int64_t j = 0; j < cnt; j++) atomic_max_f(&hpart[idx[j]], d[j]);
        }
        free(idx); free(d);
    }
}

void orc_fill_h_zeros(orc_knn* t, int64_t np, const float* x, const float* y,
                      const float* z, int32_t k, float* hpart) {
    #pragma omp parallel
    {
        int64_t* idx = (int64_t*)malloc(sizeof(int64_t) * k);
        float* d2 = (float*)malloc(sizeof(float) * k);
        #pragma omp for schedule(dynamic, 256)
        for (int64_t p = 0; p < np; p++) {
            if (hpart[p] > 0.0f) continue;
            int32_t c = orc_knn_query(t, x[p], y[p], z[p], k, idx, d2);
            hpart[p] = sqrtf(d2[c - 1]);
        }
        free(idx); free(d2);
    }
}

/* par = 0: particles in order, as the reference with one thread; par = 1:
 * particles spread over OpenMP threads with atomic FP64 deposits, as the
 * reference's own OMP ATOMIC (particles.f90:599-602), so the last bits of the
 * sums depend on the schedule */
static void ddens_sph(const orc_grid* g, int64_t np, const float* x, const float* y,
                      const float* z, const float* hpart, const float* m,
                      float* mass_out, double* mass64_out, int par) {
    int32_t n = g->n;
    float dx = g->dx, rx1 = g->rx1;
    size_t ncell = (size_t)n * n * n;
    double* u8 = (double*)calloc(ncell, sizeof(double));
    #pragma omp parallel for schedule(dynamic, 256) if (par)
    for (int64_t i = 0; i < np; i++) {
        int32_t ix = (int32_t)((x[i] - rx1) / dx) + 1;   /* particles.f90:489-491 */
        int32_t jy = (int32_t)((y[i] - rx1) / dx) + 1;
        int32_t kz = (int32_t)((z[i] - rx1) / dx) + 1;
        float h = hpart[i];
        float inv_h = 2.0f / h;
        float h2max = (4.0f * h) * h;
        int32_t inc = (int32_t)((2.0f * h) / dx + 0.5f);
        float norm = 0.0f;
        for (int pass = 0; pass < 2; pass++) {
            for (int32_t k2 = kz - inc; k2 <= kz + inc; k2++) {
                if (k2 < 1 || k2 > n) continue;
                float tz = z[i] - orc_grid_rad(g, k2); float dz2 = tz * tz;
                for (int32_t j2 = jy - inc; j2 <= jy + inc; j2++) {
                    if (j2 < 1 || j2 > n) continue;
                    float ty = y[i] - orc_grid_rad(g, j2); float dy2 = ty * ty;
                    for (int32_t i2 = ix - inc; i2 <= ix + inc; i2++) {
                        if (i2 < 1 || i2 > n) continue;
                        float tx = x[i] - orc_grid_rad(g, i2); float dx2 = tx * tx;
                        float d2 = (dx2 + dy2) + dz2;
                        if (d2 > h2max) continue;
                        float q = sqrtf(d2) * inv_h;
                        float wgt = m4(q);
                        if (pass == 0) {
                            norm = norm + wgt;
                        } else {
                            double* dst = &u8[(size_t)(i2 - 1) + (size_t)n * ((j2 - 1) + (size_t)n * (k2 - 1))];
                            double val = (double)((wgt * m[i]) / norm);
                            if (par) {
                                #pragma omp atomic
                                *dst += val;
                            } else {
                                *dst += val;
                            }
                        }
                    }
                }
            }
            if (pass == 0 && norm == 0.0f) {   /* particles.f90:549-557 */
                if (ix >= 1 && ix <= n && jy >= 1 && jy <= n && kz >= 1 && kz <= n) {
                    double* dst = &u8[(size_t)(ix - 1) + (size_t)n * ((jy - 1) + (size_t)n * (kz - 1))];
                    if (par) {
                        #pragma omp atomic
                        *dst += (double)m[i];
                    } else {
                        *dst += (double)m[i];
                    }
                }
                break;
            }
        }
    }
    for (size_t c = 0; c < ncell; c++) {
        mass_out[c] = (float)u8[c];
        if (mass64_out) mass64_out[c] = u8[c];
    }
    free(u8);
}

void orc_ddens_sph(const orc_grid* g, int64_t np, const float* x, const float* y,
                   const float* z, const float* hpart, const float* m,
                   float* mass_out, double* mass64_out) {
    ddens_sph(g, np, x, y, z, hpart, m, mass_out, mass64_out, 0);
}

void orc_ddens_sph_par(const orc_grid* g, int64_t np, const float* x, const float* y,
                       const float* z, const float* hpart, const float* m,
                       float* mass_out, double* mass64_out) {
    ddens_sph(g, np, x, y, z, hpart, m, mass_out, mass64_out, 1);
}

/* TSC weights (particles.f90:77-96) */
static void tsc_w(const orc_grid* g, float p, int32_t* I, float* V) {
    float dx = g->dx;
    *I = (int32_t)(((p - orc_grid_rad(g, 1)) / dx) + 0.49999f) + 1;
    float bas = fabsf(orc_grid_rad(g, *I - 1) - p);
    float t = 1.5f - bas / dx; V[0] = 0.5f * (t * t);
    bas = fabsf(orc_grid_rad(g, *I) - p);
    t = bas / dx; V[1] = 0.75f - t * t;
    bas = fabsf(orc_grid_rad(g, *I + 1) - p);
    t = 1.5f - bas / dx; V[2] = 0.5f * (t * t);
}

static float tsc_wrap_pos(const orc_grid* g, float p) {
    float dx = g->dx;
    float lo = orc_grid_rad(g, 1) - 0.5f * dx, hi = orc_grid_rad(g, g->n) + 0.5f * dx;
    float leng = hi - lo;
    if (g->periodic) {
        if (p < lo) p = p + leng;
        if (p >= hi) p = p - leng;
    }
    return p;
}

/* cell index after the TSC periodic wrap (particles.f90:104-113); -1 outside
 * (the reference writes out of bounds there in the non-periodic case) */
static inline int32_t tsc_cell(const orc_grid* g, int32_t i) {
    int32_t n = g->n;
    if (g->periodic) {
        if (i < 1) i += n;
        if (i > n) i -= n;
    }
    return (i < 1 || i > n) ? -1 : i;
}

void orc_dens_tsc(const orc_grid* g, int64_t np, const float* x, const float* y,
                  const float* z, const float* m, float* mass_out) {
    int32_t n = g->n;
    size_t ncell = (size_t)n * n * n;
    double* u8 = (double*)calloc(ncell, sizeof(double));
    for (int64_t p = 0; p < np; p++) {
        int32_t I, J, K; float VX[3], VY[3], VZ[3];
        tsc_w(g, tsc_wrap_pos(g, x[p]), &I, VX);
        tsc_w(g, tsc_wrap_pos(g, y[p]), &J, VY);
        tsc_w(g, tsc_wrap_pos(g, z[p]), &K, VZ);
        for (int a = -1; a <= 1; a++)
            for (int b = -1; b <= 1; b++)
                for (int c = -1; c <= 1; c++) {
                    int32_t i3 = tsc_cell(g, I + c), j3 = tsc_cell(g, J + b), k3 = tsc_cell(g, K + a);
                    if (i3 < 0 || j3 < 0 || k3 < 0) continue;
                    float inc = ((m[p] * VX[c + 1]) * VY[b + 1]) * VZ[a + 1];
                    u8[(size_t)(i3 - 1) + (size_t)n * ((j3 - 1) + (size_t)n * (k3 - 1))] += (double)inc;
                }
    }
    for (size_t c = 0; c < ncell; c++) mass_out[c] = (float)u8[c];
    free(u8);
}

void orc_vel_tsc(const orc_grid* g, int64_t np, const float* x, const float* y,
                 const float* z, const float* vx, const float* vy, const float* vz,
                 int32_t exact, float* u, float* v, float* w) {
    /* exact == 0: the reference's REAL*4 sums in its one-thread order (it adds
     * under OMP ATOMIC, particles.f90:226-249, so with more threads its last bits
     * depend on the schedule); exact != 0: FP64 sums, the order-free value */
    int32_t n = g->n;
    size_t ncell = (size_t)n * n * n;
    float* wt = (float*)calloc(ncell, sizeof(float));
    double* a8 = exact ? (double*)calloc(4 * ncell, sizeof(double)) : NULL;
    for (size_t c = 0; c < ncell; c++) { u[c] = 0.0f; v[c] = 0.0f; w[c] = 0.0f; }
    for (int64_t p = 0; p < np; p++) {
        int32_t I, J, K; float VX[3], VY[3], VZ[3];
        tsc_w(g, tsc_wrap_pos(g, x[p]), &I, VX);
        tsc_w(g, tsc_wrap_pos(g, y[p]), &J, VY);
        tsc_w(g, tsc_wrap_pos(g, z[p]), &K, VZ);
        for (int a = -1; a <= 1; a++)
            for (int b = -1; b <= 1; b++)
                for (int c = -1; c <= 1; c++) {
                    int32_t i3 = tsc_cell(g, I + c), j3 = tsc_cell(g, J + b), k3 = tsc_cell(g, K + a);
                    if (i3 < 0 || j3 < 0 || k3 < 0) continue;
                    float inc_wt = (VX[c + 1] * VY[b + 1]) * VZ[a + 1];   /* particles.f90:232 */
                    size_t o = (size_t)(i3 - 1) + (size_t)n * ((j3 - 1) + (size_t)n * (k3 - 1));
                    if (exact) {
                        a8[o] += (double)(vx[p] * inc_wt);
                        a8[ncell + o] += (double)(vy[p] * inc_wt);
                        a8[2 * ncell + o] += (double)(vz[p] * inc_wt);
                        a8[3 * ncell + o] += (double)inc_wt;
                    } else {
                        u[o] = u[o] + vx[p] * inc_wt;
                        v[o] = v[o] + vy[p] * inc_wt;
                        w[o] = w[o] + vz[p] * inc_wt;
                        wt[o] = wt[o] + inc_wt;
                    }
                }
    }
    for (size_t c = 0; c < ncell; c++) {   /* particles.f90:255-267 */
        if (exact) {
            u[c] = (float)a8[c]; v[c] = (float)a8[ncell + c]; w[c] = (float)a8[2 * ncell + c];
            wt[c] = (float)a8[3 * ncell + c];
        }
        float wei = wt[c];
        if (wei == 0.0f) wei = 1.0f;
        u[c] = u[c] / wei; v[c] = v[c] / wei; w[c] = w[c] / wei;
    }
    free(wt);
    free(a8);
}

/* Velocities not available (FLAG_DIV = -1, avism.f90:1646-1650):
 * DIVERCO = -OMEGA0**0.6 * RETE * HTE * (U1CO - 1), Fortran's left-to-right
 * products in REAL*4, the unary minus over the whole product. */
void orc_div_linear(int64_t ncell, const float* u1co, float omega0, float rete, float hte, float* div) {
    const float lin = (powf(omega0, 0.6f) * rete) * hte;
    for (int64_t c = 0; c < ncell; c++) div[c] = -(lin * (u1co[c] - 1.0f));
}

float orc_normalise(const orc_grid* g, float ache, float omega0, float zeta,
                    const float* mass, float* u1co, double* sum_out) {
    float rote, rete, hte;
    orc_cosmo(ache, omega0, zeta, &rote, &rete, &hte);
    float dx = g->dx;
    float den = ((dx * dx) * dx) * ((rete * rete) * rete);   /* avism.f90:1249 */
    size_t ncell = (size_t)g->n * g->n * g->n;
    double s = 0.0;
    for (size_t c = 0; c < ncell; c++) {
        float u = mass[c] / den;
        u = u / rote;                                         /* avism.f90:1250 */
        u = u * rote;                                         /* avism.f90:1386 */
        u1co[c] = u;
        s += (double)u;                                       /* avism.f90:1400 */
    }
    float meandens = (float)(s / (double)ncell);              /* avism.f90:1405 */
    for (size_t c = 0; c < ncell; c++) u1co[c] = u1co[c] / meandens;
    if (sum_out) *sum_out = s;
    float ul3 = (UL_CONST * UL_CONST) * UL_CONST;
    return meandens * (UM_CONST / ul3);                       /* avism.f90:1409 */
}

/* Survey mode (FLAG_SURVEY = 1, avism.f90:1386, 1543-1590, 1646-1650):
 * U1CO = the galaxies' density (mass / (dx^3 RETE^3) / ROTE * ROTE); on
 * observed cells U1CO = D/(alpha R) with alpha = DSUM/RSUM (FP64 sums in
 * Fortran order), clipped to [RATIOFLOOR, RATIOMAX] = [1e-4f, 1000] in FP64;
 * MEANRATIO = REAL*4 of the FP64 mean over observed cells; U1CO /= MEANRATIO
 * there; DIVERCO = -OMEGA0**0.6 RETE HTE (U1CO - 1).  urand: the randoms'
 * cell masses as DDENS (or TSC) leaves them.  Returns MEANDENS. */
float orc_survey(const orc_grid* g, float ache, float omega0, float zeta, const float* mass,
                 const float* urand, const int8_t* smask, float* u1co, float* div) {
    float rote, rete, hte;
    orc_cosmo(ache, omega0, zeta, &rote, &rete, &hte);
    float dx = g->dx;
    float den = ((dx * dx) * dx) * ((rete * rete) * rete);
    size_t ncell = (size_t)g->n * g->n * g->n;
    double dsum = 0.0, rsum = 0.0;
    for (size_t c = 0; c < ncell; c++) {
        float u = mass[c] / den;
        u = u / rote;
        u1co[c] = u * rote;
        if (smask[c] == 0) continue;
        dsum += (double)u1co[c];
        rsum += (double)urand[c];
    }
    const double alpha = dsum / rsum;
    const float floor4 = 1e-4f, max4 = 1000.0f;
    for (size_t c = 0; c < ncell; c++) {
        if (smask[c] == 0) continue;
        double b = urand[c] > 0.0f ? (double)u1co[c] / (alpha * (double)urand[c]) : (double)floor4;
        if (b < (double)floor4) b = (double)floor4;
        if (b > (double)max4) b = (double)max4;
        u1co[c] = (float)b;
    }
    double msum = 0.0;
    int64_t cnt = 0;
    for (size_t c = 0; c < ncell; c++)
        if (smask[c] != 0) {
            msum += (double)u1co[c];
            cnt++;
        }
    const float meanratio = (float)(msum / (double)cnt);
    for (size_t c = 0; c < ncell; c++)
        if (smask[c] != 0) u1co[c] = u1co[c] / meanratio;
    orc_div_linear((int64_t)ncell, u1co, omega0, rete, hte, div);
    float ul3 = (UL_CONST * UL_CONST) * UL_CONST;
    return rote * (UM_CONST / ul3);
}

void orc_diver_uniform(int32_t n, float dx, const float* u, const float* v,
                       const float* w, int32_t periodic, float* div) {
#define F(a, i, j, k) a[(size_t)((i) - 1) + (size_t)n * (((j) - 1) + (size_t)n * ((k) - 1))]
    for (int32_t k = 1; k <= n; k++)
        for (int32_t j = 1; j <= n; j++)
            for (int32_t i = 1; i <= n; i++) {
                float fx, fy, fz;
                /* x */
                if (i >= 2 && i <= n - 1) fx = (F(u, i + 1, j, k) - F(u, i - 1, j, k)) / 2.0f;
                else if (periodic) fx = (i == 1) ? (F(u, 2, j, k) - F(u, n, j, k)) / 2.0f
                                                 : (F(u, 1, j, k) - F(u, n - 1, j, k)) / 2.0f;
                else if (i == 1) fx = 2.0f * (F(u, 2, j, k) - F(u, 1, j, k)) - (F(u, 3, j, k) - F(u, 1, j, k)) / 2.0f;
                else fx = 2.0f * (F(u, n, j, k) - F(u, n - 1, j, k)) - (F(u, n, j, k) - F(u, n - 2, j, k)) / 2.0f;
                /* y */
                if (j >= 2 && j <= n - 1) fy = (F(v, i, j + 1, k) - F(v, i, j - 1, k)) / 2.0f;
                else if (periodic) fy = (j == 1) ? (F(v, i, 2, k) - F(v, i, n, k)) / 2.0f
                                                 : (F(v, i, 1, k) - F(v, i, n - 1, k)) / 2.0f;
                else if (j == 1) fy = 2.0f * (F(v, i, 2, k) - F(v, i, 1, k)) - (F(v, i, 3, k) - F(v, i, 1, k)) / 2.0f;
                else fy = 2.0f * (F(v, i, n, k) - F(v, i, n - 1, k)) - (F(v, i, n, k) - F(v, i, n - 2, k)) / 2.0f;
                /* z */
                if (k >= 2 && k <= n - 1) fz = (F(w, i, j, k + 1) - F(w, i, j, k - 1)) / 2.0f;
                else if (periodic) fz = (k == 1) ? (F(w, i, j, 2) - F(w, i, j, n)) / 2.0f
                                                 : (F(w, i, j, 1) - F(w, i, j, n - 1)) / 2.0f;
                else if (k == 1) fz = 2.0f * (F(w, i, j, 2) - F(w, i, j, 1)) - (F(w, i, j, 3) - F(w, i, j, 1)) / 2.0f;
                else fz = 2.0f * (F(w, i, j, n) - F(w, i, j, n - 1)) - (F(w, i, j, n) - F(w, i, j, n - 2)) / 2.0f;
                F(div, i, j, k) = (fx / dx + fy / dx) + fz / dx;   /* divergence.f90:46 */
            }
#undef F
}

/* ============================================================= void stage */
#define EXT(g, i, j, k) ((size_t)((i) - (g)->low1) + (size_t)(g)->next * \
                         ((size_t)((j) - (g)->low1) + (size_t)(g)->next * (size_t)((k) - (g)->low1)))

/* REAL_MAKE_PERIODIC (num.f90:89-133) applied to U1CO(1:N) = inner */
void orc_extend_f32(const orc_grid* g, const float* inner, float* ext) {
    int32_t n = g->n;
    for (int32_t k = g->low1; k <= g->low2; k++)
        for (int32_t j = g->low1; j <= g->low2; j++)
            for (int32_t i = g->low1; i <= g->low2; i++) {
                int32_t a = i, b = j, c = k;
                if (a < 1) a += n; else if (a > n) a -= n;
                if (b < 1) b += n; else if (b > n) b -= n;
                if (c < 1) c += n; else if (c > n) c -= n;
                ext[EXT(g, i, j, k)] = inner[(size_t)(a - 1) + (size_t)n * ((b - 1) + (size_t)n * (c - 1))];
            }
}

void orc_extend_i32(const orc_grid* g, const int32_t* inner, int32_t* ext) {
    int32_t n = g->n;
    for (int32_t k = g->low1; k <= g->low2; k++)
        for (int32_t j = g->low1; j <= g->low2; j++)
            for (int32_t i = g->low1; i <= g->low2; i++) {
                int32_t a = i, b = j, c = k;
                if (a < 1) a += n; else if (a > n) a -= n;
                if (b < 1) b += n; else if (b > n) b -= n;
                if (c < 1) c += n; else if (c > n) c -= n;
                ext[EXT(g, i, j, k)] = inner[(size_t)(a - 1) + (size_t)n * ((b - 1) + (size_t)n * (c - 1))];
            }
}

/* INDEXX (num.f90:2-84): Numerical Recipes indexx with the reference's
 * median-of-three variant; M=7 insertion sort; NSTACK=50.  1-based output. */
int32_t orc_indexx(int64_t n, const float* arr, int32_t* indx) {
    const int M = 7, NSTACK = 50;
    int64_t istack[51];
    for (int64_t j = 1; j <= n; j++) indx[j - 1] = (int32_t)j;
    if (n <= 0) return 0;
#define IX(i) indx[(i) - 1]
#define A(i) arr[(i) - 1]
    int64_t jstack = 0, l = 1, ir = n;
    for (;;) {
        if (ir - l < M) {
            for (int64_t j = l + 1; j <= ir; j++) {
                int32_t indxt = IX(j);
                float a = A(indxt);
                int64_t i;
                for (i = j - 1; i >= 1; i--) {
                    if (A(IX(i)) <= a) break;
                    IX(i + 1) = IX(i);
                }
                IX(i + 1) = indxt;
            }
            if (jstack == 0) return 0;
            ir = istack[jstack];
            l = istack[jstack - 1];
            jstack -= 2;
        } else {
            int64_t k = (l + ir) / 2;
            int32_t it = IX(k); IX(k) = IX(l + 1); IX(l + 1) = it;
            if (A(IX(l + 1)) > A(IX(ir))) { it = IX(l + 1); IX(l + 1) = IX(ir); IX(ir) = it; }
            if (A(IX(l)) > A(IX(ir))) { it = IX(l); IX(l) = IX(ir); IX(ir) = it; }
            if (A(IX(l + 1)) > A(IX(l))) { it = IX(l + 1); IX(l + 1) = IX(l); IX(l) = it; }
            int64_t i = l + 1, j = ir;
            int32_t indxt = IX(l);
            float a = A(indxt);
            for (;;) {
                do i++; while (A(IX(i)) < a);
                do j--; while (A(IX(j)) > a);
                if (j < i) break;
                it = IX(i); IX(i) = IX(j); IX(j) = it;
            }
            IX(l) = IX(j);
            IX(j) = indxt;
            jstack += 2;
            if (jstack > NSTACK) return -1;   /* 'NSTACK too small in indexx' */
            if (ir - i + 1 >= j - l) {
                istack[jstack] = ir; istack[jstack - 1] = i; ir = j - 1;
            } else {
                istack[jstack] = j - 1; istack[jstack - 1] = l; l = i;
            }
        }
    }
#undef IX
#undef A
}

int32_t orc_voidfind(const orc_grid* g, const orc_void_params* p,
                     const float* u1co, const float* diver,
                     const int32_t* marcap2, const int8_t* smask,
                     const float* reqp, int32_t* ci, float* cf, int32_t* father,
                     int64_t* stats, int64_t ncand_cap, int64_t* cand_lin,
                     int32_t* cand_r) {
    const int32_t L1 = g->low1, L2 = g->low2;
    const size_t ne3 = (size_t)g->next * g->next * g->next;
    const int32_t NV = p->nvoid_max;
    int32_t *INIX = ci, *INIY = ci + NV, *INIZ = ci + 2 * NV;
    int32_t *FINX = ci + 3 * NV, *FINY = ci + 4 * NV, *FINZ = ci + 5 * NV;
    int32_t *ICX = ci + 6 * NV, *ICY = ci + 7 * NV, *ICZ = ci + 8 * NV;
    float *RINX = cf, *RINY = cf + NV, *RINZ = cf + 2 * NV;
    float *RFIX = cf + 3 * NV, *RFIY = cf + 4 * NV, *RFIZ = cf + 5 * NV, *VOL = cf + 6 * NV;
    float DDX = g->dx, RX1 = g->rx1;
    for (int32_t i = 0; i < 9 * NV; i++) ci[i] = 0;

    /* MARK_ALL (voidfinding.f90:13-55) */
    int8_t* flagv = (int8_t*)calloc(ne3, 1);
    int8_t* fsub = (int8_t*)calloc(ne3, 1);
    int64_t nflagv = 0, nfsub = 0;
    float dthr1 = p->dens_thre + 1.0f, dthr2 = p->dens_thre2 + 1.0f;
    for (size_t c = 0; c < ne3; c++) {
        if (smask[c] == 0) continue;
        if (marcap2[c] != 0) {
            fsub[c] = 1; nfsub++;
            if (u1co[c] < dthr1 && diver[c] > 0.0f) { flagv[c] = 1; nflagv++; }
        }
    }
    /* candidates: DIVERCO2 = DIVERCO*FLAGV > 0, Fortran order (161-176) */
    int64_t nxyz = 0;
    for (size_t c = 0; c < ne3; c++) if (diver[c] * (float)flagv[c] > 0.0f) nxyz++;
    float* ddd = (float*)malloc(sizeof(float) * (nxyz > 0 ? nxyz : 1));
    int64_t* dlin = (int64_t*)malloc(sizeof(int64_t) * (nxyz > 0 ? nxyz : 1));
    int64_t ii = 0;
    for (size_t c = 0; c < ne3; c++) {
        float dijk = diver[c] * (float)flagv[c];
        if (dijk > 0.0f) { ddd[ii] = dijk; dlin[ii] = (int64_t)c; ii++; }
    }
    int32_t* ind2 = (int32_t*)malloc(sizeof(int32_t) * (nxyz > 0 ? nxyz : 1));
    if (orc_indexx(nxyz, ddd, ind2) < 0) { free(flagv); free(fsub); free(ddd); free(dlin); free(ind2); return -2; }
    /* reverse: largest divergence first (189-201) */
    int64_t* order = (int64_t*)malloc(sizeof(int64_t) * (nxyz > 0 ? nxyz : 1));
    for (int64_t i = 0; i < nxyz; i++) order[nxyz - 1 - i] = dlin[ind2[i] - 1];
    if (stats) { stats[0] = nxyz; stats[1] = nflagv; stats[2] = nfsub; stats[3] = 0; }

    /* flag grids (227-432); bit layout: 1 X1,2 X2,4 Y1,8 Y2,16 Z1,32 Z2,64 FLAG1 */
    uint8_t* fl = (uint8_t*)calloc(ne3, 1);
    const int32_t ne = g->next;
    for (int32_t kz = L1; kz <= L2; kz++)
        for (int32_t jy = L1; jy <= L2; jy++)
            for (int32_t ix = L1; ix <= L2; ix++) {
                size_t c = EXT(g, ix, jy, kz);
                uint8_t f = 0;
                int32_t indv = marcap2[c];
                if (indv > 0 && reqp) { if (reqp[indv - 1] < p->rmin_sub) f |= 64; }
                /* six directions in reference order; a masked neighbour ends the
                 * cell (CYCLE, 283/308/...): later directions stay 0 */
                int stop = 0;
                for (int dir = 0; dir < 6 && !stop; dir++) {
                    int axis = dir / 2, plus = (dir == 0 || dir == 2 || dir == 4);
                    /* reference order: +X, -X, +Y, -Y, +Z, -Z */
                    uint8_t bit = (uint8_t)(axis == 0 ? (plus ? 2 : 1) : axis == 1 ? (plus ? 8 : 4) : (plus ? 32 : 16));
                    int32_t pos = axis == 0 ? ix : axis == 1 ? jy : kz;
                    int flag = 0;
                    if (plus) { if (pos >= L2 - NSEC) flag = 1; }
                    else { if (pos <= L1 + NSEC) flag = 1; }
                    if (!flag) {
                        int32_t s = plus ? 1 : -1;
                        int32_t II = pos + s;
                        int32_t i1 = ix, j1 = jy, k1 = kz;
                        if (axis == 0) i1 = II; else if (axis == 1) j1 = II; else k1 = II;
                        size_t cn = EXT(g, i1, j1, k1);
                        if (smask[cn] == 0) { stop = 1; break; }
                        int32_t edge = plus ? L2 : L1;
                        if (II == edge) {
                            if (diver[cn] < p->div_thre) flag = 1;
                            if (u1co[cn] >= dthr2) flag = 1;
                        } else {
                            /* outward central difference at the neighbour */
                            int32_t ia = i1, ja = j1, ka = k1, ib = i1, jb = j1, kb = k1;
                            if (axis == 0) { ia += s; ib -= s; } else if (axis == 1) { ja += s; jb -= s; } else { ka += s; kb -= s; }
                            float dd = (u1co[EXT(g, ia, ja, ka)] - u1co[EXT(g, ib, jb, kb)]) / (2.0f * DDX);
                            if (dd >= p->grad_thre) flag = 1;
                            int nextok = plus ? (II < L2 - 1) : (II > L1 + 1);
                            if (flag && nextok) {
                                int32_t i2 = i1, j2 = j1, k2 = k1;
                                if (axis == 0) i2 += 2 * s; else if (axis == 1) j2 += 2 * s; else k2 += 2 * s;
                                float dd2 = (u1co[EXT(g, i2, j2, k2)] - u1co[cn]) / (2.0f * DDX);
                                if (dd2 < p->grad_thre) flag = 0;
                            }
                            if (diver[cn] < p->div_thre) flag = 1;
                            if (u1co[cn] >= dthr2) flag = 1;
                        }
                    }
                    if (flag) f |= bit;
                }
                fl[c] = f;
            }
    (void)ne;

    /* growth loop (449-704) */
    int32_t* aux = (int32_t*)calloc(ne3, sizeof(int32_t));
    int32_t nvoid = 0;
    for (int64_t l1 = 0; l1 < nxyz; l1++) {
        size_t c0 = (size_t)order[l1];
        int32_t IXc = (int32_t)(c0 % ne) + L1;
        int32_t JYc = (int32_t)((c0 / ne) % ne) + L1;
        int32_t KZc = (int32_t)(c0 / ((size_t)ne * ne)) + L1;
        if (cand_lin && l1 < ncand_cap) cand_lin[l1] = (int64_t)c0;
        if (cand_r && l1 < ncand_cap) cand_r[l1] = -1;
        int flag1 = 0;
        if (aux[c0] > 0) flag1 = 1;
        if (IXc <= L1 + NSEC || JYc <= L1 + NSEC || KZc <= L1 + NSEC ||
            IXc >= L2 - NSEC || JYc >= L2 - NSEC || KZc >= L2 - NSEC) flag1 = 1;
        if (flag1) continue;
        nvoid++;
        if (nvoid > NV) { free(flagv); free(fsub); free(ddd); free(dlin); free(ind2); free(order); free(fl); free(aux); return -3; }
        int32_t v = nvoid - 1;
        int FLAG = 0, FX1 = 0, FX2 = 0, FY1 = 0, FY2 = 0, FZ1 = 0, FZ2 = 0;
        INIX[v] = FINX[v] = IXc; INIY[v] = FINY[v] = JYc; INIZ[v] = FINZ[v] = KZc;
        int32_t indv = marcap2[c0];
        father[v] = indv;
        if (fl[c0] & 64) FLAG = 1;
        int FLAG2 = 0;
        while (FLAG == 0) {
            for (int32_t kzz = INIZ[v]; kzz <= FINZ[v]; kzz++)
                for (int32_t jyy = INIY[v]; jyy <= FINY[v]; jyy++) {
                    if (FX1 == 0 || FX2 == 0) {
                        if (FINX[v] >= L2 - NSEC) { FX2 = 1; FLAG2 = 1; }
                        else {
                            int32_t I = FINX[v]; size_t cn = EXT(g, I + 1, jyy, kzz);
                            if (fsub[cn] == 0 || marcap2[cn] != indv) FX2 = 1;
                            if (FX2 == 0) FX2 = (fl[EXT(g, I, jyy, kzz)] & 2) ? 1 : 0;
                        }
                        if (INIX[v] <= L1 + NSEC) { FX1 = 1; FLAG2 = 1; }
                        else {
                            int32_t I = INIX[v]; size_t cn = EXT(g, I - 1, jyy, kzz);
                            if (fsub[cn] == 0 || marcap2[cn] != indv) FX1 = 1;
                            if (FX1 == 0) FX1 = (fl[EXT(g, I, jyy, kzz)] & 1) ? 1 : 0;
                        }
                    }
                }
            for (int32_t kzz = INIZ[v]; kzz <= FINZ[v]; kzz++)
                for (int32_t ixx = INIX[v]; ixx <= FINX[v]; ixx++) {
                    if (FY1 == 0 || FY2 == 0) {
                        if (FINY[v] >= L2 - NSEC) { FY2 = 1; FLAG2 = 1; }
                        else {
                            int32_t I = FINY[v]; size_t cn = EXT(g, ixx, I + 1, kzz);
                            if (fsub[cn] == 0 || marcap2[cn] != indv) FY2 = 1;
                            if (FY2 == 0) FY2 = (fl[EXT(g, ixx, I, kzz)] & 8) ? 1 : 0;
                        }
                        if (INIY[v] <= L1 + NSEC) { FY1 = 1; FLAG2 = 1; }
                        else {
                            int32_t I = INIY[v]; size_t cn = EXT(g, ixx, I - 1, kzz);
                            if (fsub[cn] == 0 || marcap2[cn] != indv) FY1 = 1;
                            if (FY1 == 0) FY1 = (fl[EXT(g, ixx, I, kzz)] & 4) ? 1 : 0;
                        }
                    }
                }
            for (int32_t jyy = INIY[v]; jyy <= FINY[v]; jyy++)
                for (int32_t ixx = INIX[v]; ixx <= FINX[v]; ixx++) {
                    if (FZ1 == 0 || FZ2 == 0) {
                        if (FINZ[v] >= L2 - NSEC) { FZ2 = 1; FLAG2 = 1; }
                        else {
                            int32_t I = FINZ[v]; size_t cn = EXT(g, ixx, jyy, I + 1);
                            if (fsub[cn] == 0 || marcap2[cn] != indv) FZ2 = 1;
                            if (FZ2 == 0) FZ2 = (fl[EXT(g, ixx, jyy, I)] & 32) ? 1 : 0;
                        }
                        if (INIZ[v] <= L1 + NSEC) { FZ1 = 1; FLAG2 = 1; }
                        else {
                            int32_t I = INIZ[v]; size_t cn = EXT(g, ixx, jyy, I - 1);
                            if (fsub[cn] == 0 || marcap2[cn] != indv) FZ1 = 1;
                            if (FZ1 == 0) FZ1 = (fl[EXT(g, ixx, jyy, I)] & 16) ? 1 : 0;
                        }
                    }
                }
            if (!FX1 && !FX2 && !FY1 && !FY2 && !FZ1 && !FZ2) {
                INIX[v]--; FINX[v]++; INIY[v]--; FINY[v]++; INIZ[v]--; FINZ[v]++;
            } else {
                FX1 = FX2 = FY1 = FY2 = FZ1 = FZ2 = 1;
            }
            if (INIX[v] <= L1 + NSEC) FX1 = 1;
            if (FINX[v] >= L2 - NSEC) FX2 = 1;
            if (INIY[v] <= L1 + NSEC) FY1 = 1;
            if (FINY[v] >= L2 - NSEC) FY2 = 1;
            if (INIZ[v] <= L1 + NSEC) FZ1 = 1;
            if (FINZ[v] >= L2 - NSEC) FZ2 = 1;
            if (FX1 && FX2 && FY1 && FY2 && FZ1 && FZ2) FLAG = 1;
        }
        if (INIX[v] >= FINX[v] && INIY[v] >= FINY[v] && INIZ[v] >= FINZ[v]) FLAG2 = 1;
        if (INIX[v] > FINX[v] || INIY[v] > FINY[v] || INIZ[v] > FINZ[v]) FLAG2 = 1;
        float A = (float)(FINX[v] - INIX[v] + 1) * DDX;
        float B = (float)(FINY[v] - INIY[v] + 1) * DDX;
        float C = (float)(FINZ[v] - INIZ[v] + 1) * DDX;
        float VV = (A * B) * C;
        /* SIDE_MIN = 0 (commondata.f90:12): never triggers */
        if (INIX[v] == L1 || FINX[v] == L2 || INIY[v] == L1 || FINY[v] == L2 ||
            INIZ[v] == L1 || FINZ[v] == L2) FLAG2 = 1;
        if (cand_r && l1 < ncand_cap) cand_r[l1] = (FINX[v] - INIX[v]) / 2;
        if (FLAG2 == 0) {
            ICX[v] = IXc; ICY[v] = JYc; ICZ[v] = KZc;
            RINX[v] = (RX1 + (float)(INIX[v] - 1) * DDX) - 0.5f * DDX;
            RFIX[v] = (RX1 + (float)(FINX[v] - 1) * DDX) + 0.5f * DDX;
            RINY[v] = (RX1 + (float)(INIY[v] - 1) * DDX) - 0.5f * DDX;
            RFIY[v] = (RX1 + (float)(FINY[v] - 1) * DDX) + 0.5f * DDX;
            RINZ[v] = (RX1 + (float)(INIZ[v] - 1) * DDX) - 0.5f * DDX;
            RFIZ[v] = (RX1 + (float)(FINZ[v] - 1) * DDX) + 0.5f * DDX;
            VOL[v] = VV;
            for (int32_t kzz = INIZ[v]; kzz <= FINZ[v]; kzz++)
                for (int32_t jyy = INIY[v]; jyy <= FINY[v]; jyy++)
                    for (int32_t ixx = INIX[v]; ixx <= FINX[v]; ixx++) aux[EXT(g, ixx, jyy, kzz)] = nvoid;
        } else {
            INIX[v] = INIY[v] = INIZ[v] = FINX[v] = FINY[v] = FINZ[v] = 0;
            nvoid--;
        }
    }
    free(flagv); free(fsub); free(ddd); free(dlin); free(ind2); free(order); free(fl); free(aux);
    return nvoid;
}

/* ============================================================ merge + post */
/* JACOBI (shape.f90:46-134), float with the reference's promotions */
static void jacobi3(float A[3][3], float D[3]) {
    float B[3], Z[3];
    for (int ip = 0; ip < 3; ip++) { B[ip] = A[ip][ip]; D[ip] = B[ip]; Z[ip] = 0.0f; }
    for (int i = 1; i <= 50; i++) {
        float sm = 0.0f;
        for (int ip = 0; ip < 2; ip++)
            for (int iq = ip + 1; iq < 3; iq++) sm = sm + fabsf(A[ip][iq]);
        if (sm == 0.0f) return;
        float tresh = (i < 4) ? (float)(0.2 * (double)(sm * sm)) : 0.0f;
        for (int ip = 0; ip < 2; ip++)
            for (int iq = ip + 1; iq < 3; iq++) {
                float g = (float)(100.0 * (double)fabsf(A[ip][iq]));
                if (i > 4 && fabsf(D[ip]) + g == fabsf(D[ip]) && fabsf(D[iq]) + g == fabsf(D[iq])) {
                    A[ip][iq] = 0.0f;
                } else if (fabsf(A[ip][iq]) > tresh) {
                    float h = D[iq] - D[ip], t;
                    if (fabsf(h) + g == fabsf(h)) t = A[ip][iq] / h;
                    else {
                        float theta = (float)(0.5 * (double)h / (double)A[ip][iq]);
                        t = (float)(1.0 / ((double)fabsf(theta) + (double)sqrtf((float)(1.0 + (double)(theta * theta)))));
                        if (theta < 0.0f) t = -t;
                    }
                    float c = (float)(1.0 / (double)sqrtf((float)(1.0 + (double)(t * t))));
                    float s = t * c;
                    float tau = (float)((double)s / (1.0 + (double)c));
                    h = t * A[ip][iq];
                    Z[ip] = Z[ip] - h; Z[iq] = Z[iq] + h;
                    D[ip] = D[ip] - h; D[iq] = D[iq] + h;
                    A[ip][iq] = 0.0f;
                    for (int j = 0; j < ip; j++) {
                        float gg = A[j][ip], hh = A[j][iq];
                        A[j][ip] = gg - s * (hh + gg * tau); A[j][iq] = hh + s * (gg - hh * tau);
                    }
                    for (int j = ip + 1; j < iq; j++) {
                        float gg = A[ip][j], hh = A[j][iq];
                        A[ip][j] = gg - s * (hh + gg * tau); A[j][iq] = hh + s * (gg - hh * tau);
                    }
                    for (int j = iq + 1; j < 3; j++) {
                        float gg = A[ip][j], hh = A[iq][j];
                        A[ip][j] = gg - s * (hh + gg * tau); A[iq][j] = hh + s * (gg - hh * tau);
                    }
                }
            }
        for (int ip = 0; ip < 3; ip++) { B[ip] = B[ip] + Z[ip]; D[ip] = B[ip]; Z[ip] = 0.0f; }
    }
}

typedef struct { float d; int32_t id; } dcube;
static int cmp_dcube(const void* a, const void* b) {
    const dcube* x = (const dcube*)a; const dcube* y = (const dcube*)b;
    if (x->d < y->d) return -1;
    if (x->d > y->d) return 1;
    return (x->id > y->id) - (x->id < y->id);
}

/* kd-tree metric for cube centres (periodic length LADO0PLUS, voidfinding.f90:780-788) */
static float cube_d2(const orc_grid* g, float ax, float ay, float az, float bx, float by, float bz) {
    float d[3] = {ax - bx, ay - by, az - bz};
    float s = 0.0f;
    for (int i = 0; i < 3; i++) {
        float v = d[i];
        if (g->periodic) { float a = fabsf(v), b = g->lado0plus - fabsf(v); v = a < b ? a : b; }
        s = s + v * v;
    }
    return s;
}

int32_t orc_merge_post(const orc_grid* g, const orc_void_params* p, int32_t nvoid,
                       const int32_t* ci, const float* cf, const int32_t* father,
                       const float* u1co, float meandens, float rete,
                       int32_t* indice, int32_t* uvoid, int32_t* marca,
                       float* vf, int32_t* ncellv, int32_t* stats_i, double* stats_d,
                       int32_t stop_after) {
    const int32_t NV = p->nvoid_max;
    const int32_t *INIX = ci, *INIY = ci + NV, *INIZ = ci + 2 * NV;
    const int32_t *FINX = ci + 3 * NV, *FINY = ci + 4 * NV, *FINZ = ci + 5 * NV;
    const int32_t *ICX = ci + 6 * NV, *ICY = ci + 7 * NV, *ICZ = ci + 8 * NV;
    const float *RINX = cf, *RINY = cf + NV, *RINZ = cf + 2 * NV;
    const float *RFIX = cf + 3 * NV, *RFIY = cf + 4 * NV, *RFIZ = cf + 5 * NV, *VOL = cf + 6 * NV;
    const int32_t L1 = g->low1, L2 = g->low2, n = g->n;
    const size_t ne3 = (size_t)g->next * g->next * g->next;
    float DXX = g->dx, R1 = g->rx1;
    float *GXC = vf, *GYC = vf + nvoid, *GZC = vf + 2 * nvoid;
    float *XC = vf + 3 * nvoid, *YC = vf + 4 * nvoid, *ZC = vf + 5 * nvoid;
    float *VOLNEW = vf + 6 * nvoid, *UMEAN = vf + 7 * nvoid, *MTOT = vf + 8 * nvoid;
    float *EPS = vf + 9 * nvoid, *IPR = vf + 10 * nvoid, *REQ = vf + 11 * nvoid;
    for (int32_t i = 0; i < 13 * nvoid; i++) vf[i] = 0.0f;
    for (size_t c = 0; c < ne3; c++) marca[c] = 0;
#define UV(id) uvoid[(id) - 1]

    /* cube volume order (avism.f90:1738-1745) */
    int32_t* ind2 = (int32_t*)malloc(sizeof(int32_t) * (nvoid > 0 ? nvoid : 1));
    if (orc_indexx(nvoid, VOL, ind2) < 0) { free(ind2); return -2; }
    for (int32_t i = 0; i < nvoid; i++) indice[nvoid - 1 - i] = ind2[i];
    free(ind2);

    /* MERGE_VOID (voidfinding.f90:718-999) */
    for (int32_t i = 0; i < nvoid; i++) {
        GXC[i] = R1 + (float)(ICX[i] - 1) * DXX;
        GYC[i] = R1 + (float)(ICY[i] - 1) * DXX;
        GZC[i] = R1 + (float)(ICZ[i] - 1) * DXX;
        VOLNEW[i] = ((RFIX[i] - RINX[i]) * (RFIY[i] - RINY[i])) * (RFIZ[i] - RINZ[i]);
        uvoid[i] = -1;
    }
    int32_t nvoid2 = nvoid;
    int8_t* major = (int8_t*)calloc(nvoid > 0 ? nvoid : 1, 1);
    /* bucket grid of cube centres for the ball query */
    int32_t rmax = 0;
    for (int32_t i = 0; i < nvoid; i++) { int32_t r = (FINX[i] - INIX[i]) / 2; if (r > rmax) rmax = r; }
    const int32_t bs = 8;
    int32_t nbk = (g->next + bs - 1) / bs;
    int32_t* bstart = (int32_t*)calloc((size_t)nbk * nbk * nbk + 1, sizeof(int32_t));
    int32_t* bid = (int32_t*)malloc(sizeof(int32_t) * (nvoid > 0 ? nvoid : 1));
    int32_t* bfill = (int32_t*)calloc((size_t)nbk * nbk * nbk, sizeof(int32_t));
#define BK(i) (((ICX[i] - L1) / bs) + nbk * (((ICY[i] - L1) / bs) + nbk * ((ICZ[i] - L1) / bs)))
    for (int32_t i = 0; i < nvoid; i++) bstart[BK(i) + 1]++;
    for (int32_t b = 0; b < nbk * nbk * nbk; b++) bstart[b + 1] += bstart[b];
    for (int32_t i = 0; i < nvoid; i++) { int32_t b = BK(i); bid[bstart[b] + bfill[b]++] = i; }
    free(bfill);
    dcube* cand = (dcube*)malloc(sizeof(dcube) * (nvoid > 0 ? nvoid : 1));

    for (int32_t I = 0; I < nvoid; I++) {
        int32_t IND1 = indice[I], IND0, INDP;
        int32_t a = IND1 - 1;
        if (UV(IND1) > 0) {
            IND0 = UV(IND1);
            INDP = father[IND0 - 1];
        } else {
            IND0 = IND1;
            INDP = father[IND0 - 1];
            int32_t ix1 = INIX[a] - 1, ix2 = FINX[a] + 1, jy1 = INIY[a] - 1, jy2 = FINY[a] + 1;
            int32_t kz1 = INIZ[a] - 1, kz2 = FINZ[a] + 1;
            if (ix1 < L1) ix1 = L1; if (ix2 > L2) ix2 = L2;
            if (jy1 < L1) jy1 = L1; if (jy2 > L2) jy2 = L2;
            if (kz1 < L1) kz1 = L1; if (kz2 > L2) kz2 = L2;
            int64_t dist = 2147483647;
            for (int32_t kz = kz1; kz <= kz2; kz++)
                for (int32_t jy = jy1; jy <= jy2; jy++)
                    for (int32_t ix = ix1; ix <= ix2; ix++) {
                        int32_t mm = marca[EXT(g, ix, jy, kz)];
                        if (mm > 0) {
                            if (father[mm - 1] != INDP) continue;
                            int64_t d2 = (int64_t)(ICX[a] - ix) * (ICX[a] - ix) + (int64_t)(ICY[a] - jy) * (ICY[a] - jy) +
                                         (int64_t)(ICZ[a] - kz) * (ICZ[a] - kz);
                            if (d2 < dist) { IND0 = mm; dist = d2; }
                        }
                    }
            if (IND0 != IND1) {
                int32_t b0 = IND0 - 1;
                UV(IND1) = IND0;
                float s = VOLNEW[b0] + VOLNEW[a];
                GXC[b0] = (GXC[b0] * VOLNEW[b0] + GXC[a] * VOLNEW[a]) / s;
                GYC[b0] = (GYC[b0] * VOLNEW[b0] + GYC[a] * VOLNEW[a]) / s;
                GZC[b0] = (GZC[b0] * VOLNEW[b0] + GZC[a] * VOLNEW[a]) / s;
                VOLNEW[b0] = VOLNEW[b0] + VOLNEW[a];
                for (int32_t kz = INIZ[a]; kz <= FINZ[a]; kz++)
                    for (int32_t jy = INIY[a]; jy <= FINY[a]; jy++)
                        for (int32_t ix = INIX[a]; ix <= FINX[a]; ix++) {
                            size_t c = EXT(g, ix, jy, kz);
                            if (marca[c] == 0) marca[c] = IND0;
                        }
                nvoid2--;
            } else {
                for (int32_t kz = INIZ[a]; kz <= FINZ[a]; kz++)
                    for (int32_t jy = INIY[a]; jy <= FINY[a]; jy++)
                        for (int32_t ix = INIX[a]; ix <= FINX[a]; ix++) {
                            size_t c = EXT(g, ix, jy, kz);
                            if (marca[c] == 0) marca[c] = IND0;
                        }
                major[IND0 - 1] = 1;
            }
        }
        /* ball query around the cube's centre; the integer overlap test decides */
        int32_t ix1 = INIX[a] - NBUFF, ix2 = FINX[a] + NBUFF, jy1 = INIY[a] - NBUFF, jy2 = FINY[a] + NBUFF;
        int32_t kz1 = INIZ[a] - NBUFF, kz2 = FINZ[a] + NBUFF;
        float tx = GXC[a], ty = GYC[a], tz = GZC[a];
        int32_t r1 = (FINX[a] - INIX[a]) / 2;
        int32_t reach = r1 + rmax + 2 * NBUFF;
        int32_t bx0 = (ICX[a] - reach - L1) / bs, bx1 = (ICX[a] + reach - L1) / bs;
        int32_t by0 = (ICY[a] - reach - L1) / bs, by1 = (ICY[a] + reach - L1) / bs;
        int32_t bz0 = (ICZ[a] - reach - L1) / bs, bz1 = (ICZ[a] + reach - L1) / bs;
        if (ICX[a] - reach - L1 < 0) bx0 = 0; if (ICY[a] - reach - L1 < 0) by0 = 0; if (ICZ[a] - reach - L1 < 0) bz0 = 0;
        if (bx1 >= nbk) bx1 = nbk - 1; if (by1 >= nbk) by1 = nbk - 1; if (bz1 >= nbk) bz1 = nbk - 1;
        int32_t nc = 0;
        for (int32_t bz = bz0; bz <= bz1; bz++)
            for (int32_t by = by0; by <= by1; by++)
                for (int32_t bx = bx0; bx <= bx1; bx++) {
                    int32_t b = bx + nbk * (by + nbk * bz);
                    for (int32_t q = bstart[b]; q < bstart[b + 1]; q++) {
                        int32_t c2 = bid[q];
                        if (c2 == a) continue;
                        int32_t ix3 = INIX[c2] - NBUFF, ix4 = FINX[c2] + NBUFF, jy3 = INIY[c2] - NBUFF, jy4 = FINY[c2] + NBUFF;
                        int32_t kz3 = INIZ[c2] - NBUFF, kz4 = FINZ[c2] + NBUFF;
                        if (!(ix1 <= ix4 && ix3 <= ix2 && jy1 <= jy4 && jy3 <= jy2 && kz1 <= kz4 && kz3 <= kz2)) continue;
                        cand[nc].d = cube_d2(g, GXC[c2], GYC[c2], GZC[c2], tx, ty, tz);
                        cand[nc].id = c2 + 1;
                        nc++;
                    }
                }
        qsort(cand, (size_t)nc, sizeof(dcube), cmp_dcube);
        for (int32_t q = 0; q < nc; q++) {
            int32_t IND2 = cand[q].id, b2 = IND2 - 1;
            if (major[b2] == 1) continue;
            if (UV(IND2) >= 0) continue;
            if (INDP != father[b2]) continue;
            int32_t b0 = IND0 - 1;
            UV(IND2) = IND0;
            float s = VOLNEW[b0] + VOLNEW[b2];
            GXC[b0] = (GXC[b0] * VOLNEW[b0] + GXC[b2] * VOLNEW[b2]) / s;
            GYC[b0] = (GYC[b0] * VOLNEW[b0] + GYC[b2] * VOLNEW[b2]) / s;
            GZC[b0] = (GZC[b0] * VOLNEW[b0] + GZC[b2] * VOLNEW[b2]) / s;
            VOLNEW[b0] = VOLNEW[b0] + VOLNEW[b2];
            for (int32_t kz = INIZ[b2]; kz <= FINZ[b2]; kz++)
                for (int32_t jy = INIY[b2]; jy <= FINY[b2]; jy++)
                    for (int32_t ix = INIX[b2]; ix <= FINX[b2]; ix++) {
                        size_t c = EXT(g, ix, jy, kz);
                        if (marca[c] == 0) marca[c] = IND0;
                    }
            nvoid2--;
        }
    }
    free(cand); free(bstart); free(bid); free(major);
#undef BK
    if (stats_i) { stats_i[0] = nvoid2; stats_i[1] = 0; stats_i[2] = 0; stats_i[3] = 0; }
    if (stop_after == 1) return 0;

    /* ---- post-processing (avism.f90:1796-1973) */
    for (int32_t i = 0; i < nvoid; i++) ncellv[i] = 0;
    for (size_t c = 0; c < ne3; c++) {
        int32_t id = marca[c];
        if (id > 0 && UV(id) == -1) ncellv[id - 1]++;
    }
    /* remove a main void everywhere (the WHERE blocks of 1817-1826).  A
     * removal changes only the removed void's own cells and members, so the
     * removals decided in one pass over INDICE are applied together after it
     * (KILL_APPLY), with the same result as one at a time. */
    uint8_t* dead = (uint8_t*)calloc(nvoid > 0 ? nvoid : 1, 1);
#define KILL(id) do { int32_t _id = (id); dead[_id - 1] = 1; UV(_id) = 0;     \
        VOLNEW[_id - 1] = 0.0f; ncellv[_id - 1] = 0; } while (0)
#define KILL_APPLY() do {                                                     \
        for (size_t c_ = 0; c_ < ne3; c_++) {                                 \
            int32_t m_ = marca[c_]; if (m_ > 0 && dead[m_ - 1]) marca[c_] = 0; } \
        for (int32_t v_ = 0; v_ < nvoid; v_++) {                              \
            int32_t u_ = uvoid[v_]; if (u_ > 0 && dead[u_ - 1]) uvoid[v_] = 0; } \
    } while (0)
    for (int32_t I = 0; I < nvoid; I++) {
        int32_t id = indice[I];
        if (UV(id) >= 0) continue;
        if (ncellv[id - 1] < NCELL_MIN) KILL(id);
    }
    KILL_APPLY();
    /* VOID_SHAPE (shape.f90:140-225) */
    float* inert = (float*)calloc((size_t)9 * (nvoid > 0 ? nvoid : 1), sizeof(float));
    for (int32_t kz = L1; kz <= L2; kz++)
        for (int32_t jy = L1; jy <= L2; jy++)
            for (int32_t ix = L1; ix <= L2; ix++) {
                int32_t id = marca[EXT(g, ix, jy, kz)];
                if (id <= 0) continue;
                float rk[3] = {orc_grid_rad(g, ix) - GXC[id - 1], orc_grid_rad(g, jy) - GYC[id - 1],
                               orc_grid_rad(g, kz) - GZC[id - 1]};
                float rr = sqrtf((rk[0] * rk[0] + rk[1] * rk[1]) + rk[2] * rk[2]);
                float* T = inert + 9 * (size_t)(id - 1);
                for (int i2 = 0; i2 < 3; i2++)
                    for (int i1 = 0; i1 < 3; i1++) {
                        float delta = (i1 == i2) ? 1.0f : 0.0f;
                        T[i1 + 3 * i2] = T[i1 + 3 * i2] + ((delta * rr) * rr - rk[i1] * rk[i2]);
                    }
            }
    for (int32_t I = 0; I < nvoid; I++) {
        int32_t id = indice[I];
        if (UV(id) != -1) continue;
        float A3[3][3], D[3];
        float* T = inert + 9 * (size_t)(id - 1);
        for (int i2 = 0; i2 < 3; i2++)
            for (int i1 = 0; i1 < 3; i1++) A3[i1][i2] = T[i1 + 3 * i2] / (float)ncellv[id - 1];
        D[0] = D[1] = D[2] = 0.0f;
        jacobi3(A3, D);
        float ax[3] = {sqrtf(2.5f * fabsf(D[1] + D[2] - D[0])), sqrtf(2.5f * fabsf(D[2] + D[0] - D[1])),
                       sqrtf(2.5f * fabsf(D[0] + D[1] - D[2]))};
        /* SORT (shape.f90:2-29): selection, largest first */
        for (int i = 0; i < 2; i++) {
            int kk = i; float pp = ax[i];
            for (int j = i + 1; j < 3; j++) if (ax[j] >= pp) { kk = j; pp = ax[j]; }
            if (kk != i) { ax[kk] = ax[i]; ax[i] = pp; }
        }
        EPS[id - 1] = 1.0f - ax[2] / ax[0];
        float volm = (((float)ncellv[id - 1] * DXX) * DXX) * DXX;
        float vell = (float)(((((double)(4.0f / 3.0f) * PI_D) * (double)ax[0]) * (double)ax[1]) * (double)ax[2]);
        IPR[id - 1] = volm / vell;
        if (IPR[id - 1] > 1.0f) { IPR[id - 1] = -99.0f; EPS[id - 1] = -99.0f; }
    }
    free(inert);
    /* VOLNEW from NCELLV, then the RMIN filter (1834-1859) */
    double rmin3 = (double)(4.0f / 3.0f) * PI_D * (double)((p->rmin * p->rmin) * p->rmin);
    for (int32_t i = 0; i < nvoid; i++) VOLNEW[i] = 0.0f;
    for (int32_t I = 0; I < nvoid; I++) {
        int32_t id = indice[I];
        if (UV(id) >= 0) continue;
        VOLNEW[id - 1] = (float)((double)ncellv[id - 1] * (double)((DXX * DXX) * DXX));
    }
    for (int32_t I = 0; I < nvoid; I++) {
        int32_t id = indice[I];
        if (UV(id) >= 0) continue;
        if ((double)VOLNEW[id - 1] < rmin3) KILL(id);
    }
    KILL_APPLY();
    for (int32_t I = 0; I < nvoid; I++) {
        int32_t id = indice[I];
        if (UV(id) >= 0) continue;
        XC[id - 1] = R1 + (float)(ICX[id - 1] - 1) * DXX;
        YC[id - 1] = R1 + (float)(ICY[id - 1] - 1) * DXX;
        ZC[id - 1] = R1 + (float)(ICZ[id - 1] - 1) * DXX;
    }
    if (stop_after == 2) return 0;

    /* SIMPLE_CONNECTION_FAST (voidfinding.f90:1007-1161) */
    {
        int8_t* aux = (int8_t*)calloc(ne3, 1);
        for (size_t c = 0; c < ne3; c++) if (marca[c] > 0) aux[c] = 1;
        int32_t* qu = (int32_t*)malloc(sizeof(int32_t) * 3 * ne3);
        size_t head = 0, tail = 0;
        qu[0] = L1; qu[1] = L1; qu[2] = L1; tail = 1;
        aux[EXT(g, L1, L1, L1)] = 2;
        while (head < tail) {
            int32_t ix = qu[3 * head], jy = qu[3 * head + 1], kz = qu[3 * head + 2];
            head++;
            int32_t nb[6][3] = {{ix - 1, jy, kz}, {ix + 1, jy, kz}, {ix, jy - 1, kz},
                                {ix, jy + 1, kz}, {ix, jy, kz - 1}, {ix, jy, kz + 1}};
            for (int q = 0; q < 6; q++) {
                int32_t a = nb[q][0], b = nb[q][1], c = nb[q][2];
                if (a < L1 || a > L2 || b < L1 || b > L2 || c < L1 || c > L2) continue;
                size_t o = EXT(g, a, b, c);
                if (aux[o] == 0) {
                    aux[o] = 2;
                    qu[3 * tail] = a; qu[3 * tail + 1] = b; qu[3 * tail + 2] = c; tail++;
                }
            }
        }
        free(qu);
        for (int32_t kz = L1; kz <= L2; kz++)
            for (int32_t jy = L1; jy <= L2; jy++)
                for (int32_t ix = L1; ix <= L2; ix++) {
                    if (aux[EXT(g, ix, jy, kz)] > 0) continue;
                    int32_t I = ix;
                    while (I <= L2 && aux[EXT(g, I, jy, kz)] == 0) I++;
                    marca[EXT(g, ix, jy, kz)] = (I <= L2) ? marca[EXT(g, I, jy, kz)] : 0;
                }
        free(aux);
    }
    if (stop_after == 3) return 0;

    /* VOID_PERIODIC (voidfinding.f90:1367-1568) */
    if (g->periodic) {
        int32_t* lind = (int32_t*)malloc(sizeof(int32_t) * (2 * (size_t)nvoid + 1));
        int32_t nloc = 0;
        /* main voids with no cell in the inner box are removed; a removal
         * touches only the removed void's cells (all outside the inner box),
         * so presence is taken in one pass beforehand */
        const int32_t ne = g->next;
        uint8_t* present = (uint8_t*)calloc(nvoid > 0 ? nvoid : 1, 1);
        for (int32_t kz = 1; kz <= n; kz++)
            for (int32_t jy = 1; jy <= n; jy++)
                for (int32_t ix = 1; ix <= n; ix++) {
                    int32_t m = marca[EXT(g, ix, jy, kz)];
                    if (m > 0) present[m - 1] = 1;
                }
        for (int32_t I = 0; I < nvoid; I++) {
            int32_t id = indice[I];
            if (UV(id) >= 0) continue;
            lind[nloc++] = id;
            if (!present[id - 1]) {
                VOLNEW[id - 1] = 0.0f;
                UV(id) = 0;
                dead[id - 1] = 1;
            } else {
                lind[nloc++] = id;
            }
        }
        KILL_APPLY();
        free(present);
        /* the cells of each label (one counting sort of the ext grid); when
         * CHANGE_BELONG moves label IND to IND2, IND's member labels are
         * appended to IND2's, so a count visits exactly the cells carrying
         * the label at that moment (the counts do not depend on the order) */
        int32_t* cst = (int32_t*)calloc((size_t)nvoid + 2, sizeof(int32_t));
        for (size_t c = 0; c < ne3; c++) if (marca[c] > 0) cst[marca[c] + 1]++;
        for (int32_t i = 1; i <= nvoid + 1; i++) cst[i] += cst[i - 1];
        int32_t* cur = (int32_t*)malloc(sizeof(int32_t) * ((size_t)nvoid + 2));
        memcpy(cur, cst, sizeof(int32_t) * ((size_t)nvoid + 2));
        uint32_t* clist = (uint32_t*)malloc(sizeof(uint32_t) * ((size_t)cst[nvoid + 1] + 1));
        for (size_t c = 0; c < ne3; c++) if (marca[c] > 0) clist[cur[marca[c]]++] = (uint32_t)c;
        free(cur);
        int32_t* mhead = (int32_t*)malloc(sizeof(int32_t) * ((size_t)nvoid + 1));
        int32_t* mnext = (int32_t*)malloc(sizeof(int32_t) * ((size_t)nvoid + 1));
        int32_t* mtail = (int32_t*)malloc(sizeof(int32_t) * ((size_t)nvoid + 1));
        for (int32_t i = 0; i <= nvoid; i++) { mhead[i] = i; mnext[i] = 0; mtail[i] = i; }
        float half = g->L0 / 2.0f;
        float* lc = (float*)malloc(sizeof(float) * 3 * (size_t)(nloc > 0 ? nloc : 1));
        for (int32_t j = 0; j < nloc; j++) {
            lc[3 * j] = GXC[lind[j] - 1]; lc[3 * j + 1] = GYC[lind[j] - 1]; lc[3 * j + 2] = GZC[lind[j] - 1];
        }
        for (int32_t j = 0; j < nloc; j++) {
            int32_t id = lind[j];
            float xc1 = lc[3 * j], yc1 = lc[3 * j + 1], zc1 = lc[3 * j + 2];
            int32_t sx = 0, sy = 0, sz = 0;
            if (xc1 < -half) sx = 1; if (xc1 > half) sx = -1;
            if (yc1 < -half) sy = 1; if (yc1 > half) sy = -1;
            if (zc1 < -half) sz = 1; if (zc1 > half) sz = -1;
            if (sx == 0 && sy == 0 && sz == 0) continue;
            int32_t igx = (int32_t)((GXC[id - 1] - R1) / DXX) + 1;
            int32_t igy = (int32_t)((GYC[id - 1] - R1) / DXX) + 1;
            int32_t igz = (int32_t)((GZC[id - 1] - R1) / DXX) + 1;
            int32_t ax = igx + sx * n, ay = igy + sy * n, az = igz + sz * n;
            if (ax < L1 || ax > L2 || ay < L1 || ay > L2 || az < L1 || az > L2) continue;  /* reference: out of bounds */
            int32_t id2 = marca[EXT(g, ax, ay, az)];
            if (id2 == 0) continue;
            int64_t cnt = 0, cnt2 = 0;
            for (int32_t m = mhead[id]; m > 0; m = mnext[m])
                for (int32_t q = cst[m]; q < cst[m + 1]; q++) {
                    size_t c = clist[q];
                    if (marca[c] != id) continue;
                    int32_t ix = (int32_t)(c % (size_t)ne) + L1;
                    int32_t jy = (int32_t)((c / (size_t)ne) % (size_t)ne) + L1;
                    int32_t kz = (int32_t)(c / ((size_t)ne * ne)) + L1;
                    int32_t a2 = ix + sx * n, b2 = jy + sy * n, c2 = kz + sz * n;
                    if (a2 >= L1 && a2 <= L2 && b2 >= L1 && b2 <= L2 && c2 >= L1 && c2 <= L2) {
                        cnt2++;
                        if (marca[EXT(g, a2, b2, c2)] == id2) cnt++;
                    }
                }
            float inter = (cnt2 == 0) ? 0.0f : (float)cnt / (float)cnt2;
            if (inter > 0.4f) {  /* CHANGE_BELONG(IND2, IND) */
                UV(id) = id2;
                for (int32_t v = 0; v < nvoid; v++) if (uvoid[v] == id) uvoid[v] = id2;
                for (int32_t m = mhead[id]; m > 0; m = mnext[m])
                    for (int32_t q = cst[m]; q < cst[m + 1]; q++)
                        if (marca[clist[q]] == id) marca[clist[q]] = id2;
                if (mhead[id] > 0) {
                    if (mhead[id2] > 0) { mnext[mtail[id2]] = mhead[id]; mtail[id2] = mtail[id]; }
                    else { mhead[id2] = mhead[id]; mtail[id2] = mtail[id]; }
                    mhead[id] = 0;
                }
                VOLNEW[id - 1] = 0.0f;
            }
        }
        free(lind); free(lc); free(cst); free(clist); free(mhead); free(mnext); free(mtail);
    }

    /* final properties over the inner box (1918-1973) */
    for (int32_t i = 0; i < nvoid; i++) { ncellv[i] = 0; UMEAN[i] = 0.0f; MTOT[i] = 0.0f; }
    double volt = 0.0;
    float rete3 = (rete * rete) * rete;
    int32_t ncells_void = 0;
    for (int32_t kz = 1; kz <= n; kz++)
        for (int32_t jy = 1; jy <= n; jy++)
            for (int32_t ix = 1; ix <= n; ix++) {
                size_t c = EXT(g, ix, jy, kz);
                int32_t id = marca[c];
                if (id > 0 && UV(id) == -1) {
                    ncellv[id - 1]++;
                    UMEAN[id - 1] = UMEAN[id - 1] + u1co[c];
                    MTOT[id - 1] = MTOT[id - 1] + ((((u1co[c] * meandens) * DXX) * DXX) * DXX) * rete3;
                    volt = volt + (double)((DXX * DXX) * DXX);
                    ncells_void++;
                }
            }
    for (int32_t i = 0; i < nvoid; i++) if (ncellv[i] > 0) UMEAN[i] = UMEAN[i] / (float)ncellv[i];
    for (int32_t i = 0; i < nvoid; i++) VOLNEW[i] = 0.0f;
    for (int32_t I = 0; I < nvoid; I++) {
        int32_t id = indice[I];
        if (UV(id) >= 0) continue;
        VOLNEW[id - 1] = (float)((double)ncellv[id - 1] * (double)((DXX * DXX) * DXX));
    }
    for (int32_t I = 0; I < nvoid; I++) {
        int32_t id = indice[I];
        if (UV(id) >= 0) continue;
        if ((double)VOLNEW[id - 1] < rmin3) KILL(id);
    }
    KILL_APPLY();
    free(dead);
    int32_t nvoidt = 0;
    for (int32_t i = 0; i < nvoid; i++) if (uvoid[i] == -1) nvoidt++;
    for (int32_t I = 0; I < nvoid; I++) {   /* catalogue radius (2074) */
        int32_t id = indice[I];
        if (UV(id) != -1) continue;
        double volm = (double)VOLNEW[id - 1];
        REQ[id - 1] = (float)pow((3.0 * volm) / (4.0 * PI_D), (double)(1.0f / 3.0f));
    }
    if (stats_i) { stats_i[1] = nvoidt; stats_i[2] = ncells_void; }
    if (stats_d) {
        float l03 = (g->L0 * g->L0) * g->L0;
        stats_d[0] = volt; stats_d[1] = volt / (double)l03;
    }
#undef KILL
#undef UV
    return 0;
}

/* ------------------------------------------------------------------------
 * MASCLET AMR gas onto a level's uniform grid (FLAG_DATA = 0):
 *   READ_MASCLET adds 1 to the density contrasts (reader.f90:392-398);
 *   level IR < 0: SMOOTH_2 (amr2uniform.f90:62-121), IR = 0: the base grid
 *   (avism.f90:1077-1084), then DIVER_UNIFORM (avism.f90:1129-1146);
 *   IR = 1: DIVER_FINA_GAS (divergence.f90:55-337) and VMESH
 *   (amr2uniform.f90:130-248), U1GCO = UGR, DIVERGCO = DIVR (avism.f90:1100-1119).
 * ------------------------------------------------------------------------ */
int32_t orc_amr_gas(int32_t nb, float L0, int32_t periodic, int32_t ir, const float* base, int32_t npatch,
                    const int32_t* pint, const float* preal, const float* pdata, float* u1gco, float* divgco) {
    if (ir < -3 || ir > 1) return -1;
    const size_t nb3 = (size_t)nb * nb * nb;
    const int32_t n = ir >= 0 ? nb << ir : nb >> (-ir);
    const size_t n3 = (size_t)n * n * n;
    orc_grid* g0 = orc_grid_new(nb, L0, 0);
    orc_grid* gl = orc_grid_new(n, L0, periodic);
    const float dx0 = g0->dx, dxl = gl->dx, rx1 = orc_grid_rad(gl, 1);
    float* U1G = malloc(nb3 * sizeof(float));
    for (size_t c = 0; c < nb3; c++) U1G[c] = base[c] + 1.0f;
    const float *U2G = base + nb3, *U3G = base + 2 * nb3, *U4G = base + 3 * nb3;
#define B0(a, i, j, k) (a)[(size_t)((i) - 1) + (size_t)nb * (((j) - 1) + (size_t)nb * ((k) - 1))]
    int32_t rc = 0;
    if (ir <= 0) {
        float* u = calloc(4 * n3, sizeof(float));
        if (ir == 0) {
            memcpy(u, U1G, n3 * sizeof(float));
            memcpy(u + n3, U2G, 3 * n3 * sizeof(float));
        } else {  /* SMOOTH_2: base cells summed into the coarse cell holding them, in loop order */
            const float wei = ((dx0 * dx0) * dx0) / ((dxl * dxl) * dxl);
            const float* src[4] = {U1G, U2G, U3G, U4G};
            for (int32_t kk = 1; kk <= nb; kk++)
                for (int32_t jj = 1; jj <= nb; jj++)
                    for (int32_t ii = 1; ii <= nb; ii++) {
                        const int32_t I = (int32_t)(((orc_grid_rad(g0, ii) - rx1) / dxl) + 0.5f) + 1;
                        const int32_t J = (int32_t)(((orc_grid_rad(g0, jj) - rx1) / dxl) + 0.5f) + 1;
                        const int32_t K = (int32_t)(((orc_grid_rad(g0, kk) - rx1) / dxl) + 0.5f) + 1;
                        const size_t c = (size_t)(I - 1) + (size_t)n * ((J - 1) + (size_t)n * (K - 1));
                        for (int a = 0; a < 4; a++) u[a * n3 + c] = u[a * n3 + c] + B0(src[a], ii, jj, kk) * wei;
                    }
        }
        memcpy(u1gco, u, n3 * sizeof(float));
        orc_diver_uniform(n, dxl, u + n3, u + 2 * n3, u + 3 * n3, periodic, divgco);
        free(u);
    } else {
        /* DIVER0 on the base grid: interior differences over 2*DX0, edges copied
         * from the neighbouring plane in x, then y, then z (= the interior cell
         * at the clamped index) */
        float* div0 = malloc(nb3 * sizeof(float));
        for (int32_t k = 1; k <= nb; k++)
            for (int32_t j = 1; j <= nb; j++)
                for (int32_t i = 1; i <= nb; i++) {
                    const int32_t a = i < 2 ? 2 : (i > nb - 1 ? nb - 1 : i);
                    const int32_t b = j < 2 ? 2 : (j > nb - 1 ? nb - 1 : j);
                    const int32_t c = k < 2 ? 2 : (k > nb - 1 ? nb - 1 : k);
                    float b21 = B0(U2G, a + 1, b, c) - B0(U2G, a - 1, b, c);
                    b21 = b21 / (2.0f * dx0);
                    float b32 = B0(U3G, a, b + 1, c) - B0(U3G, a, b - 1, c);
                    b32 = b32 / (2.0f * dx0);
                    float b43 = B0(U4G, a, b, c + 1) - B0(U4G, a, b, c - 1);
                    b43 = b43 / (2.0f * dx0);
                    B0(div0, i, j, k) = (b21 + b32) + b43;
                }
        float* ugr = calloc(n3, sizeof(float));
        float* divr = calloc(n3, sizeof(float));
        int32_t* flagamr = malloc(n3 * sizeof(int32_t));
        uint8_t* flagr = calloc(nb3, 1);
        for (size_t c = 0; c < n3; c++) flagamr[c] = 1;
        const float dxr = dx0 / 2.0f;
        size_t off = 0;
        for (int32_t p = 0; p < npatch && rc == 0; p++) {
            const int32_t n1 = pint[6 * p], n2 = pint[6 * p + 1], n3p = pint[6 * p + 2];
            const int32_t L1 = pint[6 * p + 3], L2 = pint[6 * p + 4], L3 = pint[6 * p + 5];
            const size_t pc = (size_t)n1 * n2 * n3p;
            const float *P1 = pdata + off, *P2 = P1 + pc, *P3 = P2 + pc, *P4 = P3 + pc;
            off += 4 * pc;
#define PA(a, i, j, k) (a)[(size_t)((i) - 1) + (size_t)n1 * (((j) - 1) + (size_t)n2 * ((k) - 1))]
            const int32_t ixr1 = (int32_t)((((preal[3 * p] - 0.5f * dxl) - rx1) / dxl) + 0.5f);
            const int32_t jyr1 = (int32_t)((((preal[3 * p + 1] - 0.5f * dxl) - rx1) / dxl) + 0.5f);
            const int32_t kzr1 = (int32_t)((((preal[3 * p + 2] - 0.5f * dxl) - rx1) / dxl) + 0.5f);
            for (int32_t kz = 1; kz <= n3p && rc == 0; kz++)
                for (int32_t jy = 1; jy <= n2 && rc == 0; jy++)
                    for (int32_t ix = 1; ix <= n1; ix++) {
                        /* DIVER of the patch cell (divergence.f90:182-223, 228-334) */
                        const int border = ix == 1 || ix == n1 || jy == 1 || jy == n2 || kz == 1 || kz == n3p;
                        const int near = ix == 2 || ix == n1 - 1 || jy == 2 || jy == n2 - 1 || kz == 2 || kz == n3p - 1;
                        float dv = 0.0f, aux = 0.0f;
                        if (!border) {
                            float b21 = PA(P2, ix + 1, jy, kz) - PA(P2, ix - 1, jy, kz);
                            b21 = b21 / (2.0f * dxr);
                            float b32 = PA(P3, ix, jy + 1, kz) - PA(P3, ix, jy - 1, kz);
                            b32 = b32 / (2.0f * dxr);
                            float b43 = PA(P4, ix, jy, kz + 1) - PA(P4, ix, jy, kz - 1);
                            b43 = b43 / (2.0f * dxr);
                            dv = (b21 + b32) + b43;
                        }
                        if (border || near) {
                            const int32_t c1 = (ix + 1) / 2 + L1 - 1, c2 = (jy + 1) / 2 + L2 - 1, c3 = (kz + 1) / 2 + L3 - 1;
                            if (c1 < 2 || c1 > nb - 1 || c2 < 2 || c2 > nb - 1 || c3 < 2 || c3 > nb - 1) {
                                rc = -2;
                                break;
                            }
                            float b21 = B0(U2G, c1 + 1, c2, c3) - B0(U2G, c1 - 1, c2, c3);
                            b21 = b21 / (2.0f * dx0);
                            float b32 = B0(U3G, c1, c2 + 1, c3) - B0(U3G, c1, c2 - 1, c3);
                            b32 = b32 / (2.0f * dx0);
                            float b43 = B0(U4G, c1, c2, c3 + 1) - B0(U4G, c1, c2, c3 - 1);
                            b43 = b43 / (2.0f * dx0);
                            aux = (b21 + b32) + b43;
                        }
                        if (border) dv = aux;
                        else if (near) dv = 0.5f * (aux + dv);
                        /* VMESH: sums over the patches, the cell counted from 1 */
                        const int32_t X = ixr1 + ix, Y = jyr1 + jy, Z = kzr1 + kz;
                        if (X < 1 || X > n || Y < 1 || Y > n || Z < 1 || Z > n) {
                            rc = -3;
                            break;
                        }
                        const size_t c = (size_t)(X - 1) + (size_t)n * ((Y - 1) + (size_t)n * (Z - 1));
                        flagamr[c] += 1;
                        ugr[c] = ugr[c] + (PA(P1, ix, jy, kz) + 1.0f);
                        divr[c] = divr[c] + dv;
                        const int32_t x0 = (int32_t)((float)(X - 1) / 2.0f) + 1, y0 = (int32_t)((float)(Y - 1) / 2.0f) + 1,
                                      z0 = (int32_t)((float)(Z - 1) / 2.0f) + 1;
                        B0(flagr, x0, y0, z0) = 1;
                    }
#undef PA
        }
        if (rc == 0) {
            for (size_t c = 0; c < n3; c++) {
                ugr[c] = ugr[c] / (float)flagamr[c];
                divr[c] = divr[c] / (float)flagamr[c];
            }
            for (int32_t k = 1; k <= n; k++)
                for (int32_t j = 1; j <= n; j++)
                    for (int32_t i = 1; i <= n; i++) {
                        const int32_t a = (i - 1) / 2 + 1, b = (j - 1) / 2 + 1, c0 = (k - 1) / 2 + 1;
                        if (B0(flagr, a, b, c0)) continue;
                        const size_t c = (size_t)(i - 1) + (size_t)n * ((j - 1) + (size_t)n * (k - 1));
                        ugr[c] = B0(U1G, a, b, c0);
                        divr[c] = B0(div0, a, b, c0);
                    }
            memcpy(u1gco, ugr, n3 * sizeof(float));
            memcpy(divgco, divr, n3 * sizeof(float));
        }
        free(div0); free(ugr); free(divr); free(flagamr); free(flagr);
    }
#undef B0
    free(U1G);
    orc_grid_free(g0);
    orc_grid_free(gl);
    return rc;
}
