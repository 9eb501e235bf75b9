/*
 * tdvref — CPU oracle for the stereo-matching hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP implementation. Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
 *
 * Written from the algorithm descriptions of the reference files cited per
 * function; conventions are in tdvref.h.
 */
#include "tdvref.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static inline int clampi(int v, int lo, int hi)
{
    return v < lo ? lo : (v > hi ? hi : v);
}
static inline int min2(int a, int b) { return a < b ? a : b; }
static inline int max2(int a, int b) { return a > b ? a : b; }
static inline int min3(int a, int b, int c) { return min2(a, min2(b, c)); }
static inline int max3(int a, int b, int c) { return max2(a, max2(b, c)); }

/* ------------------------------------------------------------------------ */
/* Birchfield-Tomasi (birchfieldcost.cu:14-59)                               */
/* ------------------------------------------------------------------------ */

/* 2x-unit sampling-insensitive terms of pixel x of a row: the value itself
 * (doubled) and the min/max over it and the two half-way interpolants. */
static void bt_terms(const uint8_t *row, int W, int x, int *v2, int *vmin,
                     int *vmax)
{
    const int c = row[x];
    const int a = c + row[clampi(x - 1, 0, W - 1)];
    const int b = c + row[clampi(x + 1, 0, W - 1)];
    *v2 = 2 * c;
    *vmin = min3(2 * c, a, b);
    *vmax = max3(2 * c, a, b);
}

void tdvref_cost_bt2(const uint8_t *L, const uint8_t *R, int W, int H, int D,
                     uint16_t *C)
{
    for (int y = 0; y < H; y++) {
        const uint8_t *lrow = L + (size_t)y * W;
        const uint8_t *rrow = R + (size_t)y * W;
        for (int x = 0; x < W; x++) {
            int l2, lmin, lmax;
            bt_terms(lrow, W, x, &l2, &lmin, &lmax);
            uint16_t *out = C + ((size_t)y * W + x) * D;
            for (int d = 0; d < D; d++) {
                if (x - d < 0) {
                    out[d] = TDVREF_INV16;
                    continue;
                }
                int r2, rmin, rmax;
                bt_terms(rrow, W, x - d, &r2, &rmin, &rmax);
                const int a = max3(0, l2 - rmax, rmin - l2);
                const int b = max3(0, r2 - lmax, lmin - r2);
                out[d] = (uint16_t)min2(a, b);
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* SSD (ssd.cu:15-50)                                                        */
/* ------------------------------------------------------------------------ */

void tdvref_cost_ssd(const uint8_t *L, const uint8_t *R, int W, int H, int D,
                     int radius, uint32_t *C)
{
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            uint32_t *out = C + ((size_t)y * W + x) * D;
            for (int d = 0; d < D; d++) {
                if (x - d < 0) {
                    out[d] = TDVREF_INV32;
                    continue;
                }
                uint32_t s = 0;
                for (int dy = -radius; dy <= radius; dy++) {
                    const int yy = clampi(y + dy, 0, H - 1);
                    for (int dx = -radius; dx <= radius; dx++) {
                        const int lv = L[(size_t)yy * W + clampi(x + dx, 0, W - 1)];
                        const int rv = R[(size_t)yy * W + clampi(x + dx - d, 0, W - 1)];
                        s += (uint32_t)((lv - rv) * (lv - rv));
                    }
                }
                out[d] = s;
            }
        }
}

/* Box sum of a padded (W+2r) x (H+2r) field via a 2D prefix table.
 * P has (PW+1)*(PH+1) entries, P[0][*] = P[*][0] = 0. */
static void prefix2d(const int64_t *F, int PW, int PH, int64_t *P)
{
    memset(P, 0, sizeof(int64_t) * (size_t)(PW + 1));
    for (int v = 0; v < PH; v++) {
        int64_t run = 0;
        int64_t *prow = P + (size_t)(v + 1) * (PW + 1);
        const int64_t *pprev = P + (size_t)v * (PW + 1);
        prow[0] = 0;
        for (int u = 0; u < PW; u++) {
            run += F[(size_t)v * PW + u];
            prow[u + 1] = pprev[u + 1] + run;
        }
    }
}

static inline int64_t boxsum(const int64_t *P, int PW, int u0, int v0, int n)
{
    const size_t s = (size_t)(PW + 1);
    return P[(size_t)(v0 + n) * s + (u0 + n)] - P[(size_t)v0 * s + (u0 + n)] -
           P[(size_t)(v0 + n) * s + u0] + P[(size_t)v0 * s + u0];
}

void tdvref_cost_ssd_box(const uint8_t *L, const uint8_t *R, int W, int H,
                         int D, int radius, uint32_t *C)
{
    const int n = 2 * radius + 1;
    const int PW = W + 2 * radius, PH = H + 2 * radius;
    int64_t *F = (int64_t *)malloc(sizeof(int64_t) * (size_t)PW * PH);
    int64_t *P = (int64_t *)malloc(sizeof(int64_t) * (size_t)(PW + 1) * (PH + 1));
    for (int d = 0; d < D; d++) {
        /* padded coordinate (u, v) <-> image coordinate (u - r, v - r) */
        for (int v = 0; v < PH; v++) {
            const int yy = clampi(v - radius, 0, H - 1);
            for (int u = 0; u < PW; u++) {
                const int xi = u - radius;
                const int lv = L[(size_t)yy * W + clampi(xi, 0, W - 1)];
                const int rv = R[(size_t)yy * W + clampi(xi - d, 0, W - 1)];
                F[(size_t)v * PW + u] = (int64_t)(lv - rv) * (lv - rv);
            }
        }
        prefix2d(F, PW, PH, P);
        for (int y = 0; y < H; y++)
            for (int x = 0; x < W; x++) {
                uint32_t *out = C + ((size_t)y * W + x) * D + d;
                *out = (x - d < 0) ? TDVREF_INV32
                                   : (uint32_t)boxsum(P, PW, x, y, n);
            }
    }
    free(F);
    free(P);
}

/* ------------------------------------------------------------------------ */
/* NCC (crosscorrelation.cu:11-49), FP64 from exact integer window sums      */
/* ------------------------------------------------------------------------ */

void tdvref_cost_ncc(const uint8_t *L, const uint8_t *R, int W, int H, int D,
                     int radius, int zero_mean, double *C, double *rho_out)
{
    const int n = 2 * radius + 1;
    const int64_t N = (int64_t)n * n;
    const int PW = W + 2 * radius, PH = H + 2 * radius;
    const size_t PS = (size_t)(PW + 1) * (PH + 1);
    int64_t *F = (int64_t *)malloc(sizeof(int64_t) * (size_t)PW * PH);
    int64_t *PL = (int64_t *)malloc(sizeof(int64_t) * PS);
    int64_t *PL2 = (int64_t *)malloc(sizeof(int64_t) * PS);
    int64_t *PR = (int64_t *)malloc(sizeof(int64_t) * PS);
    int64_t *PR2 = (int64_t *)malloc(sizeof(int64_t) * PS);
    int64_t *PLR = (int64_t *)malloc(sizeof(int64_t) * PS);

    /* window sums of L, L^2, R, R^2 (right window centred on x - d) */
    for (int pass = 0; pass < 4; pass++) {
        const uint8_t *img = (pass < 2) ? L : R;
        const int sq = pass & 1;
        for (int v = 0; v < PH; v++) {
            const int yy = clampi(v - radius, 0, H - 1);
            for (int u = 0; u < PW; u++) {
                const int64_t iv = img[(size_t)yy * W + clampi(u - radius, 0, W - 1)];
                F[(size_t)v * PW + u] = sq ? iv * iv : iv;
            }
        }
        prefix2d(F, PW, PH, pass == 0 ? PL : pass == 1 ? PL2 : pass == 2 ? PR : PR2);
    }

    for (int d = 0; d < D; d++) {
        for (int v = 0; v < PH; v++) {
            const int yy = clampi(v - radius, 0, H - 1);
            for (int u = 0; u < PW; u++) {
                const int xi = u - radius;
                const int64_t lv = L[(size_t)yy * W + clampi(xi, 0, W - 1)];
                const int64_t rv = R[(size_t)yy * W + clampi(xi - d, 0, W - 1)];
                F[(size_t)v * PW + u] = lv * rv;
            }
        }
        prefix2d(F, PW, PH, PLR);
        for (int y = 0; y < H; y++)
            for (int x = 0; x < W; x++) {
                const size_t o = ((size_t)y * W + x) * D + d;
                if (x - d < 0) {
                    C[o] = INFINITY;
                    if (rho_out) rho_out[o] = NAN;
                    continue;
                }
                const int64_t slr = boxsum(PLR, PW, x, y, n);
                const int64_t sl = boxsum(PL, PW, x, y, n);
                const int64_t sl2 = boxsum(PL2, PW, x, y, n);
                const int64_t sr = boxsum(PR, PW, x - d, y, n);
                const int64_t sr2 = boxsum(PR2, PW, x - d, y, n);
                int64_t num, vl, vr;
                if (zero_mean) {
                    num = N * slr - sl * sr;
                    vl = N * sl2 - sl * sl;
                    vr = N * sr2 - sr * sr;
                } else {
                    num = slr;
                    vl = sl2;
                    vr = sr2;
                }
                if (vl == 0 || vr == 0) {
                    C[o] = INFINITY;
                    if (rho_out) rho_out[o] = NAN;
                    continue;
                }
                const double rho = (double)num / sqrt((double)vl * (double)vr);
                C[o] = 1.0 - rho;
                if (rho_out) rho_out[o] = rho;
            }
    }
    free(F);
    free(PL);
    free(PL2);
    free(PR);
    free(PR2);
    free(PLR);
}

/* ------------------------------------------------------------------------ */
/* SGM, standard geometry (semiglobalcpu.cpp:56-126 recurrence)              */
/* ------------------------------------------------------------------------ */

const int tdvref_dir_dx[8] = {1, -1, 0, 0, 1, -1, -1, 1};
const int tdvref_dir_dy[8] = {0, 0, 1, -1, 1, -1, 1, -1};

void tdvref_sgm_lr(const uint16_t *C, int W, int H, int D, int P1, int P2,
                   int dir, uint16_t *Lr)
{
    const int dx = tdvref_dir_dx[dir], dy = tdvref_dir_dy[dir];
    /* minimum over d of Lr at each pixel, for the predecessor lookups */
    uint32_t *mins = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)W * H);
    /* Visit order: the predecessor p - r is always visited before p. */
    const int y0 = dy >= 0 ? 0 : H - 1, ystep = dy >= 0 ? 1 : -1;
    const int x0 = dx >= 0 ? 0 : W - 1, xstep = dx >= 0 ? 1 : -1;
    for (int yi = 0, y = y0; yi < H; yi++, y += ystep)
        for (int xi = 0, x = x0; xi < W; xi++, x += xstep) {
            const size_t p = (size_t)y * W + x;
            const uint16_t *c = C + p * D;
            uint16_t *l = Lr + p * D;
            const int qx = x - dx, qy = y - dy;
            uint32_t mn = 0xFFFFFFFFu;
            /* path start: L = C (semiglobalcpu.cpp:71-84); a predecessor
             * with no valid cell restarts the path the same way */
            if (qx < 0 || qx >= W || qy < 0 || qy >= H ||
                mins[(size_t)qy * W + qx] == 0xFFFFFFFFu) {
                for (int d = 0; d < D; d++) {
                    l[d] = c[d];
                    if (c[d] != TDVREF_INV16 && c[d] < mn) mn = c[d];
                }
            } else {
                const size_t q = (size_t)qy * W + qx;
                const uint16_t *lp = Lr + q * D;
                const int m = (int)mins[q];
                for (int d = 0; d < D; d++) {
                    if (c[d] == TDVREF_INV16) {
                        l[d] = TDVREF_INV16;
                        continue;
                    }
                    /* invalid predecessors drop out of the minimum */
                    int best = m + P2;
                    if (lp[d] != TDVREF_INV16) best = min2(best, lp[d]);
                    if (d > 0 && lp[d - 1] != TDVREF_INV16)
                        best = min2(best, lp[d - 1] + P1);
                    if (d + 1 < D && lp[d + 1] != TDVREF_INV16)
                        best = min2(best, lp[d + 1] + P1);
                    const int v = c[d] + best - m;
                    l[d] = (uint16_t)v;
                    if ((uint32_t)v < mn) mn = (uint32_t)v;
                }
            }
            mins[p] = mn;
        }
    free(mins);
}

void tdvref_sgm_std(const uint16_t *C, int W, int H, int D, int P1, int P2,
                    int dir_mask, uint16_t *S)
{
    const size_t n = (size_t)W * H * D;
    uint16_t *Lr = (uint16_t *)malloc(sizeof(uint16_t) * n);
    for (size_t i = 0; i < n; i++) S[i] = C[i] == TDVREF_INV16 ? TDVREF_INV16 : 0;
    for (int dir = 0; dir < 8; dir++) {
        if (!(dir_mask & (1 << dir))) continue;
        tdvref_sgm_lr(C, W, H, D, P1, P2, dir, Lr);
        for (size_t i = 0; i < n; i++)
            if (S[i] != TDVREF_INV16) S[i] = (uint16_t)(S[i] + Lr[i]);
    }
    free(Lr);
}

/* ------------------------------------------------------------------------ */
/* SGM over uint32 and float volumes (semiglobalcpu.cpp:87-125 recurrence,   */
/* any DSIMem as semiglobaldev.cpp:16-28 accepts)                            */
/* ------------------------------------------------------------------------ */

static inline uint32_t sat_add32(uint32_t a, uint32_t b)
{
    const uint64_t s = (uint64_t)a + b;
    return s > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)s;
}
static inline uint32_t umin32(uint32_t a, uint32_t b) { return a < b ? a : b; }

/* One direction; visit order as tdvref_sgm_lr. For each pixel: L = C at a
 * path start or after a predecessor with no valid cell; otherwise
 *   t = min(L'(d), min(L'(d-1), L'(d+1)) + P1), E = min(t - m, P2),
 *   L = C + E
 * with saturating additions (0xFFFFFFFF is invalid and absorbs). */
static void sgm_u32_dir(const uint32_t *C, int W, int H, int D, uint32_t P1, uint32_t P2,
                        int dir, uint32_t *Lr, uint32_t *mins)
{
    const int dx = tdvref_dir_dx[dir], dy = tdvref_dir_dy[dir];
    const int y0 = dy >= 0 ? 0 : H - 1, ystep = dy >= 0 ? 1 : -1;
    const int x0 = dx >= 0 ? 0 : W - 1, xstep = dx >= 0 ? 1 : -1;
    for (int yi = 0, y = y0; yi < H; yi++, y += ystep)
        for (int xi = 0, x = x0; xi < W; xi++, x += xstep) {
            const size_t p = (size_t)y * W + x;
            const uint32_t *c = C + p * D;
            uint32_t *l = Lr + p * D;
            const int qx = x - dx, qy = y - dy;
            const int start = qx < 0 || qx >= W || qy < 0 || qy >= H ||
                              mins[(size_t)qy * W + qx] == TDVREF_INV32;
            uint32_t mn = TDVREF_INV32;
            if (start) {
                for (int d = 0; d < D; d++) l[d] = c[d];
            } else {
                const size_t q = (size_t)qy * W + qx;
                const uint32_t *lp = Lr + q * D;
                const uint32_t m = mins[q];
                for (int d = 0; d < D; d++) {
                    const uint32_t nm = d > 0 ? lp[d - 1] : TDVREF_INV32;
                    const uint32_t np = d + 1 < D ? lp[d + 1] : TDVREF_INV32;
                    const uint32_t t = umin32(lp[d], sat_add32(umin32(nm, np), P1));
                    l[d] = sat_add32(c[d], umin32(t - m, P2));
                }
            }
            for (int d = 0; d < D; d++) mn = umin32(mn, l[d]);
            mins[p] = mn;
        }
}

void tdvref_sgm_u32(const uint32_t *C, int W, int H, int D, uint32_t P1, uint32_t P2,
                    int dir_mask, uint32_t *S)
{
    const size_t n = (size_t)W * H * D;
    uint32_t *Lr = (uint32_t *)malloc(sizeof(uint32_t) * n);
    uint32_t *mins = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)W * H);
    int first = 1;
    for (int dir = 0; dir < 8; dir++) {
        if (!(dir_mask & (1 << dir))) continue;
        sgm_u32_dir(C, W, H, D, P1, P2, dir, Lr, mins);
        for (size_t i = 0; i < n; i++) S[i] = first ? Lr[i] : sat_add32(S[i], Lr[i]);
        first = 0;
    }
    free(Lr);
    free(mins);
}

/* A float cost as an operand: finite values as they are, +inf otherwise. */
static inline float fcost(float v) { return (v < INFINITY && v > -INFINITY) ? v : INFINITY; }
static inline float fmin2(float a, float b) { return a < b ? a : b; }

/* The float recurrence of sgm_u32_dir with IEEE additions (+inf invalid). */
static void sgm_f32_dir(const float *C, int W, int H, int D, float P1, float P2, int dir,
                        float *Lr, float *mins)
{
    const int dx = tdvref_dir_dx[dir], dy = tdvref_dir_dy[dir];
    const int y0 = dy >= 0 ? 0 : H - 1, ystep = dy >= 0 ? 1 : -1;
    const int x0 = dx >= 0 ? 0 : W - 1, xstep = dx >= 0 ? 1 : -1;
    for (int yi = 0, y = y0; yi < H; yi++, y += ystep)
        for (int xi = 0, x = x0; xi < W; xi++, x += xstep) {
            const size_t p = (size_t)y * W + x;
            const float *c = C + p * D;
            float *l = Lr + p * D;
            const int qx = x - dx, qy = y - dy;
            const int start = qx < 0 || qx >= W || qy < 0 || qy >= H ||
                              !(mins[(size_t)qy * W + qx] < INFINITY);
            float mn = INFINITY;
            if (start) {
                for (int d = 0; d < D; d++) l[d] = fcost(c[d]);
            } else {
                const size_t q = (size_t)qy * W + qx;
                const float *lp = Lr + q * D;
                const float m = mins[q];
                for (int d = 0; d < D; d++) {
                    const float nm = d > 0 ? lp[d - 1] : INFINITY;
                    const float np = d + 1 < D ? lp[d + 1] : INFINITY;
                    const float t = fmin2(lp[d], fmin2(nm, np) + P1);
                    const float e = fmin2(t - m, P2);
                    l[d] = fcost(c[d]) + e;
                }
            }
            for (int d = 0; d < D; d++) mn = fmin2(mn, l[d]);
            mins[p] = mn;
        }
}

void tdvref_sgm_f32(const float *C, int W, int H, int D, float P1, float P2, int dir_mask,
                    float *S)
{
    const size_t n = (size_t)W * H * D;
    float *Lr = (float *)malloc(sizeof(float) * n);
    float *mins = (float *)malloc(sizeof(float) * (size_t)W * H);
    int first = 1;
    for (int dir = 0; dir < 8; dir++) {
        if (!(dir_mask & (1 << dir))) continue;
        sgm_f32_dir(C, W, H, D, P1, P2, dir, Lr, mins);
        for (size_t i = 0; i < n; i++) S[i] = first ? Lr[i] : S[i] + Lr[i];
        first = 0;
    }
    free(Lr);
    free(mins);
}

/* ------------------------------------------------------------------------ */
/* SGM, reference geometry (semiglobal.cu:57-230, semiglobalcpu.cpp:56-230)  */
/* ------------------------------------------------------------------------ */

size_t tdvref_sgm_path_count(int W, int H)
{
    return (size_t)W + H + 2 * ((size_t)W + H - 1);
}

static void set_path(tdvref_sgpath *p, int sx, int sy, int ex, int ey, int dx,
                     int dy, int size)
{
    p->sx = (uint16_t)sx;
    p->sy = (uint16_t)sy;
    p->ex = (uint16_t)ex;
    p->ey = (uint16_t)ey;
    p->dx = (int16_t)dx;
    p->dy = (int16_t)dy;
    p->size = (uint16_t)size;
    p->pad = 0;
}

size_t tdvref_sgm_paths(int W, int H, tdvref_sgpath *paths)
{
    const int lastX = W - 1, lastY = H - 1;
    size_t k = 0;
    /* rows, left to right (semiglobal.cu:116-130) */
    for (int i = 0; i < H; i++) set_path(&paths[k++], 0, i, lastX, i, 1, 0, W);
    /* columns, top to bottom (semiglobal.cu:132-146) */
    for (int i = 0; i < W; i++) set_path(&paths[k++], i, 0, i, lastY, 0, 1, H);
    /* direction (1,1): starts on the top row, then on the left column
     * (semiglobal.cu:154-185). The end points are the table's own, which
     * are off the true diagonal end when W != H. */
    for (int i = 0; i < W; i++) {
        const int ey = min2(i, lastY);
        set_path(&paths[k++], lastX - i, 0, lastX, ey, 1, 1, ey + 1);
    }
    for (int i = 0; i < H - 1; i++) {
        const int ex = min2(lastY - (i + 1), lastX);
        set_path(&paths[k++], 0, i + 1, ex, lastY, 1, 1, ex + 1);
    }
    /* direction (-1,1): top row, then right column (semiglobal.cu:191-230) */
    for (int i = 0; i < W; i++) {
        const int size = min2(i + 1, H);
        set_path(&paths[k++], i, 0, 0, size - 1, -1, 1, size);
    }
    for (int i = 0; i < H - 1; i++) {
        const int sy = i + 1;
        const int size = min2(H - sy, W);
        set_path(&paths[k++], lastX, sy, lastX - size + 1, sy + size - 1, -1, 1,
                 size);
    }
    return k;
}

/* One traversal of costPath (semiglobalcpu.cpp:56-126) in the integer
 * profile: int32 values, TDVREF_INF_I32 for invalid. */
static void ref_traverse(const uint16_t *C, int W, int H, int D, int P1,
                         int P2, int flags, int sx, int sy, int dx, int dy,
                         int size, int32_t *S, int32_t *last, int32_t *next)
{
    (void)H;
    int x = sx, y = sy;
    {
        const size_t p = (size_t)y * W + x;
        for (int d = 0; d < D; d++) {
            const int32_t c = C[p * D + d] == TDVREF_INV16 ? TDVREF_INF_I32
                                                          : (int32_t)C[p * D + d];
            int32_t *s = &S[p * D + d];
            if (flags & TDVREF_SGM_START_ASSIGN)
                *s = c;
            else if (*s != TDVREF_INF_I32)
                *s = (c == TDVREF_INF_I32) ? TDVREF_INF_I32
                                           : (int32_t)((uint32_t)*s + (uint32_t)c);
            last[d] = c;
        }
    }
    const int mlimit = (flags & TDVREF_SGM_MIN_SKIP_LAST) ? D - 1 : D;
    for (int k = 1; k < size; k++) {
        x += dx;
        y += dy;
        int32_t m = TDVREF_INF_I32;
        for (int d = 0; d < mlimit; d++)
            if (last[d] < m) m = last[d];
        const size_t p = (size_t)y * W + x;
        for (int d = 0; d < D; d++) {
            const uint16_t cu = C[p * D + d];
            int32_t *s = &S[p * D + d];
            if (cu == TDVREF_INV16) {
                next[d] = TDVREF_INF_I32;
                *s = TDVREF_INF_I32;
                continue;
            }
            int64_t best = (int64_t)m + P2;
            if (last[d] != TDVREF_INF_I32 && last[d] < best) best = last[d];
            if (d > 0 && last[d - 1] != TDVREF_INF_I32 &&
                (int64_t)last[d - 1] + P1 < best)
                best = (int64_t)last[d - 1] + P1;
            if (d + 1 < D && last[d + 1] != TDVREF_INF_I32 &&
                (int64_t)last[d + 1] + P1 < best)
                best = (int64_t)last[d + 1] + P1;
            /* path costs and sums are int32 with two's-complement wrap (the
             * device's atomic adds wrap; values only leave the int32 range
             * when the minimum skipping the last disparity covers no valid
             * cell) */
            const int32_t v = (int32_t)(uint32_t)((int64_t)cu + best - m);
            next[d] = v;
            if (*s != TDVREF_INF_I32) *s = (int32_t)((uint32_t)*s + (uint32_t)v);
        }
        memcpy(last, next, sizeof(int32_t) * (size_t)D);
    }
}

void tdvref_sgm_ref(const uint16_t *C, int W, int H, int D, int P1, int P2,
                    int flags, int32_t *S)
{
    const size_t np = tdvref_sgm_path_count(W, H);
    tdvref_sgpath *paths = (tdvref_sgpath *)malloc(sizeof(tdvref_sgpath) * np);
    tdvref_sgm_paths(W, H, paths);
    const size_t n = (size_t)W * H * D;
    /* zeroVolume restated as a full clear (its x/y swap is not reproduced) */
    for (size_t i = 0; i < n; i++) S[i] = C[i] == TDVREF_INV16 ? TDVREF_INF_I32 : 0;
    int32_t *last = (int32_t *)malloc(sizeof(int32_t) * (size_t)D);
    int32_t *next = (int32_t *)malloc(sizeof(int32_t) * (size_t)D);
    for (size_t i = 0; i < np; i++) {
        const tdvref_sgpath *p = &paths[i];
        ref_traverse(C, W, H, D, P1, P2, flags, p->sx, p->sy, p->dx, p->dy,
                     p->size, S, last, next);
        ref_traverse(C, W, H, D, P1, P2, flags, p->ex, p->ey, -p->dx, -p->dy,
                     p->size, S, last, next);
    }
    free(last);
    free(next);
    free(paths);
}

static int dir_index(int dx, int dy)
{
    for (int i = 0; i < 8; i++)
        if (tdvref_dir_dx[i] == dx && tdvref_dir_dy[i] == dy) return i;
    return -1;
}

void tdvref_sgm_ref_visits(int W, int H, int32_t *visits)
{
    const size_t np = tdvref_sgm_path_count(W, H);
    tdvref_sgpath *paths = (tdvref_sgpath *)malloc(sizeof(tdvref_sgpath) * np);
    tdvref_sgm_paths(W, H, paths);
    memset(visits, 0, sizeof(int32_t) * 8 * (size_t)W * H);
    for (size_t i = 0; i < np; i++) {
        const tdvref_sgpath *p = &paths[i];
        for (int b = 0; b < 2; b++) {
            const int dx = b ? -p->dx : p->dx, dy = b ? -p->dy : p->dy;
            int x = b ? p->ex : p->sx, y = b ? p->ey : p->sy;
            const int di = dir_index(dx, dy);
            for (int k = 0; k < p->size; k++, x += dx, y += dy)
                visits[(size_t)di * W * H + (size_t)y * W + x]++;
        }
    }
    free(paths);
}

/* ------------------------------------------------------------------------ */
/* WTA (wtacpu.cpp:14-39; tree variant wta.cu:62-78)                         */
/* ------------------------------------------------------------------------ */

#define WTA_BODY(T, INVTEST)                                              \
    for (size_t p = 0; p < (size_t)W * H; p++) {                          \
        const T *s = S + p * D;                                           \
        int best = 0;                                                     \
        int have = 0;                                                     \
        T bv = 0;                                                         \
        for (int d = 0; d < D; d++) {                                     \
            const T v = s[d];                                             \
            if (INVTEST) continue;                                        \
            if (!have || v < bv) {                                        \
                bv = v;                                                   \
                best = d;                                                 \
                have = 1;                                                 \
            }                                                             \
        }                                                                 \
        disp[p] = (int16_t)best;                                          \
    }

void tdvref_wta_u16(const uint16_t *S, int W, int H, int D, int16_t *disp)
{
    WTA_BODY(uint16_t, v == TDVREF_INV16)
}
void tdvref_wta_u32(const uint32_t *S, int W, int H, int D, int16_t *disp)
{
    WTA_BODY(uint32_t, v == TDVREF_INV32)
}
void tdvref_wta_i32(const int32_t *S, int W, int H, int D, int16_t *disp)
{
    WTA_BODY(int32_t, v == TDVREF_INF_I32)
}
void tdvref_wta_f64(const double *S, int W, int H, int D, int16_t *disp)
{
    WTA_BODY(double, !(v < INFINITY))
}
void tdvref_wta_f32(const float *S, int W, int H, int D, int16_t *disp)
{
    WTA_BODY(float, !(v < INFINITY))
}

void tdvref_wta_tree_u16(const uint16_t *S, int W, int H, int D, int16_t *disp)
{
    int *val = (int *)malloc(sizeof(int) * (size_t)D);
    int *idx = (int *)malloc(sizeof(int) * (size_t)D);
    for (size_t p = 0; p < (size_t)W * H; p++) {
        for (int d = 0; d < D; d++) {
            val[d] = S[p * D + d] == TDVREF_INV16 ? 0x7FFFFFFF : S[p * D + d];
            idx[d] = d;
        }
        for (int i = D >> 1; i != 0; i >>= 1)
            for (int z = 0; z < i; z++) {
                /* v1 < v2 keeps the left operand, otherwise the right */
                if (!(val[z] < val[z + i])) {
                    val[z] = val[z + i];
                    idx[z] = idx[z + i];
                }
            }
        disp[p] = (int16_t)idx[0];
    }
    free(val);
    free(idx);
}

/* ------------------------------------------------------------------------ */
/* Scanline DP (dynamicprogcpu.cpp:13-92)                                    */
/* ------------------------------------------------------------------------ */

#define DP_INF UINT64_MAX

static inline uint64_t sat_add64(uint64_t a, uint64_t b)
{
    return (a == DP_INF || b == DP_INF) ? DP_INF : a + b;
}

void tdvref_dp(const uint32_t *C, int W, int H, int D, uint32_t p_occ,
               int flags, int16_t *disp, uint64_t *acc_out)
{
    uint64_t *prev = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)D);
    uint64_t *cur = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)D);
    int8_t *path = (int8_t *)malloc((size_t)W * D);
    for (int y = 0; y < H; y++) {
        const uint32_t *crow = C + (size_t)y * W * D;
        for (int d = 0; d < D; d++) {
            prev[d] = crow[d] == TDVREF_INV32 ? DP_INF : crow[d];
            path[d] = 0;
        }
        for (int x = 1; x < W; x++) {
            for (int d = 0; d < D; d++) {
                const uint32_t c0 = crow[(size_t)x * D + d];
                const uint64_t c1 = d > 0 ? sat_add64(prev[d - 1], p_occ) : DP_INF;
                const uint64_t c2 = prev[d];
                const uint64_t c3 = d < D - 1 ? sat_add64(prev[d + 1], p_occ) : DP_INF;
                uint64_t m;
                int8_t p;
                if (c1 < c2 && c1 < c3) {
                    m = c1;
                    p = -1;
                } else if (c2 < c3) {
                    m = c2;
                    p = 0;
                } else {
                    m = c3;
                    p = 1;
                }
                cur[d] = (c0 == TDVREF_INV32) ? DP_INF : sat_add64(m, c0);
                path[(size_t)x * D + d] = p;
            }
            uint64_t *t = prev;
            prev = cur;
            cur = t;
        }
        if (acc_out) memcpy(acc_out + (size_t)y * D, prev, sizeof(uint64_t) * D);
        /* start: lowest-index argmin of the last column */
        int z = 0;
        if (!(flags & TDVREF_DP_START_ZERO)) {
            uint64_t bv = DP_INF;
            for (int d = 0; d < D; d++)
                if (prev[d] < bv) {
                    bv = prev[d];
                    z = d;
                }
        }
        int16_t *out = disp + (size_t)y * W;
        if (flags & TDVREF_DP_REF_BACKTRACK) {
            for (int x = W - 1; x >= 0; x--) {
                const int nz = z + path[(size_t)x * D + z];
                if (nz >= 0 && nz < D) z = nz;
                out[x] = (int16_t)z;
            }
        } else {
            for (int x = W - 1; x >= 0; x--) {
                out[x] = (int16_t)z;
                const int nz = z + path[(size_t)x * D + z];
                if (x > 0 && nz >= 0 && nz < D) z = nz;
            }
        }
    }
    free(prev);
    free(cur);
    free(path);
}

/* Float DP (the recurrence of tdvref_dp on float costs): +inf invalid,
 * c1 = A(d-1) + P, c3 = A(d+1) + P and A = C + chosen each one IEEE addition,
 * the same tie rule, start and backtrack. */
void tdvref_dp_f32(const float *C, int W, int H, int D, float p_occ, int flags, int16_t *disp)
{
    float *prev = (float *)malloc(sizeof(float) * (size_t)D);
    float *cur = (float *)malloc(sizeof(float) * (size_t)D);
    int8_t *path = (int8_t *)malloc((size_t)W * D);
    for (int y = 0; y < H; y++) {
        const float *crow = C + (size_t)y * W * D;
        for (int d = 0; d < D; d++) {
            prev[d] = fcost(crow[d]);
            path[d] = 0;
        }
        for (int x = 1; x < W; x++) {
            for (int d = 0; d < D; d++) {
                const float c1 = d > 0 ? prev[d - 1] + p_occ : INFINITY;
                const float c2 = prev[d];
                const float c3 = d < D - 1 ? prev[d + 1] + p_occ : INFINITY;
                float m;
                int8_t p;
                if (c1 < c2 && c1 < c3) {
                    m = c1;
                    p = -1;
                } else if (c2 < c3) {
                    m = c2;
                    p = 0;
                } else {
                    m = c3;
                    p = 1;
                }
                cur[d] = fcost(crow[(size_t)x * D + d]) + m;
                path[(size_t)x * D + d] = p;
            }
            float *t = prev;
            prev = cur;
            cur = t;
        }
        int z = 0;
        if (!(flags & TDVREF_DP_START_ZERO)) {
            float bv = INFINITY;
            for (int d = 0; d < D; d++)
                if (prev[d] < bv) {
                    bv = prev[d];
                    z = d;
                }
        }
        int16_t *out = disp + (size_t)y * W;
        if (flags & TDVREF_DP_REF_BACKTRACK) {
            for (int x = W - 1; x >= 0; x--) {
                const int nz = z + path[(size_t)x * D + z];
                if (nz >= 0 && nz < D) z = nz;
                out[x] = (int16_t)z;
            }
        } else {
            for (int x = W - 1; x >= 0; x--) {
                out[x] = (int16_t)z;
                const int nz = z + path[(size_t)x * D + z];
                if (x > 0 && nz >= 0 && nz < D) z = nz;
            }
        }
    }
    free(prev);
    free(cur);
    free(path);
}

/* ------------------------------------------------------------------------ */
/* Fused SSD + WTA (fastwta.cu:32-53)                                        */
/* ------------------------------------------------------------------------ */

void tdvref_fast_wta_ssd(const uint8_t *L, const uint8_t *R, int W, int H,
                         int D, int radius, int16_t *disp)
{
    uint32_t *C = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)W * H * D);
    tdvref_cost_ssd_box(L, R, W, H, D, radius, C);
    tdvref_wta_u32(C, W, H, D, disp);
    free(C);
}
