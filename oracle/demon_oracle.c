/*
 * demon_oracle.c -- CPU oracle for the DEMON Yukawa + RK4 hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see demon_oracle.h).  Built with
 * -O2 -ffp-contract=off -fno-fast-math so that every + - * / is a single
 * IEEE binary64/binary32 operation in source order, as in the reference's
 * SSE2 code (VectorCompatibility.h:99-105: fmadd_pd is mul then add).
 *
 * All citations are to AU-PSL/demon_simulation_code.
 */
#define _GNU_SOURCE 1 /* M_PI */
#include "demon_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

double oracle_coulomb_const(void)
{
    /* ShieldedCoulombForce.cpp:21 with Cloud.cpp:16 */
    return 1.0 / (4.0 * M_PI * 8.8541878E-12);
}

/* ------------------------------------------------------------------------ */
/* Yukawa (shielded Coulomb) pair force                                     */
/* ------------------------------------------------------------------------ */

/* Intra-pair term for particles (p, p+1): ShieldedCoulombForce.cpp:138-159.
 * forceC = ((C*qq)*(1+u)) / (((d*d)*d)*exp(u)); F_p += forceC*d, F_p+1 -= it. */
static void yk_intra(int64_t p, const double *xs, const double *ys, const double *q,
                     double kappa, double coul, double *fx, double *fy)
{
    const double dX = xs[p] - xs[p + 1];          /* _mm_hsub_pd lane 0 */
    const double dY = ys[p] - ys[p + 1];          /* _mm_hsub_pd lane 1 */
    const double d = sqrt(dX * dX + dY * dY);
    const double u = d * kappa;
    if (u < 10.0) {
        const double qq = q[p] * q[p + 1];
        const double fc = coul * qq * (1.0 + u) / (d * d * d * exp(u));
        const double fX = fc * dX;
        const double fY = fc * dY;
        fx[p] += fX;       fx[p + 1] += -fX;      /* _mm_set_pd(-fX, fX) */
        fy[p] += fY;       fy[p + 1] += -fY;
    }
}

/* One 2-lane cross evaluation (ShieldedCoulombForce.cpp:172-198 and 211-237).
 * Lane l pairs outer particle o[l] with inner particle in[l].  If no lane is
 * inside the cutoff nothing is added; otherwise both lanes are added, a lane
 * outside the cutoff carrying exp = 0 (exp_pd, :279-292). */
static void yk_cross2(const int64_t o[2], const int64_t in[2],
                      const double *xs, const double *ys, const double *q,
                      double kappa, double coul, double *fx, double *fy)
{
    double dX[2], dY[2], d[2], u[2];
    int live[2];
    for (int l = 0; l < 2; ++l) {
        dX[l] = xs[o[l]] - xs[in[l]];
        dY[l] = ys[o[l]] - ys[in[l]];
        d[l] = sqrt(dX[l] * dX[l] + dY[l] * dY[l]);
        u[l] = d[l] * kappa;
        live[l] = u[l] < 10.0;
    }
    if (!live[0] && !live[1])
        return;
    for (int l = 0; l < 2; ++l) {
        const double e = live[l] ? exp(-u[l]) : 0.0;
        const double qq = q[o[l]] * q[in[l]];
        const double fc = coul * qq * (1.0 + u[l]) * e / (d[l] * d[l] * d[l]);
        const double fX = fc * dX[l];
        const double fY = fc * dY[l];
        fx[o[l]] += fX;
        fy[o[l]] += fY;
        fx[in[l]] -= fX;
        fy[in[l]] -= fY;
    }
}

int oracle_yukawa_add(int64_t n, const double *xs, const double *ys,
                      const double *q, double kappa, double *fx, double *fy)
{
    if (n < 2 || (n & 1))
        return -1;
    const double coul = oracle_coulomb_const();
    /* Outer loop over particle pairs (ShieldedCoulombForce.cpp:43). */
    for (int64_t p = 0; p < n - 1; p += 2) {
        yk_intra(p, xs, ys, q, kappa, coul, fx, fy);                    /* :51 */
        for (int64_t i = p + 2; i < n; i += 2) {                         /* :52 */
            const int64_t o[2] = {p, p + 1};
            const int64_t straight[2] = {i, i + 1};                      /* force  :54 */
            const int64_t swapped[2] = {i + 1, i};                       /* forcer :55 */
            yk_cross2(o, straight, xs, ys, q, kappa, coul, fx, fy);
            yk_cross2(o, swapped, xs, ys, q, kappa, coul, fx, fy);
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Cell-list form of oracle_yukawa_add (same results, pairs in range only)  */
/* ------------------------------------------------------------------------ */

typedef struct cell_grid {
    double x0, y0, h;
    int64_t nx, ny;
    int64_t *start;  /* [nx*ny + 1] */
    int64_t *idx;    /* particles sorted by cell, ascending index within a cell */
} cell_grid;

static void cell_of(const cell_grid *g, double x, double y, int64_t *cx, int64_t *cy)
{
    int64_t a = (int64_t)floor((x - g->x0) / g->h), b = (int64_t)floor((y - g->y0) / g->h);
    *cx = a < 0 ? 0 : a >= g->nx ? g->nx - 1 : a;
    *cy = b < 0 ? 0 : b >= g->ny ? g->ny - 1 : b;
}

/* Cells of side h over the bounding box (at most about 4 n cells). */
static int cell_build(cell_grid *g, int64_t n, const double *x, const double *y, double h)
{
    double x0 = INFINITY, x1 = -INFINITY, y0 = INFINITY, y1 = -INFINITY;
    for (int64_t i = 0; i < n; ++i) {
        if (x[i] < x0) x0 = x[i];
        if (x[i] > x1) x1 = x[i];
        if (y[i] < y0) y0 = y[i];
        if (y[i] > y1) y1 = y[i];
    }
    const double lim = 2.0 * sqrt((double)n) + 2.0;
    if ((x1 - x0) / h > lim || (y1 - y0) / h > lim)
        h = fmax(x1 - x0, y1 - y0) / lim;
    g->x0 = x0; g->y0 = y0; g->h = h;
    g->nx = (int64_t)((x1 - x0) / h) + 1;
    g->ny = (int64_t)((y1 - y0) / h) + 1;
    const int64_t nc = g->nx * g->ny;
    g->start = (int64_t *)calloc((size_t)nc + 1, sizeof(int64_t));
    g->idx = (int64_t *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int64_t));
    int64_t *cell = (int64_t *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int64_t));
    if (!g->start || !g->idx || !cell) {
        free(g->start); free(g->idx); free(cell);
        return -1;
    }
    for (int64_t i = 0; i < n; ++i) {
        int64_t cx, cy;
        cell_of(g, x[i], y[i], &cx, &cy);
        cell[i] = cy * g->nx + cx;
        g->start[cell[i] + 1] += 1;
    }
    for (int64_t c = 0; c < nc; ++c) g->start[c + 1] += g->start[c];
    int64_t *fill = (int64_t *)malloc((size_t)(nc > 0 ? nc : 1) * sizeof(int64_t));
    if (!fill) {
        free(g->start); free(g->idx); free(cell);
        return -1;
    }
    memcpy(fill, g->start, (size_t)nc * sizeof(int64_t));
    for (int64_t i = 0; i < n; ++i) g->idx[fill[cell[i]]++] = i;
    free(fill);
    free(cell);
    return 0;
}

static void cell_free(cell_grid *g)
{
    free(g->start);
    free(g->idx);
}

/* One live pair (a < b) as the reference computes it: the intra-pair formula
 * for a pair (p, p+1) of one load (ShieldedCoulombForce.cpp:138-159), the
 * cross-pair one otherwise (:172-198); dX = x_a - x_b, added to a and
 * subtracted from b.  Returns 0 if the pair is beyond the cutoff. */
static int yk_pair(int64_t a, int64_t b, const double *xs, const double *ys, const double *q,
                   double kappa, double coul, double *fX, double *fY)
{
    const double dX = xs[a] - xs[b];
    const double dY = ys[a] - ys[b];
    const double d = sqrt(dX * dX + dY * dY);
    const double u = d * kappa;
    if (!(u < 10.0))
        return 0;
    const double qq = q[a] * q[b];
    double fc;
    if ((a >> 1) == (b >> 1))
        fc = coul * qq * (1.0 + u) / (d * d * d * exp(u));
    else
        fc = coul * qq * (1.0 + u) * exp(-u) / (d * d * d);
    *fX = fc * dX;
    *fY = fc * dY;
    return 1;
}

/* Order of particle k's contributions in the reference's serial loop:
 * by the partner's pair index j/2, and within one partner pair the straight
 * lane (j % 2 == k % 2) before the swapped one (ShieldedCoulombForce.cpp:52-55). */
static int64_t yk_order(int64_t k, int64_t j) { return 2 * (j >> 1) + ((j & 1) != (k & 1)); }

static int cmp_i64(const void *a, const void *b)
{
    const int64_t u = *(const int64_t *)a, v = *(const int64_t *)b;
    return (u > v) - (u < v);
}

int oracle_yukawa_cells_add(int64_t n, const double *xs, const double *ys,
                            const double *q, double kappa, double *fx, double *fy)
{
    if (n < 2 || (n & 1))
        return -1;
    const double coul = oracle_coulomb_const();
    cell_grid g;
    /* cells wider than the cutoff radius 10/kappa (the margin covers the
       rounding of u): every live pair lies in neighbouring cells */
    if (cell_build(&g, n, xs, ys, (10.0 / kappa) * (1.0 + 1e-6)))
        return -1;
    int failed = 0;
#pragma omp parallel
    {
        int64_t cap = 1024, len = 0;
        int64_t *key = (int64_t *)malloc((size_t)cap * sizeof(int64_t));
        if (!key) {
#pragma omp atomic write
            failed = 1;
        }
#pragma omp for schedule(dynamic, 64)
        for (int64_t k = 0; k < n; ++k) {
            if (!key)
                continue;
            int64_t cx, cy;
            cell_of(&g, xs[k], ys[k], &cx, &cy);
            len = 0;
            for (int64_t by = cy - 1; by <= cy + 1; ++by) {
                if (by < 0 || by >= g.ny)
                    continue;
                for (int64_t bx = cx - 1; bx <= cx + 1; ++bx) {
                    if (bx < 0 || bx >= g.nx)
                        continue;
                    const int64_t c = by * g.nx + bx;
                    for (int64_t s = g.start[c]; s < g.start[c + 1]; ++s) {
                        const int64_t j = g.idx[s];
                        if (j == k)
                            continue;
                        double fX, fY;
                        if (!yk_pair(k < j ? k : j, k < j ? j : k, xs, ys, q, kappa, coul, &fX, &fY))
                            continue;
                        if (len == cap) {
                            int64_t *nk = (int64_t *)realloc(key, (size_t)(2 * cap) * sizeof(int64_t));
                            if (!nk) {
#pragma omp atomic write
                                failed = 1;
                                break;
                            }
                            key = nk;
                            cap *= 2;
                        }
                        /* order key in the high bits, partner index in the low 32 */
                        key[len++] = (yk_order(k, j) << 32) | j;
                    }
                }
            }
            qsort(key, (size_t)len, sizeof(int64_t), cmp_i64);
            double ax = fx[k], ay = fy[k];
            for (int64_t e = 0; e < len; ++e) {
                const int64_t j = key[e] & 0xFFFFFFFFll;
                double fX = 0.0, fY = 0.0;
                if (k < j) {
                    yk_pair(k, j, xs, ys, q, kappa, coul, &fX, &fY);
                    ax += fX;
                    ay += fY;
                } else {
                    yk_pair(j, k, xs, ys, q, kappa, coul, &fX, &fY);
                    ax -= fX;
                    ay -= fY;
                }
            }
            fx[k] = ax;
            fy[k] = ay;
        }
        free(key);
    }
    cell_free(&g);
    return failed ? -1 : 0;
}

float oracle_dtcheck_min_s_cells(int64_t n, const double *x, const double *y, double cell)
{
    cell_grid g;
    if (cell_build(&g, n, x, y, cell))
        return NAN;
    /* the double distance of any pair the FP32 test can find within 1e-4 m
       is below 1.01e-4 m (coordinate rounding is at most 2^-24 |x|) */
    const double r2 = 1.01e-4 * 1.01e-4;
    float best = INFINITY;
#pragma omp parallel for schedule(dynamic, 256) reduction(min : best)
    for (int64_t i = 0; i < n; ++i) {
        int64_t cx, cy;
        cell_of(&g, x[i], y[i], &cx, &cy);
        for (int64_t by = cy - 1; by <= cy + 1; ++by) {
            if (by < 0 || by >= g.ny)
                continue;
            for (int64_t bx = cx - 1; bx <= cx + 1; ++bx) {
                if (bx < 0 || bx >= g.nx)
                    continue;
                const int64_t c = by * g.nx + bx;
                for (int64_t s = g.start[c]; s < g.start[c + 1]; ++s) {
                    const int64_t j = g.idx[s];
                    if (j <= i)
                        continue;
                    const double ex = x[i] - x[j], ey = y[i] - y[j];
                    if (ex * ex + ey * ey > r2 || !oracle_dtcheck_covers(i, j))
                        continue;
                    const float sx = (float)x[i] - (float)x[j];
                    const float sy = (float)y[i] - (float)y[j];
                    const float s2 = sx * sx + sy * sy;
                    if (s2 < best)
                        best = s2;
                }
            }
        }
    }
    cell_free(&g);
    return best;
}

double oracle_modify_time_step_cells(int64_t n, const double *x, const double *y, float dist, double dt,
                                     double cell)
{
    const float s = oracle_dtcheck_min_s_cells(n, x, y, cell);
    for (int k = 0; sqrtf(s) <= dist; ++k) {
        if (k > 60)
            return NAN; /* coincident particles: the reference loops forever */
        dist /= 10.0f;
        dt /= 10.0f;
    }
    return dt;
}

void oracle_yukawa_gather_add(int64_t n, const double *xs, const double *ys,
                              const double *q, double kappa, double *fx, double *fy)
{
    const double coul = oracle_coulomb_const();
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < n; ++i) {
        double ax = fx[i], ay = fy[i];
        for (int64_t j = 0; j < n; ++j) {
            if (j == i)
                continue;
            const double dX = xs[i] - xs[j];
            const double dY = ys[i] - ys[j];
            const double d = sqrt(dX * dX + dY * dY);
            const double u = d * kappa;
            if (u < 10.0) {
                const double fc = coul * (q[i] * q[j]) * (1.0 + u) * exp(-u) / (d * d * d);
                ax += fc * dX;
                ay += fc * dY;
            }
        }
        fx[i] = ax;
        fy[i] = ay;
    }
}

void oracle_yukawa_rows(int64_t n, const double *xs, const double *ys, const double *q,
                        double kappa, const int64_t *rows, int64_t nrows, double *fx, double *fy)
{
    const double coul = oracle_coulomb_const();
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t k = 0; k < nrows; ++k) {
        const int64_t i = rows[k];
        double ax = 0.0, ay = 0.0;
        for (int64_t j = 0; j < n; ++j) {
            if (j == i)
                continue;
            const double dX = xs[i] - xs[j];
            const double dY = ys[i] - ys[j];
            const double d = sqrt(dX * dX + dY * dY);
            const double u = d * kappa;
            if (u < 10.0) {
                const double fc = coul * (q[i] * q[j]) * (1.0 + u) * exp(-u) / (d * d * d);
                ax += fc * dX;
                ay += fc * dY;
            }
        }
        fx[k] = ax;
        fy[k] = ay;
    }
}

void oracle_yukawa_longdouble(int64_t n, const double *xs, const double *ys,
                              const double *q, double kappa, double *fx, double *fy)
{
    const long double coul = 1.0L / (4.0L * 3.14159265358979323846264338327950288L * 8.8541878E-12L);
    for (int64_t i = 0; i < n; ++i) {
        long double ax = 0.0L, ay = 0.0L;
        for (int64_t j = 0; j < n; ++j) {
            if (j == i)
                continue;
            const long double dX = (long double)xs[i] - (long double)xs[j];
            const long double dY = (long double)ys[i] - (long double)ys[j];
            const long double d = sqrtl(dX * dX + dY * dY);
            const long double u = d * (long double)kappa;
            /* classify with the reference's double rounding so that the
             * set of contributing pairs is identical */
            const double dxd = xs[i] - xs[j], dyd = ys[i] - ys[j];
            const double ud = sqrt(dxd * dxd + dyd * dyd) * kappa;
            if (ud < 10.0) {
                const long double fc = coul * (long double)q[i] * (long double)q[j] *
                                       (1.0L + u) * expl(-u) / (d * d * d);
                ax += fc * dX;
                ay += fc * dY;
            }
        }
        fx[i] = (double)ax;
        fy[i] = (double)ay;
    }
}

int64_t oracle_near_pairs(int64_t n, const double *xs, const double *ys, double kappa)
{
    int64_t c = 0;
    for (int64_t i = 0; i < n; ++i)
        for (int64_t j = 0; j < n; ++j) {
            if (i == j)
                continue;
            const double dX = xs[i] - xs[j], dY = ys[i] - ys[j];
            if (sqrt(dX * dX + dY * dY) * kappa < 10.0)
                ++c;
        }
    return c;
}

/* ------------------------------------------------------------------------ */
/* O(N) forces                                                              */
/* ------------------------------------------------------------------------ */

void oracle_confinement_add(int64_t n, const double *xs, const double *ys,
                            const double *q, double confine, double *fx, double *fy)
{
    for (int64_t i = 0; i < n; ++i) {
        const double cv = q[i] * confine;       /* ConfinementForce.cpp:50 */
        fx[i] += cv * xs[i];                    /* :52 */
        fy[i] += cv * ys[i];                    /* :53 */
    }
}

void oracle_drag_add(int64_t n, const double *vxs, const double *vys,
                     const double *m, double gamma, double *fx, double *fy)
{
    const double drag_const = -gamma;           /* DragForce.h:17 */
    for (int64_t i = 0; i < n; ++i) {
        const double dr = m[i] * drag_const;    /* DragForce.cpp:144 */
        fx[i] += dr * vxs[i];                   /* :146 */
        fy[i] += dr * vys[i];                   /* :147 */
    }
}

/* ------------------------------------------------------------------------ */
/* Counter-based RNG and the thermal force                                  */
/* ------------------------------------------------------------------------ */

static inline void philox_mulhilo(uint32_t a, uint32_t b, uint32_t *hi, uint32_t *lo)
{
    const uint64_t p = (uint64_t)a * (uint64_t)b;
    *hi = (uint32_t)(p >> 32);
    *lo = (uint32_t)p;
}

void oracle_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    /* Philox4x32 with 10 rounds: multipliers 0xD2511F53 / 0xCD9E8D57,
     * Weyl key increments 0x9E3779B9 / 0xBB67AE85 (Salmon et al. 2011). */
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0, lo0, hi1, lo1;
        philox_mulhilo(0xD2511F53u, c0, &hi0, &lo0);
        philox_mulhilo(0xCD9E8D57u, c2, &hi1, &lo1);
        const uint32_t n0 = hi1 ^ c1 ^ k0;
        const uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static inline double u53(uint32_t lo, uint32_t hi)
{
    const uint64_t v = (((uint64_t)hi << 32) | lo) >> 11;
    return (double)v * 0x1.0p-53;
}

/* One thermal kick of particle i: U and theta from the Philox block keyed
 * (seed; i, step, domain << 24 | stage << 16 | step >> 32).  Domains: 1
 * ThermalForce, 3 ThermalForceLocalized, 4 TimeVaryingThermalForce. */
static void thermal_kick(int64_t i, double heat, uint64_t seed, uint64_t step, uint32_t domain,
                         int stage, double *fx, double *fy)
{
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    const uint32_t ctr[4] = {(uint32_t)i, (uint32_t)((uint64_t)i >> 32), (uint32_t)step,
                             (domain << 24) | ((uint32_t)stage << 16) |
                                 ((uint32_t)(step >> 32) & 0xFFFFu)};
    uint32_t o[4];
    oracle_philox4x32(ctr, key, o);
    const double uu = u53(o[0], o[1]);                  /* RandCache::r  in [0,1) */
    const double th = u53(o[2], o[3]) * (2.0 * M_PI);   /* RandCache::r1 in [0,2pi) */
    const double tv = uu * heat;                        /* ThermalForce.cpp:119 */
    *fx += tv * cos(th);                                /* :120 */
    *fy += tv * sin(th);                                /* :121 */
}

void oracle_thermal_add(int64_t n, double heat, uint64_t seed, uint64_t step,
                        int stage, double *fx, double *fy)
{
    for (int64_t i = 0; i < n; ++i)
        thermal_kick(i, heat, seed, step, 1u, stage, &fx[i], &fy[i]);
}

void oracle_thermal_localized_add(int64_t n, const double *xs, const double *ys,
                                  double heat_in, double heat_out, double radius,
                                  uint64_t seed, uint64_t step, int stage, double *fx,
                                  double *fy)
{
    for (int64_t i = 0; i < n; ++i) {
        const double r = sqrt(xs[i] * xs[i] + ys[i] * ys[i]);     /* ThermalForceLocalized.cpp:126 */
        const double heat = r < radius ? heat_in : heat_out;      /* :127-128 */
        thermal_kick(i, heat, seed, step, 3u, stage, &fx[i], &fy[i]);
    }
}

void oracle_tvthermal_add(int64_t n, double scale, double offset, double t, uint64_t seed,
                          uint64_t step, int stage, double *fx, double *fy)
{
    const double heat = scale * t + offset;     /* TimeVaryingThermalForce.cpp:40-42 */
    for (int64_t i = 0; i < n; ++i)
        thermal_kick(i, heat, seed, step, 4u, stage, &fx[i], &fy[i]);
}

/* ------------------------------------------------------------------------ */
/* FFTAnalysis -a                                                           */
/* ------------------------------------------------------------------------ */

void oracle_average_fft(const double *series, int64_t rows, int64_t stride, int64_t nuse, double *amp)
{
    const int64_t half = rows / 2;
    double *re = (double *)calloc((size_t)(half > 0 ? half : 1), sizeof(double));
    double *im = (double *)calloc((size_t)(half > 0 ? half : 1), sizeof(double));
    if (!re || !im) {
        free(re);
        free(im);
        return;
    }
    /* FFTAnalysis.cpp:241-257: one transform per particle, summed */
    for (int64_t p = 0; p < nuse; ++p) {
        for (int64_t k = 0; k < half; ++k) {
            double sr = 0.0, si = 0.0;
            for (int64_t t = 0; t < rows; ++t) {
                const int64_t m = (k * t) % rows;
                const double a = (2.0 * M_PI) * (double)m / (double)rows;
                const double x = series[t * stride + p];      /* :245-247 */
                sr += x * cos(a);
                si -= x * sin(a);
            }
            re[k] += sr;                                      /* :254-255 */
            im[k] += si;
        }
    }
    /* FFTAnalysis.cpp:270-279: the printed amplitude is |sum| / nuse (the
     * zeroing at :274-277 comes after amp is formed and does not change it) */
    for (int64_t k = 0; k < half; ++k)
        amp[k] = sqrt(re[k] * re[k] + im[k] * im[k]) / (double)nuse;
    free(re);
    free(im);
}

/* ------------------------------------------------------------------------ */
/* Integrator::modifyTimeStep                                               */
/* ------------------------------------------------------------------------ */

/* Lanes of loadFloatVector (Integrator.cpp:162-168): _mm_set_ps puts x[o] in
 * lane 3, so lane k holds particle o+3-k. */
static inline int64_t lane_particle(int64_t o, int k) { return o + 3 - k; }

/* tryToReduceTimeStep (Integrator.cpp:144-155) on a 4-lane vector of pairs. */
static void try_reduce(const int64_t a[4], const int64_t b[4], const double *x,
                       const double *y, float *dist, double *dt)
{
    float s[4];
    for (int k = 0; k < 4; ++k) {
        const float sx = (float)x[a[k]] - (float)x[b[k]];
        const float sy = (float)y[a[k]] - (float)y[b[k]];
        s[k] = sx * sx + sy * sy;                  /* add_ps(mul_ps, mul_ps) */
    }
    for (;;) {
        int within = 0;
        for (int k = 0; k < 4; ++k)
            within |= sqrtf(s[k]) <= *dist;         /* isWithInDistance :181 */
        if (!within)
            break;
        *dist /= 10.0f;
        *dt /= 10.0f;
    }
}

double oracle_modify_time_step(int64_t n, const double *x, const double *y,
                               float dist, double dt)
{
    for (int64_t o = 0; o < n - 1; o += 4) {                       /* :90 */
        int64_t A[4];
        for (int k = 0; k < 4; ++k)
            A[k] = lane_particle(o, k);
        /* hsub(A, shuffle(A,A,(1,2,0,3))) = (A0-A1, A2-A3, A3-A0, A2-A1)  :96-98 */
        {
            const int64_t a[4] = {A[0], A[2], A[3], A[2]};
            const int64_t b[4] = {A[1], A[3], A[0], A[1]};
            try_reduce(a, b, x, y, &dist, &dt);
        }
        /* hsub(A, shuffle(A,A,(1,3,0,2))) = (A0-A1, A2-A3, A2-A0, A3-A1)  :102-104 */
        {
            const int64_t a[4] = {A[0], A[2], A[2], A[3]};
            const int64_t b[4] = {A[1], A[3], A[0], A[1]};
            try_reduce(a, b, x, y, &dist, &dt);
        }
        for (int64_t in = o + 4; in < n; in += 4) {                  /* :107 */
            int64_t B[4];
            for (int k = 0; k < 4; ++k)
                B[k] = lane_particle(in, k);
            /* A - B                                                     :112-114 */
            {
                const int64_t b[4] = {B[0], B[1], B[2], B[3]};
                try_reduce(A, b, x, y, &dist, &dt);
            }
            /* A - shuffle(B,(3,2,1,0)) : identity shuffle, same pairs    :117-119 */
            {
                const int64_t b[4] = {B[0], B[1], B[2], B[3]};
                try_reduce(A, b, x, y, &dist, &dt);
            }
            /* A - shuffle(B,(0,3,2,1)) = A - (B1,B2,B3,B0)               :122-124 */
            {
                const int64_t b[4] = {B[1], B[2], B[3], B[0]};
                try_reduce(A, b, x, y, &dist, &dt);
            }
            /* A - shuffle(B,(1,0,3,2)) = A - (B2,B3,B0,B1)               :127-129 */
            {
                const int64_t b[4] = {B[2], B[3], B[0], B[1]};
                try_reduce(A, b, x, y, &dist, &dt);
            }
        }
    }
    return dt;
}

int oracle_dtcheck_covers(int64_t i, int64_t j)
{
    if (i == j)
        return 0;
    if (i > j) {
        const int64_t t = i; i = j; j = t;
    }
    const int64_t gi = i >> 2, gj = j >> 2;
    if (gi == gj)
        return 1;                       /* all 6 intra-group pairs are tested */
    /* cross-group: outer lane i%4, inner lane j%4; the pattern j = i+1 (mod 4)
     * is never tested (derived from the three distinct shuffles above) */
    return ((j & 3) != ((i + 1) & 3));
}

float oracle_dtcheck_min_s(int64_t n, const double *x, const double *y)
{
    float best = INFINITY;
    for (int64_t i = 0; i < n; ++i)
        for (int64_t j = i + 1; j < n; ++j) {
            if (!oracle_dtcheck_covers(i, j))
                continue;
            const float sx = (float)x[i] - (float)x[j];
            const float sy = (float)y[i] - (float)y[j];
            const float s = sx * sx + sy * sy;
            if (s < best)
                best = s;
        }
    return best;
}

/* ------------------------------------------------------------------------ */
/* Runge_Kutta4::moveParticles                                              */
/* ------------------------------------------------------------------------ */

void oracle_magnetic_add(int64_t n, const double *vxs, const double *vys,
                         const double *q, double bfield, double *fx, double *fy)
{
    for (int64_t i = 0; i < n; ++i) {
        const double qb = q[i] * bfield;           /* MagneticForce.cpp:46 */
        fx[i] = fx[i] + qb * vys[i];               /* :47 */
        fy[i] = fy[i] - qb * vxs[i];               /* :48 */
    }
}

void oracle_tvdrag_add(int64_t n, const double *vxs, const double *vys, const double *m,
                       double scale, double offset, double t, double *fx, double *fy)
{
    const double drag_const = -(scale * t + offset);  /* TimeVaryingDragForce.cpp:37-39 */
    for (int64_t i = 0; i < n; ++i) {
        const double d = m[i] * drag_const;            /* DragForce.cpp:50 */
        fx[i] = fx[i] + d * vxs[i];                    /* :52 */
        fy[i] = fy[i] + d * vys[i];                    /* :53 */
    }
}

void oracle_rect_add(int64_t n, const double *xs, const double *ys, const double *q,
                     double cx, double cy, double *fx, double *fy)
{
    for (int64_t i = 0; i < n; ++i) {
        fx[i] = fx[i] + (q[i] * cx) * xs[i];      /* RectConfinementForce.cpp:55-56 */
        fy[i] = fy[i] + (q[i] * cy) * ys[i];      /* :57-58 */
    }
}

void oracle_void_add(int64_t n, const double *xs, const double *ys, const double *q,
                     double confine, double decay, double *fx, double *fy)
{
    oracle_confinement_add(n, xs, ys, q, confine, fx, fy);   /* ConfinementForceVoid.cpp:14 */
    for (int64_t i = 0; i < n; ++i) {
        const double dv = q[i] * (-decay);                    /* :49 */
        const double r = sqrt(xs[i] * xs[i] + ys[i] * ys[i]); /* :50 (length_pd) */
        const double er = exp(r * (-decay)) / r;              /* :51 */
        fx[i] = fx[i] + (dv * er) * xs[i];                    /* :53 */
        fy[i] = fy[i] + (dv * er) * ys[i];                    /* :54 */
    }
}

void oracle_rotational_add(int64_t n, const double *xs, const double *ys, double rmin,
                           double rmax, double c, double *fx, double *fy)
{
    for (int64_t i = 0; i < n; ++i) {
        const double r = sqrt(xs[i] * xs[i] + ys[i] * ys[i]); /* RotationalForce.cpp:53 */
        if (!(r > rmin && r < rmax))                          /* :56-60 */
            continue;
        fx[i] = fx[i] - (c * ys[i]) / r;                      /* :64 */
        fy[i] = fy[i] + (c * xs[i]) / r;                      /* :65 */
    }
}

void oracle_driving_add(int64_t n, const double *xs, double t, double damp_const,
                        double amplitude, double shift, double *fx)
{
    const double wave_num = 2.0 * M_PI / 0.002;   /* DrivingForce.cpp:14 */
    const double ang_freq = 2.0 * M_PI * 10.0;    /* :15 */
    const double drive_const = -damp_const;       /* DrivingForce.h:18 */
    for (int64_t i = 0; i < n; ++i) {
        const double dist = xs[i] - shift;                          /* DrivingForce.cpp:55 */
        const double arg = xs[i] * wave_num - t * ang_freq;         /* :56 */
        const double earg = (dist - dist) / (-drive_const);         /* :57 */
        fx[i] = fx[i] + (sin(arg) * exp(earg)) * amplitude;         /* :59-60 */
    }
}

void oracle_electric_add(int64_t n, const double *xs, const double *ys, const double *q,
                         double e_field, double radius, double *fx, double *fy)
{
    const double rad = radius * (-1);             /* ElectricForce.cpp:55 */
    for (int64_t i = 0; i < n; ++i) {
        const double cv = q[i] * e_field;                         /* :51 */
        const double rr = sqrt(xs[i] * xs[i] + ys[i] * ys[i]);    /* :52-54 */
        const double ev = cv * exp(rr / rad);                     /* :56 */
        fx[i] = fx[i] + ev * xs[i];                               /* :59 */
        fy[i] = fy[i] + ev * ys[i];                               /* :60 */
    }
}

void oracle_vert_electric_add(int64_t n, const double *ys, const double *q, double e_field,
                              double decay, double *fy)
{
    for (int64_t i = 0; i < n; ++i) {
        const double cv = q[i] * e_field;                 /* VertElectricForce.cpp:49 */
        fy[i] = fy[i] + cv * exp(ys[i] / decay);          /* :50-52 */
    }
}

void oracle_gravity_add(int64_t n, const double *m, double g, double *fy)
{
    for (int64_t i = 0; i < n; ++i)
        fy[i] = fy[i] + m[i] * g;                         /* GravitationalForce.cpp:47-49 */
}

int oracle_stage_force(int64_t n, const double *xs, const double *ys,
                       const double *vxs, const double *vys, const double *q,
                       const double *m, const oracle_params *p, uint64_t step,
                       int stage, double t_stage, double *fx, double *fy)
{
    /* driver.cpp:372-402 construction order = summation order */
    const int f = p->forces;
    if (f & ORACLE_CONFINEMENT)
        oracle_confinement_add(n, xs, ys, q, p->confine, fx, fy);
    if (f & ORACLE_CONFINEMENT_VOID)
        oracle_void_add(n, xs, ys, q, p->confine, p->void_decay, fx, fy);
    if (f & ORACLE_DRAG)
        oracle_drag_add(n, vxs, vys, m, p->gamma, fx, fy);
    if (f & ORACLE_DRIVING)
        oracle_driving_add(n, xs, t_stage, p->drive_const, p->drive_amplitude, p->drive_shift, fx);
    if (f & ORACLE_MAGNETIC)
        oracle_magnetic_add(n, vxs, vys, q, p->magnetic_field, fx, fy);
    if (f & ORACLE_RECT_CONFINEMENT)
        oracle_rect_add(n, xs, ys, q, p->confine_x, p->confine_y, fx, fy);
    if (f & ORACLE_ROTATIONAL)
        oracle_rotational_add(n, xs, ys, p->rot_inner, p->rot_outer, p->rot_const, fx, fy);
    if (p->forces & ORACLE_SHIELDED_COULOMB) {
        if (p->gather == 2 && !(n & 1)) {
            if (oracle_yukawa_cells_add(n, xs, ys, q, p->shielding, fx, fy))
                return -1;
        } else if ((n & 1) || p->gather)
            oracle_yukawa_gather_add(n, xs, ys, q, p->shielding, fx, fy);
        else if (oracle_yukawa_add(n, xs, ys, q, p->shielding, fx, fy))
            return -1;
    }
    if (f & ORACLE_THERMAL)
        oracle_thermal_add(n, p->heat_val, p->seed, step, stage, fx, fy);
    if (f & ORACLE_THERMAL_LOCALIZED)
        oracle_thermal_localized_add(n, xs, ys, p->heat_in, p->heat_out, p->heat_radius, p->seed,
                                     step, stage, fx, fy);
    if (f & ORACLE_TIME_VARYING_DRAG)
        oracle_tvdrag_add(n, vxs, vys, m, p->drag_scale, p->gamma, t_stage, fx, fy);
    if (f & ORACLE_TIME_VARYING_THERMAL)
        oracle_tvthermal_add(n, p->heat_scale, p->heat_offset, t_stage, p->seed, step, stage, fx, fy);
    if (f & ORACLE_ELECTRIC)
        oracle_electric_add(n, xs, ys, q, p->electric_field, p->plasma_radius, fx, fy);
    if (f & ORACLE_GRAVITATIONAL)
        oracle_gravity_add(n, m, p->gravity, fy);
    if (f & ORACLE_VERT_ELECTRIC)
        oracle_vert_electric_add(n, ys, q, p->vert_electric_field, p->vert_decay, fy);
    return 0;
}

typedef struct rk_work {
    double *k[4], *l[4], *mm[4], *nn[4];
    double *fx, *fy, *xc, *yc, *vxc, *vyc;
    double *block;
} rk_work;

static int rk_alloc(rk_work *w, int64_t n)
{
    w->block = (double *)calloc((size_t)(n > 0 ? n : 1) * 22, sizeof(double));
    if (!w->block)
        return -1;
    double *b = w->block;
    for (int s = 0; s < 4; ++s) {
        w->k[s] = b; b += n;
        w->l[s] = b; b += n;
        w->mm[s] = b; b += n;
        w->nn[s] = b; b += n;
    }
    w->fx = b; b += n;
    w->fy = b; b += n;
    w->xc = b; b += n;
    w->yc = b; b += n;
    w->vxc = b; b += n;
    w->vyc = b;
    return 0;
}

/* (a1 + ((a2 + a3)*2 + a4)) / 6   Runge_Kutta4.cpp:147-151 */
static inline double rk_da(double a1, double a2, double a3, double a4)
{
    return (a1 + ((a2 + a3) * 2.0 + a4)) / 6.0;
}

static int rk4_one_step(int64_t n, double *x, double *y, double *vx, double *vy,
                        const double *q, const double *m, const oracle_params *p,
                        double init_dt, double *t, uint64_t *step_counter, rk_work *w)
{
    /* Runge_Kutta4.cpp:25 */
    double dt = init_dt;
    if (p->dt_check)
        dt = p->gather == 2 ? oracle_modify_time_step_cells(n, x, y, 1.0e-4f, init_dt, 10.0 / p->shielding)
                            : oracle_modify_time_step(n, x, y, 1.0e-4f, init_dt);
    if (!(dt > 0.0))
        return -1;
    const uint64_t step = *step_counter;
    for (int s = 0; s < 4; ++s) {
        const double *xs, *ys, *vxs, *vys;
        if (s == 0) {
            /* operation1 is a no-op; stage-1 getters read x, Vx (Cloud.cpp:277,415) */
            xs = x; ys = y; vxs = vx; vys = vy;
        } else {
            /* CacheOperator::operation2..4 (CacheOperator.cpp:28-73) */
            const double *pl = w->l[s - 1], *pn = w->nn[s - 1];
            const double *pk = w->k[s - 1], *pm = w->mm[s - 1];
            for (int64_t i = 0; i < n; ++i) {
                if (s < 3) {
                    w->xc[i] = 0.5 * pl[i] + x[i];
                    w->yc[i] = 0.5 * pn[i] + y[i];
                    w->vxc[i] = 0.5 * pk[i] + vx[i];
                    w->vyc[i] = 0.5 * pm[i] + vy[i];
                } else {
                    w->xc[i] = x[i] + pl[i];
                    w->yc[i] = y[i] + pn[i];
                    w->vxc[i] = vx[i] + pk[i];
                    w->vyc[i] = vy[i] + pm[i];
                }
            }
            xs = w->xc; ys = w->yc; vxs = w->vxc; vys = w->vyc;
        }
        /* stage times t, t + dt/2, t + dt/2, t + dt (Runge_Kutta4.cpp:31,50,69,88) */
        const double ts = s == 0 ? *t : s == 3 ? *t + dt : *t + dt / 2.0;
        if (oracle_stage_force(n, xs, ys, vxs, vys, q, m, p, step, s + 1, ts, w->fx, w->fy))
            return -1;
        /* stage loop, Runge_Kutta4.cpp:32-47 */
        for (int64_t i = 0; i < n; ++i) {
            w->k[s][i] = (dt * w->fx[i]) / m[i];
            w->l[s][i] = dt * vxs[i];
            w->mm[s][i] = (dt * w->fy[i]) / m[i];
            w->nn[s][i] = dt * vys[i];
            w->fx[i] = 0.0;
            w->fy[i] = 0.0;
        }
    }
    /* combine, Runge_Kutta4.cpp:107-133 */
    for (int64_t i = 0; i < n; ++i) {
        vx[i] += rk_da(w->k[0][i], w->k[1][i], w->k[2][i], w->k[3][i]);
        x[i] += rk_da(w->l[0][i], w->l[1][i], w->l[2][i], w->l[3][i]);
        vy[i] += rk_da(w->mm[0][i], w->mm[1][i], w->mm[2][i], w->mm[3][i]);
        y[i] += rk_da(w->nn[0][i], w->nn[1][i], w->nn[2][i], w->nn[3][i]);
    }
    *t += dt;                                   /* :135 */
    *step_counter = step + 1;
    return 0;
}

int oracle_rk4_steps(int64_t n, double *x, double *y, double *vx, double *vy,
                     const double *q, const double *m, const oracle_params *p,
                     double init_dt, int64_t steps, double *t, uint64_t *step_counter)
{
    if (n < 1)
        return -1;
    if (p->dt_check && (n & 3))
        return -1;
    rk_work w;
    if (rk_alloc(&w, n))
        return -1;
    int rc = 0;
    for (int64_t s = 0; s < steps && !rc; ++s)
        rc = rk4_one_step(n, x, y, vx, vy, q, m, p, init_dt, t, step_counter, &w);
    free(w.block);
    return rc;
}

/* Runge_Kutta2::moveParticles loop body (Runge_Kutta2.cpp:24-77). */
static int rk2_one_step(int64_t n, double *x, double *y, double *vx, double *vy,
                        const double *q, const double *m, const oracle_params *p,
                        double init_dt, double *t, uint64_t *step_counter, rk_work *w)
{
    const double dt = p->dt_check ? oracle_modify_time_step(n, x, y, 1.0e-4f, init_dt) : init_dt; /* :25 */
    const uint64_t step = *step_counter;
    /* stage 1 on (x, Vx) at t (:29-46) */
    if (oracle_stage_force(n, x, y, vx, vy, q, m, p, step, 1, *t, w->fx, w->fy))
        return -1;
    for (int64_t i = 0; i < n; ++i) {
        w->k[0][i] = (dt * w->fx[i]) / m[i];
        w->l[0][i] = dt * vx[i];
        w->mm[0][i] = (dt * w->fy[i]) / m[i];
        w->nn[0][i] = dt * vy[i];
        w->fx[i] = 0.0;
        w->fy[i] = 0.0;
    }
    /* operation2 (CacheOperator.cpp:28-44), stage 2 at t + dt/2 (:49-65) */
    for (int64_t i = 0; i < n; ++i) {
        w->xc[i] = 0.5 * w->l[0][i] + x[i];
        w->yc[i] = 0.5 * w->nn[0][i] + y[i];
        w->vxc[i] = 0.5 * w->k[0][i] + vx[i];
        w->vyc[i] = 0.5 * w->mm[0][i] + vy[i];
    }
    if (oracle_stage_force(n, w->xc, w->yc, w->vxc, w->vyc, q, m, p, step, 2, *t + dt / 2.0, w->fx, w->fy))
        return -1;
    for (int64_t i = 0; i < n; ++i) {
        const double k2 = (dt * w->fx[i]) / m[i];
        const double l2 = dt * w->vxc[i];
        const double m2 = (dt * w->fy[i]) / m[i];
        const double n2 = dt * w->vyc[i];
        w->fx[i] = 0.0;
        w->fy[i] = 0.0;
        vx[i] += k2;                             /* :68-74 */
        x[i] += l2;
        vy[i] += m2;
        y[i] += n2;
    }
    *t += dt;                                    /* :76 */
    *step_counter = step + 1;
    return 0;
}

int oracle_rk2_steps(int64_t n, double *x, double *y, double *vx, double *vy,
                     const double *q, const double *m, const oracle_params *p,
                     double init_dt, int64_t steps, double *t, uint64_t *step_counter)
{
    if (n < 1)
        return -1;
    if (p->dt_check && (n & 3))
        return -1;
    rk_work w;
    if (rk_alloc(&w, n))
        return -1;
    int rc = 0;
    for (int64_t s = 0; s < steps && !rc; ++s)
        rc = rk2_one_step(n, x, y, vx, vy, q, m, p, init_dt, t, step_counter, &w);
    free(w.block);
    return rc;
}

int oracle_rk4_move(int64_t n, double *x, double *y, double *vx, double *vy,
                    const double *q, const double *m, const oracle_params *p,
                    double init_dt, double end_time, double *t,
                    uint64_t *step_counter, int64_t *steps_done)
{
    if (n < 1)
        return -1;
    if (p->dt_check && (n & 3))
        return -1;
    rk_work w;
    if (rk_alloc(&w, n))
        return -1;
    int rc = 0;
    int64_t c = 0;
    while (*t < end_time && !rc) {               /* Runge_Kutta4.cpp:24 */
        rc = rk4_one_step(n, x, y, vx, vy, q, m, p, init_dt, t, step_counter, &w);
        ++c;
    }
    free(w.block);
    if (steps_done)
        *steps_done = c;
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Discrete adjoint of the RK4 map (CPU reference gradient)                 */
/* ------------------------------------------------------------------------ */

/* Stage force F(S) = q C r - gamma m v + sum_j phi_ij d_ij (gather, cross form). */
/* r2 beyond which sqrt(r2) * kappa < 10 is certainly false: the skip only
 * saves the square root and leaves the set of contributing pairs unchanged
 * (the rounding of sqrt and of the product is 2^-52 relative at most). */
static double fit_far2(double kappa)
{
    return (10.0 / kappa) * (10.0 / kappa) * (1.0 + 1e-9);
}

static void fit_force(int64_t n, const double *xs, const double *ys, const double *vxs,
                      const double *vys, const double *q, const double *m, double kappa,
                      double confine, double gamma, double *fx, double *fy)
{
    const double coul = oracle_coulomb_const();
    const double far2 = fit_far2(kappa);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        double ax = q[i] * confine * xs[i] - gamma * m[i] * vxs[i];
        double ay = q[i] * confine * ys[i] - gamma * m[i] * vys[i];
        for (int64_t j = 0; j < n; ++j) {
            if (j == i)
                continue;
            const double dX = xs[i] - xs[j], dY = ys[i] - ys[j];
            if (dX * dX + dY * dY > far2)
                continue;
            const double r = sqrt(dX * dX + dY * dY);
            const double u = r * kappa;
            if (u < 10.0) {
                const double phi = coul * (q[i] * q[j]) * (1.0 + u) * exp(-u) / (r * r * r);
                ax += phi * dX;
                ay += phi * dY;
            }
        }
        fx[i] = ax;
        fy[i] = ay;
    }
}

int oracle_fit_loss_grad(int64_t n, const double *x0, const double *y0, const double *vx0,
                         const double *vy0, const double *q, const double *m, double kappa,
                         double confine, double gamma, double dt0, int dt_check, int64_t steps,
                         int64_t obs_every, const double *tx, const double *ty, double *loss,
                         double *gq, double *gkappa, double *gconfine, double *ggamma,
                         double *gx0, double *gy0, double *gvx0, double *gvy0)
{
    if (n < 2 || steps < 1 || obs_every < 1 || (dt_check && (n & 3)))
        return -1;
    const int64_t n_obs = steps / obs_every;
    const double coul = oracle_coulomb_const();
    /* S[step][stage][4][n]: x, y, vx, vy of each stage state */
    const size_t per_stage = (size_t)4 * n;
    double *S = (double *)malloc(sizeof(double) * per_stage * 4 * (size_t)steps);
    double *st = (double *)malloc(sizeof(double) * per_stage * 2);     /* state, next */
    double *F = (double *)malloc(sizeof(double) * 2 * n);
    double *inc = (double *)malloc(sizeof(double) * 4 * 4 * n);         /* l,k,mm,nn per stage */
    double *seed = (double *)calloc((size_t)(n_obs > 0 ? n_obs : 1) * 2 * n, sizeof(double));
    double *bar = (double *)calloc((size_t)4 * n, sizeof(double));      /* adjoint of the state */
    double *ibar = (double *)calloc((size_t)4 * 4 * n, sizeof(double)); /* adjoint of increments */
    double *sbar = (double *)calloc((size_t)4 * n, sizeof(double));     /* adjoint of a stage state */
    double *dts = (double *)malloc(sizeof(double) * (size_t)steps);     /* each step's dt */
    if (!S || !st || !F || !inc || !seed || !bar || !ibar || !sbar || !dts)
        return -1;
    memcpy(st, x0, sizeof(double) * n);
    memcpy(st + n, y0, sizeof(double) * n);
    memcpy(st + 2 * n, vx0, sizeof(double) * n);
    memcpy(st + 3 * n, vy0, sizeof(double) * n);
    double L = 0.0;
    const double norm = 1.0 / (2.0 * (double)n * (double)(n_obs > 0 ? n_obs : 1));
    for (int64_t s = 0; s < steps; ++s) {
        double *Ss = S + per_stage * 4 * (size_t)s;
        /* the step's dt as Runge_Kutta4::moveParticles picks it
         * (Runge_Kutta4.cpp:25, Integrator.cpp:68-158) */
        const double dt = dt_check ? oracle_modify_time_step(n, st, st + n, 1.0e-4f, dt0) : dt0;
        dts[s] = dt;
        for (int k = 0; k < 4; ++k) {
            double *Sk = Ss + per_stage * k;
            const double c = (k == 3) ? 1.0 : 0.5;
            for (int64_t i = 0; i < n; ++i)
                for (int a = 0; a < 4; ++a)
                    Sk[a * n + i] = k == 0 ? st[a * n + i] : st[a * n + i] + c * inc[((k - 1) * 4 + a) * n + i];
            fit_force(n, Sk, Sk + n, Sk + 2 * n, Sk + 3 * n, q, m, kappa, confine, gamma, F, F + n);
            for (int64_t i = 0; i < n; ++i) {
                inc[(k * 4 + 0) * n + i] = dt * Sk[2 * n + i];    /* l  -> x  */
                inc[(k * 4 + 1) * n + i] = dt * Sk[3 * n + i];    /* n  -> y  */
                inc[(k * 4 + 2) * n + i] = dt * F[i] / m[i];      /* k  -> vx */
                inc[(k * 4 + 3) * n + i] = dt * F[n + i] / m[i];  /* mm -> vy */
            }
        }
        for (int64_t i = 0; i < n; ++i)
            for (int a = 0; a < 4; ++a)
                st[a * n + i] += rk_da(inc[(0 * 4 + a) * n + i], inc[(1 * 4 + a) * n + i],
                                       inc[(2 * 4 + a) * n + i], inc[(3 * 4 + a) * n + i]);
        if ((s + 1) % obs_every == 0 && (s + 1) / obs_every <= n_obs) {
            const int64_t o = (s + 1) / obs_every - 1;
            for (int64_t i = 0; i < n; ++i) {
                const double ex = st[i] - tx[o * n + i], ey = st[n + i] - ty[o * n + i];
                L += ex * ex + ey * ey;
                seed[(o * 2) * n + i] = 2.0 * ex * norm;
                seed[(o * 2 + 1) * n + i] = 2.0 * ey * norm;
            }
        }
    }
    *loss = L * norm;
    for (int64_t i = 0; i < n; ++i)
        gq[i] = 0.0;
    double gk = 0.0, gc = 0.0, gg = 0.0;
    const double far2 = fit_far2(kappa);
    const double wk[4] = {1.0 / 6.0, 2.0 / 6.0, 2.0 / 6.0, 1.0 / 6.0};
    for (int64_t s = steps - 1; s >= 0; --s) {
        if ((s + 1) % obs_every == 0 && (s + 1) / obs_every <= n_obs) {
            const int64_t o = (s + 1) / obs_every - 1;
            for (int64_t i = 0; i < n; ++i) {
                bar[i] += seed[(o * 2) * n + i];
                bar[n + i] += seed[(o * 2 + 1) * n + i];
            }
        }
        /* combine: x' = x + sum_k w_k inc_k */
        for (int k = 0; k < 4; ++k)
            for (int64_t e = 0; e < 4 * n; ++e)
                ibar[k * 4 * n + e] = wk[k] * bar[e];
        const double *Ss = S + per_stage * 4 * (size_t)s;
        const double dt = dts[s];
        for (int k = 3; k >= 0; --k) {
            const double *Sk = Ss + per_stage * k;
            const double *xs = Sk, *ys = Sk + n, *vxs = Sk + 2 * n, *vys = Sk + 3 * n;
            const double *lb = ibar + (k * 4 + 0) * n, *nb = ibar + (k * 4 + 1) * n;
            const double *kb = ibar + (k * 4 + 2) * n, *mb = ibar + (k * 4 + 3) * n;
            /* inc = dt f(S_k):  l = dt vx, k = dt F_x / m  ->  stage adjoint */
#pragma omp parallel for schedule(static) reduction(+ : gk, gc, gg)
            for (int64_t i = 0; i < n; ++i) {
                const double ux = kb[i] / m[i], uy = mb[i] / m[i];
                double rx = confine * q[i] * ux, ry = confine * q[i] * uy;
                double dq = confine * (ux * xs[i] + uy * ys[i]);
                double dk = 0.0;
                for (int64_t j = 0; j < n; ++j) {
                    if (j == i)
                        continue;
                    const double dX = xs[i] - xs[j], dY = ys[i] - ys[j];
                    if (dX * dX + dY * dY > far2)
                        continue;
                    const double r = sqrt(dX * dX + dY * dY);
                    const double u = r * kappa;
                    if (!(u < 10.0))
                        continue;
                    const double e = exp(-u);
                    const double g = coul * q[j] * (1.0 + u) * e / (r * r * r);   /* phi / q_i */
                    const double phi = q[i] * g;
                    const double dphi_r = -coul * q[i] * q[j] * e * (u * u + 3.0 * u + 3.0) / (r * r * r * r * r);
                    const double dux = ux - kb[j] / m[j], duy = uy - mb[j] / m[j];
                    const double dot = dX * dux + dY * duy;
                    rx += phi * dux + dphi_r * dot * dX;
                    ry += phi * duy + dphi_r * dot * dY;
                    dq += g * dot;
                    dk += (-coul * q[i] * q[j] * kappa * e / r) * (ux * dX + uy * dY);
                }
                sbar[i] = dt * rx;
                sbar[n + i] = dt * ry;
                sbar[2 * n + i] = dt * (lb[i] - gamma * kb[i]);
                sbar[3 * n + i] = dt * (nb[i] - gamma * mb[i]);
                gq[i] += dt * dq;
                gk += dt * dk;
                gc += dt * q[i] * (ux * xs[i] + uy * ys[i]);
                gg += dt * (-m[i]) * (ux * vxs[i] + uy * vys[i]);
            }
            /* S_k = state + c_k inc_{k-1} */
            for (int64_t e = 0; e < 4 * n; ++e)
                bar[e] += sbar[e];
            if (k > 0) {
                const double c = (k == 3) ? 1.0 : 0.5;
                for (int64_t e = 0; e < 4 * n; ++e)
                    ibar[(k - 1) * 4 * n + e] += c * sbar[e];
            }
        }
    }
    *gkappa = gk;
    *gconfine = gc;
    *ggamma = gg;
    if (gx0) memcpy(gx0, bar, sizeof(double) * n);
    if (gy0) memcpy(gy0, bar + n, sizeof(double) * n);
    if (gvx0) memcpy(gvx0, bar + 2 * n, sizeof(double) * n);
    if (gvy0) memcpy(gvy0, bar + 3 * n, sizeof(double) * n);
    free(S); free(st); free(F); free(inc); free(seed); free(bar); free(ibar); free(sbar); free(dts);
    return 0;
}
