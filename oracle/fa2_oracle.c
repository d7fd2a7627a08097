/*
 * ORACLE — test infrastructure only.
 *
 * A CPU restatement, in plain C, of the ForceAtlas2 algorithm of the
 * reference R package (forceatlas2r 0.3.0). Only tests/, the smoke check of
 * __graft_entry__.py and the cpu_baseline leg of bench.py may load it, and
 * only as the checker (or as the timed CPU baseline). The product library
 * (forceatlas2_rc_amd/csrc) never links or calls it.
 *
 * Two algorithms are restated:
 *
 *  1. The CPU layout of C_forceatlas2 (src/forceatlas2.c:521-827), in FP64:
 *     Gephi quadtree (tree_build, src/forceatlas2.c:64-239), exact repulsion
 *     per node and symmetric (:278-318), Barnes-Hut with a stack (:326-368),
 *     gravity (:371-381), attraction (:385-441), adaptive speed (:445-485),
 *     swinging/traction and displacement (:758-802).  This is the FP64 oracle
 *     and the CPU baseline (OpenMP, results independent of the thread count).
 *
 *  2. The all-GPU pipeline (inst/opencl/pipeline.cl, host order
 *     src/gpu_full.c:164-532): bounding box, 16-bit-per-axis Morton codes,
 *     stable LSD radix sort, Karras tree, bottom-up summaries, stack
 *     Barnes-Hut with the left child first, CSR attraction, block swing sums,
 *     displacement.  It is instantiated three times (large_body.h):
 *       _f  float everywhere        (the reference's precision = "single")
 *       _d  double everywhere       (the reference's precision = "double",
 *           with each centre of mass kept in its bounding box)
 *       _m  codes, tree, summaries, region records and the Barnes-Hut
 *           acceptance test in float from float32 positions, the force
 *           terms, swing, speed and displacement in FP64 with FP64 position
 *           accumulators (the FP64 restatement of the precision model of the
 *           HIP single-precision path, used as its parity gate: both take
 *           the same regions as monopoles, so they differ by the rounding of
 *           the terms only).
 *
 * Compiled with -O2 -ffp-contract=off and without fast-math, so that every
 * operation is an IEEE-754 round-to-nearest operation as written.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "fa2_oracle.h"

#define ORC_LEAF_SIZE 8

double orc_now(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double) ts.tv_sec + 1e-9 * (double) ts.tv_nsec;
}

int orc_has_openmp(void)
{
#ifdef _OPENMP
    return 1;
#else
    return 0;
#endif
}

/* ------------------------------------------------------------------------ */
/* Adaptive speed: src/forceatlas2.c:445-485                                 */
/* ------------------------------------------------------------------------ */

void orc_update_speed(int n, double sw, double tr, double jt_user,
                      double *speed_io, double *eff_io)
{
    double speed = *speed_io, eff = *eff_io;
    if (!(sw > 0)) return; /* nothing swings: speed unchanged (:452) */

    const double est = 0.05 * sqrt((double) n);
    const double lo = sqrt(est), hi = 10.0;
    double jt = est * tr / ((double) n * (double) n);
    if (jt > hi) jt = hi;
    if (jt < lo) jt = lo;
    jt *= jt_user;

    if (sw > 2.0 * tr) { /* :464-467 */
        if (eff > 0.05) eff *= 0.5;
        if (jt < jt_user) jt = jt_user;
    }
    const double target = jt * eff * tr / sw; /* :469-470 */
    if (sw > jt * tr) {                       /* :472-476 */
        if (eff > 0.05) eff *= 0.7;
    } else if (speed < 1000) {
        eff *= 1.3;
    }
    const double rise = target - speed; /* :479-481 */
    speed += rise < 0.5 * speed ? rise : 0.5 * speed;
    *speed_io = speed;
    *eff_io = eff;
}

/* ------------------------------------------------------------------------ */
/* Gephi quadtree: src/forceatlas2.c:42-239                                  */
/* ------------------------------------------------------------------------ */

typedef struct {
    double mass, cx, cy, size;
    int start, count, child, nchild, skip, buf;
} g_region;

typedef struct {
    double x, y, m;
    int idx;
} g_node;

typedef struct {
    int n;
    g_region *reg;
    int nreg;
    int *perm, *cnt;
    g_node *buf[2];
    double *px, *py, *pm, *ps;
} g_tree;

static int g_quadrant(const g_node *nd, double cx, double cy)
{
    return (nd->x < cx ? 0 : 2) + (nd->y < cy ? 0 : 1); /* :43 */
}

static int g_one_quadrant(const int *c, int count)
{
    return c[0] == count || c[1] == count || c[2] == count || c[3] == count;
}

static int g_tree_alloc(g_tree *t, int n, int with_size)
{
    memset(t, 0, sizeof(*t));
    t->n = n;
    const size_t nn = n > 0 ? (size_t) n : 1;
    t->reg = (g_region *) malloc(2 * nn * sizeof(g_region));
    t->perm = (int *) malloc(nn * sizeof(int));
    t->cnt = (int *) malloc(4 * 2 * nn * sizeof(int));
    t->buf[0] = (g_node *) malloc(nn * sizeof(g_node));
    t->buf[1] = (g_node *) malloc(nn * sizeof(g_node));
    t->px = (double *) malloc(nn * sizeof(double));
    t->py = (double *) malloc(nn * sizeof(double));
    t->pm = (double *) malloc(nn * sizeof(double));
    t->ps = with_size ? (double *) malloc(nn * sizeof(double)) : NULL;
    return !t->reg || !t->perm || !t->cnt || !t->buf[0] || !t->buf[1] ||
           !t->px || !t->py || !t->pm || (with_size && !t->ps);
}

static void g_tree_free(g_tree *t)
{
    free(t->reg);
    free(t->perm);
    free(t->cnt);
    free(t->buf[0]);
    free(t->buf[1]);
    free(t->px);
    free(t->py);
    free(t->pm);
    free(t->ps);
}

/* Breadth-first construction, one level at a time (:64-239). Regions of one
 * level are independent; the result does not depend on `threads`. */
static void g_tree_build(g_tree *t, int n, const double *x, const double *y,
                         const double *mass, const double *size, int threads)
{
    g_region *R = t->reg;
    (void) threads;
    {
        g_node *nd = t->buf[0];
        double m = 0.0, sx = 0.0, sy = 0.0;
        for (int i = 0; i < n; i++) {
            nd[i].x = x[i];
            nd[i].y = y[i];
            nd[i].m = mass[i];
            nd[i].idx = i;
            m += mass[i];
            sx += x[i] * mass[i];
            sy += y[i] * mass[i];
        }
        R[0].start = 0;
        R[0].count = n;
        R[0].skip = -1;
        R[0].buf = 0;
        R[0].child = -1;
        R[0].nchild = 0;
        R[0].mass = n > 1 ? m : mass[0];
        R[0].cx = n > 1 ? sx / m : x[0];
        R[0].cy = n > 1 ? sy / m : y[0];
        R[0].size = 0.0;
        t->nreg = 1;
    }

    int lo = 0, hi = 1;
    while (lo < hi) {
        /* (1) size of each region and quadrant populations (:106-126) */
#ifdef _OPENMP
#pragma omp parallel for num_threads(threads) schedule(dynamic, 16) if (threads > 1 && hi - lo > 1)
#endif
        for (int r = lo; r < hi; r++) {
            int *c = t->cnt + 4 * (size_t) (r - lo);
            c[0] = c[1] = c[2] = c[3] = 0;
            const int count = R[r].count;
            if (count < 2) continue;
            const g_node *nd = t->buf[R[r].buf] + R[r].start;
            const double cx = R[r].cx, cy = R[r].cy;
            double d2max = 0.0;
            for (int k = 0; k < count; k++) {
                const double dx = nd[k].x - cx, dy = nd[k].y - cy;
                const double d2 = dx * dx + dy * dy;
                if (d2 > d2max) d2max = d2;
                c[g_quadrant(&nd[k], cx, cy)]++;
            }
            R[r].size = 2.0 * sqrt(d2max);
        }

        /* (2) regions of the next level, in order (:129-164) */
        for (int r = lo; r < hi; r++) {
            const int *c = t->cnt + 4 * (size_t) (r - lo);
            const int start = R[r].start, count = R[r].count;
            if (count <= ORC_LEAF_SIZE) continue;
            R[r].child = t->nreg;
            if (g_one_quadrant(c, count)) {
                /* coincident nodes: one single-node leaf each, in place */
                R[r].nchild = count;
                for (int k = start; k < start + count; k++) {
                    g_region *ch = &R[t->nreg];
                    ch->start = k;
                    ch->count = 1;
                    ch->skip = k == start ? R[r].skip : t->nreg - 1;
                    ch->buf = R[r].buf;
                    ch->child = -1;
                    ch->nchild = 0;
                    t->nreg++;
                }
                continue;
            }
            int off = start;
            for (int q = 0; q < 4; q++) {
                if (c[q] == 0) continue;
                g_region *ch = &R[t->nreg];
                ch->start = off;
                ch->count = c[q];
                ch->skip = R[r].nchild == 0 ? R[r].skip : t->nreg - 1;
                ch->buf = 1 - R[r].buf;
                ch->child = -1;
                ch->nchild = 0;
                R[r].nchild++;
                t->nreg++;
                off += c[q];
            }
        }

        /* (3) move nodes into their quadrant, child masses and centres
         * (:166-216) */
#ifdef _OPENMP
#pragma omp parallel for num_threads(threads) schedule(dynamic, 16) if (threads > 1 && hi - lo > 1)
#endif
        for (int r = lo; r < hi; r++) {
            if (R[r].nchild == 0) continue;
            const int *c = t->cnt + 4 * (size_t) (r - lo);
            const int count = R[r].count;
            const g_node *nd = t->buf[R[r].buf] + R[r].start;
            if (!g_one_quadrant(c, count)) {
                g_node *out = t->buf[1 - R[r].buf];
                const double cx = R[r].cx, cy = R[r].cy;
                double qm[4] = {0, 0, 0, 0}, qx[4] = {0, 0, 0, 0},
                       qy[4] = {0, 0, 0, 0};
                int at[4];
                at[0] = R[r].start;
                for (int q = 1; q < 4; q++) at[q] = at[q - 1] + c[q - 1];
                for (int k = 0; k < count; k++) {
                    const int q = g_quadrant(&nd[k], cx, cy);
                    out[at[q]++] = nd[k];
                    qm[q] += nd[k].m;
                    qx[q] += nd[k].x * nd[k].m;
                    qy[q] += nd[k].y * nd[k].m;
                }
                g_region *ch = &R[R[r].child];
                for (int q = 0; q < 4; q++) {
                    if (c[q] == 0) continue;
                    ch->mass = qm[q];
                    ch->cx = qx[q] / qm[q];
                    ch->cy = qy[q] / qm[q];
                    ch->size = 0.0;
                    ch++;
                }
                nd = out + R[r].start;
            }
            for (int j = 0; j < R[r].nchild; j++) {
                g_region *ch = &R[R[r].child + j];
                if (ch->count > 1) continue;
                const g_node *one = nd + (ch->start - R[r].start);
                ch->mass = one->m;
                ch->cx = one->x;
                ch->cy = one->y;
                ch->size = 0.0;
            }
        }
        lo = hi;
        hi = t->nreg;
    }

    /* nodes in tree order (:222-238) */
    for (int r = 0; r < t->nreg; r++) {
        if (R[r].nchild > 0) continue;
        const g_node *nd = t->buf[R[r].buf];
        for (int k = R[r].start; k < R[r].start + R[r].count; k++) {
            const int i = nd[k].idx;
            t->perm[k] = i;
            t->px[k] = nd[k].x;
            t->py[k] = nd[k].y;
            t->pm[k] = nd[k].m;
            if (size) t->ps[k] = size[i];
        }
    }
}

int orc_gephi_tree(int n, const double *x, const double *y, const double *mass,
                   const double *size, int *nreg_out, double *reg_d,
                   int *reg_i, int *perm, double *px, double *py, double *pm,
                   double *ps)
{
    g_tree t;
    if (n < 1) return 1;
    if (g_tree_alloc(&t, n, size != NULL)) {
        g_tree_free(&t);
        return 1;
    }
    g_tree_build(&t, n, x, y, mass, size, 1);
    *nreg_out = t.nreg;
    for (int r = 0; r < t.nreg; r++) {
        reg_d[4 * r] = t.reg[r].mass;
        reg_d[4 * r + 1] = t.reg[r].cx;
        reg_d[4 * r + 2] = t.reg[r].cy;
        reg_d[4 * r + 3] = t.reg[r].size;
        reg_i[5 * r] = t.reg[r].start;
        reg_i[5 * r + 1] = t.reg[r].count;
        reg_i[5 * r + 2] = t.reg[r].child;
        reg_i[5 * r + 3] = t.reg[r].nchild;
        reg_i[5 * r + 4] = t.reg[r].skip;
    }
    memcpy(perm, t.perm, (size_t) n * sizeof(int));
    memcpy(px, t.px, (size_t) n * sizeof(double));
    memcpy(py, t.py, (size_t) n * sizeof(double));
    memcpy(pm, t.pm, (size_t) n * sizeof(double));
    if (size && ps) memcpy(ps, t.ps, (size_t) n * sizeof(double));
    g_tree_free(&t);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* CPU forces: src/forceatlas2.c:245-441                                     */
/* ------------------------------------------------------------------------ */

typedef struct {
    int n, m;
    double *x, *y, *dx, *dy;
    const double *mass, *size, *w;
    const int *from, *to;
    double scaling, gravity, theta;
    int strong, linlog, outbound, adjust;
    const int *inc_off, *inc;
} c_state;

/* :265-275; symmetric in the two nodes */
static inline double c_rep(const c_state *s, double d2, double mm, double ss)
{
    if (s->adjust) {
        const double d = sqrt(d2) - ss;
        if (d > 0) return s->scaling * mm / (d * d);
        if (d < 0) return 100.0 * s->scaling * mm;
        return 0.0;
    }
    return d2 > 0 ? s->scaling * mm / d2 : 0.0;
}

/* :278-294 */
static void c_repulse_node_exact(const c_state *s, int i)
{
    const double xi = s->x[i], yi = s->y[i], mi = s->mass[i];
    const double si = s->adjust ? s->size[i] : 0.0;
    double fx = 0.0, fy = 0.0;
    for (int j = 0; j < s->n; j++) {
        if (j == i) continue;
        const double xd = xi - s->x[j], yd = yi - s->y[j];
        const double f = c_rep(s, xd * xd + yd * yd, mi * s->mass[j],
                               s->adjust ? si + s->size[j] : 0.0);
        fx += xd * f;
        fy += yd * f;
    }
    s->dx[i] += fx;
    s->dy[i] += fy;
}

/* :299-318: same additions in the same order as the per-node form */
static void c_repulse_all_symmetric(const c_state *s)
{
    for (int i = 0; i < s->n; i++) {
        const double xi = s->x[i], yi = s->y[i], mi = s->mass[i];
        const double si = s->adjust ? s->size[i] : 0.0;
        double fx = 0.0, fy = 0.0;
        for (int j = 0; j < i; j++) {
            const double xd = xi - s->x[j], yd = yi - s->y[j];
            const double f = c_rep(s, xd * xd + yd * yd, mi * s->mass[j],
                                   s->adjust ? si + s->size[j] : 0.0);
            fx += xd * f;
            fy += yd * f;
            s->dx[j] -= xd * f;
            s->dy[j] -= yd * f;
        }
        s->dx[i] += fx;
        s->dy[i] += fy;
    }
}

/* :326-368: explicit stack; the last child is popped first */
static void c_repulse_node_bh(const c_state *s, const g_tree *t, int i,
                              int *stack)
{
    const double xi = s->x[i], yi = s->y[i], mi = s->mass[i];
    const double si = s->adjust ? s->size[i] : 0.0;
    const double th2 = s->theta * s->theta;
    double fx = 0.0, fy = 0.0;
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        const g_region *r = &t->reg[stack[--sp]];
        if (r->count > 1) {
            const double xd = xi - r->cx, yd = yi - r->cy;
            const double d2 = xd * xd + yd * yd;
            if (d2 * th2 > r->size * r->size) {
                const double f = s->scaling * (mi * r->mass) / d2;
                fx += xd * f;
                fy += yd * f;
                continue;
            }
            if (r->nchild > 0) {
                for (int c = 0; c < r->nchild; c++) stack[sp++] = r->child + c;
                continue;
            }
        }
        for (int k = r->start; k < r->start + r->count; k++) {
            if (t->perm[k] == i) continue;
            const double xd = xi - t->px[k], yd = yi - t->py[k];
            const double f = c_rep(s, xd * xd + yd * yd, mi * t->pm[k],
                                   s->adjust ? si + t->ps[k] : 0.0);
            fx += xd * f;
            fy += yd * f;
        }
    }
    s->dx[i] += fx;
    s->dy[i] += fy;
}

/* :371-381 */
static inline void c_gravity(const c_state *s, int i)
{
    const double xi = s->x[i], yi = s->y[i];
    const double dist = sqrt(xi * xi + yi * yi);
    if (dist > 0) {
        const double f = s->strong ? s->gravity * s->mass[i]
                                   : s->gravity * s->mass[i] / dist;
        s->dx[i] -= xi * f;
        s->dy[i] -= yi * f;
    }
}

/* :385-399 */
static inline double c_attr_factor(const c_state *s, int e, double coef,
                                   double xd, double yd)
{
    const int a = s->from[e], b = s->to[e];
    double f = -coef * s->w[e];
    if (s->adjust || s->linlog) {
        double dist = sqrt(xd * xd + yd * yd);
        if (s->adjust) dist -= s->size[a] + s->size[b];
        if (!(dist > 0)) return 0.0;
        if (s->linlog) f *= log1p(dist) / dist;
    }
    if (s->outbound) f /= s->mass[a];
    return f;
}

/* :402-416 */
static void c_attract_all(const c_state *s, double coef)
{
    for (int e = 0; e < s->m; e++) {
        const int a = s->from[e], b = s->to[e];
        if (a == b) continue;
        const double xd = s->x[a] - s->x[b], yd = s->y[a] - s->y[b];
        const double f = c_attr_factor(s, e, coef, xd, yd);
        if (f == 0.0) continue;
        s->dx[a] += xd * f;
        s->dy[a] += yd * f;
        s->dx[b] -= xd * f;
        s->dy[b] -= yd * f;
    }
}

/* :422-441 */
static void c_attract_node(const c_state *s, double coef, int i)
{
    double fx = s->dx[i], fy = s->dy[i];
    for (int k = s->inc_off[i]; k < s->inc_off[i + 1]; k++) {
        const int e = s->inc[k] >> 1;
        const int a = s->from[e], b = s->to[e];
        const double xd = s->x[a] - s->x[b], yd = s->y[a] - s->y[b];
        const double f = c_attr_factor(s, e, coef, xd, yd);
        if (f == 0.0) continue;
        if (s->inc[k] & 1) {
            fx -= xd * f;
            fy -= yd * f;
        } else {
            fx += xd * f;
            fy += yd * f;
        }
    }
    s->dx[i] = fx;
    s->dy[i] = fy;
}

/*
 * The CPU layout of C_forceatlas2 (src/forceatlas2.c:521-827), with plain C
 * arrays in place of R objects. `pos` holds x[0..n) then y[0..n) and is
 * updated in place.
 * dpar  = (scaling, gravity, jitter_tolerance, theta)
 * flags = (strong, linlog, outbound, adjust, barnes_hut)
 * trace_pos (optional)  : the positions after each iteration, 2n per iteration
 * trace_force (optional): the forces (dx then dy) applied at each iteration
 * times (optional, 7)   : total, tree, repulsion, 0, 0, 0, other
 * Returns 0, or 1 on invalid input or out of memory.
 */
int orc_layout(int n, int m, const int *from, const int *to, const double *w,
               double *pos, const double *size, const int *fixed,
               int iterations, const double *dpar, const int *flags,
               int threads, double *trace_pos, double *trace_force,
               double *times)
{
    if (times)
        for (int k = 0; k < 7; k++) times[k] = 0.0;
    if (n < 0 || m < 0 || iterations < 0) return 1;
    if (m > (1 << 30) - 1) return 1;
    for (int e = 0; e < m; e++)
        if (from[e] < 0 || from[e] >= n || to[e] < 0 || to[e] >= n) return 1;
    if (threads < 1) threads = 1;
#ifndef _OPENMP
    threads = 1;
#endif
    if (n == 0 || iterations == 0) return 0;

    c_state s;
    memset(&s, 0, sizeof(s));
    s.n = n;
    s.m = m;
    s.x = pos;
    s.y = pos + n;
    s.from = from;
    s.to = to;
    s.w = w;
    s.size = size;
    s.scaling = dpar[0];
    s.gravity = dpar[1];
    const double jt_user = dpar[2];
    s.theta = dpar[3];
    s.strong = flags[0];
    s.linlog = flags[1];
    s.outbound = flags[2];
    s.adjust = flags[3];
    const int barnes_hut = flags[4];

    const size_t nn = (size_t) n;
    double *mass = (double *) malloc(nn * sizeof(double));
    double *odx = (double *) calloc(nn, sizeof(double));
    double *ody = (double *) calloc(nn, sizeof(double));
    double *dx = (double *) calloc(nn, sizeof(double));
    double *dy = (double *) calloc(nn, sizeof(double));
    double *swing = (double *) malloc(nn * sizeof(double));
    double *tract = (double *) malloc(nn * sizeof(double));
    int *off = NULL, *inc = NULL, *stacks = NULL;
    g_tree tree;
    memset(&tree, 0, sizeof(tree));
    int status = 1;
    if (!mass || !odx || !ody || !dx || !dy || !swing || !tract) goto out;
    s.dx = dx;
    s.dy = dy;
    s.mass = mass;

    /* mass = 1 + degree, self-loops counted twice (:592-603) */
    double mean_mass = 0.0;
    for (int i = 0; i < n; i++) mass[i] = 1.0;
    for (int e = 0; e < m; e++) {
        mass[from[e]] += 1.0;
        mass[to[e]] += 1.0;
    }
    for (int i = 0; i < n; i++) mean_mass += mass[i];
    mean_mass /= n;

    /* incidence lists for the node-parallel attraction (:607-631) */
    if (threads > 1 && m > 0) {
        off = (int *) calloc(nn + 1, sizeof(int));
        inc = (int *) malloc(2 * (size_t) m * sizeof(int));
        if (!off || !inc) goto out;
        for (int e = 0; e < m; e++) {
            if (from[e] == to[e]) continue;
            off[from[e] + 1]++;
            off[to[e] + 1]++;
        }
        for (int i = 0; i < n; i++) off[i + 1] += off[i];
        for (int e = 0; e < m; e++) {
            if (from[e] == to[e]) continue;
            inc[off[from[e]]++] = 2 * e;
            inc[off[to[e]]++] = 2 * e + 1;
        }
        for (int i = n; i > 0; i--) off[i] = off[i - 1];
        off[0] = 0;
        s.inc_off = off;
        s.inc = inc;
    }
    const size_t stack_len = 2 * nn;
    if (barnes_hut) {
        if (g_tree_alloc(&tree, n, s.adjust)) goto out;
        stacks = (int *) malloc(stack_len * (size_t) threads * sizeof(int));
        if (!stacks) goto out;
    }

    const double coef = s.outbound ? mean_mass : 1.0;
    double speed = 1.0, eff = 1.0;
    double t_tree = 0.0, t_rep = 0.0;
    const double t_start = orc_now();

    for (int it = 0; it < iterations; it++) {
        /* keep the previous forces, start from zero (:683-693) */
        double *tmp = odx;
        odx = s.dx;
        s.dx = tmp;
        tmp = ody;
        ody = s.dy;
        s.dy = tmp;
        memset(s.dx, 0, nn * sizeof(double));
        memset(s.dy, 0, nn * sizeof(double));

        double t0 = orc_now();
        if (barnes_hut) {
            g_tree_build(&tree, n, s.x, s.y, mass, s.adjust ? size : NULL,
                         threads);
            const double t1 = orc_now();
            t_tree += t1 - t0;
            t0 = t1;
#ifdef _OPENMP
#pragma omp parallel for num_threads(threads) schedule(dynamic, 64) if (threads > 1)
#endif
            for (int k = 0; k < n; k++) {
#ifdef _OPENMP
                int *stk = stacks + stack_len * (size_t) omp_get_thread_num();
#else
                int *stk = stacks;
#endif
                c_repulse_node_bh(&s, &tree, tree.perm[k], stk);
            }
        } else if (threads > 1) {
#ifdef _OPENMP
#pragma omp parallel for num_threads(threads) schedule(static)
#endif
            for (int i = 0; i < n; i++) c_repulse_node_exact(&s, i);
        } else {
            c_repulse_all_symmetric(&s);
        }
        t_rep += orc_now() - t0;

        if (s.gravity != 0) {
#ifdef _OPENMP
#pragma omp parallel for num_threads(threads) schedule(static) if (threads > 1)
#endif
            for (int i = 0; i < n; i++) c_gravity(&s, i);
        }
        if (s.inc) {
#ifdef _OPENMP
#pragma omp parallel for num_threads(threads) schedule(dynamic, 256)
#endif
            for (int i = 0; i < n; i++) c_attract_node(&s, coef, i);
        } else {
            c_attract_all(&s, coef);
        }

#ifdef _OPENMP
#pragma omp parallel for num_threads(threads) schedule(static) if (threads > 1)
#endif
        for (int i = 0; i < n; i++) {
            const double sx = odx[i] - s.dx[i], sy = ody[i] - s.dy[i];
            const double tx = odx[i] + s.dx[i], ty = ody[i] + s.dy[i];
            swing[i] = mass[i] * sqrt(sx * sx + sy * sy);
            tract[i] = 0.5 * mass[i] * sqrt(tx * tx + ty * ty);
        }
        double tsw = 0.0, ttr = 0.0;
        for (int i = 0; i < n; i++) {
            if (fixed && fixed[i]) continue;
            tsw += swing[i];
            ttr += tract[i];
        }
        orc_update_speed(n, tsw, ttr, jt_user, &speed, &eff);

        if (trace_force) {
            memcpy(trace_force + 2 * nn * (size_t) it, s.dx, nn * sizeof(double));
            memcpy(trace_force + 2 * nn * (size_t) it + nn, s.dy,
                   nn * sizeof(double));
        }

#ifdef _OPENMP
#pragma omp parallel for num_threads(threads) schedule(static) if (threads > 1)
#endif
        for (int i = 0; i < n; i++) {
            if (fixed && fixed[i]) continue;
            double factor;
            if (s.adjust) {
                const double df = sqrt(s.dx[i] * s.dx[i] + s.dy[i] * s.dy[i]);
                factor = 0.1 * speed / (1.0 + sqrt(speed * swing[i]));
                if (df > 0) {
                    if (factor * df > 10.0) factor = 10.0 / df;
                } else {
                    factor = 0.0;
                }
            } else {
                factor = speed / (1.0 + sqrt(speed * swing[i]));
            }
            s.x[i] += s.dx[i] * factor;
            s.y[i] += s.dy[i] * factor;
        }
        if (trace_pos)
            memcpy(trace_pos + 2 * nn * (size_t) it, pos, 2 * nn * sizeof(double));
    }
    if (times) {
        times[0] = orc_now() - t_start;
        times[1] = t_tree;
        times[2] = t_rep;
        times[6] = times[0] - t_tree - t_rep;
    }
    status = 0;
out:
    /* the force buffers were swapped at every iteration */
    free(mass);
    free(odx);
    free(ody);
    free(s.dx ? s.dx : dx);
    free(s.dy ? s.dy : dy);
    free(swing);
    free(tract);
    free(off);
    free(inc);
    free(stacks);
    g_tree_free(&tree);
    return status;
}

/* One evaluation of the CPU repulsion (exact, per node: :278-294), added to
 * dx, dy. */
void orc_repulse_exact(int n, const double *x, const double *y,
                       const double *mass, const double *size, double scaling,
                       int threads, double *dx, double *dy)
{
    c_state s;
    memset(&s, 0, sizeof(s));
    s.n = n;
    s.x = (double *) x;
    s.y = (double *) y;
    s.mass = mass;
    s.size = size;
    s.adjust = size != NULL;
    s.scaling = scaling;
    s.dx = dx;
    s.dy = dy;
    (void) threads;
#ifdef _OPENMP
#pragma omp parallel for num_threads(threads) schedule(static) if (threads > 1)
#endif
    for (int i = 0; i < n; i++) c_repulse_node_exact(&s, i);
}

/* Barnes-Hut repulsion over a Gephi tree built from (x, y, mass, size): the
 * forces of repulse_node_bh (:326-368), added to dx, dy (original order). */
int orc_repulse_bh(int n, const double *x, const double *y, const double *mass,
                   const double *size, double scaling, double theta,
                   int threads, double *dx, double *dy)
{
    g_tree t;
    if (g_tree_alloc(&t, n, size != NULL)) {
        g_tree_free(&t);
        return 1;
    }
    int *stacks = (int *) malloc(2 * (size_t) n * (size_t) (threads > 1 ? threads : 1) * sizeof(int));
    if (!stacks) {
        g_tree_free(&t);
        return 1;
    }
    g_tree_build(&t, n, x, y, mass, size, threads);
    c_state s;
    memset(&s, 0, sizeof(s));
    s.n = n;
    s.x = (double *) x;
    s.y = (double *) y;
    s.mass = mass;
    s.size = size;
    s.adjust = size != NULL;
    s.scaling = scaling;
    s.theta = theta;
    s.dx = dx;
    s.dy = dy;
    (void) threads;
#ifdef _OPENMP
#pragma omp parallel for num_threads(threads) schedule(dynamic, 64) if (threads > 1)
#endif
    for (int k = 0; k < n; k++) {
#ifdef _OPENMP
        int *stk = stacks + 2 * (size_t) n * (size_t) omp_get_thread_num();
#else
        int *stk = stacks;
#endif
        c_repulse_node_bh(&s, &t, t.perm[k], stk);
    }
    free(stacks);
    g_tree_free(&t);
    return 0;
}

/* The FP64 model of the single-precision repulsion offload over the same
 * Gephi tree (the GPU kernel fa2_bh, src/gpu.c:82-123, with the records of
 * src/gpu.c:501-519): positions, masses, sizes and region records are
 * rounded to float32 as the device receives them (cx, cy, mass and
 * size * size each rounded once); the acceptance test d^2 theta^2 > size^2 is
 * taken in float32 with each operation rounded on its own (FP_CONTRACT OFF,
 * src/gpu.c:35), so the model approximates exactly the regions the float32
 * kernel approximates; every force term is then computed in FP64 from those
 * float32 values. Node k of the tree order meets its regions in the order of
 * the stack traversal (last child first). */
static void c_repulse_node_bh_single(const c_state *s, const g_tree *t, int k, int *stack)
{
    const float xf = (float) t->px[k], yf = (float) t->py[k];
    const double xi = (double) xf, yi = (double) yf;
    const double mi = (double) (float) t->pm[k];
    const double si = s->adjust ? (double) (float) t->ps[k] : 0.0;
    const float th2 = (float) (s->theta * s->theta);
    double fx = 0.0, fy = 0.0;
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        const g_region *r = &t->reg[stack[--sp]];
        if (r->count > 1) {
            const float cxf = (float) r->cx, cyf = (float) r->cy;
            const float txd = xf - cxf, tyd = yf - cyf;
            if ((txd * txd + tyd * tyd) * th2 > (float) (r->size * r->size)) {
                const double xd = xi - (double) cxf, yd = yi - (double) cyf;
                const double f = s->scaling * (mi * (double) (float) r->mass) / (xd * xd + yd * yd);
                fx += xd * f;
                fy += yd * f;
                continue;
            }
            if (r->nchild > 0) {
                for (int c = 0; c < r->nchild; c++) stack[sp++] = r->child + c;
                continue;
            }
        }
        for (int j = r->start; j < r->start + r->count; j++) {
            if (j == k) continue;
            const double xd = xi - (double) (float) t->px[j];
            const double yd = yi - (double) (float) t->py[j];
            const double f = c_rep(s, xd * xd + yd * yd, mi * (double) (float) t->pm[j],
                                   s->adjust ? si + (double) (float) t->ps[j] : 0.0);
            fx += xd * f;
            fy += yd * f;
        }
    }
    s->dx[t->perm[k]] += fx;
    s->dy[t->perm[k]] += fy;
}

int orc_repulse_bh_single(int n, const double *x, const double *y, const double *mass,
                          const double *size, double scaling, double theta, double *dx,
                          double *dy)
{
    g_tree t;
    if (g_tree_alloc(&t, n, size != NULL)) {
        g_tree_free(&t);
        return 1;
    }
    int *stack = (int *) malloc(4 * (size_t) n * sizeof(int) + 64);
    if (!stack) {
        g_tree_free(&t);
        return 1;
    }
    g_tree_build(&t, n, x, y, mass, size, 1);
    c_state s;
    memset(&s, 0, sizeof(s));
    s.n = n;
    s.adjust = size != NULL;
    s.scaling = scaling;
    s.theta = theta;
    s.dx = dx;
    s.dy = dy;
    for (int k = 0; k < n; k++) c_repulse_node_bh_single(&s, &t, k, stack);
    free(stack);
    g_tree_free(&t);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* All-GPU pipeline restated on the CPU: inst/opencl/pipeline.cl             */
/* ------------------------------------------------------------------------ */

/* Insert a zero bit before each of the 16 low bits (pipeline.cl:106-114),
 * written here as a plain bit loop. */
uint32_t orc_spread16(uint32_t v)
{
    uint32_t r = 0;
    for (int b = 0; b < 16; b++) r |= ((v >> b) & 1u) << (2 * b);
    return r;
}

/* Stable LSD radix sort of (key, index) pairs, four passes of 8 bits
 * (pipeline.cl:137-188, driver src/gpu_full.c:419-436). The result is the
 * permutation of a stable sort of the indices by key. */
void orc_radix_sort(int n, const uint32_t *key_in, uint32_t *key_out,
                    uint32_t *perm)
{
    uint32_t *ka = (uint32_t *) malloc((size_t) n * sizeof(uint32_t));
    uint32_t *ia = (uint32_t *) malloc((size_t) n * sizeof(uint32_t));
    uint32_t *kb = (uint32_t *) malloc((size_t) n * sizeof(uint32_t));
    uint32_t *ib = (uint32_t *) malloc((size_t) n * sizeof(uint32_t));
    for (int i = 0; i < n; i++) {
        ka[i] = key_in[i];
        ia[i] = (uint32_t) i;
    }
    for (int pass = 0; pass < 4; pass++) {
        const int shift = 8 * pass;
        uint32_t count[257];
        memset(count, 0, sizeof(count));
        for (int i = 0; i < n; i++) count[((ka[i] >> shift) & 255u) + 1]++;
        for (int d = 0; d < 256; d++) count[d + 1] += count[d];
        for (int i = 0; i < n; i++) {
            const uint32_t p = count[(ka[i] >> shift) & 255u]++;
            kb[p] = ka[i];
            ib[p] = ia[i];
        }
        uint32_t *t = ka;
        ka = kb;
        kb = t;
        t = ia;
        ia = ib;
        ib = t;
    }
    memcpy(key_out, ka, (size_t) n * sizeof(uint32_t));
    memcpy(perm, ia, (size_t) n * sizeof(uint32_t));
    free(ka);
    free(ia);
    free(kb);
    free(ib);
}

static inline int clz32(uint32_t v) { return v ? __builtin_clz(v) : 32; }

/* pipeline.cl:211-217: common prefix of sorted keys i and j, the index
 * breaking ties, -1 out of range. */
static inline int k_delta(const uint32_t *key, int n, int i, int j)
{
    if (j < 0 || j >= n) return -1;
    const uint32_t a = key[i], b = key[j];
    if (a == b) return 32 + clz32((uint32_t) (i ^ j));
    return clz32(a ^ b);
}

/*
 * Karras (2012) binary radix tree over n sorted keys (pipeline.cl:226-270):
 * internal nodes 0..n-2 (0 = root), leaves n-1..2n-2.
 * child[2i], child[2i+1]: children; range[2i], range[2i+1]: first and last
 * sorted node; parent[c] for the 2n-1 nodes (parent[0] = -1).
 */
void orc_karras(int n, const uint32_t *key, int *child, int *range,
                int *parent)
{
    if (n >= 1) parent[0] = -1;
    for (int i = 0; i < n - 1; i++) {
        const int d = k_delta(key, n, i, i + 1) > k_delta(key, n, i, i - 1)
                          ? 1
                          : -1;
        const int dmin = k_delta(key, n, i, i - d);
        int lmax = 2;
        while (k_delta(key, n, i, i + lmax * d) > dmin) lmax *= 2;
        int l = 0;
        for (int t = lmax / 2; t >= 1; t /= 2)
            if (k_delta(key, n, i, i + (l + t) * d) > dmin) l += t;
        const int j = i + l * d;
        const int dnode = k_delta(key, n, i, j);
        int s = 0, t = l;
        do {
            t = (t + 1) >> 1;
            if (k_delta(key, n, i, i + (s + t) * d) > dnode) s += t;
        } while (t > 1);
        const int split = i + s * d + (d < 0 ? d : 0);
        const int first = i < j ? i : j, last = i < j ? j : i;
        const int left = first == split ? n - 1 + split : split;
        const int right = last == split + 1 ? n - 1 + split + 1 : split + 1;
        child[2 * i] = left;
        child[2 * i + 1] = right;
        range[2 * i] = first;
        range[2 * i + 1] = last;
        parent[left] = i;
        parent[right] = i;
    }
}

/* Block plan of src/gpu_full.c:193-198. */
static void block_plan(int n, int *B_out, int *S_out)
{
    int B = n / 256;
    if (B < 1) B = 1;
    if (B > 16384) B = 16384;
    const int S = (n + B - 1) / B;
    B = (n + S - 1) / S;
    *B_out = B;
    *S_out = S;
}

/* The three instantiations of the pipeline. TR: arithmetic of the codes;
 * FR: arithmetic of everything else; PR: type of the stored positions. */
#define TR float
#define SR float
#define FR float
#define PR float
#define SFX(name) name##_f
#include "large_body.h"
#undef TR
#undef SR
#undef FR
#undef PR
#undef SFX

#define ORC_CLAMP_CENTRE 1
#define TR double
#define SR double
#define FR double
#define PR double
#define SFX(name) name##_d
#include "large_body.h"
#undef TR
#undef SR
#undef FR
#undef PR
#undef SFX

#define ORC_MIXED 1
#define TR float
#define SR float
#define FR double
#define PR double
#define SFX(name) name##_m
#include "large_body.h"
#undef TR
#undef SR
#undef FR
#undef PR
#undef SFX
#undef ORC_MIXED
#undef ORC_CLAMP_CENTRE

/* Preprocessing of C_forceatlas2_large (src/forceatlas2.c:870-902): mass =
 * 1 + degree, incidence CSR in edge order (self-loops skipped), inc_coef =
 * coef * w (divided by the mass of the source with outbound). inc_off has
 * n + 1 entries; inc_node and inc_coef 2m. Returns the number of
 * incidences. */
int orc_large_prepare(int n, int m, const int *from, const int *to,
                      const double *w, int outbound, double *mass,
                      int *inc_off, int *inc_node, double *inc_coef)
{
    double mean = 0.0;
    for (int i = 0; i < n; i++) mass[i] = 1.0;
    for (int e = 0; e < m; e++) {
        mass[from[e]] += 1.0;
        mass[to[e]] += 1.0;
    }
    for (int i = 0; i < n; i++) mean += mass[i];
    mean /= n;
    const double coef = outbound ? mean : 1.0;
    memset(inc_off, 0, ((size_t) n + 1) * sizeof(int));
    for (int e = 0; e < m; e++) {
        if (from[e] == to[e]) continue;
        inc_off[from[e] + 1]++;
        inc_off[to[e] + 1]++;
    }
    for (int i = 0; i < n; i++) inc_off[i + 1] += inc_off[i];
    for (int e = 0; e < m; e++) {
        if (from[e] == to[e]) continue;
        const double c = outbound ? coef * w[e] / mass[from[e]] : coef * w[e];
        inc_node[inc_off[from[e]]] = to[e];
        inc_coef[inc_off[from[e]]++] = c;
        inc_node[inc_off[to[e]]] = from[e];
        inc_coef[inc_off[to[e]]++] = c;
    }
    for (int i = n; i > 0; i--) inc_off[i] = inc_off[i - 1];
    inc_off[0] = 0;
    return inc_off[n];
}
