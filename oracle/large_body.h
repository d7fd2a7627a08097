/*
 * ORACLE — test infrastructure only (included by fa2_oracle.c).
 *
 * One instantiation of the all-GPU pipeline of the reference
 * (inst/opencl/pipeline.cl, host sequence src/gpu_full.c:403-499) on the CPU.
 * Macros set by the includer:
 *   TR  arithmetic of the bounding box and the Morton codes
 *   SR  arithmetic of the tree summaries, the region records and the
 *       Barnes-Hut acceptance test
 *   FR  arithmetic of the forces, swing and displacement
 *   PR  type of the stored positions
 *   SFX(name)  suffix of the exported names
 *   ORC_MIXED  defined for the FP64 model of the HIP single-precision path:
 *              positions are FP64 accumulators; the codes, the summaries,
 *              the region records and the acceptance test are computed in
 *              float32 from their float32 roundings, as the device does, so
 *              that both approximate the same regions; the force terms are
 *              FP64.
 *   ORC_CLAMP_CENTRE  keep each centre of mass inside its bounding box (a
 *              no-op in exact arithmetic; the reference's rounded centre can
 *              leave the box of a cluster of coincident nodes).
 */

#ifdef ORC_MIXED
#define POS_FR(v) ((FR) (TR) (v))
#else
#define POS_FR(v) ((FR) (v))
#endif

#define ORC_MIN(a, b) ((b) < (a) ? (b) : (a))
#define ORC_MAX(a, b) ((b) > (a) ? (b) : (a))

/* Bounding box and Morton codes (pipeline.cl:61-130). x and y are the
 * positions; the arithmetic is TR, as in the reference kernels. */
void SFX(orc_morton)(int n, const PR *x, const PR *y, uint32_t *key)
{
    TR minx = (TR) x[0], maxx = minx, miny = (TR) y[0], maxy = miny;
    for (int i = 1; i < n; i++) {
        const TR px = (TR) x[i], py = (TR) y[i];
        minx = ORC_MIN(minx, px);
        maxx = ORC_MAX(maxx, px);
        miny = ORC_MIN(miny, py);
        maxy = ORC_MAX(maxy, py);
    }
    const TR ex = maxx - minx, ey = maxy - miny;
    const TR extent = ORC_MAX(ex, ey);
    const TR scale = extent > (TR) 0 ? (TR) 65535 / extent : (TR) 0;
    for (int i = 0; i < n; i++) {
        TR fx = ((TR) x[i] - minx) * scale;
        TR fy = ((TR) y[i] - miny) * scale;
        fx = ORC_MIN(ORC_MAX(fx, (TR) 0), (TR) 65535);
        fy = ORC_MIN(ORC_MAX(fy, (TR) 0), (TR) 65535);
        key[i] = orc_spread16((uint32_t) fx) | (orc_spread16((uint32_t) fy) << 1);
    }
}

/* Codes, sorted permutation and Karras topology from positions. */
void SFX(orc_large_tree)(int n, const PR *x, const PR *y, uint32_t *key,
                         uint32_t *key_sorted, uint32_t *order, int *child,
                         int *range, int *parent)
{
    SFX(orc_morton)(n, x, y, key);
    orc_radix_sort(n, key, key_sorted, order);
    orc_karras(n, key_sorted, child, range, parent);
}

/* Bottom-up summaries (pipeline.cl:276-305): nd[8c] = (min x, min y, max x,
 * max y, mass, sum mx, sum my) of node c; a region merges left then right. */
static void SFX(agg_node)(int n, int r, const int *child, SR *nd)
{
    /* post-order without recursion: the tree can be 64 levels deep */
    int stack[200];
    int sp = 0;
    stack[sp++] = r;
    while (sp > 0) {
        const int c = stack[sp - 1];
        const int a = child[2 * c], b = child[2 * c + 1];
        int pushed = 0;
        /* nd[8*x + 7] marks a completed region (leaves are always ready) */
        if (a < n - 1 && nd[8 * a + 7] == (SR) 0) {
            stack[sp++] = a;
            pushed = 1;
        }
        if (b < n - 1 && nd[8 * b + 7] == (SR) 0) {
            stack[sp++] = b;
            pushed = 1;
        }
        if (pushed) continue;
        sp--;
        nd[8 * c] = ORC_MIN(nd[8 * a], nd[8 * b]);
        nd[8 * c + 1] = ORC_MIN(nd[8 * a + 1], nd[8 * b + 1]);
        nd[8 * c + 2] = ORC_MAX(nd[8 * a + 2], nd[8 * b + 2]);
        nd[8 * c + 3] = ORC_MAX(nd[8 * a + 3], nd[8 * b + 3]);
        nd[8 * c + 4] = nd[8 * a + 4] + nd[8 * b + 4];
        nd[8 * c + 5] = nd[8 * a + 5] + nd[8 * b + 5];
        nd[8 * c + 6] = nd[8 * a + 6] + nd[8 * b + 6];
        nd[8 * c + 7] = (SR) 1;
    }
}

/* The traversal records of the regions (pipeline.cl:276-323) on the tree of
 * the positions: ni[4r..4r+3] = (centre x, centre y, mass, size^2) of region
 * r in the reference numbering; mass[i] is the mass of node i. */
int SFX(orc_large_regions)(int n, const PR *x, const PR *y, const double *mass,
                           double *ni_out)
{
    if (n < 2) return 1;
    const size_t nn = (size_t) n;
    uint32_t *key = (uint32_t *) malloc(nn * sizeof(uint32_t));
    uint32_t *ks = (uint32_t *) malloc(nn * sizeof(uint32_t));
    uint32_t *order = (uint32_t *) malloc(nn * sizeof(uint32_t));
    int *child = (int *) malloc(2 * (nn - 1) * sizeof(int));
    int *range = (int *) malloc(2 * (nn - 1) * sizeof(int));
    int *parent = (int *) malloc((2 * nn - 1) * sizeof(int));
    SR *nd = (SR *) malloc(8 * (2 * nn - 1) * sizeof(SR));
    int status = 1;
    if (!key || !ks || !order || !child || !range || !parent || !nd) goto out;
    SFX(orc_large_tree)(n, x, y, key, ks, order, child, range, parent);
    for (int k = 0; k < n; k++) {
        const uint32_t i = order[k];
        SR *L = nd + 8 * (nn - 1 + (size_t) k);
        const SR px = (SR) POS_FR(x[i]), py = (SR) POS_FR(y[i]), pm = (SR) mass[i];
        L[0] = px;
        L[1] = py;
        L[2] = px;
        L[3] = py;
        L[4] = pm;
        L[5] = px * pm;
        L[6] = py * pm;
        L[7] = (SR) 1;
    }
    for (int r = 0; r < n - 1; r++) nd[8 * r + 7] = (SR) 0;
    SFX(agg_node)(n, 0, child, nd);
    for (int r = 0; r < n - 1; r++) {
        const SR m = nd[8 * r + 4];
        const SR w = nd[8 * r + 2] - nd[8 * r];
        const SR h = nd[8 * r + 3] - nd[8 * r + 1];
#ifdef ORC_CLAMP_CENTRE
        ni_out[4 * r] = (double) ORC_MIN(ORC_MAX(nd[8 * r + 5] / m, nd[8 * r]), nd[8 * r + 2]);
        ni_out[4 * r + 1] = (double) ORC_MIN(ORC_MAX(nd[8 * r + 6] / m, nd[8 * r + 1]), nd[8 * r + 3]);
#else
        ni_out[4 * r] = (double) (nd[8 * r + 5] / m);
        ni_out[4 * r + 1] = (double) (nd[8 * r + 6] / m);
#endif
        ni_out[4 * r + 2] = (double) m;
        ni_out[4 * r + 3] = (double) ((SR) 2 * (w * w + h * h));
    }
    status = 0;
out:
    free(key);
    free(ks);
    free(order);
    free(child);
    free(range);
    free(parent);
    free(nd);
    return status;
}

static inline FR SFX(fsqrt)(FR v) { return sizeof(FR) == 4 ? (FR) sqrtf((float) v) : (FR) sqrt((double) v); }
static inline FR SFX(flog1p)(FR v) { return sizeof(FR) == 4 ? (FR) log1pf((float) v) : (FR) log1p((double) v); }

/* Repulsion factor (pipeline.cl:46-55). */
static inline FR SFX(rep_factor)(FR d2, FR mm, FR ss, FR scaling, int adjust)
{
    if (adjust) {
        const FR d = SFX(fsqrt)(d2) - ss;
        if (d > (FR) 0) return scaling * mm / (d * d);
        if (d < (FR) 0) return (FR) 100 * scaling * mm;
        return (FR) 0;
    }
    return d2 > (FR) 0 ? scaling * mm / d2 : (FR) 0;
}


/*
 * The layout of fa2_full_run (src/gpu_full.c:164-532) on the CPU.
 * pos: x[0..n) then y[0..n), in/out.  Optional outputs:
 *   trace_pos   (iterations * 2n) positions after each iteration
 *   trace_force (iterations * 2n) force (x then y) applied at each iteration
 *   trace_speed (iterations * 4)  total swinging, total traction, speed,
 *                                 efficiency after the update
 *   trace_rep   (iterations * 2n) repulsion alone (before gravity and
 *                                 attraction)
 * Returns 0, 1 on bad input / memory, 2 when the totals are not numbers
 * (src/gpu_full.c:486-491).
 */
int SFX(orc_large_run)(const orc_large_params *p, double *pos, int iterations,
                       double *trace_pos, double *trace_force,
                       double *trace_speed, double *trace_rep)
{
    const int n = p->n;
    if (n < 2) return 1;
    const size_t nn = (size_t) n;
    int B, S;
    block_plan(n, &B, &S);

    PR *x = (PR *) malloc(nn * sizeof(PR));
    PR *y = (PR *) malloc(nn * sizeof(PR));
    FR *f = (FR *) calloc(2 * nn, sizeof(FR));
    FR *fo = (FR *) calloc(2 * nn, sizeof(FR));
    FR *sw = (FR *) malloc(nn * sizeof(FR));
    FR *ms = (FR *) malloc(nn * sizeof(FR));
    FR *sz = (FR *) malloc(nn * sizeof(FR));
    FR *psort = (FR *) malloc(4 * nn * sizeof(FR));
    SR *nd = (SR *) malloc(8 * (2 * nn - 1) * sizeof(SR));
    SR *ni = (SR *) malloc(4 * (nn - 1) * sizeof(SR));
    FR *part = (FR *) malloc(2 * (size_t) B * sizeof(FR));
    uint32_t *key = (uint32_t *) malloc(nn * sizeof(uint32_t));
    uint32_t *ks = (uint32_t *) malloc(nn * sizeof(uint32_t));
    uint32_t *order = (uint32_t *) malloc(nn * sizeof(uint32_t));
    int *child = (int *) malloc(2 * (nn - 1) * sizeof(int));
    int *range = (int *) malloc(2 * (nn - 1) * sizeof(int));
    int *parent = (int *) malloc((2 * nn - 1) * sizeof(int));
    int status = 1;
    if (!x || !y || !f || !fo || !sw || !ms || !sz || !psort || !nd || !ni ||
        !part || !key || !ks || !order || !child || !range || !parent)
        goto out;

    for (int i = 0; i < n; i++) {
        x[i] = (PR) pos[i];
        y[i] = (PR) pos[nn + i];
        ms[i] = (FR) p->mass[i];
        sz[i] = p->adjust ? (FR) p->size[i] : (FR) 0;
    }
    const FR scaling = (FR) p->scaling;
    const SR theta2 = (SR) (p->theta * p->theta);
    const FR gravity = (FR) p->gravity;
    double speed = 1.0, eff = 1.0;

    for (int it = 0; it < iterations; it++) {
        /* 1-3: codes, sort, tree */
        SFX(orc_large_tree)(n, x, y, key, ks, order, child, range, parent);

        /* 4: gather (pipeline.cl:195-207) and summaries */
        for (int k = 0; k < n; k++) {
            const uint32_t i = order[k];
            psort[4 * k] = POS_FR(x[i]);
            psort[4 * k + 1] = POS_FR(y[i]);
            psort[4 * k + 2] = ms[i];
            psort[4 * k + 3] = sz[i];
        }
        for (int k = 0; k < n; k++) {
            SR *L = nd + 8 * (nn - 1 + (size_t) k);
            const SR px = (SR) psort[4 * k], py = (SR) psort[4 * k + 1],
                     pm = (SR) psort[4 * k + 2];
            L[0] = px;
            L[1] = py;
            L[2] = px;
            L[3] = py;
            L[4] = pm;
            L[5] = px * pm;
            L[6] = py * pm;
            L[7] = (SR) 1;
        }
        for (int r = 0; r < n - 1; r++) nd[8 * r + 7] = (SR) 0;
        SFX(agg_node)(n, 0, child, nd);
        for (int r = 0; r < n - 1; r++) { /* pipeline.cl:312-323 */
            const SR m = nd[8 * r + 4];
            const SR w = nd[8 * r + 2] - nd[8 * r];
            const SR h = nd[8 * r + 3] - nd[8 * r + 1];
#ifdef ORC_CLAMP_CENTRE
            /* the centre of mass stays in the bounding box of the region, as
             * in exact arithmetic (the HIP path does the same) */
            ni[4 * r] = ORC_MIN(ORC_MAX(nd[8 * r + 5] / m, nd[8 * r]), nd[8 * r + 2]);
            ni[4 * r + 1] = ORC_MIN(ORC_MAX(nd[8 * r + 6] / m, nd[8 * r + 1]), nd[8 * r + 3]);
#else
            ni[4 * r] = nd[8 * r + 5] / m;
            ni[4 * r + 1] = nd[8 * r + 6] / m;
#endif
            ni[4 * r + 2] = m;
            ni[4 * r + 3] = (SR) 2 * (w * w + h * h);
        }

        /* 5: repulsion (pipeline.cl:332-383) */
        for (int k = 0; k < n; k++) {
            const FR xi = psort[4 * k], yi = psort[4 * k + 1];
            const FR mi = psort[4 * k + 2], si = psort[4 * k + 3];
            FR ax = 0, ay = 0;
            int stack[200];
            int sp = 0;
            stack[sp++] = 0;
            while (sp > 0) {
                const int r = stack[--sp];
                int first, last;
                if (r >= n - 1) {
                    first = last = r - (n - 1);
                } else {
                    /* the acceptance test in the summaries' arithmetic, the
                     * term in the forces' (the same when SR is FR) */
                    const SR txd = (SR) xi - ni[4 * r], tyd = (SR) yi - ni[4 * r + 1];
                    if ((txd * txd + tyd * tyd) * theta2 > ni[4 * r + 3]) {
                        const FR xd = xi - (FR) ni[4 * r], yd = yi - (FR) ni[4 * r + 1];
                        const FR d2 = xd * xd + yd * yd;
                        const FR g = scaling * (mi * (FR) ni[4 * r + 2]) / d2;
                        ax += xd * g;
                        ay += yd * g;
                        continue;
                    }
                    first = range[2 * r];
                    last = range[2 * r + 1];
                    if (last - first >= ORC_LEAF_SIZE) {
                        stack[sp++] = child[2 * r + 1];
                        stack[sp++] = child[2 * r];
                        continue;
                    }
                }
                for (int j = first; j <= last; j++) {
                    if (j == k) continue;
                    const FR xd = xi - psort[4 * j], yd = yi - psort[4 * j + 1];
                    const FR g = SFX(rep_factor)(xd * xd + yd * yd,
                                                 mi * psort[4 * j + 2],
                                                 si + psort[4 * j + 3],
                                                 scaling, p->adjust);
                    ax += xd * g;
                    ay += yd * g;
                }
            }
            const uint32_t i = order[k];
            f[2 * i] = ax;
            f[2 * i + 1] = ay;
        }
        if (trace_rep)
            for (int i = 0; i < n; i++) {
                trace_rep[2 * nn * (size_t) it + i] = (double) f[2 * i];
                trace_rep[2 * nn * (size_t) it + nn + i] = (double) f[2 * i + 1];
            }

        /* 6: gravity and attraction (pipeline.cl:392-429) */
        for (int i = 0; i < n; i++) {
            const FR xi = POS_FR(x[i]), yi = POS_FR(y[i]);
            FR ax = f[2 * i], ay = f[2 * i + 1];
            if (gravity != (FR) 0) {
                const FR dist = SFX(fsqrt)(xi * xi + yi * yi);
                if (dist > (FR) 0) {
                    const FR g = p->strong ? gravity * ms[i] : gravity * ms[i] / dist;
                    ax -= xi * g;
                    ay -= yi * g;
                }
            }
            const FR si = sz[i];
            for (int e = p->inc_off[i]; e < p->inc_off[i + 1]; e++) {
                const int j = p->inc_node[e];
                const FR xd = xi - POS_FR(x[j]), yd = yi - POS_FR(y[j]);
                FR g = -(FR) p->inc_coef[e];
                if (p->adjust || p->linlog) {
                    FR dist = SFX(fsqrt)(xd * xd + yd * yd);
                    if (p->adjust) dist -= si + sz[j];
                    if (!(dist > (FR) 0)) continue;
                    if (p->linlog) g *= SFX(flog1p)(dist) / dist;
                }
                ax += xd * g;
                ay += yd * g;
            }
            f[2 * i] = ax;
            f[2 * i + 1] = ay;
        }

        /* 7: swinging and traction by blocks (pipeline.cl:438-474) */
        for (int b = 0; b < B; b++) {
            const int lo = b * S, hi = ORC_MIN(n, lo + S);
            FR ts = 0, tt = 0;
            for (int i = lo; i < hi; i++) {
                const FR sx = fo[2 * i] - f[2 * i], sy = fo[2 * i + 1] - f[2 * i + 1];
                const FR tx = fo[2 * i] + f[2 * i], ty = fo[2 * i + 1] + f[2 * i + 1];
                const FR s = ms[i] * SFX(fsqrt)(sx * sx + sy * sy);
                sw[i] = s;
                if (p->any_fixed && p->fixed[i]) continue;
                ts += s;
                tt += (FR) 0.5 * ms[i] * SFX(fsqrt)(tx * tx + ty * ty);
            }
            part[2 * b] = ts;
            part[2 * b + 1] = tt;
        }
        FR tsw = 0, ttr = 0;
        for (int b = 0; b < B; b++) {
            tsw += part[2 * b];
            ttr += part[2 * b + 1];
        }
        if (!(tsw == tsw) || !(ttr == ttr)) {
            status = 2;
            goto out;
        }
        orc_update_speed(n, (double) tsw, (double) ttr, p->jitter_tolerance,
                         &speed, &eff);
        if (trace_speed) {
            trace_speed[4 * (size_t) it] = (double) tsw;
            trace_speed[4 * (size_t) it + 1] = (double) ttr;
            trace_speed[4 * (size_t) it + 2] = speed;
            trace_speed[4 * (size_t) it + 3] = eff;
        }
        if (trace_force)
            for (int i = 0; i < n; i++) {
                trace_force[2 * nn * (size_t) it + i] = (double) f[2 * i];
                trace_force[2 * nn * (size_t) it + nn + i] = (double) f[2 * i + 1];
            }

        /* 8: displacement (pipeline.cl:480-506) */
        const FR sp = (FR) speed;
        for (int i = 0; i < n; i++) {
            const FR fx = f[2 * i], fy = f[2 * i + 1];
            fo[2 * i] = fx;
            fo[2 * i + 1] = fy;
            if (p->any_fixed && p->fixed[i]) continue;
            FR factor;
            if (p->adjust) {
                const FR df = SFX(fsqrt)(fx * fx + fy * fy);
                factor = (FR) 0.1 * sp / ((FR) 1 + SFX(fsqrt)(sp * sw[i]));
                if (df > (FR) 0) {
                    if (factor * df > (FR) 10) factor = (FR) 10 / df;
                } else {
                    factor = (FR) 0;
                }
            } else {
                factor = sp / ((FR) 1 + SFX(fsqrt)(sp * sw[i]));
            }
            x[i] += (PR) (fx * factor);
            y[i] += (PR) (fy * factor);
        }
        if (trace_pos)
            for (int i = 0; i < n; i++) {
                trace_pos[2 * nn * (size_t) it + i] = (double) x[i];
                trace_pos[2 * nn * (size_t) it + nn + i] = (double) y[i];
            }
    }
    for (int i = 0; i < n; i++) {
        pos[i] = (double) x[i];
        pos[nn + i] = (double) y[i];
    }
    status = 0;
out:
    free(x);
    free(y);
    free(f);
    free(fo);
    free(sw);
    free(ms);
    free(sz);
    free(psort);
    free(nd);
    free(ni);
    free(part);
    free(key);
    free(ks);
    free(order);
    free(child);
    free(range);
    free(parent);
    return status;
}

#undef POS_FR
#undef ORC_MIN
#undef ORC_MAX
