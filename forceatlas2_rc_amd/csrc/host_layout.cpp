// Host drivers of the two GPU paths: the plain-C restatement of the .Call
// entries of the reference (C_forceatlas2, src/forceatlas2.c:521-827 and
// C_forceatlas2_large, :835-947), the Gephi quadtree the repulsion offload
// traverses (tree_build, :64-239) and the adaptive speed (:445-485).
//
// The repulsion, where the time goes, always runs on the GPU here: a layout
// without a device is refused, there is no host fallback for it.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "message.h"
#include "../../include/fa2.h"
#include "../../include/fa2_hip.h"

using namespace fa2;

// ---------------------------------------------------------------------------
// Adaptive speed (src/forceatlas2.c:445-485). Weak: the package's own
// definition wins when both are linked.
// ---------------------------------------------------------------------------

extern "C" __attribute__((weak)) void fa2_update_speed(int n, double total_swinging,
                                                       double total_traction,
                                                       double jitter_tolerance, double *speed_,
                                                       double *speed_efficiency_)
{
    if (!(total_swinging > 0)) return; // nothing moves: keep the speed
    double speed = *speed_, eff = *speed_efficiency_;
    const double est = 0.05 * std::sqrt((double) n);
    double jt = est * total_traction / ((double) n * (double) n);
    jt = jt > 10.0 ? 10.0 : jt;
    jt = jt < std::sqrt(est) ? std::sqrt(est) : jt;
    jt *= jitter_tolerance;
    if (total_swinging > 2.0 * total_traction) { // erratic: slow down
        if (eff > 0.05) eff *= 0.5;
        if (jt < jitter_tolerance) jt = jitter_tolerance;
    }
    const double target = jt * eff * total_traction / total_swinging;
    if (total_swinging > jt * total_traction) {
        if (eff > 0.05) eff *= 0.7;
    } else if (speed < 1000) {
        eff *= 1.3;
    }
    const double rise = target - speed; // at most +50% per iteration
    speed += rise < 0.5 * speed ? rise : 0.5 * speed;
    *speed_ = speed;
    *speed_efficiency_ = eff;
}

// ---------------------------------------------------------------------------
// Gephi quadtree, built level by level (src/forceatlas2.c:64-239). The
// regions, their order, the skip links and the node order must be those of
// the reference: the GPU traversal of fa2_gpu_repulse_bh visits them as the
// CPU does.
// ---------------------------------------------------------------------------

namespace {

constexpr int kGephiLeaf = 8;

inline int quadrant_of(const fa2_node &p, double cx, double cy)
{
    return (p.x < cx ? 0 : 2) + (p.y < cy ? 0 : 1);
}

struct Level {
    fa2_tree *t;
    // population of each quadrant of the regions of the current level
    int *quad(int r, int lo) { return t->cnt + 4 * (size_t) (r - lo); }
};

} // namespace

extern "C" void fa2_tree_build(fa2_tree *t, int n, const double *x, const double *y,
                               const double *mass, const double *size, int threads)
{
    // the regions of a level are measured and split independently: with
    // threads > 1 they are shared among threads (src/forceatlas2.c:108, :169,
    // :225); region ids and node order do not depend on the thread count
    const int nt = threads > 1 ? threads : 1;
    fa2_region *reg = t->reg;
    // root: every node, centre of mass (mass[0] and the node itself if alone)
    double msum = 0.0, mx = 0.0, my = 0.0;
    for (int i = 0; i < n; i++) {
        t->buf[0][i] = fa2_node{x[i], y[i], mass[i], i};
        msum += mass[i];
        mx += x[i] * mass[i];
        my += y[i] * mass[i];
    }
    reg[0] = fa2_region{n > 1 ? msum : mass[0], n > 1 ? mx / msum : x[0],
                        n > 1 ? my / msum : y[0], 0.0, 0, n, -1, 0, -1, 0};
    t->nreg = 1;
    Level L{t};
    for (int lo = 0, hi = 1; lo < hi; lo = hi, hi = t->nreg) {
        // a level is shared among threads when it has regions for all of
        // them: the top levels hold few large regions, whose split stays on
        // one thread (measured faster than a parallel loop over 1..16 regions)
        const bool wide = nt > 1 && hi - lo >= 4 * nt;
        // sizes and quadrant populations
#pragma omp parallel for schedule(static) num_threads(nt) if (wide)
        for (int r = lo; r < hi; r++) {
            int *q = L.quad(r, lo);
            q[0] = q[1] = q[2] = q[3] = 0;
            fa2_region &R = reg[r];
            if (R.count < 2) continue;
            const fa2_node *nd = t->buf[R.buf] + R.start;
            double far2 = 0.0;
            for (int k = 0; k < R.count; k++) {
                const double ux = nd[k].x - R.cx, uy = nd[k].y - R.cy;
                const double d2 = ux * ux + uy * uy;
                if (d2 > far2) far2 = d2;
                q[quadrant_of(nd[k], R.cx, R.cy)]++;
            }
            R.size = 2.0 * std::sqrt(far2);
        }
        // children, in the order of the regions and of the quadrants
        for (int r = lo; r < hi; r++) {
            fa2_region &R = reg[r];
            if (R.count <= kGephiLeaf) continue;
            const int *q = L.quad(r, lo);
            const bool one = q[0] == R.count || q[1] == R.count || q[2] == R.count ||
                             q[3] == R.count;
            R.child = t->nreg;
            if (one) { // coincident nodes: single-node leaves in place
                R.nchild = R.count;
                for (int k = 0; k < R.count; k++) {
                    const int id = t->nreg++;
                    reg[id] = fa2_region{0, 0, 0, 0, R.start + k, 1, -1, 0,
                                         k == 0 ? R.skip : id - 1, R.buf};
                }
                continue;
            }
            int at = R.start;
            for (int s = 0; s < 4; s++) {
                if (!q[s]) continue;
                const int id = t->nreg++;
                reg[id] = fa2_region{0, 0, 0, 0, at, q[s], -1, 0,
                                     R.nchild == 0 ? R.skip : id - 1, 1 - R.buf};
                R.nchild++;
                at += q[s];
            }
        }
        // move the nodes to their quadrant; masses and centres of the children
#pragma omp parallel for schedule(static) num_threads(nt) if (wide)
        for (int r = lo; r < hi; r++) {
            const fa2_region &R = reg[r];
            if (R.nchild == 0) continue;
            const int *q = L.quad(r, lo);
            const bool one = q[0] == R.count || q[1] == R.count || q[2] == R.count ||
                             q[3] == R.count;
            const fa2_node *src = t->buf[R.buf] + R.start;
            if (!one) {
                fa2_node *dst = t->buf[1 - R.buf];
                double qm[4] = {0, 0, 0, 0}, qx[4] = {0, 0, 0, 0}, qy[4] = {0, 0, 0, 0};
                int put[4] = {R.start, 0, 0, 0};
                for (int s = 1; s < 4; s++) put[s] = put[s - 1] + q[s - 1];
                for (int k = 0; k < R.count; k++) {
                    const int s = quadrant_of(src[k], R.cx, R.cy);
                    dst[put[s]++] = src[k];
                    qm[s] += src[k].m;
                    qx[s] += src[k].x * src[k].m;
                    qy[s] += src[k].y * src[k].m;
                }
                int c = R.child;
                for (int s = 0; s < 4; s++) {
                    if (!q[s]) continue;
                    reg[c].mass = qm[s];
                    reg[c].cx = qx[s] / qm[s];
                    reg[c].cy = qy[s] / qm[s];
                    reg[c].size = 0.0;
                    c++;
                }
                src = dst + R.start;
            }
            for (int c = R.child; c < R.child + R.nchild; c++) {
                if (reg[c].count > 1) continue;
                const fa2_node &one_node = src[reg[c].start - R.start];
                reg[c].mass = one_node.m;
                reg[c].cx = one_node.x;
                reg[c].cy = one_node.y;
                reg[c].size = 0.0;
            }
        }
    }
    // the nodes in tree order, from the leaf regions
    const int nreg = t->nreg;
#pragma omp parallel for schedule(static) num_threads(nt) if (nt > 1)
    for (int r = 0; r < nreg; r++) {
        if (reg[r].nchild > 0) continue;
        const fa2_node *nd = t->buf[reg[r].buf];
        for (int k = reg[r].start; k < reg[r].start + reg[r].count; k++) {
            t->perm[k] = nd[k].idx;
            t->px[k] = nd[k].x;
            t->py[k] = nd[k].y;
            t->pm[k] = nd[k].m;
            if (size) t->ps[k] = size[nd[k].idx];
        }
    }
}

// ---------------------------------------------------------------------------
// C_forceatlas2 (src/forceatlas2.c:521-827) with the repulsion on the GPU
// ---------------------------------------------------------------------------

namespace {

struct HostTree {
    std::vector<fa2_region> reg;
    std::vector<int> perm, cnt;
    std::vector<fa2_node> b0, b1;
    std::vector<double> px, py, pm, ps;
    fa2_tree t;
    void alloc(int n, bool with_size)
    {
        const size_t nn = (size_t) n;
        reg.resize(2 * nn);
        perm.resize(nn);
        cnt.resize(8 * nn);
        b0.resize(nn);
        b1.resize(nn);
        px.resize(nn);
        py.resize(nn);
        pm.resize(nn);
        if (with_size) ps.resize(nn);
        t.reg = reg.data();
        t.nreg = 0;
        t.perm = perm.data();
        t.cnt = cnt.data();
        t.buf[0] = b0.data();
        t.buf[1] = b1.data();
        t.px = px.data();
        t.py = py.data();
        t.pm = pm.data();
        t.ps = with_size ? ps.data() : nullptr;
    }
};

} // namespace

extern "C" int fa2_layout_run(int n, int m, const int *from, const int *to, const double *w,
                              double *pos, const double *size, const int *fixed, int iterations,
                              const double *dpar, const int *flags, int threads, const int *gpu,
                              double *times)
{
    if (times)
        for (int k = 0; k < 7; k++) times[k] = 0.0;
    if (n < 0) {
        set_message("invalid number of nodes");
        return 1;
    }
    if (m < 0 || m > (1 << 30) - 1) {
        set_message("too many edges");
        return 1;
    }
    if (iterations < 0) {
        set_message("invalid number of iterations");
        return 1;
    }
    for (int e = 0; e < m; e++)
        if (from[e] < 0 || from[e] >= n || to[e] < 0 || to[e] >= n) {
            set_message("edge %d refers to a node that does not exist", e + 1);
            return 1;
        }
    if (n == 0 || iterations == 0) return 0;
    const int device = gpu ? gpu[0] : -1;
    if (device < 0) {
        set_message("this library computes the repulsion on a GPU: a device is required");
        return 1;
    }
    const double scaling = dpar[0], gravity = dpar[1], jt_user = dpar[2], theta = dpar[3];
    const int strong = flags[0], linlog = flags[1], outbound = flags[2], adjust = flags[3];
    const int barnes_hut = flags[4];
    const size_t nn = (size_t) n;
    double *x = pos, *y = pos + nn;

    std::vector<double> mass(nn, 1.0), fx(nn, 0.0), fy(nn, 0.0), ox(nn, 0.0), oy(nn, 0.0);
    std::vector<double> swing(nn), tract(nn);
    for (int e = 0; e < m; e++) { // mass = 1 + degree (src/forceatlas2.c:592-603)
        mass[from[e]] += 1.0;
        mass[to[e]] += 1.0;
    }
    double mean_mass = 0.0;
    for (int i = 0; i < n; i++) mean_mass += mass[i];
    mean_mass /= n;
    const double coef = outbound ? mean_mass : 1.0;

    // With threads > 1 the host loops are shared among threads, node by node
    // (src/forceatlas2.c:607-631, :720-802): the attraction walks each
    // node's incidences in edge order and repeats the edge loop's arithmetic
    // for each of them, so every node receives the same terms in the same
    // order and the layout does not depend on the thread count.
    const int nt = threads > 1 ? threads : 1;
    std::vector<int> inc_off, inc_edge;
    if (nt > 1) {
        inc_off.assign(nn + 1, 0);
        for (int e = 0; e < m; e++)
            if (from[e] != to[e]) {
                inc_off[(size_t) from[e] + 1]++;
                inc_off[(size_t) to[e] + 1]++;
            }
        for (size_t i = 0; i < nn; i++) inc_off[i + 1] += inc_off[i];
        inc_edge.resize((size_t) inc_off[nn]);
        std::vector<int> at(inc_off.begin(), inc_off.end() - 1);
        for (int e = 0; e < m; e++)
            if (from[e] != to[e]) {
                inc_edge[(size_t) at[from[e]]++] = e;
                inc_edge[(size_t) at[to[e]]++] = e;
            }
    }
    // the attraction of edge e (:385-416), from its first node's side: false
    // when it adds nothing
    auto edge_term = [&](int e, double &xd, double &yd, double &f) {
        const int a = from[e], b = to[e];
        xd = x[a] - x[b];
        yd = y[a] - y[b];
        f = -coef * w[e];
        if (adjust || linlog) {
            double d = std::sqrt(xd * xd + yd * yd);
            if (adjust) d -= size[a] + size[b];
            if (!(d > 0)) return false;
            if (linlog) f *= std::log1p(d) / d;
        }
        if (outbound) f /= mass[a];
        return f != 0.0;
    };

    HostTree tree;
    if (barnes_hut) tree.alloc(n, adjust != 0);
    fa2_gpu *g = fa2_gpu_open(device, gpu[1], n, adjust, barnes_hut, threads);
    if (!g) return 1;

    double speed = 1.0, eff = 1.0, t_tree = 0.0, t_rep = 0.0;
    const double t_start = fa2_now();
    int status = 0;
    for (int it = 0; it < iterations && status == 0; it++) {
        fx.swap(ox);
        fy.swap(oy);
        std::fill(fx.begin(), fx.end(), 0.0);
        std::fill(fy.begin(), fy.end(), 0.0);
        double t0 = fa2_now();
        if (barnes_hut) {
            fa2_tree_build(&tree.t, n, x, y, mass.data(), adjust ? size : nullptr, threads);
            const double t1 = fa2_now();
            t_tree += t1 - t0;
            t0 = t1;
            status = fa2_gpu_repulse_bh(g, &tree.t, scaling, theta, fx.data(), fy.data());
        } else {
            status = fa2_gpu_repulse_exact(g, x, y, mass.data(), adjust ? size : nullptr, scaling,
                                           fx.data(), fy.data());
        }
        t_rep += fa2_now() - t0;
        if (status) break;
        // gravity (:371-381)
        if (gravity != 0) {
#pragma omp parallel for schedule(static) num_threads(nt) if (nt > 1)
            for (int i = 0; i < n; i++) {
                const double d = std::sqrt(x[i] * x[i] + y[i] * y[i]);
                if (d > 0) {
                    const double f = strong ? gravity * mass[i] : gravity * mass[i] / d;
                    fx[i] -= x[i] * f;
                    fy[i] -= y[i] * f;
                }
            }
        }
        // attraction (:385-416): edge by edge, or node by node over the
        // incidences in edge order
        if (nt > 1) {
#pragma omp parallel for schedule(dynamic, 256) num_threads(nt)
            for (int i = 0; i < n; i++) {
                for (int k = inc_off[(size_t) i]; k < inc_off[(size_t) i + 1]; k++) {
                    const int e = inc_edge[(size_t) k];
                    double xd, yd, f;
                    if (!edge_term(e, xd, yd, f)) continue;
                    if (from[e] == i) {
                        fx[i] += xd * f;
                        fy[i] += yd * f;
                    } else {
                        fx[i] -= xd * f;
                        fy[i] -= yd * f;
                    }
                }
            }
        } else {
            for (int e = 0; e < m; e++) {
                if (from[e] == to[e]) continue;
                double xd, yd, f;
                if (!edge_term(e, xd, yd, f)) continue;
                fx[from[e]] += xd * f;
                fy[from[e]] += yd * f;
                fx[to[e]] -= xd * f;
                fy[to[e]] -= yd * f;
            }
        }
        // swinging and traction, summed in node order (:758-778)
        double tsw = 0.0, ttr = 0.0;
#pragma omp parallel for schedule(static) num_threads(nt) if (nt > 1)
        for (int i = 0; i < n; i++) {
            const double sx = ox[i] - fx[i], sy = oy[i] - fy[i];
            const double tx = ox[i] + fx[i], ty = oy[i] + fy[i];
            swing[i] = mass[i] * std::sqrt(sx * sx + sy * sy);
            tract[i] = 0.5 * mass[i] * std::sqrt(tx * tx + ty * ty);
        }
        for (int i = 0; i < n; i++) {
            if (fixed && fixed[i]) continue;
            tsw += swing[i];
            ttr += tract[i];
        }
        fa2_update_speed(n, tsw, ttr, jt_user, &speed, &eff);
        // displacement (:780-802)
#pragma omp parallel for schedule(static) num_threads(nt) if (nt > 1)
        for (int i = 0; i < n; i++) {
            if (fixed && fixed[i]) continue;
            double factor;
            if (adjust) {
                const double df = std::sqrt(fx[i] * fx[i] + fy[i] * fy[i]);
                factor = 0.1 * speed / (1.0 + std::sqrt(speed * swing[i]));
                if (df > 0) {
                    if (factor * df > 10.0) factor = 10.0 / df;
                } else {
                    factor = 0.0;
                }
            } else {
                factor = speed / (1.0 + std::sqrt(speed * swing[i]));
            }
            x[i] += fx[i] * factor;
            y[i] += fy[i] * factor;
        }
    }
    double gt[3] = {0, 0, 0};
    fa2_gpu_times(g, gt);
    fa2_gpu_close(g);
    if (times) {
        times[0] = fa2_now() - t_start;
        times[1] = t_tree;
        times[2] = t_rep;
        times[3] = gt[0];
        times[4] = gt[1];
        times[5] = gt[2];
        times[6] = times[0] - t_tree - t_rep;
    }
    return status;
}

// ---------------------------------------------------------------------------
// C_forceatlas2_large (src/forceatlas2.c:835-947)
// ---------------------------------------------------------------------------

extern "C" int fa2_layout_large_run(int n, int m, const int *from, const int *to,
                                    const double *w, double *pos, const double *size,
                                    const int *fixed, int iterations, const double *dpar,
                                    const int *flags, const int *gpu, double *times)
{
    double tloc[FA2_FULL_NTIMES];
    double *t = times ? times : tloc;
    for (int k = 0; k < FA2_FULL_NTIMES; k++) t[k] = 0.0;
    if (n < 2) {
        set_message("invalid number of nodes");
        return 1;
    }
    if (m < 0 || m > (1 << 30) - 1) {
        set_message("too many edges");
        return 1;
    }
    if (iterations < 0) {
        set_message("invalid number of iterations");
        return 1;
    }
    for (int e = 0; e < m; e++)
        if (from[e] < 0 || from[e] >= n || to[e] < 0 || to[e] >= n) {
            set_message("edge %d refers to a node that does not exist", e + 1);
            return 1;
        }
    const int outbound = flags[2];
    const size_t nn = (size_t) n;
    std::vector<double> mass(nn, 1.0);
    for (int e = 0; e < m; e++) {
        mass[from[e]] += 1.0;
        mass[to[e]] += 1.0;
    }
    double mean_mass = 0.0;
    for (int i = 0; i < n; i++) mean_mass += mass[i];
    mean_mass /= n;
    const double coef = outbound ? mean_mass : 1.0;
    // incidence lists in edge order, self-loops left out (:882-902)
    std::vector<int> off(nn + 1, 0), node(2 * (size_t) m + 1);
    std::vector<double> cf(2 * (size_t) m + 1);
    for (int e = 0; e < m; e++)
        if (from[e] != to[e]) {
            off[from[e] + 1]++;
            off[to[e] + 1]++;
        }
    for (int i = 0; i < n; i++) off[i + 1] += off[i];
    std::vector<int> put(off.begin(), off.end() - 1);
    for (int e = 0; e < m; e++) {
        if (from[e] == to[e]) continue;
        const double c = outbound ? coef * w[e] / mass[from[e]] : coef * w[e];
        node[put[from[e]]] = to[e];
        cf[put[from[e]]++] = c;
        node[put[to[e]]] = from[e];
        cf[put[to[e]]++] = c;
    }
    std::vector<int> fixed_i;
    int any_fixed = 0;
    if (fixed)
        for (int i = 0; i < n; i++) any_fixed |= fixed[i] != 0;
    fa2_full_params p;
    memset(&p, 0, sizeof(p));
    p.n = n;
    p.inc_off = off.data();
    p.inc_node = node.data();
    p.inc_coef = cf.data();
    p.mass = mass.data();
    p.size = size;
    p.fixed = fixed;
    p.any_fixed = any_fixed;
    p.scaling = dpar[0];
    p.gravity = dpar[1];
    p.jitter_tolerance = dpar[2];
    p.theta = dpar[3];
    p.strong = flags[0];
    p.linlog = flags[1];
    p.adjust = flags[3];
    p.iterations = iterations;
    p.device = gpu ? gpu[0] : 0;
    p.use_double = gpu ? gpu[1] != 0 : 0;
    p.profile = gpu ? gpu[2] != 0 : 0;
    if (iterations == 0) return 0;
    return fa2_full_run(&p, pos, t);
}
