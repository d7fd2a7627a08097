// Repulsion offload sessions of libfa2hip.so: fa2_gpu_open / _close /
// _times / _repulse_exact / _repulse_bh, the backend of forceatlas2_gpu()
// (reference: src/gpu.c:357-704, kernels src/gpu.c:28-123).
//
// Design for CDNA4:
//  * The host arrays (FP64) are uploaded as they are and converted on the
//    device, so no host threads are needed for the conversion.
//  * Exact repulsion: every block owns 2x256 target nodes and one slice of
//    the source nodes; sources are staged through LDS in tiles of 512 nodes
//    and read back as broadcasts. Target blocks x source slices fill the 256
//    CUs even for small n. The per-slice partial forces are summed in a
//    fixed order, in FP64, by a second kernel: results are reproducible.
//  * The bulk arithmetic is FP32 with v_rcp_f32 for the 1/d^2 (1 ulp).
//    Double-precision sessions use the kernels of partial_f64.hip, which
//    repeat the reference's CPU operations in order.
//  * Barnes-Hut: one lane per node in tree order, stackless traversal along
//    the skip links of the host tree, regions packed as two 16-byte records.
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>

#include "common.h"
#include "kernels_common.h"
#include "partial_f64.h"
#include "../../include/fa2.h"

#include <type_traits>

using namespace fa2;

namespace {

constexpr int kExactThreads = 256;
constexpr int kExactTargets = 4;    // targets per thread
constexpr int kExactTile = 512;     // sources per LDS tile
constexpr int kBhThreads = 256;

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

// p[i] = (x, y, mass, size) in the session precision.
template <typename R>
__global__ void k_pack_nodes(int n, const double *__restrict__ x, const double *__restrict__ y,
                             const double *__restrict__ m, const double *__restrict__ s,
                             typename Vec<R>::v4 *__restrict__ p)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    typename Vec<R>::v4 v;
    v.x = (R) x[i];
    v.y = (R) y[i];
    v.z = (R) m[i];
    v.w = s ? (R) s[i] : (R) 0;
    p[i] = v;
}

// Exact repulsion of targets [blockIdx.x * 1024, +1024) from the sources of
// slice blockIdx.y; partial[slice][i] receives the sum over the slice of
// m_j * k(d^2) * (x_i - x_j, y_i - y_j), the factor scaling * m_i of every
// term being applied once per target by k_sum_slices. Without sizes the
// weight is m_j / (d^2 + 1e-30): a source at the target's own position has a
// zero offset (its finite weight times 0 adds 0), and any two distinct
// float32 positions are so far apart that adding 1e-30 leaves their d^2
// unchanged, so this equals the reference's factor (0 for d^2 = 0, else
// scaling m_i m_j / d^2) term by term.
template <typename R, bool ADJUST>
__global__ void __launch_bounds__(kExactThreads)
k_exact(int n, int slice_len, const typename Vec<R>::v4 *__restrict__ p,
        typename Vec<R>::v2 *__restrict__ partial)
{
    using V4 = typename Vec<R>::v4;
    __shared__ V4 tile[kExactTile];
    const int i0 = blockIdx.x * (kExactThreads * kExactTargets) + threadIdx.x;
    R xi[kExactTargets], yi[kExactTargets], si[kExactTargets];
    R ax[kExactTargets], ay[kExactTargets];
#pragma unroll
    for (int t = 0; t < kExactTargets; t++) {
        const int i = i0 + t * kExactThreads;
        V4 v = i < n ? p[i] : V4{(R) 0, (R) 0, (R) 0, (R) 0};
        xi[t] = v.x;
        yi[t] = v.y;
        si[t] = v.w;
        ax[t] = (R) 0;
        ay[t] = (R) 0;
    }
    const int j_begin = blockIdx.y * slice_len;
    const int j_end = min(n, j_begin + slice_len);
    for (int j0 = j_begin; j0 < j_end; j0 += kExactTile) {
        const int cnt = min(kExactTile, j_end - j0);
        __syncthreads();
        for (int t = threadIdx.x; t < kExactTile; t += kExactThreads)
            tile[t] = t < cnt ? p[j0 + t] : V4{(R) 0, (R) 0, (R) 0, (R) 0};
        __syncthreads();
        // Padding sources have mass 0: their contribution is exactly 0.
#pragma unroll 2
        for (int jj = 0; jj < kExactTile; jj++) {
            const V4 s = tile[jj];
#pragma unroll
            for (int t = 0; t < kExactTargets; t++) {
                const R xd = xi[t] - s.x, yd = yi[t] - s.y;
                R f;
                if (ADJUST) {
                    f = pair_factor_adjust<R>(xd, yd, s.z, si[t], s.w);
                } else {
                    // d^2 + 1e-30 in two fused multiply-adds
                    const R d2 = fma_r(xd, xd, fma_r(yd, yd, (R) 1e-30));
                    f = s.z * frcp_r(d2);
                }
                ax[t] += xd * f;
                ay[t] += yd * f;
            }
        }
    }
    typename Vec<R>::v2 *out = partial + (size_t) blockIdx.y * n;
#pragma unroll
    for (int t = 0; t < kExactTargets; t++) {
        const int i = i0 + t * kExactThreads;
        if (i < n) out[i] = typename Vec<R>::v2{ax[t], ay[t]};
    }
}

// f[i] = sum over slices, in slice order, in FP64. With ADD the force is
// added into dx[i], dy[i] instead, rounded as the host adds it to its
// arrays (a product, then a sum: no fused multiply-add).
template <typename R, bool ADD = false>
__global__ void k_sum_slices(int n, int slices, const typename Vec<R>::v2 *__restrict__ partial,
                             const double *__restrict__ mass, double scaling,
                             double2 *__restrict__ f, double *__restrict__ dx = nullptr,
                             double *__restrict__ dy = nullptr)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double sx = 0.0, sy = 0.0;
    for (int s = 0; s < slices; s++) {
        const typename Vec<R>::v2 v = partial[(size_t) s * n + i];
        sx += (double) v.x;
        sy += (double) v.y;
    }
    const double sm = scaling * mass[i]; // the factor scaling * m_i of every term
    if (ADD) {
        dx[i] += sx * sm;
        dy[i] += sy * sm;
    } else {
        f[i] = make_double2(sx * sm, sy * sm);
    }
}

// dx[i] += f[i].x, dy[i] += f[i].y.
__global__ void k_add_forces(int n, const double2 *__restrict__ f, double *__restrict__ dx,
                             double *__restrict__ dy)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double2 v = f[i];
    dx[i] += v.x;
    dy[i] += v.y;
}

// Work split of the exact repulsion of n nodes: target blocks of 1024 nodes
// x source slices, at least about 2048 blocks in all, slices a whole number
// of LDS tiles.
inline void exact_plan(int n, int *slices_out, int *slice_len_out)
{
    const size_t nn = (size_t) n;
    const int tblocks = (int) ceil_div(nn, kExactThreads * kExactTargets);
    int slices = (int) ceil_div(2048, tblocks);
    if (slices > 64) slices = 64;
    int slice_len = (int) ceil_div(nn, slices);
    slice_len = (int) (ceil_div(slice_len, kExactTile) * kExactTile);
    *slices_out = (int) ceil_div(nn, slice_len);
    *slice_len_out = slice_len;
}

// Regions of the host tree: rr = (cx, cy, mass, size^2), ri = (start, count,
// last child or -1, skip) (the reference layout of src/gpu.c:501-519).
template <typename R>
__global__ void k_pack_regions(int nreg, const fa2_region *__restrict__ reg,
                               typename Vec<R>::v4 *__restrict__ rr, int4 *__restrict__ ri)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nreg) return;
    const fa2_region g = reg[r];
    typename Vec<R>::v4 a;
    a.x = (R) g.cx;
    a.y = (R) g.cy;
    a.z = (R) g.mass;
    a.w = (R) (g.size * g.size);
    rr[r] = a;
    ri[r] = make_int4(g.start, g.count, g.nchild > 0 ? g.child + g.nchild - 1 : -1, g.skip);
}

// One region of the host tree as the walk reads it: (cx, cy, mass, size^2)
// and (first node, node count, last child or -1, skip or -1), 32 bytes, one
// scalar load.
struct alignas(32) HostRec {
    float4 c;
    int4 l;
};

// Region records of the single-precision walk, rounded to float32 as the
// reference's converter does (src/gpu.c:501-519: cx, cy, mass and size^2
// each rounded once).
__global__ void k_pack_host_records(int nreg, const fa2_region *__restrict__ reg,
                                    HostRec *__restrict__ rec)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nreg) return;
    const fa2_region g = reg[r];
    HostRec h;
    h.c = make_float4((float) g.cx, (float) g.cy, (float) g.mass, (float) (g.size * g.size));
    h.l = make_int4(g.start, g.count, g.nchild > 0 ? g.child + g.nchild - 1 : -1, g.skip);
    rec[r] = h;
}

// Barnes-Hut over the host Gephi tree, wave-synchronous. The 64 lanes of a
// wave are 64 consecutive nodes of the tree order, and the wave walks the
// union of their traversals (the reference's fa2_bh, src/gpu.c:82-123, walks
// one node per work item): one region per step, read with one uniform
// scalar load, tested by every lane at once.
//
// A lane takes part in a region only where its own traversal goes. The
// traversal visits a region's children from the last to the first, and the
// children split the region's nodes in that order of positions, so what a
// lane has already covered (regions it accepted or whose nodes it visited)
// is always the positions from `low` up to the end of the tree order: a
// region [s, e] lies on the lane's own path exactly when e < low. Each lane
// therefore meets its regions in the reference's order and sums the same
// terms in that order.
//
// The acceptance test d^2 theta^2 > size^2 is taken in float32 on the
// float32 records with each operation rounded on its own, as the
// reference's single-precision kernel takes it; each term enters the lane's
// FP64 sum through exact FP64 multiply-adds, and the factor scaling * m_i is
// applied once at the end. With node sizes, a pair's term is computed in
// FP64 from the float32 inputs (the force jumps as two discs touch).
// `p` holds kHostLeafLoad records of padding after the last node: the nodes
// of a leaf region are loaded together. The forces leave in FP64.
constexpr int kHostLeafLoad = 8;

template <bool ADJUST>
__global__ void __launch_bounds__(kBhThreads)
k_bh_walk(int n, int nreg, float scaling, float theta2, const float4 *__restrict__ p,
          const HostRec *__restrict__ rec, double2 *__restrict__ f)
{
    const int k0 = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = k0 < n ? k0 : n - 1;
    const bool valid = k0 < n;
    const float4 me = p[k];
    double sx = 0.0, sy = 0.0;
    int low = valid ? INT_MAX : INT_MIN; // lanes past the end never take part
    int c = 0;
    // a valid tree visits each region at most once; the bound ends the walk
    // of a malformed one
    for (int steps = 0; c >= 0 && c < nreg && steps < nreg; steps++) {
        c = __builtin_amdgcn_readfirstlane(c);
        const HostRec h = rec[c];
        const int s = h.l.x, cnt = h.l.y;
        const uint64_t am = __builtin_amdgcn_ballot_w64(s + cnt - 1 < low);
        if (am == 0) { // no lane walks this region
            c = h.l.w;
            continue;
        }
        uint64_t tm = am;
        if (cnt > 1) {
            const float xd = me.x - h.c.x, yd = me.y - h.c.y;
            const float d2 = sumsq_rn(xd, yd);
            const uint64_t fm = __builtin_amdgcn_ballot_w64(d2 * theta2 > h.c.w);
            if (__builtin_amdgcn_inverse_ballot_w64(am & fm)) {
                const double w = (double) (h.c.z * frcp_r(d2));
                sx = __builtin_fma((double) xd, w, sx);
                sy = __builtin_fma((double) yd, w, sy);
                low = s;
            }
            tm = am & ~fm;
            if (tm == 0) {
                c = h.l.w;
                continue;
            }
            if (h.l.z >= 0) { // some lanes open it: its last child next
                c = h.l.z;
                continue;
            }
        }
        // the region's nodes one by one (a leaf of at most 8 nodes in a tree
        // of fa2_tree_build; longer lists in pieces of 8)
        for (int b = s; b < s + cnt; b += kHostLeafLoad) {
            float4 o[kHostLeafLoad];
#pragma unroll
            for (int u = 0; u < kHostLeafLoad; u++) o[u] = p[b + u];
            if (__builtin_amdgcn_inverse_ballot_w64(tm)) {
#pragma unroll
                for (int u = 0; u < kHostLeafLoad; u++) {
                    if (b + u >= s + cnt) break;
                    if (ADJUST) {
                        // the pair in FP64 (src/gpu.c:38-47), a lane with
                        // itself adding zero (a zero offset)
                        const double dx = (double) me.x - (double) o[u].x;
                        const double dy = (double) me.y - (double) o[u].y;
                        const double d = __builtin_sqrt(sumsq_rn(dx, dy)) -
                                         ((double) me.w + (double) o[u].w);
                        const double mo = (double) o[u].z;
                        double g = d > 0.0 ? mo / (d * d) : (d < 0.0 ? 100.0 * mo : 0.0);
                        g = b + u == k ? 0.0 : g;
                        sx = __builtin_fma(dx, g, sx);
                        sy = __builtin_fma(dy, g, sy);
                    } else {
                        const float xd = me.x - o[u].x, yd = me.y - o[u].y;
                        const float d2 = xd * xd + yd * yd;
                        // the lane itself, and coincident nodes: d2 = 0, no term
                        float w = d2 > 0.0f ? o[u].z * frcp_r(d2) : 0.0f;
                        asm volatile("" : "+v"(w));
                        sx = __builtin_fma((double) xd, (double) w, sx);
                        sy = __builtin_fma((double) yd, (double) w, sy);
                    }
                }
                low = s;
            }
        }
        c = h.l.w;
    }
    if (valid) {
        const double sm = (double) scaling * (double) me.z;
        f[k] = make_double2(sx * sm, sy * sm);
    }
}

} // namespace

// ---------------------------------------------------------------------------
// Sessions
// ---------------------------------------------------------------------------

struct fa2_gpu {
    int device, n, use_double, adjust, barnes_hut, threads;
    hipStream_t stream;
    hipEvent_t ev[2] = {nullptr, nullptr}; // around the last repulsion kernel
    int timed = 0;
    double t_send, t_compute, t_fetch;
    int slices, slice_len;
    DevBuf<double> d_in;       // x, y, mass, size (4n) or px, py, pm, ps
    DevBuf<unsigned char> d_p; // packed nodes, float4 or double4
    DevBuf<unsigned char> d_partial;
    DevBuf<double2> d_f64;
    DevBuf<unsigned char> d_f; // BH forces in tree order
    DevBuf<fa2_region> d_reg;
    DevBuf<unsigned char> d_rr; // double sessions: region records (cx, cy, mass, size^2)
    DevBuf<int4> d_ri;          // and (start, count, last child, skip)
    DevBuf<HostRec> d_hrec;     // single sessions: both in one 32-byte record
    PinBuf<double> h_in;
    PinBuf<double2> h_f64;
    PinBuf<unsigned char> h_f;
    PinBuf<fa2_region> h_reg;
};

extern "C" void fa2_gpu_close(fa2_gpu *g)
{
    if (!g) return;
    if (g->stream) (void) hipStreamSynchronize(g->stream);
    for (hipEvent_t e : g->ev)
        if (e) (void) hipEventDestroy(e);
    delete g;
}

extern "C" int fa2_gpu_kernel_time(const fa2_gpu *g, double *ms)
{
    if (!g || !g->timed) {
        set_message("no repulsion kernel has run in this session");
        return 1;
    }
    float t = 0.0f;
    const hipError_t e = hipEventElapsedTime(&t, g->ev[0], g->ev[1]);
    if (e != hipSuccess) {
        set_hip_error("could not read the kernel time", e);
        return 1;
    }
    *ms = (double) t;
    return 0;
}

extern "C" void fa2_gpu_times(const fa2_gpu *g, double times[3])
{
    times[0] = g->t_send;
    times[1] = g->t_compute;
    times[2] = g->t_fetch;
}

extern "C" fa2_gpu *fa2_gpu_open(int device, int use_double, int n, int adjust,
                                 int barnes_hut, int threads)
{
    if (device < 0 || device >= device_count()) {
        set_message("there is no GPU device %d", device + 1);
        return nullptr;
    }
    if (n < 1 || n > (1 << 27)) {
        set_message("the GPU backend handles at most %d nodes", 1 << 27);
        return nullptr;
    }
    hipStream_t st = device_stream(device);
    if (!st) return nullptr;
    fa2_gpu *g = new (std::nothrow) fa2_gpu();
    if (!g) {
        set_message("out of memory");
        return nullptr;
    }
    g->device = device;
    g->n = n;
    g->use_double = use_double ? 1 : 0;
    g->adjust = adjust ? 1 : 0;
    g->barnes_hut = barnes_hut ? 1 : 0;
    g->threads = threads > 1 ? threads : 1;
    g->stream = st;
    const size_t nn = (size_t) n;
    const size_t rs = g->use_double ? sizeof(double) : sizeof(float);

    int slices, slice_len;
    exact_plan(n, &slices, &slice_len);
    g->slices = slices;
    g->slice_len = slice_len;

    HipStatus hs;
    hs.check(g->d_in.alloc(4 * nn), "alloc");
    // packed nodes, with the padding records that the Barnes-Hut walk loads
    // after the last node (zero)
    hs.check(g->d_p.alloc(4 * (nn + kHostLeafLoad) * rs), "alloc");
    if (hs.ok()) hs.check(hipMemset(g->d_p.p, 0, 4 * (nn + kHostLeafLoad) * rs), "clear");
    hs.check(g->h_in.alloc(4 * nn), "alloc");
    if (g->barnes_hut) {
        hs.check(g->d_f.alloc(2 * nn * sizeof(double)), "alloc"); // FP64 forces
        hs.check(g->h_f.alloc(2 * nn * sizeof(double)), "alloc");
        hs.check(g->d_reg.alloc(2 * nn), "alloc");
        hs.check(g->h_reg.alloc(2 * nn), "alloc");
        if (g->use_double) {
            hs.check(g->d_rr.alloc(2 * nn * 4 * rs), "alloc");
            hs.check(g->d_ri.alloc(2 * nn), "alloc");
        } else {
            hs.check(g->d_hrec.alloc(2 * nn), "alloc");
        }
    } else {
        if (!g->use_double) hs.check(g->d_partial.alloc((size_t) slices * nn * 2 * rs), "alloc");
        hs.check(g->d_f64.alloc(nn), "alloc");
        hs.check(g->h_f64.alloc(nn), "alloc");
    }
    if (!hs.ok()) {
        set_hip_error("could not allocate memory on the device", hs.err);
        fa2_gpu_close(g);
        return nullptr;
    }
    return g;
}

template <typename R>
static hipError_t launch_exact(fa2_gpu *g, double scaling)
{
    using V4 = typename Vec<R>::v4;
    using V2 = typename Vec<R>::v2;
    const int n = g->n;
    const double *in = g->d_in.p;
    if (!g->ev[0]) {
        (void) hipEventCreate(&g->ev[0]);
        (void) hipEventCreate(&g->ev[1]);
    }
    if (std::is_same<R, double>::value) { // the reference's operations, in order
        (void) hipEventRecord(g->ev[0], g->stream);
        const hipError_t e =
            launch_exact_f64(g->stream, n, scaling, g->adjust, in, in + n, in + 2 * (size_t) n,
                             g->adjust ? in + 3 * (size_t) n : nullptr, g->d_f64.p);
        (void) hipEventRecord(g->ev[1], g->stream);
        g->timed = 1;
        return e;
    }
    k_pack_nodes<R><<<ceil_div(n, 256), 256, 0, g->stream>>>(
        n, in, in + n, in + 2 * (size_t) n, g->adjust ? in + 3 * (size_t) n : nullptr,
        (V4 *) g->d_p.p);
    dim3 grid(ceil_div(n, kExactThreads * kExactTargets), g->slices);
    (void) hipEventRecord(g->ev[0], g->stream);
    if (g->adjust)
        k_exact<R, true><<<grid, kExactThreads, 0, g->stream>>>(
            n, g->slice_len, (const V4 *) g->d_p.p, (V2 *) g->d_partial.p);
    else
        k_exact<R, false><<<grid, kExactThreads, 0, g->stream>>>(
            n, g->slice_len, (const V4 *) g->d_p.p, (V2 *) g->d_partial.p);
    (void) hipEventRecord(g->ev[1], g->stream);
    g->timed = 1;
    k_sum_slices<R><<<ceil_div(n, 256), 256, 0, g->stream>>>(
        n, g->slices, (const V2 *) g->d_partial.p, in + 2 * (size_t) n, scaling, g->d_f64.p);
    return hipGetLastError();
}

extern "C" int fa2_gpu_repulse_exact(fa2_gpu *g, const double *x, const double *y,
                                     const double *mass, const double *size,
                                     double scaling, double *dx, double *dy)
{
    if (!g || g->barnes_hut) {
        set_message("this GPU session does not compute the exact repulsion");
        return 1;
    }
    const size_t n = (size_t) g->n;
    HipStatus hs;
    double t0 = fa2_now(), t1;
    (void) hipSetDevice(g->device);
    memcpy(g->h_in.p, x, n * sizeof(double));
    memcpy(g->h_in.p + n, y, n * sizeof(double));
    memcpy(g->h_in.p + 2 * n, mass, n * sizeof(double));
    if (g->adjust && size) memcpy(g->h_in.p + 3 * n, size, n * sizeof(double));
    else if (g->adjust) memset(g->h_in.p + 3 * n, 0, n * sizeof(double));
    hs.check(hipMemcpyAsync(g->d_in.p, g->h_in.p, 4 * n * sizeof(double),
                            hipMemcpyHostToDevice, g->stream), "upload");
    t1 = fa2_now();
    g->t_send += t1 - t0;
    t0 = t1;
    if (hs.ok()) hs.check(g->use_double ? launch_exact<double>(g, scaling)
                                        : launch_exact<float>(g, scaling), "launch");
    if (hs.ok()) hs.check(hipStreamSynchronize(g->stream), "compute");
    t1 = fa2_now();
    g->t_compute += t1 - t0;
    t0 = t1;
    if (hs.ok()) hs.check(hipMemcpyAsync(g->h_f64.p, g->d_f64.p, n * sizeof(double2),
                                         hipMemcpyDeviceToHost, g->stream), "fetch");
    if (hs.ok()) hs.check(hipStreamSynchronize(g->stream), "fetch");
    if (!hs.ok()) {
        set_hip_error("the computation on the GPU failed", hs.err);
        return 1;
    }
    for (size_t i = 0; i < n; i++) {
        dx[i] += g->h_f64.p[i].x;
        dy[i] += g->h_f64.p[i].y;
    }
    g->t_fetch += fa2_now() - t0;
    return 0;
}

template <typename R>
static hipError_t launch_bh(fa2_gpu *g, int nreg, double scaling, double theta)
{
    using V4 = typename Vec<R>::v4;
    const int n = g->n;
    const double *in = g->d_in.p;
    k_pack_nodes<R><<<ceil_div(n, 256), 256, 0, g->stream>>>(
        n, in, in + n, in + 2 * (size_t) n, g->adjust ? in + 3 * (size_t) n : nullptr,
        (V4 *) g->d_p.p);
    if (std::is_same<R, double>::value)
        k_pack_regions<R><<<ceil_div(nreg, 256), 256, 0, g->stream>>>(
            nreg, g->d_reg.p, (V4 *) g->d_rr.p, g->d_ri.p);
    else
        k_pack_host_records<<<ceil_div(nreg, 256), 256, 0, g->stream>>>(nreg, g->d_reg.p,
                                                                        g->d_hrec.p);
    if (!g->ev[0]) {
        (void) hipEventCreate(&g->ev[0]);
        (void) hipEventCreate(&g->ev[1]);
    }
    (void) hipEventRecord(g->ev[0], g->stream);
    if (std::is_same<R, double>::value) { // the reference's operations, in order
        const hipError_t e =
            launch_bh_f64(g->stream, n, scaling, theta * theta, g->adjust, in, in + n,
                          in + 2 * (size_t) n, g->adjust ? in + 3 * (size_t) n : nullptr,
                          (const double4 *) g->d_rr.p, g->d_ri.p, (double2 *) g->d_f.p);
        (void) hipEventRecord(g->ev[1], g->stream);
        g->timed = 1;
        return e;
    }
    const float th2 = (float) (theta * theta);
    const unsigned blocks = (unsigned) ceil_div(n, kBhThreads);
    if (g->adjust)
        k_bh_walk<true><<<blocks, kBhThreads, 0, g->stream>>>(
            n, nreg, (float) scaling, th2, (const float4 *) g->d_p.p, g->d_hrec.p,
            (double2 *) g->d_f.p);
    else
        k_bh_walk<false><<<blocks, kBhThreads, 0, g->stream>>>(
            n, nreg, (float) scaling, th2, (const float4 *) g->d_p.p, g->d_hrec.p,
            (double2 *) g->d_f.p);
    (void) hipEventRecord(g->ev[1], g->stream);
    g->timed = 1;
    return hipGetLastError();
}

// The regions of a tree passed to fa2_gpu_repulse_bh: node ranges inside
// [0, n), links inside [0, nreg) or -1. Returns the first bad region, or -1.
static int bad_region(const fa2_tree *t, int n)
{
    for (int r = 0; r < t->nreg; r++) {
        const fa2_region &g = t->reg[r];
        if (g.start < 0 || g.count < 1 || g.start > n - g.count || g.nchild < 0 ||
            (g.nchild > 0 && (g.child < 0 || g.child > t->nreg - g.nchild)) || g.skip < -1 ||
            g.skip >= t->nreg)
            return r;
    }
    return -1;
}

extern "C" int fa2_gpu_repulse_bh(fa2_gpu *g, const fa2_tree *t, double scaling,
                                  double theta, double *dx, double *dy)
{
    if (!g || !g->barnes_hut) {
        set_message("this GPU session does not compute the Barnes-Hut repulsion");
        return 1;
    }
    const size_t n = (size_t) g->n;
    const int nreg = t->nreg;
    if (nreg < 1 || (size_t) nreg > 2 * n) {
        set_message("invalid Barnes-Hut tree (%d regions for %d nodes)", nreg, g->n);
        return 1;
    }
    const int bad = bad_region(t, g->n);
    if (bad >= 0) {
        set_message("invalid Barnes-Hut tree (region %d)", bad);
        return 1;
    }
    HipStatus hs;
    double t0 = fa2_now(), t1;
    (void) hipSetDevice(g->device);
    memcpy(g->h_in.p, t->px, n * sizeof(double));
    memcpy(g->h_in.p + n, t->py, n * sizeof(double));
    memcpy(g->h_in.p + 2 * n, t->pm, n * sizeof(double));
    if (g->adjust && t->ps) memcpy(g->h_in.p + 3 * n, t->ps, n * sizeof(double));
    else if (g->adjust) memset(g->h_in.p + 3 * n, 0, n * sizeof(double));
    memcpy(g->h_reg.p, t->reg, (size_t) nreg * sizeof(fa2_region));
    hs.check(hipMemcpyAsync(g->d_in.p, g->h_in.p, 4 * n * sizeof(double),
                            hipMemcpyHostToDevice, g->stream), "upload");
    hs.check(hipMemcpyAsync(g->d_reg.p, g->h_reg.p, (size_t) nreg * sizeof(fa2_region),
                            hipMemcpyHostToDevice, g->stream), "upload");
    t1 = fa2_now();
    g->t_send += t1 - t0;
    t0 = t1;
    if (hs.ok()) hs.check(g->use_double ? launch_bh<double>(g, nreg, scaling, theta)
                                        : launch_bh<float>(g, nreg, scaling, theta), "launch");
    if (hs.ok()) hs.check(hipStreamSynchronize(g->stream), "compute");
    t1 = fa2_now();
    g->t_compute += t1 - t0;
    t0 = t1;
    if (hs.ok()) hs.check(hipMemcpyAsync(g->h_f.p, g->d_f.p, 2 * n * sizeof(double),
                                         hipMemcpyDeviceToHost, g->stream), "fetch");
    if (hs.ok()) hs.check(hipStreamSynchronize(g->stream), "fetch");
    if (!hs.ok()) {
        set_hip_error("the computation on the GPU failed", hs.err);
        return 1;
    }
    const double *f = (const double *) g->h_f.p; // FP64 in both precisions
    for (size_t k = 0; k < n; k++) {
        const int i = t->perm[k];
        dx[i] += f[2 * k];
        dy[i] += f[2 * k + 1];
    }
    g->t_fetch += fa2_now() - t0;
    return 0;
}

// ---------------------------------------------------------------------------
// The exact repulsion without allocation: fa2_gpu_repulse_exact_noalloc
// ---------------------------------------------------------------------------

namespace {

// Largest node count, and the bytes of scratch it needs: in single
// precision the packed nodes (16 bytes each) and the per-slice partial
// forces (8 bytes per node and slice, about 2048 x 1024 entries at most);
// in double precision the FP64 forces (16 bytes per node). Up to 2^18 nodes
// this is at most 21.9 MiB (261,120 nodes in 9 slices).
constexpr int kNoAllocMaxNodes = 1 << 18;
constexpr size_t kScratchBytes = (size_t) 24 << 20;
constexpr int kScratchDevices = 64;

// The scratch of every call on a device. A file-scope device array: each
// GPU holds its own copy, from the first call on it to the end of the
// process. In host code `scratch` only names the symbol: kernels get the
// device address that hipGetSymbolAddress gives for the current device.
__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
unsigned char *g_scratch_ptr[kScratchDevices] = {};

// The device address of `scratch` on `device`, which must be current: looked
// up once per device, then kept.
unsigned char *scratch_address(int device)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            set_hip_error("could not find the scratch memory of the device", e);
            return nullptr;
        }
        g_scratch_ptr[device] = (unsigned char *) ptr;
    }
    return g_scratch_ptr[device];
}

inline size_t round256(size_t b) { return (b + 255) & ~(size_t) 255; }

} // namespace

extern "C" int fa2_gpu_repulse_exact_noalloc(int device, int use_double, int n, int adjust,
                                             const double *x, const double *y,
                                             const double *mass, const double *size,
                                             double scaling, double *dx, double *dy,
                                             void *stream)
{
    if (device < 0 || device >= device_count() || device >= kScratchDevices) {
        set_message("there is no GPU device %d", device + 1);
        return 1;
    }
    if (n < 1 || n > kNoAllocMaxNodes) {
        set_message("the exact repulsion without allocation handles 1 to %d nodes (got %d)",
                    kNoAllocMaxNodes, n);
        return 1;
    }
    if (!x || !y || !mass || !dx || !dy || (adjust && !size)) {
        set_message("missing device arrays (x, y, mass, dx, dy, and size with prevent_overlap)");
        return 1;
    }
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) {
        set_hip_error("could not select the GPU device", e);
        return 1;
    }
    unsigned char *base = scratch_address(device);
    if (!base) return 1;
    const hipStream_t st = (hipStream_t) stream;
    const size_t nn = (size_t) n;
    const unsigned nb = (unsigned) ceil_div(nn, 256);
    if (use_double) { // the reference's operations, in order, as fa2_gpu_repulse_exact
        double2 *f = (double2 *) base;
        e = launch_exact_f64(st, n, scaling, adjust != 0, x, y, mass, adjust ? size : nullptr, f);
        if (e == hipSuccess) {
            k_add_forces<<<nb, 256, 0, st>>>(n, f, dx, dy);
            e = hipGetLastError();
        }
    } else {
        int slices, slice_len;
        exact_plan(n, &slices, &slice_len);
        const size_t p_bytes = round256(nn * sizeof(float4));
        if (p_bytes + (size_t) slices * nn * sizeof(float2) > kScratchBytes) {
            set_message("the exact repulsion of %d nodes needs more scratch than the %zu bytes "
                        "held for it", n, kScratchBytes);
            return 1;
        }
        float4 *p = (float4 *) base;
        float2 *partial = (float2 *) (base + p_bytes);
        k_pack_nodes<float><<<nb, 256, 0, st>>>(n, x, y, mass, adjust ? size : nullptr, p);
        dim3 grid(ceil_div(n, kExactThreads * kExactTargets), slices);
        if (adjust)
            k_exact<float, true><<<grid, kExactThreads, 0, st>>>(n, slice_len, p, partial);
        else
            k_exact<float, false><<<grid, kExactThreads, 0, st>>>(n, slice_len, p, partial);
        k_sum_slices<float, true><<<nb, 256, 0, st>>>(n, slices, partial, mass, scaling,
                                                      nullptr, dx, dy);
        e = hipGetLastError();
    }
    if (e != hipSuccess) {
        set_hip_error("the computation on the GPU failed", e);
        return 1;
    }
    return 0;
}
