// Device-side preprocessing of the all-GPU layout: the work of
// C_forceatlas2_large before fa2_full_run (src/forceatlas2.c:866-902),
// for the graphs of a batch at once, straight into the batch's packed
// arrays.
//
//   mass[i]  = 1 + degree(i): every edge counts, a self-loop twice;
//   coef     = mean mass = (n + 2m) / n with dissuade_hubs, else 1;
//   row i    = the incidences of node i in edge order, self-loops left out;
//              edge e gives (from -> to) in row from and (to -> from) in
//              row to, both with the coefficient
//   inc_coef = coef * w[e] / mass[from[e]] with dissuade_hubs, else coef * w[e].
//
// The rows come from one stable radix sort of the 2m incidence entries by
// global node id (the batch sort's own kernels, over one segment): entries
// are generated in edge order, so a stable sort keeps each row in edge
// order. The row offsets are an exclusive scan of the per-node counts, by
// blocks: each block scans its 1024 counts in LDS and leaves its total, one
// workgroup scans the totals, and each block adds the sum before it.
// Degrees are integer atomics, so every output is deterministic and equal,
// bit for bit, to the host preprocessing.
#include <algorithm>
#include <climits>

#include "common.h"
#include "ingest.h"

namespace fa2 {

namespace {

// the graph of packed edge e: eoff[g] <= e < eoff[g + 1]
__device__ inline int edge_graph(long long e, const IngestGraph *__restrict__ gr, int G)
{
    int lo = 0, hi = G - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (gr[mid].eoff <= e) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ void k_ingest_count(long long M, int G, int N, const IngestGraph *__restrict__ gr,
                               const int *__restrict__ from, const int *__restrict__ to,
                               int *deg, int *cnt, uint32_t *key, uint32_t *val, int *bad)
{
    const long long e = (long long) blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= M) return;
    const IngestGraph g = gr[edge_graph(e, gr, G)];
    const int f = from[e], t = to[e];
    uint32_t ka = (uint32_t) N, kb = (uint32_t) N; // self-loops and bad edges sort last
    if (f < 0 || f >= g.n || t < 0 || t >= g.n) {
        atomicOr(bad, 1);
    } else {
        const int a = g.noff + f, b = g.noff + t;
        atomicAdd(&deg[a], 1);
        atomicAdd(&deg[b], 1);
        if (a != b) {
            atomicAdd(&cnt[a], 1);
            atomicAdd(&cnt[b], 1);
            ka = (uint32_t) a;
            kb = (uint32_t) b;
        }
    }
    key[2 * e] = ka;
    key[2 * e + 1] = kb;
    val[2 * e] = (uint32_t) (2 * e);
    val[2 * e + 1] = (uint32_t) (2 * e + 1);
}

template <typename R>
__global__ void k_ingest_mass(int N, const int *__restrict__ deg, R *mass)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) mass[i] = (R) (1.0 + (double) deg[i]);
}

template <typename R>
__global__ void k_ingest_fill(long long E, int G, const IngestGraph *__restrict__ gr,
                              const int *__restrict__ from, const int *__restrict__ to,
                              const double *__restrict__ w, const int *__restrict__ deg,
                              const uint32_t *__restrict__ val, const int *__restrict__ inc_off,
                              int N, int *inc_node, R *inc_coef)
{
    const long long k = (long long) blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= E || k >= inc_off[N]) return;
    const uint32_t v = val[k];
    const long long e = v >> 1;
    const IngestGraph g = gr[edge_graph(e, gr, G)];
    const int a = g.noff + from[e], b = g.noff + to[e];
    inc_node[k] = (v & 1) ? a : b;
    const double we = w ? w[e] : 1.0;
    const double c = g.outbound ? g.coef * we / (1.0 + (double) deg[a]) : g.coef * we;
    inc_coef[k] = (R) c;
}

constexpr int kScanThreads = 256, kScanItems = 4;
constexpr int kScanBlock = kScanThreads * kScanItems; // counts per block

// Exclusive scan of 256 values, one per thread of the workgroup; returns
// the thread's exclusive prefix and sets `total`.
__device__ inline long long block_exclusive(long long v, long long *lds, long long &total)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    long long s = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const long long u = __shfl_up(s, o, 64);
        if (lane >= o) s += u;
    }
    if (lane == 63) lds[w] = s;
    __syncthreads();
    long long before = 0;
    for (int k = 0; k < w; k++) before += lds[k];
    total = lds[0] + lds[1] + lds[2] + lds[3];
    __syncthreads();
    return before + s - v;
}

// out[i] = sum of in[0..i) within each block of kScanBlock values, and the
// block's total in sums[block].
__global__ void __launch_bounds__(kScanThreads)
k_scan_blocks(int n, const int *__restrict__ in, int *__restrict__ out,
              long long *__restrict__ sums)
{
    __shared__ long long lds[4];
    const int base = blockIdx.x * kScanBlock + threadIdx.x * kScanItems;
    int v[kScanItems];
    long long t = 0;
#pragma unroll
    for (int u = 0; u < kScanItems; u++) {
        v[u] = base + u < n ? in[base + u] : 0;
        t += v[u];
    }
    long long total;
    long long run = block_exclusive(t, lds, total);
#pragma unroll
    for (int u = 0; u < kScanItems; u++) {
        if (base + u < n) out[base + u] = (int) run;
        run += v[u];
    }
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// sums[b] = sum of the totals of blocks 0..b-1 (one workgroup).
__global__ void __launch_bounds__(kScanThreads)
k_scan_sums(int nb, long long *__restrict__ sums)
{
    __shared__ long long lds[4];
    long long carry = 0;
    for (int b0 = 0; b0 < nb; b0 += kScanThreads) {
        const int b = b0 + (int) threadIdx.x;
        const long long v = b < nb ? sums[b] : 0;
        long long total;
        const long long ex = block_exclusive(v, lds, total);
        if (b < nb) sums[b] = carry + ex;
        carry += total;
    }
}

__global__ void __launch_bounds__(kScanThreads)
k_scan_add(int n, const long long *__restrict__ sums, int *__restrict__ out)
{
    const int base = blockIdx.x * kScanBlock + threadIdx.x * kScanItems;
    const long long add = sums[blockIdx.x];
#pragma unroll
    for (int u = 0; u < kScanItems; u++)
        if (base + u < n) out[base + u] = (int) (out[base + u] + add);
}

int bits_for(uint32_t v)
{
    int b = 1;
    while (b < 32 && (v >> b) != 0) b++;
    return b;
}

} // namespace

template <typename R>
int ingest_graphs(hipStream_t st, hipStream_t ast, int N, const std::vector<IngestGraph> &graphs,
                  const std::vector<const fa2_edges *> &edges, DevBuf<R> &mass,
                  DevBuf<int> &inc_off, DevBuf<int> &inc_node, DevBuf<R> &inc_coef,
                  std::vector<int> &h_inc_off)
{
    const int G = (int) graphs.size();
    long long M = 0;
    bool any_w = false;
    for (int g = 0; g < G; g++) {
        M += edges[g]->m;
        any_w = any_w || edges[g]->w;
    }
    if (2 * M >= (long long) INT_MAX) {
        set_message("too many edges for one batch");
        return 1;
    }
    // packed endpoints (local ids) and weights
    std::vector<int> hf((size_t) M), ht((size_t) M);
    std::vector<double> hw(any_w ? (size_t) M : 0);
    for (int g = 0; g < G; g++) {
        const fa2_edges *E = edges[g];
        const size_t o = (size_t) graphs[g].eoff;
        if (E->m > 0) {
            std::copy(E->from, E->from + E->m, hf.begin() + o);
            std::copy(E->to, E->to + E->m, ht.begin() + o);
            if (any_w) {
                if (E->w) std::copy(E->w, E->w + E->m, hw.begin() + o);
                else std::fill(hw.begin() + o, hw.begin() + o + E->m, 1.0);
            }
        }
    }
    HipStatus hs;
    DevBuf<IngestGraph> d_gr;
    DevBuf<int> d_from, d_to, d_deg, d_cnt, d_bad;
    DevBuf<double> d_w;
    DevBuf<uint32_t> d_key[2], d_val[2];
    const size_t E2 = 2 * (size_t) M;
    hs.check(d_gr.alloc(G, ast), "alloc");
    hs.check(d_from.alloc((size_t) M, ast), "alloc");
    hs.check(d_to.alloc((size_t) M, ast), "alloc");
    hs.check(d_w.alloc(any_w ? (size_t) M : 0, ast), "alloc");
    hs.check(d_deg.alloc((size_t) N, ast), "alloc");
    hs.check(d_cnt.alloc((size_t) N + 1, ast), "alloc");
    hs.check(d_bad.alloc(1, ast), "alloc");
    for (int k = 0; k < 2; k++) {
        hs.check(d_key[k].alloc(E2, ast), "alloc");
        hs.check(d_val[k].alloc(E2, ast), "alloc");
    }
    hs.check(mass.alloc((size_t) N, ast), "alloc");
    hs.check(inc_off.alloc((size_t) N + 1, ast), "alloc");
    hs.check(inc_node.alloc(E2, ast), "alloc");
    hs.check(inc_coef.alloc(E2, ast), "alloc");
    auto up = [&](void *dst, const void *src, size_t bytes) {
        if (hs.ok() && bytes) hs.check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st), "upload");
    };
    up(d_gr.p, graphs.data(), G * sizeof(IngestGraph));
    up(d_from.p, hf.data(), (size_t) M * sizeof(int));
    up(d_to.p, ht.data(), (size_t) M * sizeof(int));
    if (any_w) up(d_w.p, hw.data(), (size_t) M * sizeof(double));
    if (hs.ok()) hs.check(hipMemsetAsync(d_deg.p, 0, (size_t) N * sizeof(int), st), "memset");
    if (hs.ok()) hs.check(hipMemsetAsync(d_cnt.p, 0, ((size_t) N + 1) * sizeof(int), st), "memset");
    if (hs.ok()) hs.check(hipMemsetAsync(d_bad.p, 0, sizeof(int), st), "memset");
    if (hs.ok() && M > 0) {
        k_ingest_count<<<(unsigned) ((M + 255) / 256), 256, 0, st>>>(
            M, G, N, d_gr.p, d_from.p, d_to.p, d_deg.p, d_cnt.p, d_key[0].p, d_val[0].p,
            d_bad.p);
        hs.check(hipGetLastError(), "launch");
    }
    if (hs.ok()) {
        k_ingest_mass<R><<<ceil_div(N, 256), 256, 0, st>>>(N, d_deg.p, mass.p);
        hs.check(hipGetLastError(), "launch");
    }
    // rows: exclusive sum of the incidence counts, then the stable sort
    const int nscan = N + 1, nb = (int) ceil_div((size_t) nscan, kScanBlock);
    DevBuf<long long> d_sums;
    hs.check(d_sums.alloc((size_t) nb, ast), "alloc");
    if (hs.ok()) {
        k_scan_blocks<<<nb, kScanThreads, 0, st>>>(nscan, d_cnt.p, inc_off.p, d_sums.p);
        k_scan_sums<<<1, kScanThreads, 0, st>>>(nb, d_sums.p);
        k_scan_add<<<nb, kScanThreads, 0, st>>>(nscan, d_sums.p, inc_off.p);
        hs.check(hipGetLastError(), "launch");
    }
    int sorted = 0;
    if (hs.ok() && E2 > 1) {
        uint32_t *kk[2] = {d_key[0].p, d_key[1].p}, *vv[2] = {d_val[0].p, d_val[1].p};
        if (radix_sort_pairs_u32(st, ast, (int) E2, bits_for((uint32_t) N), kk, vv, &sorted))
            return 1;
    }
    if (hs.ok() && E2 > 0) {
        k_ingest_fill<R><<<(unsigned) ((E2 + 255) / 256), 256, 0, st>>>(
            (long long) E2, G, d_gr.p, d_from.p, d_to.p, any_w ? d_w.p : nullptr, d_deg.p,
            d_val[sorted].p, inc_off.p, N, inc_node.p, inc_coef.p);
        hs.check(hipGetLastError(), "launch");
    }
    int bad = 0;
    h_inc_off.resize((size_t) N + 1);
    if (hs.ok())
        hs.check(hipMemcpyAsync(h_inc_off.data(), inc_off.p, ((size_t) N + 1) * sizeof(int),
                                hipMemcpyDeviceToHost, st),
                 "download");
    if (hs.ok())
        hs.check(hipMemcpyAsync(&bad, d_bad.p, sizeof(int), hipMemcpyDeviceToHost, st), "download");
    if (hs.ok()) hs.check(hipStreamSynchronize(st), "ingest");
    if (!hs.ok()) {
        set_hip_error("could not prepare the graphs on the device", hs.err);
        return 1;
    }
    if (bad) {
        set_message("edge refers to a node that does not exist");
        return 1;
    }
    return 0;
}

template int ingest_graphs<float>(hipStream_t, hipStream_t, int, const std::vector<IngestGraph> &,
                                  const std::vector<const fa2_edges *> &, DevBuf<float> &,
                                  DevBuf<int> &, DevBuf<int> &, DevBuf<float> &,
                                  std::vector<int> &);
template int ingest_graphs<double>(hipStream_t, hipStream_t, int, const std::vector<IngestGraph> &,
                                   const std::vector<const fa2_edges *> &, DevBuf<double> &,
                                   DevBuf<int> &, DevBuf<int> &, DevBuf<double> &,
                                   std::vector<int> &);

} // namespace fa2
