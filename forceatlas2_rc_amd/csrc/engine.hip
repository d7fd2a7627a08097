// The all-GPU ForceAtlas2 engine of libfa2hip.so: one or many graphs laid
// out side by side in HBM, every stage of every iteration on the device.
//
// It replaces the pipeline of the reference all-GPU layout
// (inst/opencl/pipeline.cl, host loop src/gpu_full.c:164-532). A single
// graph (fa2_full_run) is a batch of one; the batch entry points of
// fa2_hip.h pack many graphs of uneven size into the same launches.
//
// Layout in HBM (N nodes of G graphs, graph g owning the node range
// [noff_g, noff_g + n_g), the same range before and after sorting):
//   pos    R2   positions used by the kernels (float32 copies in single
//               precision, the positions themselves in double precision)
//   acc    f64x2 position accumulators (single precision only)
//   F[2]   f64x2 force of this iteration / of the previous one (ping-pong)
//   Fa     f64x2 the iteration's attraction (computed on the second stream
//               beside the traversal; k_swing adds it to the force)
//   keys/vals   Morton codes and node ids, two buffers for the radix sort
//   psort  R4   (x, y, mass, radius) in sorted order
//   tn     {R4 (cx, cy, mass, size^2), int4 (left, skip, first, last)} for
//               the n_g - 1 regions of each graph, stored at
//               [noff_g - g, noff_g - g + n_g - 1)
//   CSR    inc_off (N+1), inc_node (I), inc_coef (I, R)
// Node references inside a tree: r >= 0 a region, r < 0 the leaf at sorted
// position ~r. The traversal's own links are record indices (record_of).
//
// Every reduction is made in a fixed order: results are reproducible run
// to run, and a graph gives the same result alone and in a batch as long
// as the tiles of its nodes do not change (they depend only on the graph).
#include <algorithm>
#include <atomic>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <tuple>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "common.h"
#include "kernels_common.h"
#include "engine.h"
#include "ingest.h"
#include "../../include/fa2.h"

namespace fa2 {

// The traversal reads one record per region and per leaf: region r at r,
// the leaf at sorted position p at NI + p. Its links (left child, skip) are
// record indices, and kNoRecord ends a walk.
constexpr int kNoRecord = -1;
__device__ inline int record_of(int r, int NI) { return r >= 0 ? r : NI + ~r; }
constexpr int kLeafSize = 8; // regions of at most 8 nodes are visited node by node
// the lengths of the summary unit lists (see k_karras) are counted with
// atomics from the whole device: one counter per 128-byte line, so that they
// spread over the L2 channels
constexpr int kCountStride = 32;

// ---------------------------------------------------------------------------
// Device descriptors
// ---------------------------------------------------------------------------

// GF_UNIFORM: every node's incidences carry one coefficient (equal weights
// without dissuade_hubs), so the attraction reads it once per node
enum : int { GF_STRONG = 1, GF_LINLOG = 2, GF_ADJUST = 4, GF_FIXED = 8, GF_UNIFORM = 16 };

struct GraphDesc {
    int noff, n, ioff, toff;
    int ntiles, flags, iters, pad;
    double scaling, gravity, theta2, jt;
};

struct GraphState {
    double speed, eff, tsw, ttr;
    int done;   // iterations applied
    int status; // 0, or 2 when the totals were not numbers
    int active;  // 1 when this iteration moves the graph
    int retired; // 1 once the host took the graph out: no forces computed
};

struct Tile {
    int g, lo, hi, j; // graph, node range [lo, hi), tile index in the graph
};

// Tile b of a launch. When the batch is a single graph of N1 nodes (N1 > 0),
// its tiles are those of one range, computed here rather than loaded: the
// kernel's first loads then wait on no other.
__device__ inline Tile tile_of(const Tile *__restrict__ tiles, int b, int N1, int per_tile)
{
    if (N1 > 0) return Tile{0, b * per_tile, min(N1, (b + 1) * per_tile), b};
    return tiles[b];
}

template <typename R> struct alignas(32) TNode {
    typename Vec<R>::v4 c; // centre x, centre y, mass, size^2
    int4 l;                // left child, skip, first, last (sorted positions)
};

// ---------------------------------------------------------------------------
// Codes: bounding boxes and Morton keys (pipeline.cl:61-130)
// ---------------------------------------------------------------------------

template <typename R>
__device__ inline void block_minmax4(R &a, R &b, R &c, R &d, R *lds /* 4*4 */)
{
    a = wave_min(a);
    b = wave_min(b);
    c = wave_max(c);
    d = wave_max(d);
    const int w = threadIdx.x >> 6;
    if (lane_id() == 0) {
        lds[4 * w] = a;
        lds[4 * w + 1] = b;
        lds[4 * w + 2] = c;
        lds[4 * w + 3] = d;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < (int) (blockDim.x >> 6); k++) {
            a = min(a, lds[4 * k]);
            b = min(b, lds[4 * k + 1]);
            c = max(c, lds[4 * k + 2]);
            d = max(d, lds[4 * k + 3]);
        }
    }
}

// Bounding box of each tile (used once before the first iteration; later
// the displacement kernel produces it).
template <typename R>
__global__ void __launch_bounds__(256)
k_tile_bbox(const Tile *__restrict__ tiles, const typename Vec<R>::v2 *__restrict__ pos,
            typename Vec<R>::v4 *__restrict__ tbox)
{
    __shared__ R lds[16];
    const Tile t = tiles[blockIdx.x];
    R a = (R) INFINITY, b = (R) INFINITY, c = (R) -INFINITY, d = (R) -INFINITY;
    for (int i = t.lo + threadIdx.x; i < t.hi; i += blockDim.x) {
        const typename Vec<R>::v2 p = pos[i];
        a = min(a, p.x);
        b = min(b, p.y);
        c = max(c, p.x);
        d = max(d, p.y);
    }
    block_minmax4<R>(a, b, c, d, lds);
    if (threadIdx.x == 0) tbox[blockIdx.x] = mk4<R>(a, b, c, d);
}

__device__ inline uint32_t spread16(uint32_t v)
{
    v &= 0x0000FFFFu;
    v = (v | (v << 8)) & 0x00FF00FFu;
    v = (v | (v << 4)) & 0x0F0F0F0Fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}

__device__ inline float div_rn(float a, float b) { return __fdiv_rn(a, b); }
__device__ inline double div_rn(double a, double b) { return __ddiv_rn(a, b); }

constexpr int kSortThreads = 256;
constexpr int kSortItems = 8;
static_assert(kSortThreads == 256, "the sort kernels hold one digit per thread, four waves");
constexpr int kTile = kSortThreads * kSortItems; // 2048 nodes

// Morton code of every node: 16 bits per axis, x in the even bits, computed
// with IEEE arithmetic as written (the codes must match the reference bit for
// bit, pipeline.cl:116-130).
//
// The graph's bounding box (pipeline.cl:61-99) comes first: every workgroup
// of the graph reduces the boxes of its tiles, which the last displacement
// (or k_tile_bbox) made, itself (a minimum and a maximum are the same in
// any order), while its own positions load. The tile's counts of the sort's
// first digit (the low 8 bits) are made here too, as k_sort_upsweep makes
// them for the later passes, so the first pass starts at its row scans.
template <typename R>
__global__ void __launch_bounds__(kSortThreads)
k_morton(const Tile *__restrict__ tiles, const GraphDesc *__restrict__ gd,
         const GraphState *__restrict__ gs, const typename Vec<R>::v2 *__restrict__ pos,
         const typename Vec<R>::v4 *__restrict__ tbox, uint32_t *__restrict__ key,
         uint32_t *__restrict__ val, uint32_t *__restrict__ hist, int UT, int *__restrict__ ucnt,
         int N1)
{
    using V2 = typename Vec<R>::v2;
    using V4 = typename Vec<R>::v4;
    __shared__ uint32_t h[kSortThreads]; // one digit's count per thread
    __shared__ R s_box[4][4];            // each wave's part of the graph's box
    // the lengths of the summary unit lists, filled by k_karras
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < UT; q += gridDim.x * blockDim.x)
        ucnt[q * kCountStride] = 0;
    const Tile t = tile_of(tiles, blockIdx.x, N1, kTile);
    if (gs[t.g].retired) return; // its tree is not used any more
    const GraphDesc g = gd[t.g];
    const int tid = (int) threadIdx.x;
    h[tid] = 0;
    // the thread's positions (a slot past the tile's end reads its last
    // position and takes no part)
    V2 pv[kSortItems];
#pragma unroll
    for (int r = 0; r < kSortItems; r++) pv[r] = pos[min(t.lo + tid + r * kSortThreads, t.hi - 1)];
    R a = (R) INFINITY, b = (R) INFINITY, c = (R) -INFINITY, d = (R) -INFINITY;
    // four tile boxes per thread in flight at once (a slot past the end
    // reads the last box again, which changes no minimum or maximum)
    for (int j = tid; j < g.ntiles; j += 4 * kSortThreads) {
        V4 v[4];
#pragma unroll
        for (int u = 0; u < 4; u++) v[u] = tbox[g.toff + min(j + u * kSortThreads, g.ntiles - 1)];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            a = min(a, v[u].x);
            b = min(b, v[u].y);
            c = max(c, v[u].z);
            d = max(d, v[u].w);
        }
    }
    a = wave_min(a);
    b = wave_min(b);
    c = wave_max(c);
    d = wave_max(d);
    const int w = tid >> 6;
    if (lane_id() == 0) {
        s_box[w][0] = a;
        s_box[w][1] = b;
        s_box[w][2] = c;
        s_box[w][3] = d;
    }
    __syncthreads();
    const V4 bb = mk4<R>(min(min(s_box[0][0], s_box[1][0]), min(s_box[2][0], s_box[3][0])),
                         min(min(s_box[0][1], s_box[1][1]), min(s_box[2][1], s_box[3][1])),
                         max(max(s_box[0][2], s_box[1][2]), max(s_box[2][2], s_box[3][2])),
                         max(max(s_box[0][3], s_box[1][3]), max(s_box[2][3], s_box[3][3])));
    const R ex = bb.z - bb.x, ey = bb.w - bb.y;
    const R extent = ex > ey ? ex : ey;
    const R scale = extent > (R) 0 ? div_rn((R) 65535, extent) : (R) 0;
#pragma unroll
    for (int r = 0; r < kSortItems; r++) {
        const int i = t.lo + tid + r * kSortThreads;
        if (i >= t.hi) break;
        R fx = (pv[r].x - bb.x) * scale;
        R fy = (pv[r].y - bb.y) * scale;
        fx = min(max(fx, (R) 0), (R) 65535);
        fy = min(max(fy, (R) 0), (R) 65535);
        const uint32_t k = spread16((uint32_t) fx) | (spread16((uint32_t) fy) << 1);
        key[i] = k;
        val[i] = (uint32_t) i;
        atomicAdd(&h[k & 255u], 1u);
    }
    __syncthreads();
    hist[(size_t) g.toff * 256 + (size_t) tid * g.ntiles + t.j] = h[tid];
}

// ---------------------------------------------------------------------------
// Segmented stable LSD radix sort, 4 passes of 8 bits (pipeline.cl:137-188).
// Tiles never straddle two graphs; the scan is made per (graph, digit) row,
// so every pass keeps each graph in its own node range.
// hist row (g, d) starts at toff_g*256 + d*ntiles_g.
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(kSortThreads)
k_sort_upsweep(const Tile *__restrict__ tiles, const GraphDesc *__restrict__ gd,
               const GraphState *__restrict__ gs, const uint32_t *__restrict__ key, int shift,
               uint32_t *__restrict__ hist, int N1)
{
    __shared__ uint32_t h[256];
    const Tile t = tile_of(tiles, blockIdx.x, N1, kTile);
    if (gs[t.g].retired) return;
    h[threadIdx.x] = 0;
    // the thread's keys, every load in flight at once (a slot past the
    // tile's end reads the tile's last key and is not counted)
    uint32_t k8[kSortItems];
#pragma unroll
    for (int r = 0; r < kSortItems; r++)
        k8[r] = key[min(t.lo + (int) threadIdx.x + r * kSortThreads, t.hi - 1)];
#pragma unroll
    for (int r = 0; r < kSortItems; r++) asm volatile("" : "+v"(k8[r]));
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSortItems; r++)
        if (t.lo + (int) threadIdx.x + r * kSortThreads < t.hi)
            atomicAdd(&h[(k8[r] >> shift) & 255u], 1u);
    __syncthreads();
    const GraphDesc g = gd[t.g];
    hist[(size_t) g.toff * 256 + (size_t) threadIdx.x * g.ntiles + t.j] = h[threadIdx.x];
}

// The tile counts of a sort pass hold one row of ntiles counts per (graph,
// digit); each row is scanned (exclusive, in place) and its total written to
// rowtot[256 g + digit]. A graph's rows go to the kernel for its size, over a
// list of such graphs (glist): one thread per row up to kRowThreadTiles
// tiles, one wave per row up to 64, one workgroup per row beyond. (One wave
// for a row of a few tiles would make a batch of many small graphs launch
// 256 mostly idle waves per graph.)
constexpr int kRowThreadTiles = 16;

__global__ void __launch_bounds__(256)
k_sort_rowscan_thread(int rows, const int *__restrict__ glist, const GraphDesc *__restrict__ gd,
                      const GraphState *__restrict__ gs, uint32_t *__restrict__ hist,
                      uint32_t *__restrict__ rowtot)
{
    const int r = blockIdx.x * 256 + (int) threadIdx.x;
    if (r >= rows) return;
    const int g = glist[r >> 8], d = r & 255;
    if (gs[g].retired) return;
    const GraphDesc G = gd[g];
    uint32_t *row = hist + (size_t) G.toff * 256 + (size_t) d * G.ntiles;
    uint32_t v[kRowThreadTiles];
#pragma unroll
    for (int j = 0; j < kRowThreadTiles; j++) v[j] = j < G.ntiles ? row[j] : 0u;
    uint32_t carry = 0;
#pragma unroll
    for (int j = 0; j < kRowThreadTiles; j++) {
        if (j < G.ntiles) row[j] = carry;
        carry += v[j];
    }
    rowtot[g * 256 + d] = carry;
}

__global__ void __launch_bounds__(256)
k_sort_rowscan(int rows, const int *__restrict__ glist, const GraphDesc *__restrict__ gd,
               const GraphState *__restrict__ gs, uint32_t *__restrict__ hist,
               uint32_t *__restrict__ rowtot)
{
    const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= rows) return;
    const int g = glist[r >> 8], d = r & 255;
    if (gs[g].retired) return;
    const GraphDesc G = gd[g];
    uint32_t *row = hist + (size_t) G.toff * 256 + (size_t) d * G.ntiles;
    const int lane = lane_id();
    uint32_t carry = 0;
    for (int b = 0; b < G.ntiles; b += 64) {
        const int j = b + lane;
        const uint32_t v = j < G.ntiles ? row[j] : 0u;
        uint32_t s = v; // inclusive scan across the wave
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t u = __shfl_up(s, o, 64);
            if (lane >= o) s += u;
        }
        if (j < G.ntiles) row[j] = carry + s - v;
        carry += __shfl(s, 63, 64);
    }
    if (lane == 0) rowtot[g * 256 + d] = carry;
}

__global__ void __launch_bounds__(256)
k_sort_rowscan_block(const int *__restrict__ glist, const GraphDesc *__restrict__ gd,
                     const GraphState *__restrict__ gs, uint32_t *__restrict__ hist,
                     uint32_t *__restrict__ rowtot)
{
    __shared__ uint32_t wsum[4];
    const int r = blockIdx.x;
    const int g = glist[r >> 8], d = r & 255;
    if (gs[g].retired) return;
    const GraphDesc G = gd[g];
    uint32_t *row = hist + (size_t) G.toff * 256 + (size_t) d * G.ntiles;
    const int lane = lane_id(), w = threadIdx.x >> 6;
    uint32_t carry = 0;
    // chunks of 256 tiles, kRowChunks of them loaded at once (one round trip
    // for rows of up to 2048 tiles), then scanned one after the other
    constexpr int kRowChunks = 8;
    for (int b0 = 0; b0 < G.ntiles; b0 += 256 * kRowChunks) {
        uint32_t v[kRowChunks];
#pragma unroll
        for (int c = 0; c < kRowChunks; c++) {
            const int j = b0 + 256 * c + (int) threadIdx.x;
            v[c] = j < G.ntiles ? row[j] : 0u;
        }
#pragma unroll
        for (int c = 0; c < kRowChunks; c++) {
            const int j = b0 + 256 * c + (int) threadIdx.x;
            if (b0 + 256 * c >= G.ntiles) break; // uniform
            uint32_t s = v[c];
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t u = __shfl_up(s, o, 64);
                if (lane >= o) s += u;
            }
            if (lane == 63) wsum[w] = s;
            __syncthreads();
            uint32_t before = carry;
            for (int k = 0; k < w; k++) before += wsum[k];
            const uint32_t total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
            if (j < G.ntiles) row[j] = before + s - v[c];
            __syncthreads();
            carry += total;
        }
    }
    if (threadIdx.x == 0) rowtot[g * 256 + d] = carry;
}

// Stable scatter of one tile. Wave w ranks its 512 keys in 8 rounds of 64
// (ballot match on the 8 digit bits), then the waves are combined in order:
// the rank of a key counts the equal digits before it in the tile. Keys go
// through LDS so that the global writes of one digit are contiguous. (The
// node records in sorted order, pipeline.cl:195-207, are gathered by
// k_aggregate_tile, beside its other first loads.)
__global__ void __launch_bounds__(kSortThreads)
k_sort_downsweep(const Tile *__restrict__ tiles, const GraphDesc *__restrict__ gd,
                 const GraphState *__restrict__ gs, const uint32_t *__restrict__ key_in,
                 const uint32_t *__restrict__ val_in,
                 int shift, const uint32_t *__restrict__ hist, const uint32_t *__restrict__ rowtot,
                 uint32_t *__restrict__ key_out, uint32_t *__restrict__ val_out, int N1)
{
    __shared__ uint32_t wcnt[4][256];
    __shared__ int goff[256];
    __shared__ uint32_t skey[kTile];
    __shared__ uint32_t sval[kTile];
    __shared__ uint32_t scan_tmp[256];

    const Tile t = tile_of(tiles, blockIdx.x, N1, kTile);
    if (gs[t.g].retired) return;
    const GraphDesc G = gd[t.g];
    const int cnt = t.hi - t.lo;
    const int w = threadIdx.x >> 6, lane = lane_id();
    const int tid = threadIdx.x;
    for (int k = 0; k < 4; k++) wcnt[k][tid] = 0;
    __syncthreads();

    // the wave's keys and values, and the tile's row offsets and the
    // graph's digit totals used further down: every load in flight at once
    // (a slot past the tile's end reads its last key and takes no part)
    uint32_t kk[kSortItems], vv[kSortItems];
#pragma unroll
    for (int r = 0; r < kSortItems; r++) {
        const int q = min(w * (kSortItems * 64) + r * 64 + lane, cnt - 1);
        kk[r] = key_in[t.lo + q];
        vv[r] = val_in[t.lo + q];
    }
    uint32_t rt = rowtot[t.g * 256 + tid];
    uint32_t pre = hist[(size_t) G.toff * 256 + (size_t) tid * G.ntiles + t.j];
#pragma unroll
    for (int r = 0; r < kSortItems; r++) asm volatile("" : "+v"(kk[r]), "+v"(vv[r]));
    asm volatile("" : "+v"(rt), "+v"(pre));
    int dig[kSortItems], rank[kSortItems];
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
    for (int r = 0; r < kSortItems; r++) {
        const int q = w * (kSortItems * 64) + r * 64 + lane;
        const bool ok = q < cnt;
        const int d = (int) ((kk[r] >> shift) & 255u);
        dig[r] = ok ? d : -1;
        uint64_t peers = __ballot(ok);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const bool bit = (d >> b) & 1;
            const uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const int before = __popcll(peers & lt);
        // the lowest lane of each digit reserves the slots of its peers
        uint32_t base = 0;
        if (ok && before == 0) base = atomicAdd(&wcnt[w][d], (uint32_t) __popcll(peers));
        const int leader = __ffsll((unsigned long long) peers) - 1;
        base = __shfl(base, leader < 0 ? lane : leader, 64);
        rank[r] = (int) base + before;
    }
    __syncthreads();

    // per digit: counts of the waves, tile total, then the tile's digit
    // starts; and the digit bases of the graph from its 256 row totals. Both
    // are exclusive scans over the 256 digits, one digit per thread: within
    // each wave by shuffles, then across the four waves.
    const uint32_t c0 = wcnt[0][tid], c1 = wcnt[1][tid], c2 = wcnt[2][tid], c3 = wcnt[3][tid];
    const uint32_t tot = c0 + c1 + c2 + c3;
    uint32_t a = tot, b = rt; // inclusive scans within the wave
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t ua = __shfl_up(a, o, 64), ub = __shfl_up(b, o, 64);
        if (lane >= o) {
            a += ua;
            b += ub;
        }
    }
    if (lane == 63) {
        scan_tmp[w] = a;
        scan_tmp[4 + w] = b;
    }
    __syncthreads();
    uint32_t pa = 0, pb = 0;
    for (int v = 0; v < w; v++) {
        pa += scan_tmp[v];
        pb += scan_tmp[4 + v];
    }
    const uint32_t start = pa + a - tot;
    const uint32_t gbase = pb + b - rt;
    wcnt[0][tid] = start;
    wcnt[1][tid] = start + c0;
    wcnt[2][tid] = start + c0 + c1;
    wcnt[3][tid] = start + c0 + c1 + c2;
    goff[tid] = (int) (G.noff + gbase + pre) - (int) start;
    __syncthreads();

#pragma unroll
    for (int r = 0; r < kSortItems; r++) {
        if (dig[r] < 0) continue;
        const int p = (int) wcnt[w][dig[r]] + rank[r];
        skey[p] = kk[r];
        sval[p] = vv[r];
    }
    __syncthreads();
    // each thread moves kSortItems keys
    uint32_t ko[kSortItems], vo[kSortItems];
    int dsto[kSortItems];
#pragma unroll
    for (int r = 0; r < kSortItems; r++) {
        const int p = tid + r * kSortThreads;
        const bool in = p < cnt;
        ko[r] = in ? skey[p] : 0u;
        vo[r] = in ? sval[p] : 0u;
        dsto[r] = in ? goff[(ko[r] >> shift) & 255u] + p : -1;
    }
#pragma unroll
    for (int r = 0; r < kSortItems; r++) {
        const int dst = dsto[r];
        if (dst < G.noff || dst >= G.noff + G.n) continue; // past the tile's end only
        key_out[dst] = ko[r];
        val_out[dst] = vo[r];
    }
}

// Stable LSD radix sort of n (key, value) pairs on the low `bits` bits of
// the keys (ceil(bits / 8) passes of 8 bits), with the kernels above over a
// single segment: the pairs are one graph of ceil(n / 2048) tiles. Input in
// k[0], v[0]; the sorted pairs end in k[*out], v[*out]. The tile table and
// counts are allocated in the order of `ast` (synchronously when null); the
// kernels run on `st`, which the caller synchronises before it frees k, v.
int radix_sort_pairs_u32(hipStream_t st, hipStream_t ast, int n, int bits, uint32_t *k[2],
                         uint32_t *v[2], int *out)
{
    *out = 0;
    if (n <= 1) return 0;
    const int T = (int) ceil_div((size_t) n, kTile);
    std::vector<Tile> tiles(T);
    for (int j = 0; j < T; j++) tiles[j] = Tile{0, j * kTile, std::min(n, (j + 1) * kTile), j};
    GraphDesc gd;
    memset(&gd, 0, sizeof(gd));
    gd.n = n;
    gd.ntiles = T;
    GraphState gs;
    memset(&gs, 0, sizeof(gs));
    const int zero = 0;
    HipStatus hs;
    DevBuf<Tile> d_tiles;
    DevBuf<GraphDesc> d_gd;
    DevBuf<GraphState> d_gs;
    DevBuf<int> d_glist;
    DevBuf<uint32_t> d_hist, d_rowtot;
    hs.check(d_tiles.alloc(T, ast), "alloc");
    hs.check(d_gd.alloc(1, ast), "alloc");
    hs.check(d_gs.alloc(1, ast), "alloc");
    hs.check(d_glist.alloc(1, ast), "alloc");
    hs.check(d_hist.alloc((size_t) 256 * T, ast), "alloc");
    hs.check(d_rowtot.alloc(256, ast), "alloc");
    if (hs.ok())
        hs.check(hipMemcpyAsync(d_tiles.p, tiles.data(), T * sizeof(Tile), hipMemcpyHostToDevice,
                                st), "upload");
    if (hs.ok())
        hs.check(hipMemcpyAsync(d_gd.p, &gd, sizeof(gd), hipMemcpyHostToDevice, st), "upload");
    if (hs.ok())
        hs.check(hipMemcpyAsync(d_gs.p, &gs, sizeof(gs), hipMemcpyHostToDevice, st), "upload");
    if (hs.ok())
        hs.check(hipMemcpyAsync(d_glist.p, &zero, sizeof(int), hipMemcpyHostToDevice, st),
                 "upload");
    const int passes = std::max(1, (bits + 7) / 8);
    int in = 0;
    for (int pass = 0; hs.ok() && pass < passes; pass++) {
        k_sort_upsweep<<<T, kSortThreads, 0, st>>>(d_tiles.p, d_gd.p, d_gs.p, k[in], 8 * pass,
                                                    d_hist.p, n);
        if (T <= kRowThreadTiles)
            k_sort_rowscan_thread<<<1, 256, 0, st>>>(256, d_glist.p, d_gd.p, d_gs.p, d_hist.p,
                                                     d_rowtot.p);
        else if (T <= 64)
            k_sort_rowscan<<<64, 256, 0, st>>>(256, d_glist.p, d_gd.p, d_gs.p, d_hist.p,
                                               d_rowtot.p);
        else
            k_sort_rowscan_block<<<256, 256, 0, st>>>(d_glist.p, d_gd.p, d_gs.p, d_hist.p,
                                                      d_rowtot.p);
        k_sort_downsweep<<<T, kSortThreads, 0, st>>>(d_tiles.p, d_gd.p, d_gs.p, k[in], v[in],
                                                      8 * pass, d_hist.p, d_rowtot.p, k[in ^ 1],
                                                      v[in ^ 1], n);
        hs.check(hipGetLastError(), "launch");
        in ^= 1;
    }
    // the host arrays above and the device tables must outlive the kernels
    if (hs.ok()) hs.check(hipStreamSynchronize(st), "sort");
    if (!hs.ok()) {
        set_hip_error("the sort on the GPU failed", hs.err);
        return 1;
    }
    *out = in;
    return 0;
}

// ---------------------------------------------------------------------------
// Tree: gather, Karras radix tree, bottom-up summaries (pipeline.cl:195-323)
// ---------------------------------------------------------------------------

// node4[i] = (x, y, mass, radius), after positions were replaced.
template <typename R>
__global__ void k_pack_node4(int N, const typename Vec<R>::v2 *__restrict__ pos,
                             const R *__restrict__ mass, const R *__restrict__ size,
                             typename Vec<R>::v4 *__restrict__ node4)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const typename Vec<R>::v2 p = pos[i];
    node4[i] = mk4<R>(p.x, p.y, mass[i], size[i]);
}

// Common prefix of the sorted keys x (at local position a) and y (at b) of
// one graph, equal keys told apart by their positions (pipeline.cl:211-217).
__device__ inline int kdelta(uint32_t x, uint32_t y, int a, int b)
{
    if (x == y) return 32 + __clz((int) ((uint32_t) a ^ (uint32_t) b));
    return __clz((int) (x ^ y));
}

// Summaries are merged bottom-up in levels of units of sorted positions,
// aligned to the graph's first position. Level 1: blocks of agg_block<R>()
// positions (half a tile in single precision, a quarter in double), one
// workgroup each, in LDS. Level L >= 2: units of agg_block<R>() * 16^(L-1)
// positions. A region belongs to the smallest unit that holds its whole
// range; the regions of one unit of level >= 2 are listed by k_karras and
// merged by one workgroup in LDS (k_aggregate_unit), level after level.
//
// Capacity: a region of a unit of level L crosses one of the 15 boundaries
// between the unit's 16 sub-units. The regions whose ranges contain one
// boundary form a nested chain in which the common prefix of the (key,
// position) pairs strictly shrinks, so the chain has at most 64 regions, and
// a unit at most 15 * 64 = 960: the lists of kUnitCap entries never overflow.
constexpr int kUnitFan = 16;
constexpr int kUnitCap = 1024;

constexpr int kMaxLevels = 8; // units of 512 * 16^6 positions exceed any graph

template <typename R> constexpr int agg_block() { return sizeof(R) == 4 ? kTile / 2 : kTile / 4; }
template <typename R> constexpr int agg_shift() { return sizeof(R) == 4 ? 10 : 9; }
static_assert(agg_block<float>() == 1 << agg_shift<float>() &&
                  agg_block<double>() == 1 << agg_shift<double>() && kUnitFan == 16,
              "summary units are powers of two, 16 sub-units each");

struct Unit {
    int g, index, pad0, pad1; // graph, index of its list
};

// Region t of graph gi[t] (Karras 2012, as pipeline.cl:226-270), with global
// node references. Also records, for the gap after each sorted position, the
// node a traversal visits next (the right child of the region split there),
// and lists the region in its summary unit when that is of level >= 2:
// xslot[t] = level << 16 | position in the unit's list, 0 inside a block.
//
// The searches for a region's range and split compare its key with keys at
// growing distances, one dependent load after another. The workgroup of
// kKarrasThreads regions first stages the sorted keys of its positions and
// kKarrasWin more on each side in LDS (region t of graph g sits at global
// position t + g, so a workgroup's positions span at most 2 * kKarrasThreads
// - 1, with graphs of two nodes); only the probes of the few regions whose
// range reaches beyond the window load from global memory.
constexpr int kKarrasThreads = 256;
constexpr int kKarrasWin = 1024;
constexpr int kKarrasStage = 2 * kKarrasThreads + 2 * kKarrasWin;
static_assert(kKarrasStage % kKarrasThreads == 0, "whole rounds of staging");

template <typename R>
__global__ void __launch_bounds__(kKarrasThreads)
k_karras(int NI, int N, const int *__restrict__ gi, const GraphDesc *__restrict__ gd,
         const GraphState *__restrict__ gs, const uint32_t *__restrict__ ks,
         int2 *__restrict__ child, int2 *__restrict__ rng, int *__restrict__ parent_leaf,
         int *__restrict__ parent_int, int *__restrict__ escape, const int *__restrict__ ubase,
         int *__restrict__ ucnt, int *__restrict__ lists, int *__restrict__ xslot, int N1)
{
    __shared__ uint32_t s_key[kKarrasStage];
    const int t0 = (int) blockIdx.x * kKarrasThreads;
    const int t = t0 + (int) threadIdx.x;
    const int tl = min(t0 + kKarrasThreads - 1, NI - 1);
    // the region's graph and descriptor are loaded beside the keys
    // (a batch of one graph: graph 0, whose regions sit at their own positions)
    const int g = N1 > 0 ? 0 : gi[min(t, NI - 1)];
    const bool retired = gs[g].retired;
    const GraphDesc G = gd[g];
    const int lo = t0 + (N1 > 0 ? 0 : gi[t0]) - kKarrasWin; // s_key[q]: the key at position lo + q
    const int cnt = min(tl + (N1 > 0 ? 0 : gi[tl]) + kKarrasWin + 1 - lo, kKarrasStage);
    constexpr int kPer = kKarrasStage / kKarrasThreads;
    uint32_t kv[kPer];
#pragma unroll
    for (int r = 0; r < kPer; r++) { // every load in flight at once
        const int p = min(max(lo + (int) threadIdx.x + r * kKarrasThreads, 0), N - 1);
        kv[r] = ks[p];
    }
#pragma unroll
    for (int r = 0; r < kPer; r++) {
        const int q = (int) threadIdx.x + r * kKarrasThreads;
        const int p = lo + q;
        if (q < cnt) s_key[q] = p >= 0 && p < N ? kv[r] : 0u;
    }
    __syncthreads();
    if (t >= NI || retired) return;
    const int n = G.n, i = t - G.ioff, base = G.noff;
    auto key = [&](int b) -> uint32_t { // key at local position b of the graph
        const int q = base + b - lo;
        return (unsigned) q < (unsigned) cnt ? s_key[q] : ks[base + b];
    };
    const uint32_t ki = key(i);
    // δ(i, b), -1 out of the graph
    auto delta = [&](int b) -> int { return b < 0 || b >= n ? -1 : kdelta(ki, key(b), i, b); };
    const int d = delta(i + 1) > delta(i - 1) ? 1 : -1;
    const int dmin = delta(i - d);
    int lmax = 2;
    while (delta(i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int s = lmax / 2; s >= 1; s /= 2)
        if (delta(i + (l + s) * d) > dmin) l += s;
    const int j = i + l * d;
    const int dnode = delta(j);
    int s = 0, st = l;
    do {
        st = (st + 1) >> 1;
        if (delta(i + (s + st) * d) > dnode) s += st;
    } while (st > 1);
    const int split = i + s * d + min(d, 0);
    const int first = min(i, j), last = max(i, j);
    const int S = base + split;
    const int left = first == split ? ~S : S - g;
    const int right = last == split + 1 ? ~(S + 1) : S + 1 - g;
    child[t] = make_int2(left, right);
    if (left >= 0) parent_int[left] = t;
    else parent_leaf[~left] = t;
    if (right >= 0) parent_int[right] = t;
    else parent_leaf[~right] = t;
    if (i == 0) parent_int[t] = -1;
    escape[S - g] = right;
    rng[t] = make_int2(base + first, base + last);
    // units are powers of two: level L splits positions at bit
    // log2(block) + 4 (L - 1); positions stay below 2^26, so the shift stays
    // below 32
    int xs = 0;
    int sh = agg_shift<R>();
    if ((first >> sh) != (last >> sh)) {
        int L = 1;
        do {
            L++;
            sh += 4;
        } while ((first >> sh) != (last >> sh) && L < kMaxLevels - 1);
        const int ub = ubase[g * kMaxLevels + L];
        if (ub >= 0) { // always: the graph is longer than a unit of level L - 1
            const int unit = ub + (first >> sh);
            const int q = atomicAdd(&ucnt[unit * kCountStride], 1);
            if (q < kUnitCap) lists[(size_t) unit * kUnitCap + q] = t;
            xs = L << 16 | min(q, 0xFFFF);
        }
    }
    xslot[t] = xs;
}

template <typename R> struct Summ {
    R x0, y0, x1, y1, m, mx, my;
};

template <typename R>
__device__ inline Summ<R> leaf_summ(const typename Vec<R>::v4 p)
{
    return Summ<R>{p.x, p.y, p.x, p.y, p.z, p.x * p.z, p.y * p.z};
}

// Summaries of a region from those of its two children, merged left then
// right (pipeline.cl:276-305).
template <typename R> __device__ inline Summ<R> merge_summ(const Summ<R> &a, const Summ<R> &b)
{
    Summ<R> r;
    r.x0 = min(a.x0, b.x0);
    r.y0 = min(a.y0, b.y0);
    r.x1 = max(a.x1, b.x1);
    r.y1 = max(a.y1, b.y1);
    r.m = a.m + b.m;
    r.mx = a.mx + b.mx;
    r.my = a.my + b.my;
    return r;
}

// (cx, cy, M, size^2) of a region (pipeline.cl:312-323). The centre of mass
// is kept inside the bounding box of the region, as exact arithmetic
// guarantees: rounded, it can fall an ulp away from a cluster of coincident
// nodes, which would then see their own region as a far monopole.
template <typename R> __device__ inline typename Vec<R>::v4 region_record(const Summ<R> &r)
{
    const R w = r.x1 - r.x0, h = r.y1 - r.y0;
    const R cx = min(max(r.mx / r.m, r.x0), r.x1);
    const R cy = min(max(r.my / r.m, r.y0), r.y1);
    return mk4<R>(cx, cy, r.m, (R) 2 * sumsq_rn(w, h));
}

// Leaf to root, the second child to arrive at a region merging it. The
// reference climbs with atomic_inc and no fence (pipeline.cl:294-302); here
// every level runs in the LDS of one workgroup, and each kernel boundary
// publishes a level's summaries to the next: no agent-scope fence is needed.

// Level 1 in LDS. The workgroup of a block of sorted positions [lo, hi)
// (half a tile in single precision, a quarter in double: 60 and 52 KB of LDS,
// two workgroups per CU; whole tiles measured slower, at one per CU)
// first stages the block's leaf records and the links of the regions that
// may lie inside it (internal ids tbase .. tbase + (hi - lo) - 1); a region
// that leaves the block gets an arrival count that starts at kOutside, so
// that one LDS atomic both classifies a region and counts its arrivals. The
// climb then reads and writes summaries in LDS only. Written to HBM: the
// traversal record of every region merged here, and the summaries of those
// whose parent leaves the block (read by the next levels).
constexpr int kOutside = 1 << 24;

template <typename R>
__global__ void __launch_bounds__(1024)
k_aggregate_tile(const Tile *__restrict__ tiles, const GraphDesc *__restrict__ gd,
                 const GraphState *__restrict__ gs, const uint32_t *__restrict__ order,
                 const typename Vec<R>::v4 *__restrict__ node4,
                 typename Vec<R>::v4 *__restrict__ psort, const int2 *__restrict__ child,
                 const int2 *__restrict__ rng, const int *__restrict__ parent_leaf,
                 const int *__restrict__ parent_int, const int *__restrict__ escape, int NI,
                 typename Vec<R>::v4 *nd, TNode<R> *tn, int N1)
{
    using V4 = typename Vec<R>::v4;
    constexpr int B = agg_block<R>();
    constexpr int kParts = kTile / B;
    __shared__ V4 s_leaf[B];
    __shared__ int2 s_child[B];
    __shared__ int s_parent[B];
    __shared__ int s_vis[B];
    __shared__ int s_span[B]; // last - first of each region inside
    __shared__ R s_sum[7][B]; // x0, y0, x1, y1, m, mx, my
    const Tile T = tile_of(tiles, blockIdx.x / kParts, N1, kTile);
    const int g = T.g;
    const int lo = T.lo + (int) (blockIdx.x % kParts) * B;
    const int hi = min(T.hi, lo + B);
    if (lo >= hi || gs[g].retired) return;
    const GraphDesc D = gd[g];
    const int gend = D.noff + D.n;
    const int tbase = lo - g; // the regions that may lie inside: tbase + [0, hi - lo)
    // internal ids stop at noff + n - 1 - g - 1: the last position of a
    // graph has no region of its own index
    const int nslots = (hi == gend) ? hi - lo - 1 : hi - lo;
    const bool adjust = D.flags & GF_ADJUST;
    // One position per thread (a block has at most blockDim.x positions).
    // Every load of the thread is issued before the first wait: its leaf,
    // gathered from the node records in their original order (pipeline.cl:
    // 195-207: psort[p] = node4[order[p]], stored below for the traversal
    // and the unit levels), the escape link after it, the links of region
    // tbase + q and the leaf's parent; indices are clamped to valid entries
    // and unused values dropped.
    const int q = (int) threadIdx.x;
    const bool has_leaf = q < hi - lo, has_slot = q < nslots;
    const int pos = lo + min(q, hi - lo - 1);
    const int t = tbase + min(q, max(nslots - 1, 0));
    V4 leaf = node4[order[pos]];
    int esc_pos = escape[min(pos, gend - 2) - g];
    int pl = parent_leaf[pos];
    int2 rg = make_int2(0, 0); // the region's range of sorted positions
    int2 ch = make_int2(0, 0);
    int par = -1;
    if (nslots > 0) { // a block of the graph's last position alone has no region
        rg = rng[t];
        ch = child[t];
        par = parent_int[t];
    }
    asm volatile("" : "+v"(leaf.x), "+v"(leaf.y), "+v"(leaf.z), "+v"(leaf.w), "+v"(esc_pos),
                 "+v"(pl), "+v"(rg.x), "+v"(rg.y), "+v"(ch.x), "+v"(ch.y), "+v"(par));
    const int4 l = make_int4(record_of(ch.x, NI), 0, rg.x, rg.y);
    const bool inside = has_slot && l.z >= lo && l.w < hi;
    // the skip link of a region merged here, loaded now for the end
    const int esc_last = inside ? escape[min(l.w, gend - 2) - g] : 0;
    if (has_leaf) {
        psort[pos] = leaf;
        s_leaf[q] = leaf;
        if (has_slot) {
            s_child[q] = ch;
            s_parent[q] = par;
            s_span[q] = l.w - l.z;
        }
        s_vis[q] = inside ? 0 : kOutside;
    }
    __syncthreads();
    // the leaf's own traversal record (NI + sorted position): a region of one
    // node and size 0. Accepted as a monopole it gives exactly the pair term
    // without sizes; with sizes it is never accepted (size^2 infinite) and is
    // visited as a pair. A walk reaches a leaf's record only inside its
    // parent, and enters a region only when it has more than kLeafSize nodes
    // (a smaller one is visited node by node from psort and left by its skip
    // link): the record of a leaf whose parent is such a small region is
    // never read, and is not written (most leaves; a parent outside the
    // block counts as large).
    if (has_leaf) {
        const int pq = pl - tbase;
        if (!(pq >= 0 && pq < nslots && s_span[pq] < kLeafSize)) {
            TNode<R> rec;
            rec.c = mk4<R>(leaf.x, leaf.y, leaf.z, adjust ? (R) INFINITY : (R) 0);
            rec.l = make_int4(0, pos == gend - 1 ? kNoRecord : record_of(esc_pos, NI), pos, pos);
            tn[(size_t) NI + pos] = rec;
        }
    }
    if (has_leaf) {
        int t = pl;
        Summ<R> r = leaf_summ<R>(leaf);
        // a root-to-leaf path has at most 64 regions (32 key bits, then up to
        // 32 bits of position); the bound only guards against a corrupt tree
        for (int depth = 0; t >= 0 && depth < 128; depth++) {
            const int q = t - tbase;
            const int old = (q >= 0 && q < nslots)
                                ? __hip_atomic_fetch_add(&s_vis[q], 1, __ATOMIC_ACQ_REL,
                                                         __HIP_MEMORY_SCOPE_WORKGROUP)
                                : kOutside;
            if (old >= kOutside) break; // the region leaves the block: the next levels
            if (old == 0) break; // the other child is not summarised yet
            const int2 c = s_child[q];
            Summ<R> a, b;
            if (c.x < 0) {
                a = leaf_summ<R>(s_leaf[~c.x - lo]);
            } else {
                const int u = c.x - tbase;
                a = Summ<R>{s_sum[0][u], s_sum[1][u], s_sum[2][u], s_sum[3][u],
                            s_sum[4][u], s_sum[5][u], s_sum[6][u]};
            }
            if (c.y < 0) {
                b = leaf_summ<R>(s_leaf[~c.y - lo]);
            } else {
                const int u = c.y - tbase;
                b = Summ<R>{s_sum[0][u], s_sum[1][u], s_sum[2][u], s_sum[3][u],
                            s_sum[4][u], s_sum[5][u], s_sum[6][u]};
            }
            r = merge_summ<R>(a, b);
            s_sum[0][q] = r.x0;
            s_sum[1][q] = r.y0;
            s_sum[2][q] = r.x1;
            s_sum[3][q] = r.y1;
            s_sum[4][q] = r.m;
            s_sum[5][q] = r.mx;
            s_sum[6][q] = r.my;
            const int p = s_parent[q];
            if (p >= 0) {
                const int pq = p - tbase;
                if (pq < 0 || pq >= nslots || s_vis[pq] >= kOutside) {
                    // the parent leaves the block: its merge reads this summary
                    nd[2 * t] = mk4<R>(r.x0, r.y0, r.x1, r.y1);
                    nd[2 * t + 1] = mk4<R>(r.m, r.mx, r.my, (R) 0);
                }
            }
            t = p;
        }
    }
    // the traversal records of the regions merged here, whole, one per
    // thread in region order (coalesced 32-byte writes): the summary from
    // LDS, the links of k_karras and the skip link
    __syncthreads();
    if (inside) {
        const Summ<R> r{s_sum[0][q], s_sum[1][q], s_sum[2][q], s_sum[3][q],
                        s_sum[4][q], s_sum[5][q], s_sum[6][q]};
        TNode<R> rec;
        rec.c = region_record<R>(r);
        rec.l = make_int4(l.x, l.w == gend - 1 ? kNoRecord : record_of(esc_last, NI), l.z, l.w);
        tn[t] = rec;
    }
}

// One unit of level `level` >= 2: its listed regions are staged in LDS, one
// per thread, with their children's places (in the list, or summaries of
// blocks, lower units and leaves loaded from HBM up front); a region with
// both children from outside is merged at once and starts a climb, and the
// climb then runs through LDS arrival counters as in level 1. Written to
// HBM: every listed region's traversal record, and the summaries of those
// whose parent lies in a higher unit.
template <typename R>
__global__ void __launch_bounds__(1024)
k_aggregate_unit(const Unit *__restrict__ units, int level, const GraphDesc *__restrict__ gd,
                 GraphState *__restrict__ gs, const typename Vec<R>::v4 *__restrict__ psort,
                 const int2 *__restrict__ child, const int2 *__restrict__ rng,
                 const int *__restrict__ parent_int,
                 const int *__restrict__ escape, const int *__restrict__ xslot,
                 const int *__restrict__ lists, const int *__restrict__ ucnt, int NI,
                 typename Vec<R>::v4 *nd, TNode<R> *tn)
{
    using V4 = typename Vec<R>::v4;
    __shared__ int s_t[kUnitCap];
    __shared__ int2 s_cs[kUnitCap]; // children's places in the list, -1 outside
    __shared__ int s_ps[kUnitCap];  // parent's place, -1 in a higher unit, -2 none
    __shared__ int s_vis[kUnitCap];
    __shared__ int s_skip[kUnitCap];
    __shared__ R s_sum[7][kUnitCap]; // x0, y0, x1, y1, m, mx, my
    const Unit U = units[blockIdx.x];
    if (gs[U.g].retired) return;
    const GraphDesc D = gd[U.g];
    const int gend = D.noff + D.n;
    const int cnt = ucnt[U.index * kCountStride];
    if (cnt > kUnitCap) { // cannot happen (see kUnitCap): stop the graph, never run past LDS
        if (threadIdx.x == 0) gs[U.g].status = 2;
        return;
    }
    auto put = [&](int q, const Summ<R> &r) {
        s_sum[0][q] = r.x0;
        s_sum[1][q] = r.y0;
        s_sum[2][q] = r.x1;
        s_sum[3][q] = r.y1;
        s_sum[4][q] = r.m;
        s_sum[5][q] = r.mx;
        s_sum[6][q] = r.my;
    };
    auto get = [&](int q) {
        return Summ<R>{s_sum[0][q], s_sum[1][q], s_sum[2][q], s_sum[3][q],
                       s_sum[4][q], s_sum[5][q], s_sum[6][q]};
    };
    const int q = threadIdx.x;
    bool start = false;
    if (q < cnt) {
        const int t = lists[(size_t) U.index * kUnitCap + q];
        const int2 c = child[t];
        const int p = parent_int[t];
        const int last = rng[t].y;
        const int xl = c.x >= 0 ? xslot[c.x] : 0, xr = c.y >= 0 ? xslot[c.y] : 0;
        const int xp = p >= 0 ? xslot[p] : 0;
        const int esc = last == gend - 1 ? 0 : escape[last - U.g];
        // the summaries a child from outside the list brings (loaded for
        // both children; unused when the child is listed)
        V4 au, av, bu, bv;
        if (c.x < 0) au = psort[~c.x];
        else {
            au = nd[2 * c.x];
            av = nd[2 * c.x + 1];
        }
        if (c.y < 0) bu = psort[~c.y];
        else {
            bu = nd[2 * c.y];
            bv = nd[2 * c.y + 1];
        }
        const bool inl = c.x >= 0 && (xl >> 16) == level, inr = c.y >= 0 && (xr >> 16) == level;
        const Summ<R> a = c.x < 0 ? leaf_summ<R>(au) : Summ<R>{au.x, au.y, au.z, au.w, av.x, av.y, av.z};
        const Summ<R> b = c.y < 0 ? leaf_summ<R>(bu) : Summ<R>{bu.x, bu.y, bu.z, bu.w, bv.x, bv.y, bv.z};
        s_t[q] = t;
        s_cs[q] = make_int2(inl ? (xl & 0xFFFF) : -1, inr ? (xr & 0xFFFF) : -1);
        s_ps[q] = p < 0 ? -2 : ((xp >> 16) == level ? (xp & 0xFFFF) : -1);
        s_skip[q] = last == gend - 1 ? kNoRecord : record_of(esc, NI);
        // a child from outside is complete: counted as arrived; with one, its
        // summary waits in the region's own slot until the other arrives
        s_vis[q] = (inl ? 0 : 1) + (inr ? 0 : 1);
        if (!inl && !inr) {
            put(q, merge_summ<R>(a, b));
            start = true;
        } else if (!inl) {
            put(q, a);
        } else if (!inr) {
            put(q, b);
        }
    }
    __syncthreads();
    if (start) {
        int r = q;
        for (int depth = 0; depth < 128; depth++) { // a chain has at most 64 regions
            const int pr = s_ps[r];
            if (pr < 0) break;
            const int old = __hip_atomic_fetch_add(&s_vis[pr], 1, __ATOMIC_ACQ_REL,
                                                   __HIP_MEMORY_SCOPE_WORKGROUP);
            if (old != 1) break; // the other child is not summarised yet
            const int2 cs = s_cs[pr];
            const Summ<R> a = get(cs.x >= 0 ? cs.x : pr), b = get(cs.y >= 0 ? cs.y : pr);
            put(pr, merge_summ<R>(a, b));
            r = pr;
        }
    }
    __syncthreads();
    if (q < cnt) {
        const int t = s_t[q];
        const Summ<R> r = get(q);
        // the whole record in one 32-byte write
        TNode<R> rec;
        const int2 rg = rng[t];
        rec.l = make_int4(record_of(child[t].x, NI), s_skip[q], rg.x, rg.y);
        rec.c = region_record<R>(r);
        tn[t] = rec;
        if (s_ps[q] == -1) { // merged by a higher unit
            nd[2 * t] = mk4<R>(r.x0, r.y0, r.x1, r.y1);
            nd[2 * t + 1] = mk4<R>(r.m, r.mx, r.my, (R) 0);
        }
    }
}

// Workgroup b of a launch runs on XCD b mod 8, each XCD with its own L2.
// With run > 0, workgroup b takes the block that gives each XCD runs of
// `run` consecutive blocks (a bijection of [0, nb)): neighbouring nodes walk
// mostly the same tree records, which one L2 then holds for the whole run.
// Which hardware runs a block changes nothing in what the block computes.
__device__ inline int xcd_block(int b, int nb, int run)
{
    constexpr int kXcds = 8;
    if (run <= 0) return b;
    const int span = kXcds * run, full = nb / span * span;
    if (b >= full) return b;
    const int within = b % span;
    return b - within + (within % kXcds) * run + within / kXcds;
}

// ---------------------------------------------------------------------------
// Repulsion: Barnes-Hut over the Karras tree (pipeline.cl:332-383), walked
// by a whole wave at once.
//
// The 64 lanes of a wave hold 64 consecutive nodes in Morton order, which
// are neighbours in space and visit mostly the same regions. The wave walks
// the union of their visits in depth-first order (left child first, along
// the skip links), one region per step: the region is the same for every
// lane, so its record is a scalar load and every branch is uniform. A lane
// takes part in a step only when its own traversal would visit that region:
// it keeps `done`, the last sorted position covered by the last region it
// accepted (or visited node by node), and a region is skipped by the lane
// when it starts at or before `done`. The wave opens a region when at least
// one taking-part lane does not accept it. Every lane therefore sees
// exactly the regions, in exactly the order, of the reference traversal,
// and accumulates the same terms in the same order.
// ---------------------------------------------------------------------------

// Per-lane sum of the attraction's terms c_e * (x_j - x_i) (times the
// linlog factor): the weight w and the offsets are computed in the working
// precision, and each term xd * w is formed exactly in FP64 (a product of two
// FP32 values fits in 53 bits) by a fused multiply-add into the FP64 sum. A
// lane's sum depends on its own terms only, not on the other lanes of its
// wave: a graph gives the same result alone and inside any batch.
template <typename R> struct ForceSum {
    double x = 0.0, y = 0.0;
    __device__ inline void add(R xd, R yd, R w)
    {
        const double wd = (double) w;
        x = __builtin_fma((double) xd, wd, x);
        y = __builtin_fma((double) yd, wd, y);
    }
};


// With node sizes (prevent_overlap) a node pair's term is computed in FP64
// from the FP32 inputs: the force jumps by orders of magnitude as two discs
// touch, and the terms of overlapping discs are large and cancel. The term
// goes straight into an FP64 total, without the factor scaling * m_i.
template <typename R>
__device__ inline void lane_pair_adjust(typename Vec<R>::v4 me, typename Vec<R>::v4 o,
                                        double &sx, double &sy, bool on)
{
    const double dx = (double) (me.x - o.x), dy = (double) (me.y - o.y);
    const double d = __builtin_sqrt(dx * dx + dy * dy) - ((double) me.w + (double) o.w);
    const double mo = (double) o.z;
    double f = d > 0.0 ? mo * frcp_r(d * d) : (d < 0.0 ? 100.0 * mo : 0.0);
    f = on ? f : 0.0;
    sx += dx * f;
    sy += dy * f;
}

// Per-lane force of the Barnes-Hut traversal, without the common factor
// scaling * m_i. In single precision a graph adds each term xd * w by an FP32
// fused multiply-add: the bulk force arithmetic is FP32, as the precision
// model asks, and FP64 is kept for the global sums, the speed and the
// position accumulators. A lane sums a few hundred terms at theta 1.2: over
// the 2^21-node R-MAT workload the forces stay within 8e-7 of the FP64
// oracle normwise in every degree class, against a tolerance of 1e-5. Two
// kinds of graph keep an FP64 total, each term formed exactly in FP64 (a
// product of two FP32 values fits in 53 bits): a graph with node sizes,
// whose overlapping pairs give large terms that cancel, and a graph at
// theta = 0, whose lanes sum every other node (FP32 sums of 2^14 terms
// measured 5e-6). The kernels carry the FP64 totals only when some graph of
// the batch needs them (FP64); a graph that does not keeps +0 there, so it
// gets the same bits from either kernel. A lane's sum depends on its own
// terms, in the traversal's order, and not on the other lanes of its wave,
// so a graph gives the same result alone and inside any batch. In double
// precision every total is FP64.
template <typename R, bool FP64> struct LaneForce {
    static constexpr bool kWide = FP64 && sizeof(R) == 4;
    R x = (R) 0, y = (R) 0;
    double wx = 0.0, wy = 0.0; // the FP64 total of a graph that needs one
    __device__ inline void term(R xd, R yd, R w, bool wide)
    {
        if (kWide && wide) {
            const double wd = (double) w; // exact products of FP32 values
            wx = __builtin_fma((double) xd, wd, wx);
            wy = __builtin_fma((double) yd, wd, wy);
        } else {
            x = fma_r(xd, w, x);
            y = fma_r(yd, w, y);
        }
    }
    // a node pair of a graph with sizes (which always has an FP64 total)
    __device__ inline void pair_sized(typename Vec<R>::v4 me, typename Vec<R>::v4 o, bool on)
    {
        if (kWide) {
            lane_pair_adjust<R>(me, o, wx, wy, on);
        } else {
            double sx = (double) x, sy = (double) y; // R is double here
            lane_pair_adjust<R>(me, o, sx, sy, on);
            x = (R) sx;
            y = (R) sy;
        }
    }
    __device__ inline double total_x() const { return kWide ? (double) x + wx : (double) x; }
    __device__ inline double total_y() const { return kWide ? (double) y + wy : (double) y; }
};

// Whether graph G's traversal sums take the FP64 total (see LaneForce).
__device__ inline bool fp64_sums(const GraphDesc &G)
{
    return (G.flags & GF_ADJUST) || G.theta2 == 0.0;
}

// With COUNT, the kernel also counts its work for the roofline: regions
// accepted (A) and opened (O), and node pairs (P), into cnt[0..2].
// psort holds kLeafSize records of padding after the last node, so that the
// nodes of a small region are loaded at once.
template <typename R, bool FP64_SUMS, bool COUNT = false, bool WIDE = false>
__global__ void __launch_bounds__(1024)
k_repulse(int N, int NI, const int *__restrict__ gid, const GraphDesc *__restrict__ gd,
          const GraphState *__restrict__ gs, const typename Vec<R>::v4 *__restrict__ psort,
          const TNode<R> *__restrict__ tn, const uint32_t *__restrict__ order,
          double2 *__restrict__ F, int xrun, unsigned long long *cnt = nullptr)
{
    using V4 = typename Vec<R>::v4;
    unsigned long long nA = 0, nO = 0, nP = 0, nS = 0;
    const int k0 = xcd_block((int) blockIdx.x, (int) gridDim.x, xrun) * blockDim.x + threadIdx.x;
    const int k = k0 < N ? k0 : N - 1;
    const bool valid = k0 < N && !gs[gid[k]].retired;
    const int my_g = valid ? gid[k] : -1;
    const V4 me = psort[k];
    LaneForce<R, FP64_SUMS> sum;
    // the lanes of one wave may belong to several graphs of a batch: walk
    // each graph's tree in turn with its own lanes
    uint64_t todo = __ballot(valid);
    while (todo) {
        const int leader = __ffsll((unsigned long long) todo) - 1;
        const int g = __builtin_amdgcn_readlane(my_g, leader);
        const bool mine = my_g == g;
        todo &= ~__ballot(mine);
        const GraphDesc G = gd[g];
        const bool adjust = FP64_SUMS && (G.flags & GF_ADJUST);
        const bool wide = FP64_SUMS && fp64_sums(G);
        const R theta2 = (R) G.theta2;
        int done = mine ? INT_MIN + 1 : INT_MAX; // lanes of other graphs never take part
        int c = G.ioff;                          // the root (every graph has n >= 2)
        // every node of the tree is visited at most once (2n - 1 of them); the
        // bound only guards against a corrupt tree
        // One exit, at the loop's head: a walk ends at kNoRecord. It always
        // ends: a descent goes to a left child, whose range lies strictly
        // inside its parent's, and a skip link to the first node after the
        // current range, so the walk is a depth-first pass that meets every
        // node of the tree at most once. (k_karras builds a valid binary
        // radix tree from any array of keys, whatever the positions.)
        while (c >= 0) {
            c = __builtin_amdgcn_readfirstlane(c);
            if (COUNT) nS++;
            // c is the record of a region or a leaf; lane masks of the step
            // are kept in scalar registers: the comparisons write them
            // directly and the lanes are selected from them without per-lane
            // selects. The record is one scalar load, at a 32-bit byte offset
            // while the records take less than 4 GiB.
            const TNode<R> node =
                WIDE ? tn[c]
                     : *(const TNode<R> *) ((const char *) tn +
                                            (uint32_t) c * (uint32_t) sizeof(TNode<R>));
            const V4 rc = node.c;
            const int4 rl = node.l;
            // keep the record together in SGPRs: one load, one wait (else the
            // compiler loads the mass apart, inside the branch that uses it)
            asm volatile("" ::"s"(rc.x), "s"(rc.y), "s"(rc.z), "s"(rc.w), "s"(rl.x), "s"(rl.y),
                         "s"(rl.z), "s"(rl.w));
            const R xd = me.x - rc.x, yd = me.y - rc.y;
            const R d2 = sumsq_rn(xd, yd);
            const uint64_t am = __builtin_amdgcn_ballot_w64(rl.z > done);
            const uint64_t fm = __builtin_amdgcn_ballot_w64(d2 * theta2 > rc.w);
            if (__builtin_amdgcn_inverse_ballot_w64(am & fm)) {
                // the monopole, for the lanes that accept: a branch on the
                // lane mask (the lanes' sums are updated in place)
                asm volatile("");
                sum.term(xd, yd, rc.z * frcp_r(d2), wide);
                done = rl.w;
                if (COUNT) nA++;
            }
            const uint64_t tm = am & ~fm; // lanes that open the region
            if (tm == 0) { // every lane is done with this region
                c = rl.y;
                continue;
            }
            if (COUNT) nO += __builtin_amdgcn_inverse_ballot_w64(tm) && rl.z != rl.w;
            if (rl.w - rl.z >= kLeafSize) { // descend
                c = rl.x;
                continue;
            }
            // a small region or a leaf: its nodes one by one, loaded together
            // (psort is padded)
            const int j0 = rl.z, j1 = rl.w;
            V4 o[kLeafSize];
#pragma unroll
            for (int u = 0; u < kLeafSize; u++) o[u] = psort[j0 + u];
            if (sizeof(R) == 4) // issued together, one wait for all of them
                asm volatile("" ::"s"(o[0].x), "s"(o[0].y), "s"(o[0].z), "s"(o[1].x), "s"(o[1].y),
                             "s"(o[1].z), "s"(o[2].x), "s"(o[2].y), "s"(o[2].z), "s"(o[3].x),
                             "s"(o[3].y), "s"(o[3].z), "s"(o[4].x), "s"(o[4].y), "s"(o[4].z),
                             "s"(o[5].x), "s"(o[5].y), "s"(o[5].z), "s"(o[6].x), "s"(o[6].y),
                             "s"(o[6].z), "s"(o[7].x), "s"(o[7].y), "s"(o[7].z));
            if (__builtin_amdgcn_inverse_ballot_w64(tm)) {
                // a lane meets itself with xd = yd = 0, which adds nothing:
                // d2 = 0 gives a zero weight, and with sizes an overlap
                // weight times a zero offset
#pragma unroll
                for (int u = 0; u < kLeafSize; u++) {
                    if (j0 + u > j1) break;
                    if (COUNT) nP += (j0 + u) != k;
                    if (adjust) {
                        sum.pair_sized(me, o[u], true);
                    } else {
                        const R pxd = me.x - o[u].x, pyd = me.y - o[u].y;
                        const R pd2 = pxd * pxd + pyd * pyd;
                        R w = pd2 > (R) 0 ? o[u].z * frcp_r(pd2) : (R) 0;
                        // the weight selected once, in a VGPR, before its
                        // two fused multiply-adds
                        if (sizeof(R) == 4) asm volatile("" : "+v"(w));
                        sum.term(pxd, pyd, w, wide);
                    }
                }
                done = j1;
            }
            c = rl.y;
        }
    }
    if (valid) {
        // gravity (pipeline.cl:404-411): the force leaves this kernel as
        // repulsion minus gravity, the attraction kernels add the edges
        const GraphDesc G = gd[my_g];
        const double sscale = G.scaling * (double) me.z; // scaling * m_i in FP64
        double fx = sscale * sum.total_x(), fy = sscale * sum.total_y();
        const R grav = (R) G.gravity;
        if (grav != (R) 0) {
            const R dist = fsqrt_r(me.x * me.x + me.y * me.y);
            if (dist > (R) 0) {
                const R gg = (G.flags & GF_STRONG) ? grav * me.z : grav * me.z / dist;
                fx -= (double) (me.x * gg);
                fy -= (double) (me.y * gg);
            }
        }
        F[order[k]] = make_double2(fx, fy);
    }
    if (COUNT) {
        atomicAdd(&cnt[0], nA);
        atomicAdd(&cnt[1], nO);
        atomicAdd(&cnt[2], nP);
        if (lane_id() == 0) atomicAdd(&cnt[3], nS);
    }
}

// One step of a wave's walk (the body of k_repulse's walk, without the
// counters) over the record `node`; returns the next record, -1 at the end.
template <typename R, bool FP64_SUMS>
__device__ __forceinline__ int walk_step(const TNode<R> &node, typename Vec<R>::v4 me,
                                         R theta2, bool adjust, bool wide, int &done,
                                         LaneForce<R, FP64_SUMS> &sum,
                                         const typename Vec<R>::v4 *__restrict__ psort)
{
    using V4 = typename Vec<R>::v4;
    const V4 rc = node.c;
    const int4 rl = node.l;
    asm volatile("" ::"s"(rc.x), "s"(rc.y), "s"(rc.z), "s"(rc.w), "s"(rl.x), "s"(rl.y),
                 "s"(rl.z), "s"(rl.w));
    const R xd = me.x - rc.x, yd = me.y - rc.y;
    const R d2 = sumsq_rn(xd, yd);
    const uint64_t am = __builtin_amdgcn_ballot_w64(rl.z > done);
    const uint64_t fm = __builtin_amdgcn_ballot_w64(d2 * theta2 > rc.w);
    if (__builtin_amdgcn_inverse_ballot_w64(am & fm)) {
        asm volatile("");
        sum.term(xd, yd, rc.z * frcp_r(d2), wide);
        done = rl.w;
    }
    const uint64_t tm = am & ~fm;
    if (tm == 0) return rl.y;
    if (rl.w - rl.z >= kLeafSize) return rl.x;
    const int j0 = rl.z, j1 = rl.w;
    V4 o[kLeafSize];
#pragma unroll
    for (int u = 0; u < kLeafSize; u++) o[u] = psort[j0 + u];
    if (sizeof(R) == 4)
        asm volatile("" ::"s"(o[0].x), "s"(o[0].y), "s"(o[0].z), "s"(o[1].x), "s"(o[1].y),
                     "s"(o[1].z), "s"(o[2].x), "s"(o[2].y), "s"(o[2].z), "s"(o[3].x),
                     "s"(o[3].y), "s"(o[3].z), "s"(o[4].x), "s"(o[4].y), "s"(o[4].z),
                     "s"(o[5].x), "s"(o[5].y), "s"(o[5].z), "s"(o[6].x), "s"(o[6].y),
                     "s"(o[6].z), "s"(o[7].x), "s"(o[7].y), "s"(o[7].z));
    if (__builtin_amdgcn_inverse_ballot_w64(tm)) {
#pragma unroll
        for (int u = 0; u < kLeafSize; u++) {
            if (j0 + u > j1) break;
            if (FP64_SUMS && adjust) {
                sum.pair_sized(me, o[u], true);
            } else {
                const R pxd = me.x - o[u].x, pyd = me.y - o[u].y;
                const R pd2 = pxd * pxd + pyd * pyd;
                R w = pd2 > (R) 0 ? o[u].z * frcp_r(pd2) : (R) 0;
                if (sizeof(R) == 4) asm volatile("" : "+v"(w));
                sum.term(pxd, pyd, w, wide);
            }
        }
        done = j1;
    }
    return rl.y;
}

// The force that leaves the traversal: repulsion minus gravity
// (pipeline.cl:404-411), written in the original order.
template <typename R, bool FP64>
__device__ __forceinline__ void repulse_out(const GraphDesc &G, typename Vec<R>::v4 me,
                                            const LaneForce<R, FP64> &sum, uint32_t dst,
                                            double2 *__restrict__ F)
{
    const double sscale = G.scaling * (double) me.z;
    double fx = sscale * sum.total_x(), fy = sscale * sum.total_y();
    const R grav = (R) G.gravity;
    if (grav != (R) 0) {
        const R dist = fsqrt_r(me.x * me.x + me.y * me.y);
        if (dist > (R) 0) {
            const R gg = (G.flags & GF_STRONG) ? grav * me.z : grav * me.z / dist;
            fx -= (double) (me.x * gg);
            fy -= (double) (me.y * gg);
        }
    }
    st_stream(F + dst, fx, fy);
}

// The traversal with two walks per wave: lanes hold node k of the first 64
// sorted positions of the wave's 128 and node k + 64 of the second, and the
// wave walks both unions in one loop. Each iteration issues the records of
// both walks before either step, so one load latency serves two steps and
// either walk's arithmetic covers the other's wait. Each lane's terms, and
// their order, are those of k_repulse. A wave whose 128 nodes do not all
// belong to one graph walks its two halves one after the other, graph by
// graph, as k_repulse does. (Three or four walks per wave, with every record
// load issued before any step, measured slower at 2^21 nodes: 0.76 and
// 0.83 ms against 0.68 ms for two walks in that form.)
template <typename R, bool FP64_SUMS, bool WIDE>
__global__ void __launch_bounds__(1024)
k_repulse2(int N, int NI, const int *__restrict__ gid, const GraphDesc *__restrict__ gd,
           const GraphState *__restrict__ gs, const typename Vec<R>::v4 *__restrict__ psort,
           const TNode<R> *__restrict__ tn, const uint32_t *__restrict__ order,
           double2 *__restrict__ F, int xrun)
{
    using V4 = typename Vec<R>::v4;
    auto record = [&](int r) -> TNode<R> {
        return WIDE ? tn[r]
                    : *(const TNode<R> *) ((const char *) tn +
                                           (uint32_t) r * (uint32_t) sizeof(TNode<R>));
    };
    const int lane = lane_id();
    const int wv = xcd_block((int) blockIdx.x, (int) gridDim.x, xrun) * (int) (blockDim.x >> 6) +
                   (int) (threadIdx.x >> 6);
    const int kk0[2] = {wv * 128 + lane, wv * 128 + 64 + lane};
    int kk[2], gg[2];
    bool vv[2];
    V4 me[2];
    LaneForce<R, FP64_SUMS> sum[2];
#pragma unroll
    for (int h = 0; h < 2; h++) {
        kk[h] = kk0[h] < N ? kk0[h] : N - 1;
        vv[h] = kk0[h] < N && !gs[gid[kk[h]]].retired;
        gg[h] = vv[h] ? gid[kk[h]] : -1;
        me[h] = psort[kk[h]];
    }
    const uint64_t any0 = __ballot(vv[0]), any1 = __ballot(vv[1]);
    if ((any0 | any1) == 0) return;
    const int lead = __ffsll((unsigned long long) (any0 ? any0 : any1)) - 1;
    const int g = __builtin_amdgcn_readlane(any0 ? gg[0] : gg[1], lead);
    const bool one = __ballot((vv[0] && gg[0] != g) || (vv[1] && gg[1] != g)) == 0;
    if (one && any0 && any1) {
        const GraphDesc G = gd[g];
        const bool adjust = FP64_SUMS && (G.flags & GF_ADJUST);
        const bool wide = FP64_SUMS && fp64_sums(G);
        const R theta2 = (R) G.theta2;
        int done0 = vv[0] ? INT_MIN + 1 : INT_MAX, done1 = vv[1] ? INT_MIN + 1 : INT_MAX;
        int c0 = G.ioff, c1 = G.ioff;
        // both walks, then the one that goes on alone (each is a
        // depth-first pass that meets every node of the tree at most once,
        // and ends at -1)
        while (c0 >= 0 && c1 >= 0) {
            const TNode<R> n0 = record(c0), n1 = record(c1);
            c0 = __builtin_amdgcn_readfirstlane(
                walk_step<R, FP64_SUMS>(n0, me[0], theta2, adjust, wide, done0, sum[0], psort));
            c1 = __builtin_amdgcn_readfirstlane(
                walk_step<R, FP64_SUMS>(n1, me[1], theta2, adjust, wide, done1, sum[1], psort));
        }
        while (c0 >= 0)
            c0 = __builtin_amdgcn_readfirstlane(walk_step<R, FP64_SUMS>(
                record(c0), me[0], theta2, adjust, wide, done0, sum[0], psort));
        while (c1 >= 0)
            c1 = __builtin_amdgcn_readfirstlane(walk_step<R, FP64_SUMS>(
                record(c1), me[1], theta2, adjust, wide, done1, sum[1], psort));
    } else {
        // the halves one after the other, each graph by graph
#pragma unroll
        for (int h = 0; h < 2; h++) {
            uint64_t todo = __ballot(vv[h]);
            while (todo) {
                const int leader = __ffsll((unsigned long long) todo) - 1;
                const int gh = __builtin_amdgcn_readlane(gg[h], leader);
                const bool mine = gg[h] == gh;
                todo &= ~__ballot(mine);
                const GraphDesc G = gd[gh];
                const bool adjust = FP64_SUMS && (G.flags & GF_ADJUST);
                const bool wide = FP64_SUMS && fp64_sums(G);
                const R theta2 = (R) G.theta2;
                int done = mine ? INT_MIN + 1 : INT_MAX;
                int c = G.ioff;
                while (c >= 0)
                    c = __builtin_amdgcn_readfirstlane(walk_step<R, FP64_SUMS>(
                        record(c), me[h], theta2, adjust, wide, done, sum[h], psort));
            }
        }
    }
#pragma unroll
    for (int h = 0; h < 2; h++)
        if (vv[h]) repulse_out<R, FP64_SUMS>(gd[gg[h]], me[h], sum[h], order[kk[h]], F);
}

// ---------------------------------------------------------------------------
// Attraction along the edges (pipeline.cl:392-429), added to the force that
// the repulsion kernel left (repulsion minus gravity). Balanced by degree:
//  * nodes of 1..16 incidences: one lane each;
//  * 17..256: 16 lanes each, combined with a fixed butterfly;
//  * more: work items of at most kChunk consecutive incidences, one wave
//    each; a node cut into several items gets the sum of their partial sums
//    in item order (k_attract_fix). Nodes without edges are not visited.
// A lane adds its incidences e0, e0 + step, ... in that order whatever the
// unroll U (the number of incidences whose loads are in flight at once), so
// U changes the speed and never the result.
// ---------------------------------------------------------------------------

constexpr int kChunk = 1024;

// log1p(d) / d of the linlog attraction. In single precision it is
// computed as log(u) / (u - 1) with u = 1 + d: for u != 1 this equals
// log1p(d) / d up to the roundings of log(u), of u - 1 and of the
// reciprocal (Kahan's log1p identity), with the hardware's v_log_f32 (log2)
// and v_rcp_f32; for u = 1 the ratio is 1. In double precision it is the
// library's log1p and an IEEE division.
__device__ inline float linlog_factor(float d)
{
    const float u = 1.0f + d;
    const float r = (__builtin_amdgcn_logf(u) * 0.693147180559945309f) *
                    __builtin_amdgcn_rcpf(u - 1.0f);
    return u == 1.0f ? 1.0f : r;
}
__device__ inline double linlog_factor(double d) { return log1p(d) / d; }

// The attraction of one incidence (pipeline.cl:412-427): weight
// w = -c, times log1p(d) / d with linlog; with sizes the distance is the gap
// between the discs, computed in FP64 as for the repulsion, and the edge only
// pulls while it is positive. The term w * (xd, yd) goes into the FP64 sum
// when `on`.
template <typename R>
__device__ inline void attract_term(typename Vec<R>::v2 pi, typename Vec<R>::v2 pj, R c, R si,
                                    R sj, bool linlog, bool adjust, bool on, ForceSum<R> &sum)
{
    const R xd = pi.x - pj.x, yd = pi.y - pj.y;
    R w = -c;
    if (adjust) {
        const double dx = (double) xd, dy = (double) yd;
        const double gap = __builtin_sqrt(dx * dx + dy * dy) - ((double) si + (double) sj);
        on = on && gap > 0.0;
        if (linlog) w *= linlog_factor((R) gap);
    } else if (linlog) {
        const R dist = fsqrt_r(xd * xd + yd * yd);
        on = on && dist > (R) 0;
        w *= linlog_factor(dist);
    }
    if (on) sum.add(xd, yd, w);
}

// The incidences e0, e0 + step, ... < e1 of one node, in that order. Each
// round first loads the list entries of U incidences, then their U
// positions, then adds their terms; the index of a slot past e1 is clamped
// to `last` (a valid entry: the list's last, or in the sliced layout the
// lane's own last), so every load is unconditional. With UNIFORM every
// coefficient of the node equals c0 and none is loaded.
template <typename R, int U, bool UNIFORM>
__device__ inline void attract_terms(int e0, int e1, int step, int last, bool linlog, bool adjust,
                                     typename Vec<R>::v2 pi, R si,
                                     const typename Vec<R>::v2 *__restrict__ pos,
                                     const R *__restrict__ size, const int *__restrict__ inc_node,
                                     const R *__restrict__ inc_coef, R c0, ForceSum<R> &sum)
{
    for (int e = e0; e < e1; e += U * step) {
        int j[U];
        R c[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int q = min(e + u * step, last);
            j[u] = ld_stream(inc_node + q);
            c[u] = UNIFORM ? c0 : ld_stream(inc_coef + q);
        }
        typename Vec<R>::v2 pj[U];
        R sj[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            pj[u] = pos[j[u]];
            sj[u] = adjust ? size[j[u]] : (R) 0;
        }
#pragma unroll
        for (int u = 0; u < U; u++)
            attract_term<R>(pi, pj[u], c[u], si, sj[u], linlog, adjust, e + u * step < e1, sum);
    }
}

template <int GL> __device__ inline void group_sum(double &x, double &y)
{
#pragma unroll
    for (int o = GL / 2; o >= 1; o >>= 1) {
        x += __shfl_xor(x, o, 64);
        y += __shfl_xor(y, o, 64);
    }
}

// The attraction's workgroups go to the XCDs in runs of kAttractXcdRun
// consecutive blocks (xcd_block): the nodes of a run are neighbours in the
// work lists, mostly of the same graph in a batch, so that one L2 holds the
// positions their incidences gather (512 SBM graphs of 200-20000 nodes: the
// one-lane class 0.162 -> 0.152 ms; no change on one R-MAT graph of 2^21
// nodes, whose neighbours are random). The mapping is a bijection: no result
// changes.
constexpr int kAttractXcdRun = 32;

// While the attraction runs beside another phase (forked onto the second
// stream, beside the traversal by default), each of its workgroups reserves
// 54 KB of LDS that it does not use: a CU (160 KB) then holds at most two of
// them, eight of its 32 wave slots, and the traversal keeps the rest. The
// attraction alone gets slower (0.274 -> 0.309 ms on the 2^21-node R-MAT
// graph), but it still ends well before the traversal, which it now slows
// less: the replayed iteration went from 1.073 to 1.061 ms there, and from
// 0.969 to 0.961 ms on the batch of 512 block-model graphs (two interleaved
// pairs of runs each; the tuner times the attraction's unrolls with the
// reservation too). Marking the position gathers non-temporal instead, so
// that they would not evict the traversal's records from L2, made the
// iteration 1.17 ms: the gathers reuse the lines of the hubs' positions.
constexpr unsigned kAttractBesideLds = 54 * 1024;

// GL lanes per node; the work list holds (node, first incidence, end, graph).
// The one-lane class (GL = 1) reads its lists in the sliced layout: the 64
// lanes of a wave have their r-th entries side by side (entry r of lane l at
// slice + 64 r + l), so that each load of list entries is one contiguous
// 256-byte access; its work entries hold (node, slice + l, slice + l + 64 deg,
// graph), and the lane's entries are those of its incidence list, in order.
template <typename R, int GL, int U>
__global__ void __launch_bounds__(256)
k_attract(int count, const int4 *__restrict__ work, const GraphDesc *__restrict__ gd,
          const GraphState *__restrict__ gs, const typename Vec<R>::v2 *__restrict__ pos,
          const R *__restrict__ size, const int *__restrict__ inc_node,
          const R *__restrict__ inc_coef, double2 *__restrict__ Fa)
{
    const int blk = xcd_block((int) blockIdx.x, (int) gridDim.x, kAttractXcdRun);
    const int slot = blk * (256 / GL) + (int) threadIdx.x / GL;
    const int sub = (int) threadIdx.x % GL;
    const int4 wk = work[min(slot, count - 1)];
    const int i = wk.x;
    const int flags = gd[wk.w].flags;
    const bool live = slot < count && !gs[wk.w].retired;
    constexpr int step = GL == 1 ? 64 : GL;
    const int last = GL == 1 ? wk.z - 64 : wk.z - 1;
    ForceSum<R> sum;
    if (flags & GF_UNIFORM)
        attract_terms<R, U, true>(wk.y + sub, wk.z, step, last, flags & GF_LINLOG,
                                  flags & GF_ADJUST, pos[i], size[i], pos, size, inc_node,
                                  inc_coef, inc_coef[wk.y], sum);
    else
        attract_terms<R, U, false>(wk.y + sub, wk.z, step, last, flags & GF_LINLOG,
                                   flags & GF_ADJUST, pos[i], size[i], pos, size, inc_node,
                                   inc_coef, (R) 0, sum);
    group_sum<GL>(sum.x, sum.y);
    if (!live || sub != 0) return;
    Fa[i] = make_double2(sum.x, sum.y);
}

// The one-lane class's lists copied into the sliced layout (see k_attract):
// lane p's entries go to slice + 64 r + lane, r < deg; the rest of a slice
// is never read (a lane clamps its loads to its own last entry).
template <typename R>
__global__ void __launch_bounds__(256)
k_sell_fill(int count, const int4 *__restrict__ work, const int *__restrict__ inc_off,
            const int *__restrict__ inc_node, const R *__restrict__ inc_coef,
            int *__restrict__ sell_node, R *__restrict__ sell_coef)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= count) return;
    const int4 wk = work[p];
    const int e0 = inc_off[wk.x], deg = (wk.z - wk.y) / 64;
    for (int r = 0; r < deg; r++) {
        sell_node[wk.y + 64 * r] = inc_node[e0 + r];
        sell_coef[wk.y + 64 * r] = inc_coef[e0 + r];
    }
}

// One wave per item (node, first incidence, end, partial slot or -1).
template <typename R, int U>
__global__ void __launch_bounds__(256)
k_attract_items(int count, const int4 *__restrict__ items, const int *__restrict__ gid,
                const GraphDesc *__restrict__ gd, const GraphState *__restrict__ gs,
                const typename Vec<R>::v2 *__restrict__ pos,
                const R *__restrict__ size, const int *__restrict__ inc_node,
                const R *__restrict__ inc_coef, double2 *__restrict__ Fa,
                double2 *__restrict__ part)
{
    const int w = xcd_block((int) blockIdx.x, (int) gridDim.x, kAttractXcdRun) * 4 + (threadIdx.x >> 6);
    if (w >= count) return;
    const int lane = lane_id();
    const int4 it = items[w];
    const int i = it.x;
    const int g = gid[i];
    const int flags = gd[g].flags;
    ForceSum<R> sum;
    if (flags & GF_UNIFORM)
        attract_terms<R, U, true>(it.y + lane, it.z, 64, it.z - 1, flags & GF_LINLOG,
                                  flags & GF_ADJUST, pos[i], size[i], pos, size, inc_node,
                                  inc_coef, inc_coef[it.y], sum);
    else
        attract_terms<R, U, false>(it.y + lane, it.z, 64, it.z - 1, flags & GF_LINLOG,
                                   flags & GF_ADJUST, pos[i], size[i], pos, size, inc_node,
                                   inc_coef, (R) 0, sum);
    group_sum<64>(sum.x, sum.y);
    if (lane != 0 || gs[g].retired) return;
    if (it.w < 0) {
        Fa[i] = make_double2(sum.x, sum.y);
    } else {
        part[it.w] = make_double2(sum.x, sum.y);
    }
}

// One wave per node cut into several items: (node, first slot, slots).
template <typename R>
__global__ void __launch_bounds__(256)
k_attract_fix(int count, const int4 *__restrict__ hubs, const int *__restrict__ gid,
              const GraphState *__restrict__ gs, const double2 *__restrict__ part,
              double2 *__restrict__ Fa)
{
    const int w = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (w >= count) return;
    const int lane = lane_id();
    const int4 h = hubs[w];
    if (gs[gid[h.x]].retired) return;
    double ax = 0.0, ay = 0.0;
    for (int q = lane; q < h.z; q += 64) {
        const double2 v = part[h.y + q];
        ax += v.x;
        ay += v.y;
    }
    group_sum<64>(ax, ay);
    if (lane != 0) return;
    Fa[h.x] = make_double2(ax, ay);
}

// ---------------------------------------------------------------------------
// Swinging, traction, adaptive speed and displacement
// (pipeline.cl:438-506, src/forceatlas2.c:445-485). Per-node values and all
// sums are FP64; each tile gives one partial, each graph sums its tiles in a
// fixed order.
// ---------------------------------------------------------------------------

__device__ inline void swing_node(double2 f, double2 fo, double m, double &sw, double &tr)
{
    const double sx = fo.x - f.x, sy = fo.y - f.y;
    const double tx = fo.x + f.x, ty = fo.y + f.y;
    sw = m * __builtin_sqrt(sx * sx + sy * sy);
    tr = 0.5 * m * __builtin_sqrt(tx * tx + ty * ty);
}

template <typename R>
__global__ void __launch_bounds__(256)
k_swing(const Tile *__restrict__ tiles, const GraphDesc *__restrict__ gd,
        const GraphState *__restrict__ gs, const R *__restrict__ mass,
        const uint8_t *__restrict__ fixed, double2 *__restrict__ F,
        const double2 *__restrict__ Fa, const double2 *__restrict__ Fold,
        double2 *__restrict__ part, int N1)
{
    __shared__ double red[2][4];
    const Tile t = tile_of(tiles, blockIdx.x, N1, kTile);
    const bool use_fixed = gd[t.g].flags & GF_FIXED;
    // the force of the iteration: repulsion minus gravity, plus the
    // attraction (held apart while the two were computed side by side;
    // zero for nodes without edges)
    const bool live = !gs[t.g].retired;
    double a = 0.0, b = 0.0;
    for (int i = t.lo + threadIdx.x; i < t.hi; i += blockDim.x) {
        double2 f = F[i];
        if (live) {
            const double2 fa = ld_stream(Fa + i);
            f.x += fa.x;
            f.y += fa.y;
            F[i] = f;
        }
        double sw, tr;
        swing_node(f, Fold[i], (double) mass[i], sw, tr);
        if (use_fixed && fixed[i]) continue;
        a += sw;
        b += tr;
    }
    a = wave_sum(a);
    b = wave_sum(b);
    const int w = threadIdx.x >> 6;
    if (lane_id() == 0) {
        red[0][w] = a;
        red[1][w] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0)
        part[blockIdx.x] = make_double2(((red[0][0] + red[0][1]) + red[0][2]) + red[0][3],
                                        ((red[1][0] + red[1][1]) + red[1][2]) + red[1][3]);
}

__global__ void k_retire(GraphState *gs, int g) { gs[g].retired = 1; }

// Device restatement of fa2_update_speed (src/forceatlas2.c:445-485).
__device__ inline void update_speed(int n, double sw, double tr, double jt_user, double &speed,
                                    double &eff)
{
    if (!(sw > 0)) return;
    const double est = 0.05 * __builtin_sqrt((double) n);
    const double lo = __builtin_sqrt(est), hi = 10.0;
    double jt = est * tr / ((double) n * (double) n);
    if (jt > hi) jt = hi;
    if (jt < lo) jt = lo;
    jt *= jt_user;
    if (sw > 2.0 * tr) {
        if (eff > 0.05) eff *= 0.5;
        if (jt < jt_user) jt = jt_user;
    }
    const double target = jt * eff * tr / sw;
    if (sw > jt * tr) {
        if (eff > 0.05) eff *= 0.7;
    } else if (speed < 1000) {
        eff *= 1.3;
    }
    const double rise = target - speed;
    speed += rise < 0.5 * speed ? rise : 0.5 * speed;
}

// One wave per graph: totals of its tiles, then the new speed.
__global__ void __launch_bounds__(256)
k_speed(int G, const GraphDesc *__restrict__ gd, const double2 *__restrict__ part,
        GraphState *__restrict__ gs)
{
    const int g = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (g >= G) return;
    const int lane = lane_id();
    const GraphDesc D = gd[g];
    double a = 0.0, b = 0.0;
    // eight partials per lane in flight at once, added in the order of the
    // one-at-a-time loop (tile lane, lane + 64, ...)
    for (int t = lane; t < D.ntiles; t += 64 * 8) {
        double2 v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = part[D.toff + min(t + 64 * u, D.ntiles - 1)];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            if (t + 64 * u < D.ntiles) {
                a += v[u].x;
                b += v[u].y;
            }
        }
    }
    a = wave_sum(a);
    b = wave_sum(b);
    if (lane != 0) return;
    GraphState s = gs[g];
    s.active = 0;
    if (s.status == 0 && s.done < D.iters) {
        if (!(a == a) || !(b == b)) {
            s.status = 2; // the totals are not numbers (src/gpu_full.c:486-491)
        } else {
            update_speed(D.n, a, b, D.jt, s.speed, s.eff);
            s.tsw = a;
            s.ttr = b;
            s.done += 1;
            s.active = 1;
        }
    }
    gs[g] = s;
}

template <typename R> struct Acc;
template <> struct Acc<float> {
    static constexpr bool separate = true;
};
template <> struct Acc<double> {
    static constexpr bool separate = false;
};

// Displacement (pipeline.cl:480-506) into the FP64 accumulators, float32
// copies for the next iteration, and the tile's bounding box.
template <typename R>
__global__ void __launch_bounds__(256)
k_apply(const Tile *__restrict__ tiles, const GraphDesc *__restrict__ gd,
        const GraphState *__restrict__ gs, const R *__restrict__ mass,
        const uint8_t *__restrict__ fixed, const double2 *__restrict__ F,
        const double2 *__restrict__ Fold, typename Vec<R>::v2 *__restrict__ pos,
        const R *__restrict__ size, typename Vec<R>::v4 *__restrict__ node4,
        double2 *__restrict__ acc, typename Vec<R>::v4 *__restrict__ tbox, int N1)
{
    __shared__ R lds[16];
    const Tile t = tile_of(tiles, blockIdx.x, N1, kTile);
    const GraphDesc D = gd[t.g];
    const GraphState S = gs[t.g];
    const bool move = S.active;
    const bool use_fixed = D.flags & GF_FIXED, adjust = D.flags & GF_ADJUST;
    const double speed = S.speed;
    R a = (R) INFINITY, b = (R) INFINITY, c = (R) -INFINITY, d = (R) -INFINITY;
    for (int i = t.lo + threadIdx.x; i < t.hi; i += blockDim.x) {
        typename Vec<R>::v2 p = pos[i];
        if (move && !(use_fixed && fixed[i])) {
            const double2 f = ld_stream(F + i);
            double sw, tr;
            swing_node(f, ld_stream(Fold + i), (double) mass[i], sw, tr);
            double factor;
            if (adjust) {
                const double fx = (double) f.x, fy = (double) f.y;
                const double df = __builtin_sqrt(fx * fx + fy * fy);
                factor = 0.1 * speed / (1.0 + __builtin_sqrt(speed * sw));
                if (df > 0) {
                    if (factor * df > 10.0) factor = 10.0 / df;
                } else {
                    factor = 0.0;
                }
            } else {
                factor = speed / (1.0 + __builtin_sqrt(speed * sw));
            }
            if (Acc<R>::separate) {
                double2 q = ld_stream(acc + i);
                q.x += (double) f.x * factor;
                q.y += (double) f.y * factor;
                st_stream(acc + i, q.x, q.y);
                p = mk2<R>((R) q.x, (R) q.y);
            } else {
                p.x += (R) ((double) f.x * factor);
                p.y += (R) ((double) f.y * factor);
            }
            pos[i] = p;
            node4[i] = mk4<R>(p.x, p.y, mass[i], size[i]);
        }
        a = min(a, p.x);
        b = min(b, p.y);
        c = max(c, p.x);
        d = max(d, p.y);
    }
    block_minmax4<R>(a, b, c, d, lds);
    if (threadIdx.x == 0) tbox[blockIdx.x] = mk4<R>(a, b, c, d);
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

// Host work of EngineImpl::open split over threads: f(k) for every k in
// [0, count), pieces taken in turn by up to `threads` threads (the calling
// one among them; fewer when the system gives no more).
constexpr int kHostPiece = 1 << 16;

template <typename F> static void host_parallel(int count, int threads, const F &f)
{
    threads = std::max(1, std::min(threads, count));
    std::atomic<int> next{0};
    auto work = [&] {
        for (int k = next.fetch_add(1); k < count; k = next.fetch_add(1)) f(k);
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; t++) {
        try {
            pool.emplace_back(work);
        } catch (...) {
            break;
        }
    }
    work();
    for (std::thread &th : pool) th.join();
}

// Threads for host work of about `items` elements: one below a million.
static int host_threads_for(long long items)
{
    if (items < (1LL << 20)) return 1;
    const unsigned hc = std::thread::hardware_concurrency();
    return (int) std::min(16u, std::max(1u, hc));
}

// A copy in the order of stream st, waited for: it sees the work enqueued on
// st before it and no other stream's (the null stream's synchronous copies
// would also wait for other threads' blocking streams).
static hipError_t copy_on(void *dst, const void *src, size_t bytes, hipMemcpyKind kind,
                          hipStream_t st)
{
    hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, st);
    return e == hipSuccess ? hipStreamSynchronize(st) : e;
}

template <typename R> struct EngineImpl final : Engine {
    using V2 = typename Vec<R>::v2;
    using V4 = typename Vec<R>::v4;

    int G = 0, N = 0, NI = 0, T = 0;
    long long I = 0;
    std::vector<GraphDesc> h_gd;
    std::vector<Tile> h_tiles;
    std::vector<int4> h_class[2];
    std::vector<int4> h_items, h_hubs;
    int nslots = 0;
    bool any_fp64_sums = false; // a graph with sizes or at theta 0: FP64 traversal sums
    int parity = 0; // F[parity] receives the force of the next iteration
    // N when the batch is a single graph (its tiles and regions are then
    // computed by the kernels rather than loaded, see tile_of), else 0
    int N1 = 0;

    DevBuf<GraphDesc> d_gd;
    DevBuf<GraphState> d_gs;
    DevBuf<Tile> d_tiles;
    DevBuf<int> d_gid, d_gi;
    DevBuf<V2> d_pos;
    DevBuf<V4> d_node4; // (x, y, mass, radius) in the original order
    // the graphs whose tile-count rows each rowscan kernel scans: up to
    // kRowThreadTiles tiles, up to 64, more
    std::vector<int> h_rowg[3];
    DevBuf<int> d_rowg[3];
    DevBuf<double2> d_F[2]; // forces are kept in FP64: the swing is their difference
    // the attraction of the iteration, computed beside the repulsion on a
    // second stream and added to it by k_swing (zero for nodes without edges)
    DevBuf<double2> d_Fa;
    hipStream_t s_att = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    DevBuf<double2> d_acc;
    DevBuf<R> d_mass, d_size, d_coef;
    DevBuf<uint8_t> d_fixed;
    DevBuf<int> d_inc_off, d_inc_node;
    DevBuf<int4> d_class[2];
    // the one-lane class's lists in the sliced layout (k_attract), and the
    // first entry of each wave's slice (one more at the end)
    DevBuf<int> d_sell_node, d_sell_base;
    DevBuf<R> d_sell_coef;
    DevBuf<int4> d_items, d_hubs;
    DevBuf<double2> d_hubpart;
    DevBuf<V4> d_tbox, d_psort, d_nd;
    DevBuf<uint32_t> d_keyA, d_valA, d_keyB, d_valB, d_hist, d_rowtot;
    DevBuf<int2> d_child, d_rng;
    DevBuf<TNode<R>> d_tn;
    DevBuf<int> d_parent_leaf, d_parent_int, d_escape;
    // summary units of level >= 2, level after level (units of level L at
    // level_off[L] .. + level_cnt[L]); their region lists and lengths; the
    // first unit of each (graph, level); each region's level and place
    DevBuf<Unit> d_units;
    std::vector<Unit> h_units;
    int level_off[kMaxLevels] = {0}, level_cnt[kMaxLevels] = {0};
    DevBuf<int> d_ubase, d_ucnt, d_lists, d_xslot;
    DevBuf<double2> d_part;

    // profiling: events between the phases of each iteration
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::vector<hipEvent_t>> ev_iters;
    double stage_s[FA2_FULL_NTIMES] = {0};

    // two iterations (one period of the force buffers) captured once
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;

    // launch shapes (none changes a result)
    fa2_tuning shape = {{4, 4, 4}, 256, 32, 2};

    ~EngineImpl() override
    {
        if (s_att) (void) hipStreamSynchronize(s_att);
        if (stream) (void) hipStreamSynchronize(stream);
        if (graph_exec) (void) hipGraphExecDestroy(graph_exec);
        if (graph) (void) hipGraphDestroy(graph);
        for (hipEvent_t e : ev_pool) (void) hipEventDestroy(e);
        if (ev_fork) (void) hipEventDestroy(ev_fork);
        if (ev_join) (void) hipEventDestroy(ev_join);
        if (s_att) (void) hipStreamDestroy(s_att);
    }

    int open(const EngineGraph *gr, int ngraphs) override;
    int run(int iterations) override;
    int sync() override;
    int positions(int graph, double *out) override;
    int forces(int graph, double *out) override;
    int speed_state(int graph, double out[4]) override;
    int tree(int graph, uint32_t *keys_sorted, uint32_t *order, int *child, int *range,
             int *parent) override;
    int regions(int graph, double *rec, int *skip) override;
    int stage_times(double *t) override;
    int csr(int graph, double *mass, int *inc_off, int *inc_node, double *inc_coef) override;
    int retire(int graph) override;
    int progress(int *done, int *status) override;
    int graph_status(int graph) override;
    bool capture_pair();
    int count_work(unsigned long long out[4]) override;
    int read_states(int count, const int *graphs, double *const *pos, double *const *fold,
                    double (*sp)[4]) override;
    int set_state(int graph, const double *pos, const double *fold, double speed,
                  double eff) override;

    int tune(int force, fa2_tuning *chosen, double *times) override;
    int get_tuning(fa2_tuning *t) override;
    int set_tuning(const fa2_tuning *t) override;

    void enqueue_iteration(std::vector<hipEvent_t> *ev);
    void mark(std::vector<hipEvent_t> *ev);
    // phases 1-4: codes, sort, tree of the current positions (scratch
    // buffers only); `fork` (0, 1) starts the attraction on the second
    // stream before the codes or after the sort, -1 never (enqueue_forces
    // forks it after the tree)
    void enqueue_tree(std::vector<hipEvent_t> *ev, int fork = -1);
    // phases 1-6 without profiling events: the attraction forked at `fork`
    // (0..2) on the second stream and joined after the traversal, or (3)
    // launched after the traversal on the batch's stream; the traversal
    // writes Fout
    void enqueue_forces(double2 *Fout, int fork);
    void fork_attraction();
    // the attraction kernel of class c at unroll u, writing into Fa
    void launch_attract_class(int c, int u, double2 *Fa, hipStream_t st, bool beside = false);
    // every attraction kernel of an iteration, on st
    void launch_attraction(hipStream_t st, bool beside = false);
    void drop_graph();

    // the repulsion kernel for this batch: with FP64 sums (some graph with
    // sizes or at theta 0) or not, counting or not, record offsets in 32 or
    // 64 bits; `threads` per workgroup (64 to
    // 1024) and XCD runs of `xrun` blocks (0: none); each wave walks for the
    // same 64 nodes whatever the shape
    template <bool COUNT> void launch_repulse(double2 *Fout, unsigned long long *cnt,
                                              int threads = 0, int xrun = -1)
    {
        const bool wide = ((size_t) NI + (size_t) N) * sizeof(TNode<R>) >= ((size_t) 1 << 32);
        if (threads <= 0) threads = shape.repulse_threads;
        if (xrun < 0) xrun = shape.repulse_xcd_run;
        // the timed traversal walks two halves per wave (k_repulse2), the
        // counting pass one (k_repulse): the same terms either way
        const unsigned blocks = ceil_div(N, COUNT ? threads : 2 * threads);
#define FA2_REPULSE(ADJ, WIDE)                                                                   \
    if (COUNT)                                                                                   \
        k_repulse<R, ADJ, COUNT, WIDE><<<blocks, threads, 0, stream>>>(                          \
            N, NI, d_gid.p, d_gd.p, d_gs.p, d_psort.p, d_tn.p, d_valA.p, Fout, xrun, cnt);      \
    else                                                                                         \
        k_repulse2<R, ADJ, WIDE><<<blocks, threads, 0, stream>>>(                               \
            N, NI, d_gid.p, d_gd.p, d_gs.p, d_psort.p, d_tn.p, d_valA.p, Fout, xrun)
        if (any_fp64_sums) {
            if (wide) FA2_REPULSE(true, true);
            else FA2_REPULSE(true, false);
        } else {
            if (wide) FA2_REPULSE(false, true);
            else FA2_REPULSE(false, false);
        }
#undef FA2_REPULSE
    }
};

template <typename R> int EngineImpl<R>::open(const EngineGraph *gr, int ngraphs)
{
    G = ngraphs;
    if (G < 1) {
        set_message("no graph to lay out");
        return 1;
    }
    // graph descriptors, tiles
    const bool from_edges = gr[0].edges != nullptr;
    long long nsum = 0, isum = 0;
    for (int g = 0; g < G; g++) {
        const fa2_full_params *p = gr[g].p;
        if ((gr[g].edges != nullptr) != from_edges) {
            set_message("every graph of a batch needs its edge list, or none does");
            return 1;
        }
        if (const char *why = graph_args_error(p, gr[g].edges)) {
            set_message("%s", why);
            return 1;
        }
        nsum += p->n;
        isum += from_edges ? 2LL * gr[g].edges->m : p->inc_off[p->n];
    }
    if (nsum > INT_MAX / 2 || isum > INT_MAX - 1) {
        set_message("too many nodes or edges for one batch");
        return 1;
    }
    N = (int) nsum;
    I = isum;
    NI = N - G;
    h_gd.resize(G);
    int noff = 0, toff = 0;
    for (int g = 0; g < G; g++) {
        const fa2_full_params *p = gr[g].p;
        GraphDesc &D = h_gd[g];
        memset(&D, 0, sizeof(D));
        D.noff = noff;
        D.n = p->n;
        D.ioff = noff - g;
        D.toff = toff;
        D.ntiles = (int) ceil_div(p->n, kTile);
        D.flags = (p->strong ? GF_STRONG : 0) | (p->linlog ? GF_LINLOG : 0) |
                  (p->adjust ? GF_ADJUST : 0) | (p->any_fixed ? GF_FIXED : 0);
        D.iters = p->iterations;
        D.scaling = p->scaling;
        D.gravity = p->gravity;
        D.theta2 = p->theta * p->theta;
        D.jt = p->jitter_tolerance;
        any_fp64_sums = any_fp64_sums || p->adjust || D.theta2 == 0.0;
        for (int j = 0; j < D.ntiles; j++) {
            Tile t;
            t.g = g;
            t.lo = noff + j * kTile;
            t.hi = std::min(noff + p->n, t.lo + kTile);
            t.j = j;
            h_tiles.push_back(t);
        }
        noff += p->n;
        toff += D.ntiles;
    }
    T = toff;
    N1 = G == 1 ? N : 0;
    for (int g = 0; g < G; g++) {
        const int nt = h_gd[g].ntiles;
        h_rowg[nt <= kRowThreadTiles ? 0 : (nt <= 64 ? 1 : 2)].push_back(g);
    }
    // summary units: a graph takes part in level L while it is longer than
    // a unit of level L - 1; its last level has a single unit
    std::vector<int> ubase((size_t) G * kMaxLevels, -1);
    for (int L = 2; L < kMaxLevels; L++) {
        level_off[L] = (int) h_units.size();
        const long long below = (long long) agg_block<R>() << (4 * (L - 2)), size = below * kUnitFan;
        for (int g = 0; g < G; g++) {
            const long long n = h_gd[g].n;
            if (n <= below) continue;
            ubase[(size_t) g * kMaxLevels + L] = (int) h_units.size();
            for (long long u = 0; u < (n + size - 1) / size; u++)
                h_units.push_back(Unit{g, (int) h_units.size(), 0, 0});
        }
        level_cnt[L] = (int) h_units.size() - level_off[L];
    }

    // packed host arrays, filled by pieces of at most kHostPiece nodes of one
    // graph on several host threads; the large ones are left uninitialised
    // (every element is written below)
    const size_t nn = (size_t) N;
    const size_t ni_all = from_edges ? 1 : (size_t) (I > 0 ? I : 1);
    std::vector<int> inc_off(nn + 1);
    std::unique_ptr<int[]> gid(new (std::nothrow) int[nn]);
    std::unique_ptr<int[]> gi(new (std::nothrow) int[NI > 0 ? NI : 1]);
    std::unique_ptr<int[]> inc_node(new (std::nothrow) int[ni_all]);
    std::unique_ptr<R[]> coef(new (std::nothrow) R[ni_all]);
    std::unique_ptr<R[]> mass(new (std::nothrow) R[from_edges ? 1 : nn]);
    std::unique_ptr<R[]> size(new (std::nothrow) R[nn]);
    std::unique_ptr<V2[]> pos(new (std::nothrow) V2[nn]);
    std::unique_ptr<V4[]> node4(new (std::nothrow) V4[from_edges ? 1 : nn]);
    std::unique_ptr<double2[]> acc(new (std::nothrow) double2[Acc<R>::separate ? nn : 1]);
    std::unique_ptr<uint8_t[]> fixed(new (std::nothrow) uint8_t[nn]);
    if (!gid || !gi || !inc_node || !coef || !mass || !size || !pos || !node4 || !acc || !fixed) {
        set_message("out of host memory");
        return 1;
    }
    struct Piece {
        int g, i0, i1;          // graph, its nodes [i0, i1)
        int bad = 0;            // 1: offsets out of order, 2: a node out of range
        bool nonuniform = false;
        std::vector<int4> cls[2][4], items, hubs; // attraction work, see below
        int nslot = 0;
    };
    std::vector<Piece> pieces;
    std::vector<long long> gbase(G); // each graph's first incidence
    long long ioff = 0;
    for (int g = 0; g < G; g++) {
        gbase[g] = ioff;
        if (!from_edges) ioff += gr[g].p->inc_off[gr[g].p->n];
        for (int i = 0; i < h_gd[g].n; i += kHostPiece) {
            Piece pc;
            pc.g = g;
            pc.i0 = i;
            pc.i1 = std::min(h_gd[g].n, i + kHostPiece);
            pieces.push_back(std::move(pc));
        }
    }
    const int nthreads = host_threads_for((long long) N + (from_edges ? 0 : I));
    host_parallel((int) pieces.size(), nthreads, [&](int k) {
        Piece &pc = pieces[k];
        const int g = pc.g;
        const fa2_full_params *p = gr[g].p;
        const GraphDesc &D = h_gd[g];
        const double *x = gr[g].pos, *y = gr[g].pos + p->n;
        const long long ib = gbase[g];
        for (int i = pc.i0; i < pc.i1; i++) {
            const size_t q = (size_t) D.noff + i;
            gid[q] = g;
            size[q] = (p->adjust && p->size) ? (R) p->size[i] : (R) 0;
            fixed[q] = (p->any_fixed && p->fixed) ? (p->fixed[i] != 0) : 0;
            pos[q] = mk2<R>((R) x[i], (R) y[i]);
            if (Acc<R>::separate) acc[q] = make_double2(x[i], y[i]);
            if (!from_edges) {
                mass[q] = (R) p->mass[i];
                inc_off[q] = (int) (ib + p->inc_off[i]);
                node4[q] = mk4<R>(pos[q].x, pos[q].y, mass[q], size[q]);
            }
        }
        for (int i = pc.i0; i < std::min(pc.i1, D.n - 1); i++) gi[(size_t) D.ioff + i] = g;
        if (from_edges) return;
        const int ni = p->inc_off[p->n];
        bool uniform = true; // every coefficient of a node equal to its first, bit for bit
        for (int i = pc.i0; i < pc.i1; i++) {
            const int e0 = p->inc_off[i], e1 = p->inc_off[i + 1];
            if (e0 < 0 || e1 < e0 || e1 > ni) {
                pc.bad = 1;
                return;
            }
            const R first = e1 > e0 ? (R) p->inc_coef[e0] : (R) 0;
            for (int e = e0; e < e1; e++) {
                const int j = p->inc_node[e];
                if (j < 0 || j >= p->n) {
                    pc.bad = 2;
                    return;
                }
                inc_node[(size_t) ib + e] = D.noff + j;
                const R c = (R) p->inc_coef[e];
                coef[(size_t) ib + e] = c;
                uniform = uniform && memcmp(&c, &first, sizeof(R)) == 0;
            }
        }
        pc.nonuniform = !uniform;
    });
    std::vector<char> nonuniform(G, 0);
    for (const Piece &pc : pieces) {
        if (pc.bad) {
            set_message(pc.bad == 1
                            ? "the incidence offsets must rise from 0 to the number of incidences"
                            : "edge refers to a node that does not exist");
            return 1;
        }
        nonuniform[pc.g] |= pc.nonuniform;
    }
    if (!from_edges)
        for (int g = 0; g < G; g++)
            if (!nonuniform[g]) h_gd[g].flags |= GF_UNIFORM;
    inc_off[nn] = (int) ioff;
    const hipStream_t ast = async_alloc ? stream : nullptr;
    if (from_edges) {
        // mass, incidence lists and coefficients on the device; only the
        // row offsets come back, for the attraction work lists
        std::vector<IngestGraph> ig(G);
        std::vector<const fa2_edges *> ed(G);
        long long eo = 0;
        for (int g = 0; g < G; g++) {
            const fa2_edges *E = gr[g].edges;
            const long long n = h_gd[g].n;
            // the mean mass: its sum n + 2m is an integer, exact in FP64
            ig[g] = IngestGraph{eo, h_gd[g].noff, h_gd[g].n, E->outbound ? 1 : 0, 0,
                                E->outbound ? (double) (n + 2LL * E->m) / (double) n : 1.0};
            ed[g] = E;
            eo += E->m;
            // one weight for all edges and no dissuade_hubs: a node's
            // coefficients are all equal (with dissuade_hubs both incidences
            // of an edge divide by the mass of its first node, so a list's
            // coefficients vary with the other end)
            bool same = !E->outbound;
            for (int e = 1; E->w && e < E->m && same; e++)
                same = memcmp(&E->w[e], &E->w[0], sizeof(double)) == 0;
            if (same) h_gd[g].flags |= GF_UNIFORM;
        }
        if (ingest_graphs<R>(stream, ast, N, ig, ed, d_mass, d_inc_off, d_inc_node, d_coef, inc_off))
            return 1;
        I = inc_off[nn];
    }
    // attraction work: one lane per node up to 16 incidences, 16 lanes up
    // to 256, then waves over chunks of kChunk incidences. Lanes of one wave
    // run as many rounds as its longest list: each of the first two classes
    // is ordered by degree band (4 bands, longest first), node order within
    // a band. Every piece lists its own nodes; the lists are joined in piece
    // order, which is node order.
    host_parallel((int) pieces.size(), nthreads, [&](int k) {
        Piece &pc = pieces[k];
        const int noff = h_gd[pc.g].noff;
        for (int i = noff + pc.i0; i < noff + pc.i1; i++) {
            const int e0 = inc_off[i], e1 = inc_off[i + 1], deg = e1 - e0;
            if (deg == 0) continue;
            if (deg <= 16) {
                pc.cls[0][3 - (deg - 1) / 4].push_back(make_int4(i, e0, e1, pc.g));
            } else if (deg <= 256) {
                pc.cls[1][3 - (deg - 1) / 64].push_back(make_int4(i, e0, e1, pc.g));
            } else if (deg <= kChunk) {
                pc.items.push_back(make_int4(i, e0, e1, -1));
            } else { // hub partial sums: slots numbered within the piece
                const int first = pc.nslot;
                for (int e = e0; e < e1; e += kChunk)
                    pc.items.push_back(make_int4(i, e, std::min(e1, e + kChunk), pc.nslot++));
                pc.hubs.push_back(make_int4(i, first, pc.nslot - first, 0));
            }
        }
    });
    // where each piece's lists go
    const size_t P = pieces.size();
    std::vector<size_t> at_cls(2 * 4 * P), at_items(P), at_hubs(P);
    std::vector<int> at_slot(P);
    size_t ncls[2] = {0, 0}, nitems = 0, nhubs = 0;
    nslots = 0;
    for (int c = 0; c < 2; c++)
        for (int b = 0; b < 4; b++)
            for (size_t k = 0; k < P; k++) {
                at_cls[(c * 4 + b) * P + k] = ncls[c];
                ncls[c] += pieces[k].cls[c][b].size();
            }
    for (size_t k = 0; k < P; k++) {
        at_items[k] = nitems;
        at_hubs[k] = nhubs;
        at_slot[k] = nslots;
        nitems += pieces[k].items.size();
        nhubs += pieces[k].hubs.size();
        nslots += pieces[k].nslot;
    }
    for (int c = 0; c < 2; c++) h_class[c].resize(ncls[c]);
    h_items.resize(nitems);
    h_hubs.resize(nhubs);
    host_parallel((int) P, nthreads, [&](int k) {
        Piece &pc = pieces[k];
        for (int c = 0; c < 2; c++)
            for (int b = 0; b < 4; b++)
                std::copy(pc.cls[c][b].begin(), pc.cls[c][b].end(),
                          h_class[c].begin() + at_cls[(c * 4 + b) * P + k]);
        const int sb = at_slot[k];
        for (size_t q = 0; q < pc.items.size(); q++) {
            int4 it = pc.items[q];
            if (it.w >= 0) it.w += sb;
            h_items[at_items[k] + q] = it;
        }
        for (size_t q = 0; q < pc.hubs.size(); q++) {
            int4 h = pc.hubs[q];
            h.y += sb;
            h_hubs[at_hubs[k] + q] = h;
        }
    });
    // the one-lane class in the sliced layout: the slice of wave w (work
    // entries 64 w .. 64 w + 63) holds 64 x (its longest list) entries; a
    // work entry becomes (node, slice + lane, slice + lane + 64 deg, graph)
    const size_t n0 = h_class[0].size(), nw0 = (n0 + 63) / 64;
    std::vector<int> sell_base(nw0 + 1, 0);
    {
        long long at = 0;
        for (size_t w = 0; w < nw0; w++) {
            int longest = 0;
            for (size_t p = 64 * w; p < std::min(n0, 64 * w + 64); p++)
                longest = std::max(longest, h_class[0][p].z - h_class[0][p].y);
            sell_base[w] = (int) at;
            at += 64LL * longest;
            if (at > INT_MAX - 64) {
                set_message("too many incidences for one batch");
                return 1;
            }
        }
        sell_base[nw0] = (int) at;
        for (size_t p = 0; p < n0; p++) {
            int4 &wk = h_class[0][p];
            const int deg = wk.z - wk.y;
            wk.y = sell_base[p / 64] + (int) (p % 64);
            wk.z = wk.y + 64 * deg;
        }
    }

    HipStatus hs;
    auto up = [&](void *dst, const void *src, size_t bytes) {
        if (hs.ok() && bytes) hs.check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream), "upload");
    };
    hs.check(d_gd.alloc(G, ast), "alloc");
    hs.check(d_gs.alloc(G, ast), "alloc");
    hs.check(d_tiles.alloc(T, ast), "alloc");
    hs.check(d_gid.alloc(nn, ast), "alloc");
    hs.check(d_gi.alloc(NI, ast), "alloc");
    hs.check(d_pos.alloc(nn, ast), "alloc");
    hs.check(d_node4.alloc(nn, ast), "alloc");
    hs.check(d_F[0].alloc(nn, ast), "alloc");
    hs.check(d_F[1].alloc(nn, ast), "alloc");
    hs.check(d_Fa.alloc(nn, ast), "alloc");
    hs.check(d_acc.alloc(Acc<R>::separate ? nn : 1, ast), "alloc");
    if (!from_edges) {
        hs.check(d_mass.alloc(nn, ast), "alloc");
        hs.check(d_inc_off.alloc(nn + 1, ast), "alloc");
        hs.check(d_inc_node.alloc((size_t) I, ast), "alloc");
        hs.check(d_coef.alloc((size_t) I, ast), "alloc");
    }
    hs.check(d_size.alloc(nn, ast), "alloc");
    hs.check(d_fixed.alloc(nn, ast), "alloc");
    for (int c = 0; c < 2; c++) hs.check(d_class[c].alloc(h_class[c].size(), ast), "alloc");
    hs.check(d_sell_node.alloc(sell_base[nw0], ast), "alloc");
    hs.check(d_sell_coef.alloc(sell_base[nw0], ast), "alloc");
    hs.check(d_sell_base.alloc(nw0 + 1, ast), "alloc");
    hs.check(d_items.alloc(h_items.size(), ast), "alloc");
    hs.check(d_hubs.alloc(h_hubs.size(), ast), "alloc");
    hs.check(d_hubpart.alloc(nslots, ast), "alloc");
    hs.check(d_tbox.alloc(T, ast), "alloc");
    hs.check(d_psort.alloc(nn + kLeafSize, ast), "alloc");
    hs.check(d_nd.alloc(2 * (size_t) (NI > 0 ? NI : 1), ast), "alloc");
    hs.check(d_keyA.alloc(nn, ast), "alloc");
    hs.check(d_valA.alloc(nn, ast), "alloc");
    hs.check(d_keyB.alloc(nn, ast), "alloc");
    hs.check(d_valB.alloc(nn, ast), "alloc");
    hs.check(d_hist.alloc(256 * (size_t) T, ast), "alloc");
    hs.check(d_rowtot.alloc(256 * (size_t) G, ast), "alloc");
    for (int c = 0; c < 3; c++) hs.check(d_rowg[c].alloc(h_rowg[c].size(), ast), "alloc");
    hs.check(d_child.alloc(NI, ast), "alloc");
    hs.check(d_rng.alloc(NI, ast), "alloc");
    hs.check(d_tn.alloc((size_t) NI + nn, ast), "alloc"); // regions, then one record per leaf
    hs.check(d_parent_leaf.alloc(nn, ast), "alloc");
    hs.check(d_parent_int.alloc(NI, ast), "alloc");
    hs.check(d_escape.alloc(NI, ast), "alloc");
    const size_t UT = h_units.size();
    hs.check(d_units.alloc(UT, ast), "alloc");
    hs.check(d_ubase.alloc((size_t) G * kMaxLevels, ast), "alloc");
    hs.check(d_ucnt.alloc(UT * kCountStride, ast), "alloc");
    hs.check(d_lists.alloc(UT * kUnitCap, ast), "alloc");
    hs.check(d_xslot.alloc(NI > 0 ? NI : 1, ast), "alloc");
    hs.check(d_part.alloc(T, ast), "alloc");
    if (!hs.ok()) {
        set_hip_error("could not allocate memory on the device", hs.err);
        return 1;
    }
    std::vector<GraphState> gs(G);
    bool any_fold = false;
    for (int g = 0; g < G; g++) {
        memset(&gs[g], 0, sizeof(GraphState));
        gs[g].speed = gr[g].speed;
        gs[g].eff = gr[g].eff;
        any_fold = any_fold || gr[g].fold;
    }
    up(d_gd.p, h_gd.data(), G * sizeof(GraphDesc));
    up(d_gs.p, gs.data(), G * sizeof(GraphState));
    up(d_tiles.p, h_tiles.data(), T * sizeof(Tile));
    up(d_units.p, h_units.data(), h_units.size() * sizeof(Unit));
    up(d_ubase.p, ubase.data(), ubase.size() * sizeof(int));
    up(d_gid.p, gid.get(), nn * sizeof(int));
    if (NI > 0) up(d_gi.p, gi.get(), (size_t) NI * sizeof(int));
    up(d_pos.p, pos.get(), nn * sizeof(V2));
    if (Acc<R>::separate) up(d_acc.p, acc.get(), nn * sizeof(double2));
    up(d_size.p, size.get(), nn * sizeof(R));
    up(d_fixed.p, fixed.get(), nn);
    if (from_edges) {
        if (hs.ok()) {
            k_pack_node4<R><<<ceil_div(N, 256), 256, 0, stream>>>(N, d_pos.p, d_mass.p, d_size.p,
                                                                  d_node4.p);
            hs.check(hipGetLastError(), "launch");
        }
    } else {
        up(d_node4.p, node4.get(), nn * sizeof(V4));
        up(d_mass.p, mass.get(), nn * sizeof(R));
        up(d_inc_off.p, inc_off.data(), (nn + 1) * sizeof(int));
        up(d_inc_node.p, inc_node.get(), (size_t) I * sizeof(int));
        up(d_coef.p, coef.get(), (size_t) I * sizeof(R));
    }
    for (int c = 0; c < 2; c++) up(d_class[c].p, h_class[c].data(), h_class[c].size() * sizeof(int4));
    up(d_sell_base.p, sell_base.data(), (nw0 + 1) * sizeof(int));
    if (hs.ok() && n0 > 0) { // after the incidence lists, uploaded or built above
        k_sell_fill<R><<<ceil_div((int) n0, 256), 256, 0, stream>>>(
            (int) n0, d_class[0].p, d_inc_off.p, d_inc_node.p, d_coef.p, d_sell_node.p,
            d_sell_coef.p);
        hs.check(hipGetLastError(), "launch");
    }
    up(d_items.p, h_items.data(), h_items.size() * sizeof(int4));
    up(d_hubs.p, h_hubs.data(), h_hubs.size() * sizeof(int4));
    for (int c = 0; c < 3; c++) up(d_rowg[c].p, h_rowg[c].data(), h_rowg[c].size() * sizeof(int));
    if (hs.ok())
        hs.check(hipMemsetAsync(d_psort.p + nn, 0, kLeafSize * sizeof(V4), stream), "memset");
    if (hs.ok()) hs.check(hipMemsetAsync(d_F[0].p, 0, nn * sizeof(double2), stream), "memset");
    // the first iteration reads the previous force from F[1]: zero, or the
    // force a resumed graph left
    std::vector<double2> fold;
    if (any_fold) {
        fold.assign(nn, make_double2(0.0, 0.0));
        host_parallel(G, host_threads_for(N), [&](int g) {
            const double *f = gr[g].fold;
            const int n = h_gd[g].n, o = h_gd[g].noff;
            for (int i = 0; f && i < n; i++) fold[(size_t) o + i] = make_double2(f[i], f[n + i]);
        });
        up(d_F[1].p, fold.data(), nn * sizeof(double2));
    } else if (hs.ok()) {
        hs.check(hipMemsetAsync(d_F[1].p, 0, nn * sizeof(double2), stream), "memset");
    }
    if (hs.ok()) hs.check(hipMemsetAsync(d_Fa.p, 0, nn * sizeof(double2), stream), "memset");
    // the attraction's stream (default priority) and the events that fork it
    // from the batch's stream and join it back (not made without overlap:
    // an idle stream still takes a share of the hardware queues)
    if (hs.ok() && overlap && !s_att)
        hs.check(hipStreamCreateWithFlags(&s_att, hipStreamNonBlocking), "stream");
    if (hs.ok() && !ev_fork) hs.check(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming), "event");
    if (hs.ok() && !ev_join) hs.check(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming), "event");
    if (hs.ok()) {
        k_tile_bbox<R><<<T, 256, 0, stream>>>(d_tiles.p, d_pos.p, d_tbox.p);
        hs.check(hipGetLastError(), "launch");
    }
    // the host vectors are freed on return: wait for the copies
    if (hs.ok()) hs.check(hipStreamSynchronize(stream), "upload");
    if (!hs.ok()) {
        set_hip_error("could not send the graph to the device", hs.err);
        return 1;
    }
    parity = 0;
    return 0;
}

template <typename R> void EngineImpl<R>::mark(std::vector<hipEvent_t> *ev)
{
    if (!ev) return;
    hipEvent_t e;
    if (ev_pool.empty()) {
        (void) hipEventCreate(&e);
    } else {
        e = ev_pool.back();
        ev_pool.pop_back();
    }
    (void) hipEventRecord(e, stream);
    ev->push_back(e);
}

template <typename R>
void EngineImpl<R>::launch_attract_class(int c, int u, double2 *Fa, hipStream_t st, bool beside)
{
    const int cnt = c < 2 ? (int) h_class[c].size() : (int) h_items.size();
    if (cnt == 0) return;
    const unsigned lds = beside ? kAttractBesideLds : 0;
#define FA2_ATTRACT(U)                                                                          \
    do {                                                                                        \
        if (c == 0)                                                                             \
            k_attract<R, 1, U><<<ceil_div(cnt, 256), 256, lds, st>>>(                         \
                cnt, d_class[0].p, d_gd.p, d_gs.p, d_pos.p, d_size.p, d_sell_node.p,            \
                d_sell_coef.p, Fa);                                                             \
        else if (c == 1)                                                                        \
            k_attract<R, 16, U><<<ceil_div(cnt, 16), 256, lds, st>>>(                         \
                cnt, d_class[1].p, d_gd.p, d_gs.p, d_pos.p, d_size.p, d_inc_node.p, d_coef.p, Fa); \
        else                                                                                    \
            k_attract_items<R, U><<<ceil_div(cnt, 4), 256, lds, st>>>(                        \
                cnt, d_items.p, d_gid.p, d_gd.p, d_gs.p, d_pos.p, d_size.p, d_inc_node.p,        \
                d_coef.p, Fa, d_hubpart.p);                                                     \
    } while (0)
    switch (u) {
    case 1: FA2_ATTRACT(1); break;
    case 2: FA2_ATTRACT(2); break;
    case 8: FA2_ATTRACT(8); break;
    default: FA2_ATTRACT(4); break;
    }
#undef FA2_ATTRACT
}

template <typename R> void EngineImpl<R>::launch_attraction(hipStream_t st, bool beside)
{
    // the hubs' work items first: the largest kernel gets the wave slots that
    // free up first beside the traversal, and the lists of few incidences
    // fill in after it (each kernel writes its own nodes: the order changes
    // no result)
    for (int c = 2; c >= 0; c--)
        launch_attract_class(c, shape.attract_unroll[c], d_Fa.p, st, beside);
    if (!h_hubs.empty())
        k_attract_fix<R><<<ceil_div((int) h_hubs.size(), 4), 256, 0, st>>>(
            (int) h_hubs.size(), d_hubs.p, d_gid.p, d_gs.p, d_hubpart.p, d_Fa.p);
}

template <typename R> void EngineImpl<R>::fork_attraction()
{
    (void) hipEventRecord(ev_fork, stream);
    (void) hipStreamWaitEvent(s_att, ev_fork, 0);
    launch_attraction(s_att, true);
    (void) hipEventRecord(ev_join, s_att);
}

template <typename R> void EngineImpl<R>::enqueue_forces(double2 *Fout, int fork)
{
    if (!s_att) fork = 3;
    enqueue_tree(nullptr, fork < 2 ? fork : -1);
    // 5: repulsion minus gravity. Beside the traversal (fork 2) the
    // attraction's branch starts from the same point but is launched after
    // the traversal: captured in this order, the traversal starts 6 us
    // sooner after the tree in the replayed graph (C3: 1.032 -> 1.026 ms per
    // iteration), and the attraction, which ends well before it, later.
    if (fork == 2) (void) hipEventRecord(ev_fork, stream);
    launch_repulse<false>(Fout, nullptr);
    if (fork == 2) {
        (void) hipStreamWaitEvent(s_att, ev_fork, 0);
        launch_attraction(s_att, true);
        (void) hipEventRecord(ev_join, s_att);
    }
    if (fork < 3) (void) hipStreamWaitEvent(stream, ev_join, 0);
    else launch_attraction(stream);
}

template <typename R> void EngineImpl<R>::enqueue_iteration(std::vector<hipEvent_t> *ev)
{
    double2 *Fc = d_F[parity].p, *Fo = d_F[1 - parity].p;
    // 6: the attraction needs only the positions of the last displacement.
    // Without profiling it runs on the second stream, forked at the point
    // the launch shapes name and joined before the swing. Beside the
    // traversal (the default) it overlaps best: the traversal is bound by
    // its VALU work, the attraction by the gathers of the vector memory
    // pipeline; beside the codes and the sort, which wait on memory too,
    // both slow down.
    if (!ev && overlap) {
        enqueue_forces(Fc, shape.attract_fork);
    } else {
        enqueue_tree(ev);
        launch_repulse<false>(Fc, nullptr);
        mark(ev);
        launch_attraction(stream);
    }
    mark(ev);
    // 7-8: the force (k_swing adds the attraction), swinging, speed,
    // displacement
    k_swing<R><<<T, 256, 0, stream>>>(d_tiles.p, d_gd.p, d_gs.p, d_mass.p, d_fixed.p, Fc, d_Fa.p,
                                      Fo, d_part.p, N1);
    k_speed<<<ceil_div(G, 4), 256, 0, stream>>>(G, d_gd.p, d_part.p, d_gs.p);
    k_apply<R><<<T, 256, 0, stream>>>(d_tiles.p, d_gd.p, d_gs.p, d_mass.p, d_fixed.p, Fc, Fo,
                                      d_pos.p, d_size.p, d_node4.p, d_acc.p, d_tbox.p, N1);
    mark(ev);
    parity = 1 - parity;
}

template <typename R> void EngineImpl<R>::enqueue_tree(std::vector<hipEvent_t> *ev, int fork)
{
    if (fork == 0) fork_attraction();
    mark(ev);
    // 1-2: codes
    k_morton<R><<<T, kSortThreads, 0, stream>>>(d_tiles.p, d_gd.p, d_gs.p, d_pos.p, d_tbox.p, d_keyA.p,
                                       d_valA.p, d_hist.p, (int) h_units.size(), d_ucnt.p, N1);
    mark(ev);
    // 3: sort, A -> B -> A -> B -> A (k_aggregate_tile gathers psort)
    for (int pass = 0; pass < 4; pass++) {
        uint32_t *ki = pass % 2 ? d_keyB.p : d_keyA.p, *vi = pass % 2 ? d_valB.p : d_valA.p;
        uint32_t *ko = pass % 2 ? d_keyA.p : d_keyB.p, *vo = pass % 2 ? d_valA.p : d_valB.p;
        if (pass > 0) // the first pass's counts come from k_morton
            k_sort_upsweep<<<T, kSortThreads, 0, stream>>>(d_tiles.p, d_gd.p, d_gs.p, ki, 8 * pass,
                                                            d_hist.p, N1);
        if (const int n0 = (int) h_rowg[0].size())
            k_sort_rowscan_thread<<<n0, 256, 0, stream>>>(256 * n0, d_rowg[0].p, d_gd.p, d_gs.p,
                                                          d_hist.p, d_rowtot.p);
        if (const int n1 = (int) h_rowg[1].size())
            k_sort_rowscan<<<64 * n1, 256, 0, stream>>>(256 * n1, d_rowg[1].p, d_gd.p, d_gs.p,
                                                        d_hist.p, d_rowtot.p);
        if (const int n2 = (int) h_rowg[2].size())
            k_sort_rowscan_block<<<256 * n2, 256, 0, stream>>>(d_rowg[2].p, d_gd.p, d_gs.p,
                                                               d_hist.p, d_rowtot.p);
        k_sort_downsweep<<<T, kSortThreads, 0, stream>>>(d_tiles.p, d_gd.p, d_gs.p, ki, vi, 8 * pass,
                                                          d_hist.p, d_rowtot.p, ko, vo, N1);
    }
    mark(ev);
    if (fork == 1) fork_attraction();
    // 4: tree
    if (NI > 0) {
        k_karras<R><<<ceil_div(NI, kKarrasThreads), kKarrasThreads, 0, stream>>>(
            NI, N, d_gi.p, d_gd.p, d_gs.p, d_keyA.p, d_child.p, d_rng.p, d_parent_leaf.p, d_parent_int.p,
            d_escape.p, d_ubase.p, d_ucnt.p, d_lists.p, d_xslot.p, N1);
        k_aggregate_tile<R><<<T * (kTile / agg_block<R>()), 1024, 0, stream>>>(
            d_tiles.p, d_gd.p, d_gs.p, d_valA.p, d_node4.p, d_psort.p, d_child.p, d_rng.p,
            d_parent_leaf.p, d_parent_int.p, d_escape.p, NI, d_nd.p, d_tn.p, N1);
        for (int L = 2; L < kMaxLevels; L++)
            if (level_cnt[L] > 0)
                k_aggregate_unit<R><<<level_cnt[L], 1024, 0, stream>>>(
                    d_units.p + level_off[L], L, d_gd.p, d_gs.p, d_psort.p, d_child.p, d_rng.p,
                    d_parent_int.p, d_escape.p, d_xslot.p, d_lists.p, d_ucnt.p, NI, d_nd.p,
                    d_tn.p);
    }
    mark(ev);
}

// Record two iterations into a HIP graph (no memory set or copy is part of
// an iteration). False when the capture fails: the iterations are then
// launched one kernel at a time.
template <typename R> bool EngineImpl<R>::capture_pair()
{
    if (graph_exec) return true;
    if (hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return false;
    enqueue_iteration(nullptr);
    enqueue_iteration(nullptr);
    hipGraph_t g = nullptr;
    if (hipStreamEndCapture(stream, &g) != hipSuccess || !g) {
        // the stream leaves capture whatever the outcome; the iterations then
        // run launch by launch
        if (g) (void) hipGraphDestroy(g);
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(stream, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone) {
            hipGraph_t dead = nullptr;
            if (hipStreamEndCapture(stream, &dead) == hipSuccess && dead) (void) hipGraphDestroy(dead);
        }
        (void) hipGetLastError();
        return false;
    }
    if (hipGraphInstantiate(&graph_exec, g, nullptr, nullptr, 0) != hipSuccess) {
        (void) hipGraphDestroy(g);
        graph_exec = nullptr;
        (void) hipGetLastError();
        return false;
    }
    graph = g;
    return true;
}

template <typename R> int EngineImpl<R>::run(int iterations)
{
    (void) hipSetDevice(device);
    double work = 0.0;
    if (!profile && use_graph && iterations >= 2) {
        if (parity != 0) { // the captured pair starts from F[0]
            enqueue_iteration(nullptr);
            iterations--;
        }
        if (!capture_pair()) use_graph = 0;
        for (; use_graph && iterations >= 2; iterations -= 2) {
            work += 2.0 * ((double) N + (double) I);
            if (work >= 1e6) {
                work = 0.0;
                if (interrupted && interrupted()) {
                    set_message("the layout was interrupted");
                    return 1;
                }
            }
            const hipError_t e = hipGraphLaunch(graph_exec, stream);
            if (e != hipSuccess) {
                set_hip_error("the computation on the GPU failed", e);
                return 1;
            }
        }
    }
    for (int it = 0; it < iterations; it++) {
        work += (double) N + (double) I;
        if (work >= 1e6) {
            work = 0.0;
            if (interrupted && interrupted()) {
                set_message("the layout was interrupted");
                return 1;
            }
        }
        if (profile) {
            ev_iters.emplace_back();
            enqueue_iteration(&ev_iters.back());
        } else {
            enqueue_iteration(nullptr);
        }
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            set_hip_error("the computation on the GPU failed", e);
            return 1;
        }
    }
    return 0;
}

template <typename R> int EngineImpl<R>::sync()
{
    (void) hipSetDevice(device);
    hipError_t e = hipStreamSynchronize(stream);
    if (e != hipSuccess) {
        set_hip_error("the computation on the GPU failed", e);
        return 1;
    }
    // fold the profiling events into the stage times
    static const int slot[6] = {FA2_FULL_T_CODES, FA2_FULL_T_SORT, FA2_FULL_T_TREE,
                                FA2_FULL_T_REPULSION, FA2_FULL_T_ATTRACTION, FA2_FULL_T_MOVE};
    for (auto &evs : ev_iters) {
        for (size_t k = 0; k + 1 < evs.size() && k < 6; k++) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, evs[k], evs[k + 1]) == hipSuccess)
                stage_s[slot[k]] += 1e-3 * ms;
        }
        for (hipEvent_t ev : evs) ev_pool.push_back(ev);
    }
    ev_iters.clear();
    std::vector<GraphState> gs(G);
    e = copy_on(gs.data(), d_gs.p, G * sizeof(GraphState), hipMemcpyDeviceToHost, stream);
    if (e != hipSuccess) {
        set_hip_error("the computation on the GPU failed", e);
        return 1;
    }
    for (int g = 0; g < G; g++)
        if (gs[g].status) {
            set_message("the computation on the GPU produced invalid numbers");
            return 1;
        }
    return 0;
}

template <typename R> int EngineImpl<R>::graph_status(int graph)
{
    if (graph < 0 || graph >= G) return -1;
    GraphState s;
    if (copy_on(&s, d_gs.p + graph, sizeof(s), hipMemcpyDeviceToHost, stream) != hipSuccess)
        return -1;
    return s.status;
}

template <typename R> int EngineImpl<R>::positions(int graph, double *out)
{
    if (graph < 0 || graph >= G) {
        set_message("no graph %d in this batch", graph);
        return 1;
    }
    (void) hipSetDevice(device);
    const GraphDesc &D = h_gd[graph];
    const size_t n = (size_t) D.n;
    hipError_t e;
    if (Acc<R>::separate) {
        std::vector<double2> h(n);
        e = hipMemcpyAsync(h.data(), d_acc.p + D.noff, n * sizeof(double2), hipMemcpyDeviceToHost,
                           stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        for (size_t i = 0; i < n && e == hipSuccess; i++) {
            out[i] = h[i].x;
            out[n + i] = h[i].y;
        }
    } else {
        std::vector<V2> h(n);
        e = hipMemcpyAsync(h.data(), d_pos.p + D.noff, n * sizeof(V2), hipMemcpyDeviceToHost,
                           stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        for (size_t i = 0; i < n && e == hipSuccess; i++) {
            out[i] = (double) h[i].x;
            out[n + i] = (double) h[i].y;
        }
    }
    if (e != hipSuccess) {
        set_hip_error("could not fetch the layout from the device", e);
        return 1;
    }
    return 0;
}

template <typename R> int EngineImpl<R>::forces(int graph, double *out)
{
    if (graph < 0 || graph >= G) {
        set_message("no graph %d in this batch", graph);
        return 1;
    }
    (void) hipSetDevice(device);
    const GraphDesc &D = h_gd[graph];
    const size_t n = (size_t) D.n;
    std::vector<double2> h(n);
    // the last force is in F[1 - parity]
    hipError_t e = hipMemcpyAsync(h.data(), d_F[1 - parity].p + D.noff, n * sizeof(double2),
                                  hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) {
        set_hip_error("could not fetch the forces from the device", e);
        return 1;
    }
    for (size_t i = 0; i < n; i++) {
        out[i] = h[i].x;
        out[n + i] = h[i].y;
    }
    return 0;
}

template <typename R> int EngineImpl<R>::speed_state(int graph, double out[4])
{
    if (graph < 0 || graph >= G) {
        set_message("no graph %d in this batch", graph);
        return 1;
    }
    GraphState s;
    hipError_t e = copy_on(&s, d_gs.p + graph, sizeof(s), hipMemcpyDeviceToHost, stream);
    if (e != hipSuccess) {
        set_hip_error("could not fetch the speed from the device", e);
        return 1;
    }
    out[0] = s.tsw;
    out[1] = s.ttr;
    out[2] = s.speed;
    out[3] = s.eff;
    return 0;
}

template <typename R>
int EngineImpl<R>::tree(int graph, uint32_t *keys_sorted, uint32_t *order, int *child, int *range,
                        int *parent)
{
    if (graph < 0 || graph >= G) {
        set_message("no graph %d in this batch", graph);
        return 1;
    }
    (void) hipSetDevice(device);
    const GraphDesc &D = h_gd[graph];
    const int n = D.n;
    std::vector<uint32_t> ks(n), ord(n);
    std::vector<int2> ch(n - 1);
    std::vector<TNode<R>> tn(n - 1);
    std::vector<int> pl(n), pi(n - 1);
    hipError_t e = hipStreamSynchronize(stream);
    auto down = [&](void *dst, const void *src, size_t bytes) {
        if (e == hipSuccess && bytes) e = copy_on(dst, src, bytes, hipMemcpyDeviceToHost, stream);
    };
    down(ks.data(), d_keyA.p + D.noff, n * sizeof(uint32_t));
    down(ord.data(), d_valA.p + D.noff, n * sizeof(uint32_t));
    down(ch.data(), d_child.p + D.ioff, (n - 1) * sizeof(int2));
    down(tn.data(), d_tn.p + D.ioff, (n - 1) * sizeof(TNode<R>));
    down(pl.data(), d_parent_leaf.p + D.noff, n * sizeof(int));
    down(pi.data(), d_parent_int.p + D.ioff, (n - 1) * sizeof(int));
    if (e != hipSuccess) {
        set_hip_error("could not fetch the tree from the device", e);
        return 1;
    }
    // reference numbering: regions 0..n-2, leaves n-1..2n-2 (pipeline.cl:25-27)
    auto ref = [&](int r) { return r >= 0 ? r - D.ioff : n - 1 + (~r - D.noff); };
    for (int k = 0; k < n; k++) {
        if (keys_sorted) keys_sorted[k] = ks[k];
        if (order) order[k] = ord[k] - (uint32_t) D.noff;
    }
    for (int r = 0; r < n - 1; r++) {
        if (child) {
            child[2 * r] = ref(ch[r].x);
            child[2 * r + 1] = ref(ch[r].y);
        }
        if (range) {
            range[2 * r] = tn[r].l.z - D.noff;
            range[2 * r + 1] = tn[r].l.w - D.noff;
        }
        if (parent) parent[r] = pi[r] < 0 ? -1 : pi[r] - D.ioff;
    }
    if (parent)
        for (int k = 0; k < n; k++) parent[n - 1 + k] = pl[k] - D.ioff;
    return 0;
}

template <typename R> int EngineImpl<R>::regions(int graph, double *rec, int *skip)
{
    if (graph < 0 || graph >= G) {
        set_message("no graph %d in this batch", graph);
        return 1;
    }
    (void) hipSetDevice(device);
    const GraphDesc &D = h_gd[graph];
    const int n = D.n;
    std::vector<TNode<R>> tn(n - 1);
    hipError_t e = hipStreamSynchronize(stream);
    if (e == hipSuccess)
        e = copy_on(tn.data(), d_tn.p + D.ioff, (n - 1) * sizeof(TNode<R>), hipMemcpyDeviceToHost,
                    stream);
    if (e != hipSuccess) {
        set_hip_error("could not fetch the tree from the device", e);
        return 1;
    }
    for (int r = 0; r < n - 1; r++) {
        if (rec) {
            rec[4 * r] = (double) tn[r].c.x;
            rec[4 * r + 1] = (double) tn[r].c.y;
            rec[4 * r + 2] = (double) tn[r].c.z;
            rec[4 * r + 3] = (double) tn[r].c.w;
        }
        if (skip) {
            const int s = tn[r].l.y; // a record index
            skip[r] = s < 0 ? -1 : s < NI ? s - D.ioff : n - 1 + (s - NI - D.noff);
        }
    }
    return 0;
}

template <typename R>
int EngineImpl<R>::csr(int graph, double *mass, int *inc_off, int *inc_node, double *inc_coef)
{
    if (graph < 0 || graph >= G) {
        set_message("no graph %d in this batch", graph);
        return 1;
    }
    (void) hipSetDevice(device);
    const GraphDesc &D = h_gd[graph];
    const size_t n = (size_t) D.n;
    std::vector<int> off(n + 1);
    std::vector<R> m(n);
    hipError_t e = hipStreamSynchronize(stream);
    auto down = [&](void *dst, const void *src, size_t bytes) {
        if (e == hipSuccess && bytes) e = copy_on(dst, src, bytes, hipMemcpyDeviceToHost, stream);
    };
    down(off.data(), d_inc_off.p + D.noff, (n + 1) * sizeof(int));
    down(m.data(), d_mass.p + D.noff, n * sizeof(R));
    const size_t i0 = (size_t) off[0], ni = (size_t) (off[n] - off[0]);
    std::vector<int> node(ni);
    std::vector<R> coef(ni);
    down(node.data(), d_inc_node.p + i0, ni * sizeof(int));
    down(coef.data(), d_coef.p + i0, ni * sizeof(R));
    if (e != hipSuccess) {
        set_hip_error("could not fetch the incidence lists from the device", e);
        return 1;
    }
    for (size_t i = 0; i < n; i++) {
        if (mass) mass[i] = (double) m[i];
        if (inc_off) inc_off[i] = off[i] - (int) i0;
    }
    if (inc_off) inc_off[n] = off[n] - (int) i0;
    for (size_t k = 0; k < ni; k++) {
        if (inc_node) inc_node[k] = node[k] - D.noff;
        if (inc_coef) inc_coef[k] = (double) coef[k];
    }
    return 0;
}

template <typename R> int EngineImpl<R>::retire(int graph)
{
    if (graph < 0 || graph >= G) {
        set_message("no graph %d in this batch", graph);
        return 1;
    }
    (void) hipSetDevice(device);
    k_retire<<<1, 1, 0, stream>>>(d_gs.p, graph);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_hip_error("could not update the batch", e);
        return 1;
    }
    return 0;
}

template <typename R> int EngineImpl<R>::progress(int *done, int *status)
{
    (void) hipSetDevice(device);
    std::vector<GraphState> gs(G);
    hipError_t e = hipMemcpyAsync(gs.data(), d_gs.p, G * sizeof(GraphState), hipMemcpyDeviceToHost,
                                  stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) {
        set_hip_error("could not read the progress of the batch", e);
        return 1;
    }
    for (int g = 0; g < G; g++) {
        if (done) done[g] = gs[g].done;
        if (status) status[g] = gs[g].status;
    }
    return 0;
}

template <typename R> int EngineImpl<R>::stage_times(double *t)
{
    for (int k = 0; k < FA2_FULL_NTIMES; k++) t[k] = stage_s[k];
    return 0;
}

template <typename R> int EngineImpl<R>::count_work(unsigned long long out[4])
{
    (void) hipSetDevice(device);
    DevBuf<unsigned long long> c;
    hipError_t e = c.alloc(4);
    if (e == hipSuccess) e = hipMemsetAsync(c.p, 0, 4 * sizeof(unsigned long long), stream);
    // the tree of the last iteration; the forces go to the buffer that the
    // next iteration overwrites first
    double2 *Fs = d_F[parity].p;
    if (e == hipSuccess) {
        launch_repulse<true>(Fs, c.p);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out, c.p, 4 * sizeof(unsigned long long),
                                            hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) {
        set_hip_error("could not count the work of the repulsion", e);
        return 1;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Autotuner. Candidates differ only in launch shape, never in what a lane
// adds or in which order, so any choice gives bit-identical layouts. The
// choice is cached per process by device, precision, sizes (adjust or not)
// and the binary magnitude of the node count and of each attraction class.
// ---------------------------------------------------------------------------

namespace {
struct TuneKey {
    int device, real_bytes, adjust, n, c0, c1, c2;
    bool operator<(const TuneKey &o) const
    {
        return std::tie(device, real_bytes, adjust, n, c0, c1, c2) <
               std::tie(o.device, o.real_bytes, o.adjust, o.n, o.c0, o.c1, o.c2);
    }
};
std::mutex tune_mutex;
std::map<TuneKey, fa2_tuning> tune_cache;

int magnitude(long long v)
{
    int b = 0;
    while (v > 0) {
        b++;
        v >>= 1;
    }
    return b;
}

bool valid_unroll(int u) { return u == 1 || u == 2 || u == 4 || u == 8; }
} // namespace

template <typename R> void EngineImpl<R>::drop_graph()
{
    if (graph_exec) (void) hipGraphExecDestroy(graph_exec);
    if (graph) (void) hipGraphDestroy(graph);
    graph_exec = nullptr;
    graph = nullptr;
}

template <typename R> int EngineImpl<R>::get_tuning(fa2_tuning *t)
{
    *t = shape;
    return 0;
}

template <typename R> int EngineImpl<R>::set_tuning(const fa2_tuning *t)
{
    for (int c = 0; c < 3; c++)
        if (!valid_unroll(t->attract_unroll[c])) {
            set_message("attraction unroll %d: 1, 2, 4 or 8", t->attract_unroll[c]);
            return 1;
        }
    if (t->repulse_threads != 64 && t->repulse_threads != 128 && t->repulse_threads != 256 &&
        t->repulse_threads != 512 && t->repulse_threads != 1024) {
        set_message("traversal workgroup of %d threads: 64, 128, 256, 512 or 1024",
                    t->repulse_threads);
        return 1;
    }
    if (t->repulse_xcd_run < 0 || t->repulse_xcd_run > 4096) {
        set_message("XCD run of %d blocks: 0 to 4096", t->repulse_xcd_run);
        return 1;
    }
    if (t->attract_fork < 0 || t->attract_fork > 3) {
        set_message("attraction fork %d: 0, 1, 2 or 3", t->attract_fork);
        return 1;
    }
    (void) hipSetDevice(device);
    (void) hipStreamSynchronize(stream); // a captured graph may still be running
    shape = *t;
    drop_graph(); // recaptured with the new shapes at the next run
    return 0;
}

template <typename R> int EngineImpl<R>::tune(int force, fa2_tuning *chosen, double *times)
{
    (void) hipSetDevice(device);
    const TuneKey key{device, (int) sizeof(R), any_fp64_sums ? 1 : 0, magnitude(N),
                      magnitude((long long) h_class[0].size()),
                      magnitude((long long) h_class[1].size()), magnitude((long long) h_items.size())};
    if (!force) {
        std::lock_guard<std::mutex> lock(tune_mutex);
        auto it = tune_cache.find(key);
        if (it != tune_cache.end()) {
            if (set_tuning(&it->second)) return 1;
            if (chosen) *chosen = shape;
            return 0;
        }
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t err = hipEventCreate(&e0);
    if (err == hipSuccess) err = hipEventCreate(&e1);
    // the tree of the current positions; the candidates write only the force
    // buffer that the next iteration overwrites first
    double2 *Fs = d_F[parity].p;
    if (err == hipSuccess) {
        enqueue_tree(nullptr);
        err = hipGetLastError();
    }
    auto best_of_1 = [&](auto &&launch) -> double {
        float ms = 0.f;
        if (err == hipSuccess) err = hipEventRecord(e0, stream);
        if (err == hipSuccess) {
            launch();
            err = hipGetLastError();
        }
        if (err == hipSuccess) err = hipEventRecord(e1, stream);
        if (err == hipSuccess) err = hipEventSynchronize(e1);
        if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
        return ms;
    };
    auto best_of_3 = [&](auto &&launch) -> double {
        double best = 0.0;
        for (int k = 0; k < 3 && err == hipSuccess; k++) {
            err = hipEventRecord(e0, stream);
            if (err != hipSuccess) break;
            launch();
            err = hipGetLastError();
            if (err == hipSuccess) err = hipEventRecord(e1, stream);
            if (err == hipSuccess) err = hipEventSynchronize(e1);
            float ms = 0.f;
            if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
            if (err == hipSuccess && (k == 0 || ms < best)) best = ms;
        }
        return best;
    };
    double t[FA2_TUNE_NTIMES] = {0};
    fa2_tuning pick = shape;
    static const int threads[5] = {64, 128, 256, 512, 1024}, unroll[4] = {1, 2, 4, 8};
    for (int k = 0; k < 5; k++)
        t[k] = best_of_3([&] { launch_repulse<false>(Fs, nullptr, threads[k], 0); });
    // each default (256 threads, XCD runs of 32 blocks, unroll 4) stays
    // unless another shape is faster by 2%: closer timings are noise
    // the fastest candidate replaces the default when 2% faster
    auto choose = [&](const double *tt, int n, int dflt) {
        int b = 0;
        for (int k = 1; k < n; k++)
            if (tt[k] < tt[b]) b = k;
        return tt[b] < 0.98 * tt[dflt] ? b : dflt;
    };
    pick.repulse_threads = threads[choose(t, 5, 2)];
    static const int runs[5] = {0, 8, 32, 128, 512};
    for (int k = 0; k < 5; k++)
        t[5 + k] = best_of_3(
            [&] { launch_repulse<false>(Fs, nullptr, pick.repulse_threads, runs[k]); });
    pick.repulse_xcd_run = runs[choose(t + 5, 5, 2)];
    // an attraction kernel runs after the traversal, whose records and
    // nodes fill the caches: each candidate is timed right after one, with
    // the LDS reservation it runs with beside the traversal
    for (int c = 0; c < 3; c++) {
        const size_t cnt = c < 2 ? h_class[c].size() : h_items.size();
        if (cnt == 0) continue;
        for (int k = 0; k < 4; k++) {
            double best = 0.0;
            for (int rep = 0; rep < 3 && err == hipSuccess; rep++) {
                launch_repulse<false>(Fs, nullptr, pick.repulse_threads, pick.repulse_xcd_run);
                const double tk =
                    best_of_1([&] { launch_attract_class(c, unroll[k], d_Fa.p, stream, true); });
                if (rep == 0 || tk < best) best = tk;
            }
            t[10 + 4 * c + k] = best;
        }
        pick.attract_unroll[c] = unroll[choose(t + 10 + 4 * c, 4, 2)];
    }
    // where the attraction runs beside the other phases: phases 1-6 with the
    // shapes chosen so far (without a second stream only 3 is possible)
    if (s_att && overlap) {
        const fa2_tuning keep = shape;
        shape = pick;
        for (int f = 0; f < 4; f++) t[22 + f] = best_of_3([&] { enqueue_forces(Fs, f); });
        pick.attract_fork = choose(t + 22, 4, 2);
        // the traversal's workgroup size was timed alone; beside the
        // attraction a larger one can lose what it gained (512 SBM graphs:
        // 512 threads 2.4% faster alone, phases 1-6 8% slower). Another size
        // than the default is kept only when phases 1-6 are 2% faster too.
        if (pick.repulse_threads != 256) {
            shape = pick;
            t[26] = best_of_3([&] { enqueue_forces(Fs, pick.attract_fork); });
            shape.repulse_threads = 256;
            t[27] = best_of_3([&] { enqueue_forces(Fs, pick.attract_fork); });
            if (!(t[26] < 0.98 * t[27])) pick.repulse_threads = 256;
        }
        shape = keep;
    }
    if (e0) (void) hipEventDestroy(e0);
    if (e1) (void) hipEventDestroy(e1);
    if (err != hipSuccess) {
        set_hip_error("could not time the launch shapes", err);
        return 1;
    }
    if (set_tuning(&pick)) return 1;
    {
        std::lock_guard<std::mutex> lock(tune_mutex);
        tune_cache[key] = pick;
    }
    if (chosen) *chosen = pick;
    if (times)
        for (int k = 0; k < FA2_TUNE_NTIMES; k++) times[k] = t[k];
    return 0;
}

template <typename R>
int EngineImpl<R>::read_states(int count, const int *graphs, double *const *pos,
                               double *const *fold, double (*sp)[4])
{
    int lo = N, hi = 0; // the node span of the graphs asked for
    bool want_pos = false, want_fold = false;
    for (int k = 0; k < count; k++) {
        const int g = graphs[k];
        if (g < 0 || g >= G) {
            set_message("no graph %d in this batch", g);
            return 1;
        }
        lo = std::min(lo, h_gd[g].noff);
        hi = std::max(hi, h_gd[g].noff + h_gd[g].n);
        want_pos = want_pos || (pos && pos[k]);
        want_fold = want_fold || (fold && fold[k]);
    }
    if (count <= 0) return 0;
    (void) hipSetDevice(device);
    const size_t span = (size_t) (hi - lo);
    std::vector<double2> acc_h(want_pos && Acc<R>::separate ? span : 0);
    std::vector<V2> pos_h(want_pos && !Acc<R>::separate ? span : 0);
    std::vector<double2> f_h(want_fold ? span : 0);
    std::vector<GraphState> gs_h(sp ? G : 0);
    hipError_t e = hipSuccess;
    auto down = [&](void *dst, const void *src, size_t bytes) {
        if (e == hipSuccess && bytes)
            e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream);
    };
    down(acc_h.data(), d_acc.p + lo, acc_h.size() * sizeof(double2));
    down(pos_h.data(), d_pos.p + lo, pos_h.size() * sizeof(V2));
    // the last force is in F[1 - parity]
    down(f_h.data(), d_F[1 - parity].p + lo, f_h.size() * sizeof(double2));
    down(gs_h.data(), d_gs.p, gs_h.size() * sizeof(GraphState));
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) {
        set_hip_error("could not fetch the state of the batch from the device", e);
        return 1;
    }
    for (int k = 0; k < count; k++) {
        const GraphDesc &D = h_gd[graphs[k]];
        const size_t n = (size_t) D.n, o = (size_t) (D.noff - lo);
        if (pos && pos[k]) {
            double *out = pos[k];
            for (size_t i = 0; i < n; i++) {
                out[i] = Acc<R>::separate ? acc_h[o + i].x : (double) pos_h[o + i].x;
                out[n + i] = Acc<R>::separate ? acc_h[o + i].y : (double) pos_h[o + i].y;
            }
        }
        if (fold && fold[k]) {
            double *out = fold[k];
            for (size_t i = 0; i < n; i++) {
                out[i] = f_h[o + i].x;
                out[n + i] = f_h[o + i].y;
            }
        }
        if (sp) {
            const GraphState &st = gs_h[graphs[k]];
            sp[k][0] = st.tsw;
            sp[k][1] = st.ttr;
            sp[k][2] = st.speed;
            sp[k][3] = st.eff;
        }
    }
    return 0;
}

template <typename R>
int EngineImpl<R>::set_state(int graph, const double *pos, const double *fold, double speed,
                             double eff)
{
    if (graph < 0 || graph >= G) {
        set_message("no graph %d in this batch", graph);
        return 1;
    }
    (void) hipSetDevice(device);
    const GraphDesc &D = h_gd[graph];
    const size_t n = (size_t) D.n;
    std::vector<V2> p(n);
    std::vector<double2> fo(n);
    std::vector<double2> a(Acc<R>::separate ? n : 0);
    for (size_t i = 0; i < n; i++) {
        p[i] = mk2<R>((R) pos[i], (R) pos[n + i]);
        if (Acc<R>::separate) a[i] = make_double2(pos[i], pos[n + i]);
        fo[i] = fold ? make_double2(fold[i], fold[n + i]) : make_double2(0.0, 0.0);
    }
    GraphState st;
    hipError_t e = hipStreamSynchronize(stream);
    if (e == hipSuccess)
        e = copy_on(&st, d_gs.p + graph, sizeof(st), hipMemcpyDeviceToHost, stream);
    st.speed = speed;
    st.eff = eff;
    st.status = 0;
    if (e == hipSuccess)
        e = copy_on(d_gs.p + graph, &st, sizeof(st), hipMemcpyHostToDevice, stream);
    if (e == hipSuccess)
        e = copy_on(d_pos.p + D.noff, p.data(), n * sizeof(V2), hipMemcpyHostToDevice, stream);
    if (e == hipSuccess && Acc<R>::separate)
        e = copy_on(d_acc.p + D.noff, a.data(), n * sizeof(double2), hipMemcpyHostToDevice, stream);
    // the next iteration reads the previous force from F[1 - parity]
    if (e == hipSuccess)
        e = copy_on(d_F[1 - parity].p + D.noff, fo.data(), n * sizeof(double2),
                    hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) {
        k_tile_bbox<R><<<T, 256, 0, stream>>>(d_tiles.p, d_pos.p, d_tbox.p);
        k_pack_node4<R><<<ceil_div(N, 256), 256, 0, stream>>>(N, d_pos.p, d_mass.p, d_size.p,
                                                              d_node4.p);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) {
        set_hip_error("could not set the state of the graph", e);
        return 1;
    }
    return 0;
}

// Device bytes of one graph of n nodes and `incidences` incidence entries,
// following the allocations of EngineImpl::open (per node, per region, per
// incidence, per tile; the attraction work lists at their largest).
template <typename R> static size_t footprint_of(long long n, long long incidences)
{
    using V2 = typename Vec<R>::v2;
    using V4 = typename Vec<R>::v4;
    const long long ni = n - 1, tiles = (n + kTile - 1) / kTile;
    const long long items = incidences / kChunk + n / 256 + 2;
    size_t b = 0;
    b += (size_t) n * (sizeof(int) /*gid*/ + sizeof(V2) /*pos*/ + sizeof(V4) /*node4*/ +
                       3 * sizeof(double2) /*F, Fa*/ + (Acc<R>::separate ? sizeof(double2) : 0) +
                       3 * sizeof(R) /*mass, size*/ + 1 /*fixed*/ + sizeof(int) /*inc_off*/ +
                       sizeof(int4) /*class lists*/ + sizeof(V4) /*psort*/ +
                       4 * sizeof(uint32_t) /*keys, values*/ + sizeof(int) /*parent_leaf*/);
    b += (size_t) ni * (sizeof(int) /*gi*/ + 2 * sizeof(V4) /*nd*/ + 2 * sizeof(int2) /*child, range*/ +
                        sizeof(TNode<R>) + 3 * sizeof(int) /*parent, escape, xslot*/);
    // summary units of level >= 2: at most n / (15 * block) plus one per level
    const long long units = n / (15LL * agg_block<R>()) + kMaxLevels;
    b += (size_t) units * (sizeof(Unit) + (kUnitCap + kCountStride) * sizeof(int));
    b += (size_t) n * sizeof(TNode<R>); // the leaves' traversal records
    b += (size_t) incidences * (sizeof(int) + sizeof(R));
    // the one-lane class's sliced lists: a lane's slot count is its wave's
    // longest list, at most 16, and at most 15 more than its own
    const long long sliced = std::min(16 * n, incidences + 15 * n) + 1024;
    b += (size_t) sliced * (sizeof(int) + sizeof(R)) + (size_t) (n / 64 + 2) * sizeof(int);
    b += (size_t) items * (2 * sizeof(int4) + sizeof(double2));
    b += (size_t) tiles * (sizeof(Tile) + 2 * sizeof(V4) + sizeof(double2) + 256 * sizeof(uint32_t));
    b += sizeof(GraphDesc) + sizeof(GraphState) + 256 * sizeof(uint32_t) +
         kMaxLevels * sizeof(int) + 4096;
    return b;
}

size_t engine_footprint(int use_double, long long n, long long incidences)
{
    return use_double ? footprint_of<double>(n, incidences) : footprint_of<float>(n, incidences);
}

Engine *make_engine(int use_double)
{
    if (use_double) return new (std::nothrow) EngineImpl<double>();
    return new (std::nothrow) EngineImpl<float>();
}

} // namespace fa2
