// Device-side preprocessing of the all-GPU layout (ingest.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "common.h"
#include "../../include/fa2_hip.h"

namespace fa2 {

// One graph of a batch for the preprocessing: its first packed edge, its
// first node in the batch, its node count, dissuade_hubs and its attraction
// scale (the mean mass with dissuade_hubs, else 1).
struct IngestGraph {
    long long eoff;
    int noff, n;
    int outbound, pad;
    double coef;
};

// Mass (N), packed incidence offsets (N + 1, global incidence indices),
// incidence nodes (global node ids) and coefficients of the graphs, built on
// the device from their edge lists; the offsets are also returned to the
// host. The buffers are allocated here (in the order of `ast` when it is
// not null). 0 on success, 1 with a message.
template <typename R>
int ingest_graphs(hipStream_t st, hipStream_t ast, int N, const std::vector<IngestGraph> &graphs,
                  const std::vector<const fa2_edges *> &edges, DevBuf<R> &mass,
                  DevBuf<int> &inc_off, DevBuf<int> &inc_node, DevBuf<R> &inc_coef,
                  std::vector<int> &h_inc_off);

// Stable LSD radix sort of n (key, value) pairs on the low `bits` bits of
// the keys, by the kernels of the batch sort over one segment (engine.hip):
// k[0], v[0] in, k[*out], v[*out] out. 0 on success, 1 with a message.
int radix_sort_pairs_u32(hipStream_t st, hipStream_t ast, int n, int bits, uint32_t *k[2],
                         uint32_t *v[2], int *out);

} // namespace fa2
