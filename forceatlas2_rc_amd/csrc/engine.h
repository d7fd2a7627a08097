// Host interface of the all-GPU engine (engine.hip): one object per batch of
// graphs resident on one device.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/fa2.h"
#include "../../include/fa2_hip.h"

namespace fa2 {

// One graph of a batch: its arguments (as for fa2_full_run) and its initial
// positions (x[0..n) then y[0..n)).
struct EngineGraph {
    const fa2_full_params *p;
    const double *pos;
    // the edge list, when the mass and incidence lists are to be built on
    // the device (p's mass and inc_* are then not read); else null
    const fa2_edges *edges = nullptr;
    // the state a graph resumes from (as left by earlier iterations, see
    // read_states): its previous force (x then y; null for zero, the start
    // of a layout), speed and speed efficiency
    const double *fold = nullptr;
    double speed = 1.0, eff = 1.0;
};

struct Engine {
    int device = 0;
    hipStream_t stream = nullptr;
    int profile = 0;
    int (*interrupted)(void) = nullptr;
    // allocate and free the device memory in the order of `stream`
    int async_alloc = 0;
    // replay the iterations from a captured HIP graph (without profile)
    int use_graph = 1;
    // compute the attraction on a second stream beside the tree and the
    // repulsion (without profile)
    int overlap = 1;
    virtual ~Engine() = default;
    // Upload the graphs; 0 on success (message in fa2_gpu_error()).
    virtual int open(const EngineGraph *graphs, int ngraphs) = 0;
    // Enqueue iterations; no host synchronisation.
    virtual int run(int iterations) = 0;
    // Wait for the device; 1 on a device error or invalid numbers.
    virtual int sync() = 0;
    virtual int positions(int graph, double *out) = 0;
    virtual int forces(int graph, double *out) = 0;
    virtual int speed_state(int graph, double out[4]) = 0;
    virtual int tree(int graph, uint32_t *keys_sorted, uint32_t *order, int *child, int *range,
                     int *parent) = 0;
    // Traversal records of the regions of one graph's tree, in the
    // reference numbering: (centre x, centre y, mass, size^2) and the skip
    // link (the region or leaf visited after the region's subtree, -1 at the
    // end). Either pointer may be NULL.
    virtual int regions(int graph, double *rec, int *skip) = 0;
    virtual int stage_times(double *t) = 0;
    // The incidence lists of one graph as held on the device, in the graph's
    // numbering (see fa2_batch_csr).
    virtual int csr(int graph, double *mass, int *inc_off, int *inc_node, double *inc_coef) = 0;
    // Stop computing graph `graph` (its positions and forces stay as they
    // are): the batch keeps its slot, but its repulsion and attraction are
    // skipped from the next iteration on.
    virtual int retire(int graph) = 0;
    // Iterations applied to each graph so far and its status (0, or 2 once
    // it produced invalid numbers: it then stops), ngraphs values each, read
    // from the device after the work enqueued before the call.
    virtual int progress(int *done, int *status) = 0;
    virtual int graph_status(int graph) = 0;
    // Work of the repulsion on the current tree: regions accepted, regions
    // opened, node pairs, wave steps (untimed instrumented traversal).
    virtual int count_work(unsigned long long out[4]) = 0;
    // Replace the state of one graph: positions (x then y), previous force
    // (x then y, NULL for zero), speed and speed efficiency.
    virtual int set_state(int graph, const double *pos, const double *fold, double speed,
                          double eff) = 0;
    // The state of `count` graphs at once, after the work enqueued before the
    // call, with one transfer per array: positions pos[k] (x then y), last
    // force fold[k] (x then y) and (total swinging, total traction, speed,
    // speed efficiency) sp[k] of graph graphs[k]. Any of pos, fold, sp, or
    // any of their entries, may be null.
    virtual int read_states(int count, const int *graphs, double *const *pos,
                            double *const *fold, double (*sp)[4]) = 0;
    // Launch shapes (fa2_tuning of fa2_hip.h); tune() times the candidates
    // on the current positions and keeps the fastest.
    virtual int tune(int force, fa2_tuning *chosen, double *times) = 0;
    virtual int get_tuning(fa2_tuning *t) = 0;
    virtual int set_tuning(const fa2_tuning *t) = 0;
};

// The arguments of one graph that every entry point checks before it
// touches the device: its size, and its edge list (e non-null) or its
// masses and incidence lists. Returns the message of the first problem, or
// null.
inline const char *graph_args_error(const fa2_full_params *p, const fa2_edges *e)
{
    if (!p) return "no graph to lay out";
    if (p->n < 2 || p->n > (1 << 26)) return "the all-GPU layout handles between 2 and 67 million nodes";
    if (e) {
        if (e->m < 0 || (e->m > 0 && (!e->from || !e->to)))
            return "an edge list needs m >= 0 and both endpoint arrays";
        return nullptr;
    }
    if (!p->inc_off || !p->mass || p->inc_off[0] != 0 || p->inc_off[p->n] < 0 ||
        (p->inc_off[p->n] > 0 && (!p->inc_node || !p->inc_coef)))
        return "every graph needs its masses and incidence lists (inc_off[0] = 0, "
               "inc_off[n] >= 0, inc_node and inc_coef when inc_off[n] > 0)";
    return nullptr;
}

Engine *make_engine(int use_double);

// Device memory one graph of n nodes and `incidences` incidence entries
// takes in a batch (an estimate from the engine's own allocations).
size_t engine_footprint(int use_double, long long n, long long incidences);

} // namespace fa2
