/*
 * fa2_hip.h — entry points of libfa2hip.so beyond the drop-in ABI of fa2.h.
 *
 *  - fa2_layout_run / fa2_layout_large_run: the two .Call entries of the
 *    reference (C_forceatlas2, src/forceatlas2.c:521-827, and
 *    C_forceatlas2_large, :835-947) as plain C functions over plain arrays,
 *    with the same preprocessing (mass = 1 + degree, incidence lists,
 *    attraction coefficients). They are the host drivers of the two GPU
 *    paths: the repulsion offload (forceatlas2_gpu) and the all-GPU layout
 *    (forceatlas2_large).
 *  - fa2_batch_*: many graphs of uneven size laid out together on one
 *    device, every iteration of every graph in the same kernel launches,
 *    with state resident between calls. No counterpart in the reference.
 *  - fa2_serve_*: a layout server over the batch engine (continuous
 *    batching, admission by device-memory footprint). No counterpart in the
 *    reference.
 */
#ifndef FA2_HIP_H
#define FA2_HIP_H

#include <stdint.h>

#include "fa2.h"

#ifdef __cplusplus
extern "C" {
#endif

/* C_forceatlas2 (src/forceatlas2.c:521-827) without R: n nodes, m edges
 * from[e] -- to[e] (0-based) of weight w[e]; pos = x then y, updated in
 * place; size (n radii), fixed (n flags, may be NULL); dpar = (scaling,
 * gravity, jitter_tolerance, theta); flags = (strong, linlog, outbound,
 * adjust, barnes_hut); gpu = (device, use_double), device -1 computing the
 * repulsion on the host. times (7, may be NULL) = total, tree, repulsion,
 * gpu_send, gpu_compute, gpu_fetch, other. 0 on success, 1 on error. */
int fa2_layout_run(int n, int m, const int *from, const int *to, const double *w,
                   double *pos, const double *size, const int *fixed, int iterations,
                   const double *dpar, const int *flags, int threads, const int *gpu,
                   double *times);

/* C_forceatlas2_large (src/forceatlas2.c:835-947) without R: flags =
 * (strong, linlog, outbound, adjust); gpu = (device, use_double, profile).
 * times (FA2_FULL_NTIMES, may be NULL). 0 on success, 1 on error. */
int fa2_layout_large_run(int n, int m, const int *from, const int *to, const double *w,
                         double *pos, const double *size, const int *fixed, int iterations,
                         const double *dpar, const int *flags, const int *gpu, double *times);

/* The Gephi quadtree of the reference (tree_build, src/forceatlas2.c:64-239)
 * built on the host, into a tree whose arrays have been allocated by the
 * caller: reg (2n), perm (n), cnt (8n), buf[0], buf[1] (n), px, py, pm (n),
 * ps (n, or NULL when size is NULL). */
void fa2_tree_build(fa2_tree *t, int n, const double *x, const double *y, const double *mass,
                    const double *size, int threads);

/* Device time in milliseconds of the last repulsion kernel of a
 * fa2_gpu_open session (the all-pairs or the Barnes-Hut kernel; the packing
 * and summing kernels around it excluded), measured with HIP events on the
 * session's stream. 0 on success, 1 before any repulsion call. No
 * counterpart in the reference, whose fa2_gpu_times (src/gpu.c:386-391)
 * returns host wall times. */
int fa2_gpu_kernel_time(const fa2_gpu *g, double *ms);

/* The exact all-pairs repulsion of fa2_gpu_repulse_exact without any device
 * allocation. No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). There is no session: x, y, mass (n
 * each), size (n, read only with adjust, and then required) and dx, dy (n
 * each) are DEVICE arrays of doubles on `device`, and the force is added
 * into dx, dy as fa2_gpu_repulse_exact adds it into its host arrays, with
 * the same kernels and the same result bit for bit. The work is enqueued on
 * `stream` (a hipStream_t; NULL for the null stream) and the call returns
 * without waiting: synchronise the stream before reading dx, dy. Partial
 * results live in a scratch array of 24 MiB that the library holds on each
 * device it runs on, from the first call there to the end of the process;
 * n is limited to 1..2^18 nodes (262,144). Every call on a device shares
 * that device's scratch: calls that overlap on one device (on different
 * streams, or from several host threads) race and give wrong results;
 * calls in order on one stream are safe. 0 on success, 1 on error (message
 * in fa2_gpu_error). */
int fa2_gpu_repulse_exact_noalloc(int device, int use_double, int n, int adjust,
                                  const double *x, const double *y, const double *mass,
                                  const double *size, double scaling, double *dx, double *dy,
                                  void *stream);

/* ----- batches ---------------------------------------------------------- */

typedef struct fa2_batch fa2_batch;

/* A graph given by its edge list, for the preprocessing on the device:
 * m edges from[e] -- to[e] (0-based), weights w (already transformed as by
 * R/forceatlas2.R:254-264; NULL for all 1), and dissuade_hubs. */
typedef struct {
    int m;
    const int *from, *to;
    const double *w;
    int outbound;
} fa2_edges;

/* Upload ngraphs graphs (arguments as for fa2_full_run; the device, the
 * precision and `profile` are those of params[0]) with their initial
 * positions pos[g] (x then y). params[g].iterations caps the iterations of
 * graph g: once reached, the graph stops moving. NULL on error. */
fa2_batch *fa2_batch_open(int ngraphs, const fa2_full_params *params, const double *const *pos);

/* As fa2_batch_open, with the preprocessing of C_forceatlas2_large
 * (src/forceatlas2.c:866-902: mass = 1 + degree, incidence lists in edge
 * order without self-loops, attraction coefficients) done on the device
 * from edges[g]: the mass, inc_off, inc_node and inc_coef fields of
 * params[g] are not read. The result is bit-identical to preparing the
 * graphs on the host. NULL on error. */
fa2_batch *fa2_batch_open_edges(int ngraphs, const fa2_full_params *params,
                                const fa2_edges *edges, const double *const *pos);

/* The incidence lists of graph g as the device holds them, in the graph's
 * own numbering: mass (n), inc_off (n + 1), inc_node (inc_off[n]) and
 * inc_coef (inc_off[n], in the batch's precision). Any pointer may be NULL. */
int fa2_batch_csr(fa2_batch *b, int g, double *mass, int *inc_off, int *inc_node,
                  double *inc_coef);

/* Enqueue `iterations` iterations of every graph; returns without waiting. */
int fa2_batch_run(fa2_batch *b, int iterations);

/* Wait for the device. 1 on a device error or when a graph produced
 * invalid numbers. */
int fa2_batch_sync(fa2_batch *b);

/* Positions (x then y) and last forces (x then y) of graph g. */
int fa2_batch_positions(fa2_batch *b, int g, double *pos);
int fa2_batch_forces(fa2_batch *b, int g, double *f);

/* (total swinging, total traction, speed, speed efficiency) of graph g
 * after its last iteration. */
int fa2_batch_speed(fa2_batch *b, int g, double out[4]);

/* Iterations applied to each graph so far and its status (0, or 2 once its
 * swing or traction totals stopped being numbers: it then stops moving, as
 * the reference stops its layout, src/gpu_full.c:486-491), one value per
 * graph each, after the work enqueued before the call. Either pointer may
 * be NULL. */
int fa2_batch_progress(fa2_batch *b, int *done, int *status);

/* The Morton tree of graph g at its last iteration, in the numbering of the
 * reference (regions 0..n-2, leaves n-1..2n-2): sorted keys (n), sorted
 * permutation (n), children (2(n-1)), ranges (2(n-1)), parents (2n-1). Any
 * pointer may be NULL. */
int fa2_batch_tree(fa2_batch *b, int g, uint32_t *keys_sorted, uint32_t *order, int *child,
                   int *range, int *parent);

/* The traversal records of the regions of graph g at its last iteration, in
 * the numbering of the reference: rec (4(n-1)) = (centre x, centre y, mass,
 * size^2) per region (pipeline.cl:312-323); skip (n-1) = the region (0..n-2)
 * or leaf (n-1..2n-2) that the traversal visits after the region's subtree,
 * -1 after the last. Either pointer may be NULL. The reference keeps these
 * records inside the OpenCL buffers and has no entry point for them. */
int fa2_batch_regions(fa2_batch *b, int g, double *rec, int *skip);

/* Device time of each phase (FA2_FULL_NTIMES slots, seconds, slot 0 unused)
 * summed over the iterations run with profile set. */
int fa2_batch_stage_times(fa2_batch *b, double *times);

/* Work of the Barnes-Hut repulsion on the tree of the last iteration,
 * counted by an untimed instrumented traversal: out = (regions accepted as
 * monopoles, regions opened, node pairs, steps of the waves' walks). A
 * single node met as a region's child counts as an accepted region when
 * its monopole is taken (it is the same term as the pair). */
int fa2_batch_count_work(fa2_batch *b, unsigned long long out[4]);

/* Replace the state of graph g between two runs: positions (x then y), the
 * force of the previous iteration (x then y; NULL for zero, as at the
 * start), the speed and the speed efficiency. With the positions alone
 * (f_old NULL, speed = efficiency = 1) this restarts a layout from stored
 * coordinates as the reference does; with the full state it resumes one. */
int fa2_batch_set_state(fa2_batch *b, int g, const double *pos, const double *f_old,
                        double speed, double speed_efficiency);

/* Launch shapes of a batch's kernels. None of them changes a result, only
 * the speed: attract_unroll[c] is the number of incidences whose loads each
 * lane has in flight at once (1, 2, 4 or 8) in the attraction kernel of
 * class c (c = 0: nodes of 1..16 incidences, one lane each; 1: 17..256,
 * 16 lanes each; 2: longer lists, one wave per chunk),
 * repulse_threads the workgroup size of the Barnes-Hut traversal (64, 128,
 * 256, 512 or 1024), repulse_xcd_run its mapping of blocks to XCDs and
 * attract_fork where the attraction runs beside the other phases. The
 * reference has no counterpart (its work sizes are fixed in
 * src/gpu_full.c:193-198, its kernels run one after the other). */
typedef struct fa2_tuning {
    int attract_unroll[3];
    int repulse_threads;
    /* 0, or the traversal's workgroups handed to each XCD in runs of this
     * many consecutive blocks (0..4096; default 32), so that one L2 serves
     * the tree records of neighbouring nodes */
    int repulse_xcd_run;
    /* where the attraction kernels start on the batch's second stream:
     * 0 with the iteration (beside the codes, sort and tree), 1 after the
     * sort (beside the tree), 2 after the tree (beside the traversal; the
     * default), 3 after the traversal on the batch's own stream. Iterations
     * launched with profiling events, and server batches, always use 3. */
    int attract_fork;
} fa2_tuning;

/* Timings behind fa2_batch_tune: the traversal at 64, 128, 256, 512 and
 * 1024 threads, then at XCD runs of 0, 8, 32, 128 and 512 blocks, then for
 * each attraction class the unrolls 1, 2, 4 and 8 (ms; 0 where a class is
 * empty), then phases 1-6 of an iteration (codes, sort, tree, traversal and
 * attraction) with the attraction forked at 0, 1, 2 and 3, then phases 1-6
 * with the chosen traversal workgroup size and with 256 threads (0 when the
 * chosen size is 256: a size other than 256, timed faster alone, is kept
 * only when these phases are also 2% faster). */
#define FA2_TUNE_NTIMES 28

/* Autotuner: builds the tree of the batch's current positions and times
 * every candidate launch shape on it, with HIP events on the batch's stream
 * (best of three launches each; an attraction candidate right after a
 * traversal launch, as in an iteration), then keeps for the following
 * iterations each default shape (256 threads, XCD runs of 32, unroll 4, fork
 * 2) unless another was timed at least 2% faster. The layout state is left as it was: the candidates write only
 * into the force buffer that the next iteration overwrites. Choices are
 * cached per process by device, precision and the batch's size class; with
 * force == 0 a cached choice is reused without timing (times then stay
 * untouched). chosen and times (FA2_TUNE_NTIMES) may be NULL. 0 on success. */
int fa2_batch_tune(fa2_batch *b, int force, fa2_tuning *chosen, double *times);
int fa2_batch_get_tuning(fa2_batch *b, fa2_tuning *t);
/* 1 with a message on a value out of the ranges above. */
int fa2_batch_set_tuning(fa2_batch *b, const fa2_tuning *t);

/* The HIP stream of the batch (hipStream_t), for callers that time kernels
 * with events on it. */
void *fa2_batch_stream(fa2_batch *b);

void fa2_batch_close(fa2_batch *b);

/* ----- server ----------------------------------------------------------- */

typedef struct fa2_server fa2_server;

/* A layout server on one device: requests are admitted by device-memory
 * footprint (budget_bytes, or 80% of the free memory when <= 0) into
 * batches of one precision each, up to max_batches (1..16) of them running
 * at once on their own streams, each advanced `chunk` iterations at a
 * time. When every stream is taken, waiting requests join a running batch
 * at a chunk boundary after some of its requests finished (the batch is
 * rebuilt with its remaining requests' state and the newcomers). Each
 * stream has a host thread of its own. The streams should have a hardware
 * queue each: run the process with GPU_MAX_HW_QUEUES >= max_batches (the
 * HIP runtime's default is 4). NULL on error. */
fa2_server *fa2_serve_open(int device, double budget_bytes, int max_batches, int chunk);

/* Queue one layout (arguments as for fa2_full_run, use_double choosing its
 * precision; the device and profile fields are the server's) from the
 * initial positions pos (x then y). The inputs are copied. Returns the
 * request id, or -1. */
int fa2_serve_submit(fa2_server *s, const fa2_full_params *p, const double *pos);

/* As fa2_serve_submit, with the graph given by its edge list and its
 * mass and incidence lists prepared on the device (as fa2_batch_open_edges;
 * p's mass and inc_* fields are not read). */
int fa2_serve_submit_edges(fa2_server *s, const fa2_full_params *p, const fa2_edges *edges,
                           const double *pos);

/* 0 queued, 1 running, 2 done, 3 failed; -1 for an unknown request. */
int fa2_serve_status(fa2_server *s, int job);

/* Wait for a request: its final positions (x then y, may be NULL) and
 * times (may be NULL) = (seconds queued, seconds from admission to the
 * result, seconds from submission to the result). 0 on success, 1 when the
 * request failed (message in fa2_gpu_error()). */
int fa2_serve_wait(fa2_server *s, int job, double *pos, double times[3]);

/* out = (batches formed, requests admitted, most batches running at once,
 * memory budget in bytes, batches rebuilt to take waiting requests in,
 * seconds the server spent rebuilding them, batches formed whose iterations
 * could not be captured into a HIP graph and ran launch by launch). */
int fa2_serve_stats(fa2_server *s, double out[7]);

/* Finish the queued and running requests, then release the server. */
void fa2_serve_close(fa2_server *s);

/* Device bytes a request of n nodes and `incidences` incidence entries
 * takes in a batch: the footprint used for admission. */
double fa2_serve_footprint(int use_double, int n, long long incidences);

#ifdef __cplusplus
}
#endif

#endif
