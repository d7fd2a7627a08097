// C entry points of the all-GPU path: fa2_full_run (the drop-in of
// src/gpu_full.c:164-532) and the batch API of fa2_hip.h.
#include <algorithm>
#include <cstring>
#include <ctime>
#include <new>
#include <vector>

#include "common.h"
#include "engine.h"
#include "../../include/fa2.h"
#include "../../include/fa2_hip.h"

using namespace fa2;

struct fa2_batch {
    Engine *e = nullptr;
    ~fa2_batch() { delete e; }
};

static fa2_batch *open_batch(int ngraphs, const fa2_full_params *params,
                             const fa2_edges *edges, const double *const *pos);

extern "C" fa2_batch *fa2_batch_open(int ngraphs, const fa2_full_params *params,
                                     const double *const *pos)
{
    return open_batch(ngraphs, params, nullptr, pos);
}

extern "C" fa2_batch *fa2_batch_open_edges(int ngraphs, const fa2_full_params *params,
                                           const fa2_edges *edges, const double *const *pos)
{
    if (!edges) {
        set_message("no edge lists");
        return nullptr;
    }
    return open_batch(ngraphs, params, edges, pos);
}

extern "C" int fa2_batch_csr(fa2_batch *b, int g, double *mass, int *inc_off, int *inc_node,
                             double *inc_coef)
{
    if (!b) return 1;
    return b->e->csr(g, mass, inc_off, inc_node, inc_coef);
}

static fa2_batch *open_batch(int ngraphs, const fa2_full_params *params,
                             const fa2_edges *edges, const double *const *pos)
{
    if (ngraphs < 1 || !params || !pos) {
        set_message("no graph to lay out");
        return nullptr;
    }
    for (int g = 0; g < ngraphs; g++) // before the device is touched
        if (const char *why = graph_args_error(&params[g], edges ? &edges[g] : nullptr)) {
            set_message("%s", why);
            return nullptr;
        }
    const int device = params[0].device;
    hipStream_t st = device_stream(device);
    if (!st) return nullptr;
    fa2_batch *b = new (std::nothrow) fa2_batch();
    if (!b) {
        set_message("out of memory");
        return nullptr;
    }
    b->e = make_engine(params[0].use_double);
    if (!b->e) {
        delete b;
        set_message("out of memory");
        return nullptr;
    }
    b->e->device = device;
    b->e->stream = st;
    b->e->profile = params[0].profile;
    b->e->interrupted = params[0].interrupted;
    std::vector<EngineGraph> gr(ngraphs);
    for (int g = 0; g < ngraphs; g++) {
        gr[g].p = &params[g];
        gr[g].pos = pos[g];
        gr[g].edges = edges ? &edges[g] : nullptr;
    }
    if (b->e->open(gr.data(), ngraphs)) {
        delete b;
        return nullptr;
    }
    return b;
}

extern "C" int fa2_batch_run(fa2_batch *b, int iterations)
{
    if (!b) return 1;
    return b->e->run(iterations);
}

extern "C" int fa2_batch_sync(fa2_batch *b)
{
    if (!b) return 1;
    return b->e->sync();
}

extern "C" int fa2_batch_positions(fa2_batch *b, int g, double *pos)
{
    if (!b) return 1;
    return b->e->positions(g, pos);
}

extern "C" int fa2_batch_forces(fa2_batch *b, int g, double *f)
{
    if (!b) return 1;
    return b->e->forces(g, f);
}

extern "C" int fa2_batch_speed(fa2_batch *b, int g, double out[4])
{
    if (!b) return 1;
    return b->e->speed_state(g, out);
}

extern "C" int fa2_batch_progress(fa2_batch *b, int *done, int *status)
{
    if (!b) return 1;
    return b->e->progress(done, status);
}

extern "C" int fa2_batch_tree(fa2_batch *b, int g, uint32_t *keys_sorted, uint32_t *order,
                              int *child, int *range, int *parent)
{
    if (!b) return 1;
    return b->e->tree(g, keys_sorted, order, child, range, parent);
}

extern "C" int fa2_batch_regions(fa2_batch *b, int g, double *rec, int *skip)
{
    if (!b) return 1;
    return b->e->regions(g, rec, skip);
}

extern "C" int fa2_batch_stage_times(fa2_batch *b, double *times)
{
    if (!b) return 1;
    return b->e->stage_times(times);
}

extern "C" int fa2_batch_count_work(fa2_batch *b, unsigned long long out[4])
{
    if (!b) return 1;
    return b->e->count_work(out);
}

extern "C" int fa2_batch_set_state(fa2_batch *b, int g, const double *pos, const double *f_old,
                                   double speed, double speed_efficiency)
{
    if (!b) return 1;
    return b->e->set_state(g, pos, f_old, speed, speed_efficiency);
}

extern "C" int fa2_batch_tune(fa2_batch *b, int force, fa2_tuning *chosen, double *times)
{
    if (!b) {
        set_message("no batch");
        return 1;
    }
    return b->e->tune(force, chosen, times);
}

extern "C" int fa2_batch_get_tuning(fa2_batch *b, fa2_tuning *t)
{
    if (!b || !t) {
        set_message("no batch or no tuning");
        return 1;
    }
    return b->e->get_tuning(t);
}

extern "C" int fa2_batch_set_tuning(fa2_batch *b, const fa2_tuning *t)
{
    if (!b || !t) {
        set_message("no batch or no tuning");
        return 1;
    }
    return b->e->set_tuning(t);
}

extern "C" void *fa2_batch_stream(fa2_batch *b) { return b ? (void *) b->e->stream : nullptr; }

extern "C" void fa2_batch_close(fa2_batch *b) { delete b; }

// Wait for `ev`, asking the caller's interrupt callback every 0.5 ms.
// Returns 1 (message set) if it asked to stop or the device failed.
static int wait_polling(hipEvent_t ev, int (*interrupted)(void))
{
    for (;;) {
        const hipError_t q = hipEventQuery(ev);
        if (q == hipSuccess) return 0;
        if (q != hipErrorNotReady) {
            set_hip_error("the computation on the GPU failed", q);
            return 1;
        }
        if (interrupted()) {
            set_message("the layout was interrupted");
            return 1;
        }
        struct timespec ts = {0, 500000};
        nanosleep(&ts, nullptr);
    }
}

// The iterations of fa2_full_run with an interrupt callback: chunks of
// iterations enqueued one ahead of the one waited on, so that the device
// never idles while the host polls. The engine also asks the callback as
// the reference does, once per 10^6 node and incidence visits enqueued
// (src/gpu_full.c:404-411).
static int run_interruptible(Engine *e, const fa2_full_params *p)
{
    const double per = (double) p->n + (double) (p->inc_off ? p->inc_off[p->n] : 0);
    int chunk = (int) std::min(256.0, std::max(2.0, 1e8 / per));
    chunk += chunk & 1; // whole replays of the captured pair of iterations
    hipEvent_t ev[2] = {nullptr, nullptr};
    for (hipEvent_t &x : ev)
        if (hipEventCreateWithFlags(&x, hipEventDisableTiming) != hipSuccess) {
            set_message("cannot create a HIP event");
            for (hipEvent_t y : ev)
                if (y) (void) hipEventDestroy(y);
            return 1;
        }
    int status = 0, left = p->iterations, k = 0;
    bool prev = false; // ev[k ^ 1] marks the end of a chunk not yet waited for
    while (status == 0) {
        bool cur = false;
        if (left > 0) {
            const int c = std::min(chunk, left);
            left -= c;
            if (e->run(c) != 0) {
                status = 1;
                break;
            }
            (void) hipEventRecord(ev[k], e->stream);
            cur = true;
        }
        if (prev && wait_polling(ev[k ^ 1], p->interrupted)) status = 1;
        if (!cur) break;
        prev = true;
        k ^= 1;
    }
    (void) hipStreamSynchronize(e->stream); // nothing may run when the batch closes
    for (hipEvent_t y : ev) (void) hipEventDestroy(y);
    if (status) return 1;
    return e->sync();
}

// fa2_full_run: upload, iterate, download (src/gpu_full.c:164-532). The
// adaptive speed stays on the device; without an interrupt callback the host
// waits once, at the end. Returns 1 with "the layout was interrupted" when
// the callback returns non-zero, and with "the computation on the GPU
// produced invalid numbers" when the swing or traction totals stop being
// numbers (src/gpu_full.c:486-491).
extern "C" int fa2_full_run(const fa2_full_params *p, double *pos, double *times)
{
    for (int i = 0; i < FA2_FULL_NTIMES; i++) times[i] = 0.0;
    if (!p || p->n < 2 || p->n > (1 << 26)) {
        set_message("the all-GPU layout handles between 2 and 67 million nodes");
        return 1;
    }
    const double *pp = pos;
    fa2_batch *b = fa2_batch_open(1, p, &pp);
    if (!b) return 1;
    const double start = fa2_now(); // the iterations and the download, as the reference
    int status = 1;
    const int ran = p->interrupted ? run_interruptible(b->e, p)
                                   : (b->e->run(p->iterations) || b->e->sync());
    if (ran == 0 && b->e->positions(0, pos) == 0) {
        status = 0;
        double t[FA2_FULL_NTIMES];
        b->e->stage_times(t);
        for (int i = 1; i < FA2_FULL_NTIMES; i++) times[i] = t[i];
        times[FA2_FULL_T_TOTAL] = fa2_now() - start;
    }
    fa2_batch_close(b);
    return status;
}
