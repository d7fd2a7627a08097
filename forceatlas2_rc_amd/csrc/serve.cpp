// fa2_serve_*: a layout server over the batch engine. No counterpart in the
// reference, whose forceatlas2_large() lays out one graph per call
// (src/gpu_full.c:164-532) and keeps nothing on the device between calls.
//
// Requests are copied in by fa2_serve_submit (prepared incidence lists) or
// fa2_serve_submit_edges (edge lists, prepared on the device) and wait in
// arrival order. Each stream of the server has a host thread of its own
// (a lane) that admits them by device-memory footprint: when the lane is
// free, the request at the head of the queue and every later one of the
// same precision and kind that fits in what the running batches leave of
// the budget join one new batch (engine.hip) on the lane's stream.
// Batches on different streams run at the same time, so requests that
// arrive while others are on the GPU start as soon as a lane and memory
// are free, and their work overlaps on the device; a lane busy on the host
// (building a batch, reading results) holds up only its own stream. Each
// batch advances in chunks of iterations; after each chunk its lane reads
// how far every graph has come, fetches the positions of the graphs that
// reached their iteration count, and retires them (their repulsion and
// attraction are skipped from then on). A batch's memory, allocated in its
// stream's order, is released when its last request is done.
//
// Continuous batching: when no lane can take the waiting requests (every
// lane busy, or the memory taken), a batch that has finished some of its
// requests takes waiting ones in at its
// next chunk boundary. Its remaining requests' state (positions, previous
// force, speed and speed efficiency, iterations done) is read back, the
// batch is rebuilt on its stream with them and the newcomers that fit in
// the memory budget, and each continuing graph resumes from its state.
//
// A request's result does not depend on the other requests: every kernel of
// the engine computes a graph from that graph's data alone, in a fixed
// order, and an iteration depends only on that state, so a layout is
// bit-identical alone, inside any batch, and across a rebuild.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "common.h"
#include "engine.h"
#include "../../include/fa2.h"
#include "../../include/fa2_hip.h"

using namespace fa2;

namespace {

enum { kQueued = 0, kRunning = 1, kDone = 2, kFailed = 3 };

struct Job {
    int id = 0;
    fa2_full_params p{};
    std::vector<int> inc_off, inc_node, fixed;
    std::vector<double> inc_coef, mass, size, pos;
    // an edge list instead of incidence lists: prepared on the device
    bool has_edges = false;
    fa2_edges edges{};
    std::vector<int> from, to;
    std::vector<double> w;
    double t_submit = 0, t_admit = 0, t_done = 0;
    size_t bytes = 0;
    int state = kQueued;
    std::string error;
    // iterations applied in earlier batches, and the state they left
    // (positions in pos; previous force; speed and efficiency)
    int done_before = 0;
    std::vector<double> fold;
    double speed = 1.0, eff = 1.0;
};

struct Running {
    Engine *e = nullptr;
    std::vector<Job *> jobs;
    std::vector<char> finished;
    hipStream_t st = nullptr;
    hipEvent_t ev = nullptr;
    size_t bytes = 0;
    int target = 0; // iterations enqueued so far
    int longest = 0;
};

double now_s() { return fa2_now(); }

} // namespace

struct fa2_server {
    int device = 0, chunk = 25;
    size_t budget = 0, used = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Job *> queue;
    std::vector<std::unique_ptr<Job>> jobs;
    std::vector<Running *> running;
    std::vector<hipStream_t> streams;
    std::vector<std::thread> lanes;
    int idle = 0; // lanes without a batch
    bool stop = false;
    // statistics
    long long batches = 0, admitted = 0, refills = 0;
    long long uncaptured = 0; // batches whose iterations ran launch by launch
    double refill_s = 0; // lane time spent rebuilding batches
    int peak_running = 0;

    void lane(hipStream_t st);
    bool can_admit() const;
    Running *admit(std::unique_lock<std::mutex> &lk, hipStream_t st, hipEvent_t ev);
    bool advance(Running *r);
    void finish_batch(Running *r);
    std::string open_engine(Running *r, const std::vector<Job *> &jobs, int use_double,
                            bool has_edges);
    bool refill(Running *r, const std::vector<int> &done);
};

// A new engine for `jobs` on r's stream: each job continues from its
// iterations done and saved state. An empty string on success.
std::string fa2_server::open_engine(Running *r, const std::vector<Job *> &jobs, int use_double,
                                    bool has_edges)
{
    const size_t G = jobs.size();
    std::vector<fa2_full_params> ps(G);
    std::vector<EngineGraph> gr(G);
    r->longest = 0;
    for (size_t k = 0; k < G; k++) {
        ps[k] = jobs[k]->p;
        ps[k].iterations = jobs[k]->p.iterations - jobs[k]->done_before;
        r->longest = std::max(r->longest, ps[k].iterations);
        gr[k] = EngineGraph{&ps[k], jobs[k]->pos.data(), has_edges ? &jobs[k]->edges : nullptr};
        if (jobs[k]->done_before > 0) { // resumes from its saved state
            gr[k].fold = jobs[k]->fold.data();
            gr[k].speed = jobs[k]->speed;
            gr[k].eff = jobs[k]->eff;
        }
    }
    r->e = make_engine(use_double);
    if (!r->e) return "out of memory";
    r->e->device = device;
    r->e->stream = r->st;
    r->e->async_alloc = 1;
    // batches already run side by side: one stream each (a second one per
    // batch would share the hardware queues between batches), and the
    // default launch shapes (the autotuner synchronises the device and
    // would hold the lane between chunks)
    r->e->overlap = 0;
    if (r->e->open(gr.data(), (int) G) != 0) return fa2_gpu_error();
    r->jobs = jobs;
    r->finished.assign(G, 0);
    r->target = 0;
    return std::string();
}

// At a chunk boundary of r, when no lane can admit the waiting requests
// and some of r's precision and kind wait: those that fit in the budget, with what r's
// finished requests give back, join r's remaining ones in a rebuilt batch
// on r's stream. A rebuild reads back and uploads again the remaining
// requests, so it is made only when some of them finished or when the
// newcomers hold at least a quarter as many nodes as the remaining ones.
// False when nothing joined (r goes on as it is); on a failure its
// requests fail.
bool fa2_server::refill(Running *r, const std::vector<int> &done)
{
    const int use_double = r->jobs[0]->p.use_double;
    const bool has_edges = r->jobs[0]->has_edges;
    std::vector<Job *> keep;
    std::vector<int> keep_k;
    size_t keep_bytes = 0;
    long long keep_nodes = 0;
    for (size_t k = 0; k < r->jobs.size(); k++)
        if (!r->finished[k]) {
            keep.push_back(r->jobs[k]);
            keep_k.push_back((int) k);
            keep_bytes += r->jobs[k]->bytes;
            keep_nodes += r->jobs[k]->p.n;
        }
    const bool slot_free = keep.size() < r->jobs.size();
    std::vector<Job *> take;
    size_t take_bytes = 0;
    {
        std::lock_guard<std::mutex> lk(mu);
        if (queue.empty() || can_admit()) return false;
        size_t avail = budget - (used - r->bytes) - keep_bytes;
        long long take_nodes = 0;
        std::deque<Job *> rest;
        for (Job *j : queue) {
            if (j->p.use_double != use_double || j->has_edges != has_edges || j->bytes > avail) {
                rest.push_back(j);
                continue;
            }
            avail -= j->bytes;
            take_nodes += j->p.n;
            take_bytes += j->bytes;
            take.push_back(j);
        }
        if (take.empty() || (!slot_free && 4 * take_nodes < keep_nodes)) return false;
        queue.swap(rest);
        used = used - r->bytes + keep_bytes + take_bytes;
        const double t = now_s();
        for (Job *j : take) {
            j->t_admit = t;
            j->state = kRunning;
        }
        admitted += (long long) take.size();
        refills++;
    }
    r->bytes = keep_bytes + take_bytes;
    const double t_start = now_s();
    // the remaining requests' state, in one transfer per array
    std::string err;
    const size_t K = keep.size();
    std::vector<double *> kpos(K), kfold(K);
    std::vector<double[4]> ksp(K);
    for (size_t q = 0; q < K; q++) {
        keep[q]->fold.resize(2 * (size_t) keep[q]->p.n);
        kpos[q] = keep[q]->pos.data();
        kfold[q] = keep[q]->fold.data();
    }
    if (r->e->read_states((int) K, keep_k.data(), kpos.data(), kfold.data(), ksp.data()) != 0)
        err = fa2_gpu_error();
    for (size_t q = 0; q < K && err.empty(); q++) {
        Job *j = keep[q];
        j->speed = ksp[q][2];
        j->eff = ksp[q][3];
        j->done_before += done[keep_k[q]];
    }
    delete r->e; // waits for its stream, frees in the stream's order
    r->e = nullptr;
    std::vector<Job *> all = keep;
    all.insert(all.end(), take.begin(), take.end());
    if (err.empty()) err = open_engine(r, all, use_double, has_edges);
    if (err.empty()) {
        r->target = std::min(chunk, r->longest);
        if (r->e->run(r->target) != 0 || hipEventRecord(r->ev, r->st) != hipSuccess)
            err = fa2_gpu_error();
    }
    std::lock_guard<std::mutex> lk(mu);
    refill_s += now_s() - t_start;
    if (err.empty() && !r->e->use_graph) uncaptured++;
    if (!err.empty()) {
        for (Job *j : all) {
            j->state = kFailed;
            j->error = err;
            j->t_done = now_s();
        }
        r->jobs = all;
        r->finished.assign(all.size(), 1);
        cv_done.notify_all();
    }
    return true;
}

// Whether an idle lane may start a batch now: a request waits and the one
// at the head fits in what the running batches leave of the budget (or can
// never fit, and is to be failed). Called with the lock held.
bool fa2_server::can_admit() const
{
    if (queue.empty() || idle == 0) return false;
    const size_t b = queue.front()->bytes;
    return b > budget || used + b <= budget;
}

// One new batch on stream st from the requests at the head of the queue
// that fit; nullptr when none could start. Called with the lock held, which
// is released while the batch is built and held again on return.
Running *fa2_server::admit(std::unique_lock<std::mutex> &lk, hipStream_t st, hipEvent_t ev)
{
    // a request that alone exceeds the budget can never run
    while (!queue.empty() && queue.front()->bytes > budget) {
        Job *j = queue.front();
        queue.pop_front();
        j->state = kFailed;
        j->error = "the layout needs more device memory than the server's budget";
        j->t_done = now_s();
        cv_done.notify_all();
    }
    if (queue.empty() || used + queue.front()->bytes > budget) return nullptr;
    const int use_double = queue.front()->p.use_double;
    const bool has_edges = queue.front()->has_edges;
    std::vector<Job *> take;
    size_t bytes = 0;
    for (auto it = queue.begin(); it != queue.end();) {
        Job *j = *it;
        if (j->p.use_double != use_double || j->has_edges != has_edges ||
            used + bytes + j->bytes > budget) {
            ++it;
            continue;
        }
        bytes += j->bytes;
        take.push_back(j);
        it = queue.erase(it);
    }
    used += bytes;
    const double t = now_s();
    for (Job *j : take) {
        j->t_admit = t;
        j->state = kRunning;
    }
    lk.unlock();

    // build the batch without holding the lock: the upload takes a while
    Running *r = new Running();
    r->st = st;
    r->ev = ev;
    r->bytes = bytes;
    std::string err = open_engine(r, take, use_double, has_edges);
    if (err.empty()) {
        r->target = std::min(chunk, r->longest);
        if (r->e->run(r->target) != 0 || hipEventRecord(r->ev, st) != hipSuccess)
            err = fa2_gpu_error();
    }

    lk.lock();
    if (!err.empty()) {
        for (Job *j : take) {
            j->state = kFailed;
            j->error = err;
            j->t_done = now_s();
        }
        lk.unlock();
        delete r->e; // waits for its stream, frees in the stream's order
        lk.lock();
        delete r;
        used -= bytes;
        cv_done.notify_all();
        return nullptr;
    }
    running.push_back(r);
    batches++;
    if (!r->e->use_graph) uncaptured++;
    admitted += (long long) take.size();
    peak_running = std::max(peak_running, (int) running.size());
    return r;
}

// After a completed chunk: hand out the finished requests, enqueue the next
// chunk. True when the batch is over.
bool fa2_server::advance(Running *r)
{
    const int G = (int) r->jobs.size();
    std::vector<int> done(G), status(G);
    std::string err;
    if (r->e->progress(done.data(), status.data()) != 0) err = fa2_gpu_error();
    // the positions of the graphs that reached their iteration count, in one
    // transfer
    std::vector<int> fin;
    std::vector<std::vector<double>> fpos;
    std::vector<double *> fptr;
    int frc = 0;
    if (err.empty()) {
        for (int k = 0; k < G; k++)
            if (!r->finished[k] && status[k] == 0 &&
                r->jobs[k]->done_before + done[k] >= r->jobs[k]->p.iterations)
                fin.push_back(k);
        fpos.resize(fin.size());
        for (size_t q = 0; q < fin.size(); q++) {
            fpos[q].resize(2 * (size_t) r->jobs[fin[q]]->p.n);
            fptr.push_back(fpos[q].data());
        }
        if (!fin.empty())
            frc = r->e->read_states((int) fin.size(), fin.data(), fptr.data(), nullptr, nullptr);
    }
    std::string ferr = frc ? fa2_gpu_error() : std::string();
    size_t fq = 0;
    bool all = true;
    for (int k = 0; k < G; k++) {
        if (r->finished[k]) continue;
        Job *j = r->jobs[k];
        if (!err.empty()) {
            std::lock_guard<std::mutex> lk(mu);
            j->state = kFailed;
            j->error = err;
            j->t_done = now_s();
            r->finished[k] = 1;
            continue;
        }
        if (status[k] != 0) { // invalid numbers: the graph stopped
            (void) r->e->retire(k);
            std::lock_guard<std::mutex> lk(mu);
            j->state = kFailed;
            j->error = "the computation on the GPU produced invalid numbers";
            j->t_done = now_s();
            r->finished[k] = 1;
            continue;
        }
        if (j->done_before + done[k] < j->p.iterations) {
            all = false;
            continue;
        }
        std::vector<double> &out = fpos[fq++];
        if (frc == 0) (void) r->e->retire(k);
        std::lock_guard<std::mutex> lk(mu);
        if (frc == 0) {
            j->pos.swap(out);
            j->state = kDone;
        } else {
            j->state = kFailed;
            j->error = ferr;
        }
        j->t_done = now_s();
        r->finished[k] = 1;
    }
    cv_done.notify_all();
    if (all || !err.empty()) return true;
    if (refill(r, done)) {
        for (char f : r->finished)
            if (!f) return false;
        return true; // the rebuild failed
    }
    const int step = std::min(chunk, r->longest - r->target);
    r->target += step;
    if (r->e->run(step) != 0 || hipEventRecord(r->ev, r->st) != hipSuccess) {
        std::lock_guard<std::mutex> lk(mu);
        for (int k = 0; k < G; k++)
            if (!r->finished[k]) {
                r->jobs[k]->state = kFailed;
                r->jobs[k]->error = fa2_gpu_error();
                r->jobs[k]->t_done = now_s();
                r->finished[k] = 1;
            }
        cv_done.notify_all();
        return true;
    }
    return false;
}

void fa2_server::finish_batch(Running *r)
{
    delete r->e; // waits for its stream, frees in the stream's order
    std::lock_guard<std::mutex> lk(mu);
    used -= r->bytes;
    running.erase(std::find(running.begin(), running.end(), r));
    delete r;
}

// The host thread of stream st: start a batch when requests wait and fit,
// run it chunk by chunk to its end, and again, until the server closes and
// the queue is empty.
void fa2_server::lane(hipStream_t st)
{
    (void) hipSetDevice(device);
    // this thread's calls leave the other lanes' graph captures alone
    hipStreamCaptureMode mode = hipStreamCaptureModeRelaxed;
    (void) hipThreadExchangeStreamCaptureMode(&mode);
    hipEvent_t ev = nullptr;
    // a blocking event: the lane sleeps while its chunk runs
    const bool have_ev =
        hipEventCreateWithFlags(&ev, hipEventDisableTiming | hipEventBlockingSync) == hipSuccess;
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        cv_work.wait(lk, [&] { return (stop && queue.empty()) || can_admit(); });
        if (stop && queue.empty()) break;
        if (!have_ev) { // nothing can run on this lane
            Job *j = queue.front();
            queue.pop_front();
            j->state = kFailed;
            j->error = "could not create a HIP event";
            j->t_done = now_s();
            cv_done.notify_all();
            continue;
        }
        idle--;
        Running *r = admit(lk, st, ev);
        if (r) {
            lk.unlock();
            bool over = false;
            while (!over) {
                (void) hipEventSynchronize(r->ev); // errors surface in progress()
                over = advance(r);
            }
            finish_batch(r);
            lk.lock();
        }
        idle++;
        // memory came back: another lane may start the next batch
        cv_work.notify_all();
    }
    lk.unlock();
    if (have_ev) (void) hipEventDestroy(ev);
}

extern "C" fa2_server *fa2_serve_open(int device, double budget_bytes, int max_batches,
                                      int chunk)
{
    if (device < 0 || device >= device_count()) {
        set_message("there is no GPU device %d", device + 1);
        return nullptr;
    }
    if (max_batches < 1 || max_batches > 16 || chunk < 1) {
        set_message("max_batches must be in 1..16 and chunk at least 1");
        return nullptr;
    }
    if (hipSetDevice(device) != hipSuccess) {
        set_message("could not select the GPU device");
        return nullptr;
    }
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
        set_message("could not read the device memory size");
        return nullptr;
    }
    fa2_server *s = new (std::nothrow) fa2_server();
    if (!s) {
        set_message("out of memory");
        return nullptr;
    }
    s->device = device;
    s->chunk = chunk;
    // by default 80% of the memory free now
    s->budget = budget_bytes > 0 ? (size_t) budget_bytes : (size_t) (0.8 * (double) free_b);
    for (int k = 0; k < max_batches; k++) {
        hipStream_t st = nullptr;
        // default priority on purpose: streams of different priorities are
        // never mixed in this library
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
            for (hipStream_t q : s->streams) (void) hipStreamDestroy(q);
            delete s;
            set_message("could not create a HIP stream");
            return nullptr;
        }
        s->streams.push_back(st);
    }
    s->idle = max_batches;
    for (hipStream_t st : s->streams) s->lanes.emplace_back([s, st] { s->lane(st); });
    return s;
}

static int submit(fa2_server *s, const fa2_full_params *p, const fa2_edges *e, const double *pos)
{
    if (!s || !p || !pos) {
        set_message("no request");
        return -1;
    }
    if (const char *why = graph_args_error(p, e)) {
        set_message("%s", why);
        return -1;
    }
    if (p->iterations < 1) {
        set_message("a request needs at least one iteration");
        return -1;
    }
    auto j = std::make_unique<Job>();
    const int n = p->n;
    long long I = 0;
    if (e) {
        j->has_edges = true;
        j->from.assign(e->from, e->from + e->m);
        j->to.assign(e->to, e->to + e->m);
        if (e->w) j->w.assign(e->w, e->w + e->m);
        j->edges = fa2_edges{e->m, j->from.data(), j->to.data(), e->w ? j->w.data() : nullptr,
                             e->outbound};
        I = 2LL * e->m;
    } else {
        I = p->inc_off[n];
        j->inc_off.assign(p->inc_off, p->inc_off + n + 1);
        j->inc_node.assign(p->inc_node, p->inc_node + I);
        j->inc_coef.assign(p->inc_coef, p->inc_coef + I);
        j->mass.assign(p->mass, p->mass + n);
    }
    if (p->adjust && p->size) j->size.assign(p->size, p->size + n);
    if (p->any_fixed && p->fixed) j->fixed.assign(p->fixed, p->fixed + n);
    j->pos.assign(pos, pos + 2 * (size_t) n);
    j->p = *p;
    j->p.inc_off = e ? nullptr : j->inc_off.data();
    j->p.inc_node = e ? nullptr : j->inc_node.data();
    j->p.inc_coef = e ? nullptr : j->inc_coef.data();
    j->p.mass = e ? nullptr : j->mass.data();
    j->p.size = j->size.empty() ? nullptr : j->size.data();
    j->p.fixed = j->fixed.empty() ? nullptr : j->fixed.data();
    j->p.device = s->device;
    j->p.use_double = p->use_double ? 1 : 0;
    j->p.profile = 0;
    j->p.interrupted = nullptr;
    // with an edge list, the upper bound of the incidences, plus the
    // preprocessing's own buffers (endpoints, weights, two key/value pairs
    // per edge, twice for the sort)
    j->bytes = engine_footprint(j->p.use_double, n, I) + (e ? (size_t) e->m * 48 : 0);
    std::lock_guard<std::mutex> lk(s->mu);
    j->id = (int) s->jobs.size();
    j->t_submit = now_s();
    s->queue.push_back(j.get());
    s->jobs.push_back(std::move(j));
    s->cv_work.notify_all();
    return s->jobs.back()->id;
}

extern "C" int fa2_serve_submit(fa2_server *s, const fa2_full_params *p, const double *pos)
{
    return submit(s, p, nullptr, pos);
}

extern "C" int fa2_serve_submit_edges(fa2_server *s, const fa2_full_params *p,
                                      const fa2_edges *edges, const double *pos)
{
    if (!edges) {
        set_message("no edge list");
        return -1;
    }
    return submit(s, p, edges, pos);
}

extern "C" int fa2_serve_status(fa2_server *s, int job)
{
    if (!s) return -1;
    std::lock_guard<std::mutex> lk(s->mu);
    if (job < 0 || job >= (int) s->jobs.size()) return -1;
    return s->jobs[job]->state;
}

extern "C" int fa2_serve_wait(fa2_server *s, int job, double *pos, double times[3])
{
    if (!s) {
        set_message("no server");
        return 1;
    }
    std::unique_lock<std::mutex> lk(s->mu);
    if (job < 0 || job >= (int) s->jobs.size()) {
        set_message("no request %d on this server", job);
        return 1;
    }
    Job *j = s->jobs[job].get();
    s->cv_done.wait(lk, [&] { return j->state == kDone || j->state == kFailed; });
    if (times) {
        times[0] = j->t_admit > 0 ? j->t_admit - j->t_submit : 0.0;
        times[1] = j->t_admit > 0 ? j->t_done - j->t_admit : 0.0;
        times[2] = j->t_done - j->t_submit;
    }
    if (j->state == kFailed) {
        set_message("%s", j->error.c_str());
        return 1;
    }
    if (pos) std::copy(j->pos.begin(), j->pos.end(), pos);
    return 0;
}

extern "C" int fa2_serve_stats(fa2_server *s, double out[7])
{
    if (!s) return 1;
    std::lock_guard<std::mutex> lk(s->mu);
    out[0] = (double) s->batches;
    out[1] = (double) s->admitted;
    out[2] = (double) s->peak_running;
    out[3] = (double) s->budget;
    out[4] = (double) s->refills;
    out[5] = s->refill_s;
    out[6] = (double) s->uncaptured;
    return 0;
}

extern "C" void fa2_serve_close(fa2_server *s)
{
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->cv_work.notify_all();
    }
    for (std::thread &t : s->lanes)
        if (t.joinable()) t.join();
    for (hipStream_t st : s->streams) (void) hipStreamDestroy(st);
    delete s;
}

extern "C" double fa2_serve_footprint(int use_double, int n, long long incidences)
{
    return (double) engine_footprint(use_double, n, incidences);
}
