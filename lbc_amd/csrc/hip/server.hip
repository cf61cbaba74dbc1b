// server.hip -- continuous-batching request server over one lbmx_batch
// (include/lbmx_batch.h).
//
// One worker thread owns the batch and its stream.  Its loop:
//   1. admit waiting requests in arrival order while the arena can hold them
//      (a request that could never fit, even into an empty batch, fails at once);
//   2. enqueue one chunk of steps for every lattice in the batch (a HIP-graph
//      launch); at most two chunks are in flight, so the host prepares the next
//      chunk while the device runs the current one;
//   3. when a chunk has finished on the device, every lattice whose last step was
//      in it is retired: its probe spectra are computed on a second stream (its
//      state is final and later chunks no longer touch it), the result is
//      published, and its memory is released for the next admissions.
// Latency of a request = t_done - t_submit on one steady clock.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

#include "batch_impl.h"

using namespace lbmx;

namespace {

struct Request {
    lbmx_case c;
    lbmx_tl_spec tl;
    long long tag;
    double t_submit;
    bool has_tl;   // false: a request without a spectrum (a channel solve)
};

struct Running {
    Request r;
    int id;
    double t_admit;
    long long end_epoch;    // batch epoch after the request's last step
    unsigned guard = 0;     // divergence guard bits seen at a chunk boundary
};

struct Inflight {
    hipEvent_t ev;
    long long epoch_after;
    unsigned *status;       // pinned copy of the guard words after this chunk
};

}  // namespace

struct lbmx_server {
    lbmx_batch *b = nullptr;
    int chunk = 64;
    hipStream_t aux = nullptr;
    double *spec_dev = nullptr;
    double *spec_host = nullptr;     // spectra (4 x 64 doubles), then the two totals
    double *tot_dev = nullptr;       // totals partials (2 x kTotMaxBlocks) and result (2)
    std::vector<unsigned *> status_pool;   // worker only: pinned guard-word copies
    std::thread worker;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Request> waiting;       // guarded by mu
    std::deque<lbmx_result> results;   // guarded by mu
    int n_running = 0, n_finished = 0; // guarded by mu
    bool stop = false;                 // guarded by mu
    std::vector<Running> running;      // worker only
    std::vector<hipEvent_t> ev_pool;   // worker only
    std::chrono::steady_clock::time_point t0;

    double now() const
    {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

// Spectra of one finished lattice (aux stream), then TL on the host; a request
// without a spectrum returns only its totals (nfreq = 0).
static int finish(lbmx_server *s, const Running &R, lbmx_result *out)
{
    const BatchSlot *S = lbmx_batch_slot(s->b, R.id);
    if (!S) return lbmx_fail(LBMX_ERR_STATE, "finished lattice vanished");
    if (R.r.has_tl) {
        if (int rc = lbmx_batch_spectrum_of(*S, R.r.tl, s->aux, s->spec_dev, s->spec_host, out))
            return rc;
    } else {
        out->nfreq = 0;
    }
    double *tot = s->tot_dev + 2 * kTotMaxBlocks;
    if (int rc = lbmx_batch_totals_of(s->b, R.id, s->tot_dev, tot, s->aux)) return rc;
    HIPCHK(hipMemcpyAsync(s->spec_host + 256, tot, 2 * sizeof(double), hipMemcpyDeviceToHost,
                          s->aux));
    HIPCHK(hipStreamSynchronize(s->aux));
    out->mass = s->spec_host[256];
    out->energy = s->spec_host[257];
    return LBMX_OK;
}

static void publish(lbmx_server *s, const Request &r, int status, double t_admit,
                    lbmx_result *res)
{
    res->tag = r.tag;
    res->status = status;
    res->steps = r.c.steps;
    res->t_submit = r.t_submit;
    res->t_admit = t_admit;
    res->t_done = s->now();
    std::lock_guard<std::mutex> g(s->mu);
    s->results.push_back(*res);
    s->n_finished++;
}

static void worker_main(lbmx_server *s)
{
    hipSetDevice(0);
    std::deque<Inflight> inflight;
    long long done_epoch = s->b->epoch;
    for (;;) {
        // 1. admissions
        std::vector<Request> admit;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv.wait(lk, [&] {
                return s->stop || !s->waiting.empty() || !s->running.empty() || !inflight.empty();
            });
            if (s->stop && s->waiting.empty() && s->running.empty() && inflight.empty()) break;
            while (!s->waiting.empty()) {
                Request r = s->waiting.front();
                int id = -1;
                const int rc = lbmx_batch_add(s->b, &r.c, &id);
                if (rc == LBMX_ERR_NOMEM && !(s->running.empty() && inflight.empty())) break;
                s->waiting.pop_front();
                if (rc != LBMX_OK) {
                    lk.unlock();
                    lbmx_result res{};
                    publish(s, r, rc, -1.0, &res);
                    lk.lock();
                    continue;
                }
                s->running.push_back(Running{r, id, s->now(), s->b->epoch + r.c.steps});
                s->n_running++;
            }
        }
        // 2. one chunk for everything in the batch, unless every running request
        //    already has all its steps enqueued
        bool need = false;
        for (const Running &R : s->running) need = need || R.end_epoch > s->b->epoch;
        if (need) {
            if (s->ev_pool.empty()) {
                hipEvent_t e;
                hipEventCreateWithFlags(&e, hipEventDisableTiming);
                s->ev_pool.push_back(e);
            }
            hipEvent_t e = s->ev_pool.back();
            s->ev_pool.pop_back();
            unsigned *st = s->status_pool.back();
            s->status_pool.pop_back();
            int rc = lbmx_batch_enqueue(s->b, s->chunk);
            // the guard words after this chunk, read once the chunk is done
            if (rc == LBMX_OK &&
                hipMemcpyAsync(st, s->b->status_dev, sizeof(unsigned) * s->b->slots.size(),
                               hipMemcpyDeviceToHost, s->b->stream) != hipSuccess)
                rc = LBMX_ERR_HIP;
            if (rc == LBMX_OK && hipEventRecord(e, s->b->stream) != hipSuccess) rc = LBMX_ERR_HIP;
            if (rc != LBMX_OK) {
                // fail everything running; the batch is unusable
                for (const Running &R : s->running) {
                    lbmx_result res{};
                    publish(s, R.r, rc, R.t_admit, &res);
                    lbmx_batch_remove(s->b, R.id);
                }
                s->running.clear();
                std::lock_guard<std::mutex> g(s->mu);
                s->n_running = 0;
                s->ev_pool.push_back(e);
                s->status_pool.push_back(st);
                continue;
            }
            inflight.push_back(Inflight{e, s->b->epoch, st});
        }
        // 3. keep at most two chunks in flight; drain when nothing more is queued
        while (!inflight.empty() && (inflight.size() > 1 || !need)) {
            Inflight f = inflight.front();
            inflight.pop_front();
            hipEventSynchronize(f.ev);
            // only chunks queued after a request's admission carry its guard word (its
            // slot may have held a diverged lattice before)
            for (Running &R : s->running)
                if (f.epoch_after > R.end_epoch - R.r.c.steps) R.guard |= f.status[R.id];
            s->ev_pool.push_back(f.ev);
            s->status_pool.push_back(f.status);
            done_epoch = f.epoch_after;
        }
        // 4. retire finished lattices and those whose guard tripped
        for (size_t i = 0; i < s->running.size();) {
            const Running &R = s->running[i];
            if (R.end_epoch > done_epoch && !R.guard) {
                i++;
                continue;
            }
            lbmx_result res{};
            res.guard = R.guard;
            const int rc = R.guard ? LBMX_ERR_DIVERGED : finish(s, R, &res);
            lbmx_batch_remove(s->b, R.id);
            publish(s, R.r, rc, R.t_admit, &res);
            {
                std::lock_guard<std::mutex> g(s->mu);
                s->n_running--;
            }
            s->running.erase(s->running.begin() + i);
        }
    }
    for (hipEvent_t e : s->ev_pool) hipEventDestroy(e);
    s->ev_pool.clear();
}

static int status_pool_init(lbmx_server *s)
{
    // one copy per chunk in flight (at most two) and one being queued
    for (int i = 0; i < 4; i++) {
        unsigned *p = nullptr;
        HIPCHK(hipHostMalloc((void **)&p, sizeof(unsigned) * kBatchMaxSlots, hipHostMallocDefault));
        std::memset(p, 0, sizeof(unsigned) * kBatchMaxSlots);
        s->status_pool.push_back(p);
    }
    return LBMX_OK;
}

extern "C" {

LBMX_API int lbmx_server_create(int precision, size_t budget, int chunk,
                                const lbmx_tuning *tuning, lbmx_server **out)
{
    if (!out) return lbmx_fail(LBMX_ERR_ARG, "null argument");
    *out = nullptr;
    if (chunk < 1 || chunk > 4096) return lbmx_fail(LBMX_ERR_ARG, "chunk must be in [1, 4096]");
    lbmx_server *s = new lbmx_server();
    s->chunk = chunk;
    s->t0 = std::chrono::steady_clock::now();
    int rc = lbmx_batch_create(precision, budget, &s->b);
    if (!rc) rc = lbmx_batch_set_chunk(s->b, chunk);
    if (!rc && !(chunk & 1)) rc = lbmx_batch_set_pairs(s->b, 1);
    if (!rc && tuning) rc = lbmx_batch_set_tuning(s->b, tuning);
    if (!rc && hipStreamCreateWithFlags(&s->aux, hipStreamNonBlocking) != hipSuccess)
        rc = lbmx_fail(LBMX_ERR_HIP, "aux stream");
    if (!rc && hipMalloc((void **)&s->spec_dev, 64 * 4 * 8) != hipSuccess)
        rc = lbmx_fail(LBMX_ERR_NOMEM, "spectrum buffer");
    if (!rc && hipHostMalloc((void **)&s->spec_host, (64 * 4 + 2) * 8, hipHostMallocDefault) !=
                   hipSuccess)
        rc = lbmx_fail(LBMX_ERR_NOMEM, "spectrum staging");
    if (!rc && hipMalloc((void **)&s->tot_dev, (2 * kTotMaxBlocks + 2) * 8) != hipSuccess)
        rc = lbmx_fail(LBMX_ERR_NOMEM, "totals scratch");
    if (!rc) rc = status_pool_init(s);
    if (rc) {
        if (s->b) lbmx_batch_destroy(s->b);
        if (s->aux) hipStreamDestroy(s->aux);
        if (s->spec_dev) hipFree(s->spec_dev);
        if (s->spec_host) hipHostFree(s->spec_host);
        if (s->tot_dev) hipFree(s->tot_dev);
        for (unsigned *p : s->status_pool) hipHostFree(p);
        delete s;
        return rc;
    }
    s->worker = std::thread(worker_main, s);
    *out = s;
    return LBMX_OK;
}

LBMX_API int lbmx_server_submit(lbmx_server *s, const lbmx_case *c, const lbmx_tl_spec *tl,
                                long long tag)
{
    if (!s || !c) return lbmx_fail(LBMX_ERR_ARG, "null argument");
    if (tl)
        if (int rc = lbmx_tl_check(c, tl)) return rc;
    if (c->init == LBMX_INIT_FIELDS)
        return lbmx_fail(LBMX_ERR_ARG, "served requests start from a generated state");
    Request r{*c, tl ? *tl : lbmx_tl_spec{}, tag, s->now(), tl != nullptr};
    r.c.keep_moments = 1;   // the final moments give the result's mass and energy
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (s->stop) return lbmx_fail(LBMX_ERR_STATE, "server is stopping");
        s->waiting.push_back(r);
    }
    s->cv.notify_one();
    return LBMX_OK;
}

LBMX_API int lbmx_server_poll(lbmx_server *s, lbmx_result *out, int max)
{
    if (!s || (!out && max > 0)) return lbmx_fail(LBMX_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> g(s->mu);
    int n = 0;
    while (n < max && !s->results.empty()) {
        out[n++] = s->results.front();
        s->results.pop_front();
    }
    return n;
}

LBMX_API double lbmx_server_clock(lbmx_server *s) { return s ? s->now() : 0.0; }

LBMX_API int lbmx_server_stats(lbmx_server *s, int *waiting, int *running, int *finished,
                               size_t *peak_bytes)
{
    if (!s) return lbmx_fail(LBMX_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> g(s->mu);
    if (waiting) *waiting = (int)s->waiting.size();
    if (running) *running = s->n_running;
    if (finished) *finished = s->n_finished;
    if (peak_bytes) *peak_bytes = s->b->arena.peak;
    return LBMX_OK;
}

LBMX_API void lbmx_server_destroy(lbmx_server *s)
{
    if (!s) return;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
    }
    s->cv.notify_one();
    if (s->worker.joinable()) s->worker.join();
    lbmx_batch_destroy(s->b);
    if (s->aux) hipStreamDestroy(s->aux);
    if (s->spec_dev) hipFree(s->spec_dev);
    if (s->spec_host) hipHostFree(s->spec_host);
    if (s->tot_dev) hipFree(s->tot_dev);
    for (unsigned *p : s->status_pool) hipHostFree(p);
    delete s;
}

}  // extern "C"
