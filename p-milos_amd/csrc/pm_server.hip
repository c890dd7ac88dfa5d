// pm_server.hip -- the serving loop of the C ABI (pm_server_*, pm_submit,
// pm_poll, pm_wait): scans of any size and wavelength sampling arrive at
// any time and share the GPU.
//
// Host side only.  Two threads per server:
//   stager      admits queued scans while the device bytes of the admitted
//               ones stay within the HBM budget (first queued scan that
//               fits), uploads their Stokes cube (and stray light), and
//               downloads the results of finished scans, freeing their
//               budget.  Downloads go first: they release memory.
//   dispatcher  splits every admitted scan into pixel chunks and keeps each
//               compute stream fed with at most DEPTH chunks, always taking
//               the next chunk from the admitted scan with the least work
//               left (shortest remaining work first), so a small scan that
//               arrives behind a large one is not queued behind all of it.
//               Chunks of different scans run at the same time on the
//               device: a persistent kernel's workgroups retire as its work
//               queue drains and the next chunk's take their place.
// Each compute stream has its own work-queue counters.  A pixel's result
// does not depend on which chunk, stream or workgroup computes it, so a scan
// gives the same bits alone and in the mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "pm_internal.h"
#include "../../include/pmilos.h"

namespace {

using Clock = std::chrono::steady_clock;

constexpr int DEPTH = 2;                 // chunks queued per compute stream
constexpr size_t ALIGN = 256;

size_t up(size_t n, size_t a) { return (n + a - 1) / a * a; }

// Device blocks are kept for reuse until the server ends: freeing device
// memory waits for the whole device, which would hold up every other scan.
class DevPool {
    struct Block { char *p; size_t n; bool used; };
    std::vector<Block> blocks_;

public:
    char *get(size_t n)
    {
        n = up(n, size_t(2) << 20);
        Block *best = nullptr;
        for (auto &b : blocks_)
            if (!b.used && b.n >= n && b.n <= 2 * n && (!best || b.n < best->n)) best = &b;
        if (best) { best->used = true; return best->p; }
        char *p = nullptr;
        if (hipMalloc(&p, n) != hipSuccess) {
            // give back idle blocks and try once more
            for (auto it = blocks_.begin(); it != blocks_.end();) {
                if (!it->used) { (void)hipFree(it->p); it = blocks_.erase(it); }
                else ++it;
            }
            if (hipMalloc(&p, n) != hipSuccess) return nullptr;
        }
        blocks_.push_back({p, n, true});
        return p;
    }
    void put(char *p)
    {
        for (auto &b : blocks_)
            if (b.p == p) b.used = false;
    }
    void release()
    {
        for (auto &b : blocks_) (void)hipFree(b.p);
        blocks_.clear();
    }
};

enum State { QUEUED, READY, FINISHED, DONE };

struct Req {
    long long id = 0;
    pm_ctx *ctx = nullptr;
    const float *stokes = nullptr, *slight = nullptr;
    double *models = nullptr;
    float *chi2 = nullptr, *fit = nullptr;
    int *iters = nullptr, *accepted = nullptr;
    int npix = 0, L = 0, chunk = 0;
    size_t bytes = 0;
    double cost = 1;                     // relative work per pixel
    // device buffers, carved from one pool block
    char *blk = nullptr;
    float *d_stokes = nullptr, *d_slight = nullptr, *d_chi2 = nullptr, *d_fit = nullptr;
    double *d_models = nullptr;
    int *d_iters = nullptr, *d_acc = nullptr;
    int next_px = 0, done_px = 0;
    int rc = PM_OK;
    State state = QUEUED;
    Clock::time_point t_submit, t_done;
};

struct Inflight {
    Req *r;
    int px;
    hipEvent_t ev;
};

struct Slot {
    hipStream_t s = nullptr;
    int *counters = nullptr;
    unsigned long long *fallbacks = nullptr;
    std::deque<Inflight> q;
    std::vector<hipEvent_t> spare;
};

size_t req_bytes(int npix, int L, bool slight, bool fit)
{
    size_t n = npix, cube = n * 4 * L * sizeof(float);
    return up(cube, ALIGN) + (slight ? up(cube, ALIGN) : 0) + up(n * 11 * sizeof(double), ALIGN) +
           3 * up(n * sizeof(int), ALIGN) + (fit ? up(cube, ALIGN) : 0);
}

}  // namespace

struct pm_server {
    int device = 0;
    size_t budget = 0;
    int chunk_px = 0;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Req *> queued;
    std::vector<Req *> ready;
    std::deque<Req *> finished;
    std::map<long long, Req *> all;
    size_t in_use = 0;
    bool stop = false;
    long long next_id = 1;
    hipStream_t xfer = nullptr;
    std::vector<Slot> slots;
    DevPool pool;
    std::thread stager, dispatcher;

    void stager_loop();
    void dispatcher_loop();
    bool admit_one(std::unique_lock<std::mutex> &lk);
    void finish_one(Req *r);
};

// Admits the first queued scan that fits the budget (any scan when nothing
// is admitted): allocates and uploads it.  Called with the lock held;
// returns false when nothing could be admitted.
bool pm_server::admit_one(std::unique_lock<std::mutex> &lk)
{
    auto it = std::find_if(queued.begin(), queued.end(), [&](Req *r) {
        return in_use + r->bytes <= budget || in_use == 0;
    });
    if (it == queued.end()) return false;
    Req *r = *it;
    queued.erase(it);
    in_use += r->bytes;
    lk.unlock();
    int rc = PM_OK;
    char *p = pool.get(r->bytes);
    if (!p) rc = PM_ENOMEM;
    if (rc == PM_OK) {
        size_t n = r->npix, cube = n * 4 * r->L * sizeof(float);
        r->blk = p;
        r->d_stokes = (float *)p; p += up(cube, ALIGN);
        if (r->slight) { r->d_slight = (float *)p; p += up(cube, ALIGN); }
        r->d_models = (double *)p; p += up(n * 11 * sizeof(double), ALIGN);
        r->d_chi2 = (float *)p; p += up(n * sizeof(float), ALIGN);
        r->d_iters = (int *)p; p += up(n * sizeof(int), ALIGN);
        r->d_acc = (int *)p; p += up(n * sizeof(int), ALIGN);
        if (r->fit) r->d_fit = (float *)p;
        if (hipMemcpyAsync(r->d_stokes, r->stokes, cube, hipMemcpyHostToDevice, xfer) != hipSuccess ||
            (r->slight && hipMemcpyAsync(r->d_slight, r->slight, cube, hipMemcpyHostToDevice,
                                         xfer) != hipSuccess) ||
            hipStreamSynchronize(xfer) != hipSuccess)
            rc = PM_EDEVICE;
    }
    lk.lock();
    if (rc != PM_OK) {
        r->rc = rc;
        finished.push_back(r);           // completes with the error
    } else {
        r->state = READY;
        ready.push_back(r);
    }
    cv.notify_all();
    return true;
}

// Downloads a finished scan's results and releases its budget.
void pm_server::finish_one(Req *r)
{
    if (r->rc == PM_OK && r->blk) {
        size_t n = r->npix, cube = n * 4 * r->L * sizeof(float);
        bool ok = hipMemcpyAsync(r->models, r->d_models, n * 11 * sizeof(double),
                                 hipMemcpyDeviceToHost, xfer) == hipSuccess &&
                  hipMemcpyAsync(r->chi2, r->d_chi2, n * sizeof(float), hipMemcpyDeviceToHost,
                                 xfer) == hipSuccess &&
                  hipMemcpyAsync(r->iters, r->d_iters, n * sizeof(int), hipMemcpyDeviceToHost,
                                 xfer) == hipSuccess &&
                  (!r->accepted || hipMemcpyAsync(r->accepted, r->d_acc, n * sizeof(int),
                                                  hipMemcpyDeviceToHost, xfer) == hipSuccess) &&
                  (!r->fit || hipMemcpyAsync(r->fit, r->d_fit, cube, hipMemcpyDeviceToHost,
                                             xfer) == hipSuccess) &&
                  hipStreamSynchronize(xfer) == hipSuccess;
        if (!ok) r->rc = PM_EDEVICE;
    }
    std::lock_guard<std::mutex> g(mu);
    if (r->blk) pool.put(r->blk);
    r->blk = nullptr;
    in_use -= r->bytes;
    r->t_done = Clock::now();
    r->state = DONE;
    cv.notify_all();
}

void pm_server::stager_loop()
{
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    while (true) {
        if (!finished.empty()) {
            Req *r = finished.front();
            finished.pop_front();
            lk.unlock();
            finish_one(r);
            lk.lock();
            continue;
        }
        if (!queued.empty() && admit_one(lk)) continue;
        if (stop && queued.empty() && ready.empty() && finished.empty()) break;
        cv.wait(lk);
    }
}

void pm_server::dispatcher_loop()
{
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    while (true) {
        bool busy = false;
        // reap completed chunks
        for (auto &sl : slots) {
            while (!sl.q.empty()) {
                hipError_t e = hipEventQuery(sl.q.front().ev);
                if (e == hipErrorNotReady) break;
                Inflight f = sl.q.front();
                sl.q.pop_front();
                sl.spare.push_back(f.ev);
                if (e != hipSuccess) f.r->rc = PM_EDEVICE;
                f.r->done_px += f.px;
                if (f.r->done_px == f.r->npix) {
                    f.r->state = FINISHED;
                    ready.erase(std::find(ready.begin(), ready.end(), f.r));
                    finished.push_back(f.r);
                    cv.notify_all();
                }
            }
            busy |= !sl.q.empty();
        }
        // feed the streams, least remaining work first
        for (auto &sl : slots) {
            while ((int)sl.q.size() < DEPTH) {
                Req *best = nullptr;
                double best_w = 0;
                for (Req *r : ready) {
                    if (r->next_px >= r->npix) continue;
                    double w = (r->npix - r->next_px) * r->cost;
                    if (!best || w < best_w) { best = r; best_w = w; }
                }
                if (!best) break;
                Req *r = best;
                const int p0 = r->next_px, px = std::min(r->chunk, r->npix - p0);
                const size_t o4 = (size_t)p0 * 4 * r->L;
                int rc = pm_invert_launch(r->ctx, r->d_stokes + o4, r->d_slight ? r->d_slight + o4 : nullptr,
                                          px, r->d_models + (size_t)p0 * 11, r->d_chi2 + p0,
                                          r->d_iters + p0, r->d_acc + p0,
                                          r->d_fit ? r->d_fit + o4 : nullptr, sl.counters,
                                          sl.fallbacks, sl.s);
                hipEvent_t ev = nullptr;
                if (!sl.spare.empty()) { ev = sl.spare.back(); sl.spare.pop_back(); }
                else if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) ev = nullptr;
                if (rc != PM_OK || !ev || hipEventRecord(ev, sl.s) != hipSuccess) {
                    if (rc == PM_OK) rc = PM_EDEVICE;
                    r->rc = rc;
                    if (ev) sl.spare.push_back(ev);
                    // the chunk is counted as done so the scan completes
                    r->next_px += px;
                    r->done_px += px;
                    if (r->done_px == r->npix) {
                        r->state = FINISHED;
                        ready.erase(std::find(ready.begin(), ready.end(), r));
                        finished.push_back(r);
                        cv.notify_all();
                    }
                    continue;
                }
                r->next_px += px;
                sl.q.push_back({r, px, ev});
                busy = true;
            }
        }
        if (stop && queued.empty() && ready.empty() && !busy) break;
        if (busy) {
            lk.unlock();
            std::this_thread::sleep_for(std::chrono::microseconds(50));
            lk.lock();
        } else {
            cv.wait(lk);
        }
    }
}

extern "C" int pm_server_create(const pm_server_opts *opts, pm_server **out)
{
    if (!out) return PM_EINVAL;
    *out = nullptr;
    pm_server_opts o = {};
    if (opts) o = *opts;
    if (o.streams <= 0) o.streams = 3;
    if (o.streams > 8) return PM_EINVAL;
    if (o.chunk_px <= 0) o.chunk_px = 65536;
    pm_server *s = new pm_server();
    if (hipGetDevice(&s->device) != hipSuccess) { delete s; return PM_EDEVICE; }
    if (o.hbm_budget == 0) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) { delete s; return PM_EDEVICE; }
        o.hbm_budget = fr / 2;
    }
    s->budget = o.hbm_budget;
    s->chunk_px = o.chunk_px;
    int rc = PM_OK;
    if (hipStreamCreateWithFlags(&s->xfer, hipStreamNonBlocking) != hipSuccess) rc = PM_EDEVICE;
    s->slots.resize(o.streams);
    for (auto &sl : s->slots) {
        if (rc != PM_OK) break;
        if (hipStreamCreateWithFlags(&sl.s, hipStreamNonBlocking) != hipSuccess ||
            hipMalloc(&sl.counters, pm_counters_bytes()) != hipSuccess ||
            hipMalloc(&sl.fallbacks, sizeof(unsigned long long)) != hipSuccess)
            rc = PM_EDEVICE;
    }
    if (rc != PM_OK) {
        pm_server_destroy(s);
        return rc;
    }
    s->stager = std::thread(&pm_server::stager_loop, s);
    s->dispatcher = std::thread(&pm_server::dispatcher_loop, s);
    *out = s;
    return PM_OK;
}

extern "C" void pm_server_destroy(pm_server *s)
{
    if (!s) return;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
        s->cv.notify_all();
    }
    if (s->stager.joinable()) s->stager.join();
    if (s->dispatcher.joinable()) s->dispatcher.join();
    (void)hipSetDevice(s->device);
    for (auto &sl : s->slots) {
        if (sl.s) (void)hipStreamSynchronize(sl.s);
        for (auto &f : sl.q) (void)hipEventDestroy(f.ev);
        for (auto ev : sl.spare) (void)hipEventDestroy(ev);
        if (sl.s) (void)hipStreamDestroy(sl.s);
        if (sl.counters) (void)hipFree(sl.counters);
        if (sl.fallbacks) (void)hipFree(sl.fallbacks);
    }
    if (s->xfer) (void)hipStreamDestroy(s->xfer);
    s->pool.release();
    for (auto &kv : s->all) delete kv.second;
    delete s;
}

extern "C" int pm_submit(pm_server *s, pm_ctx *ctx, const float *stokes, const float *slight,
                         int npix, double *models, float *chi2, int *iters, int *accepted,
                         float *fit, long long *ticket)
{
    if (!s || !ctx || !ticket || npix < 0) return PM_EINVAL;
    if (npix > 0 && (!stokes || !models || !chi2 || !iters)) return PM_EINVAL;
    if (pm_ctx_device(ctx) != s->device) return PM_EINVAL;
    Req *r = new Req();
    r->ctx = ctx;
    r->stokes = stokes;
    r->slight = slight;
    r->npix = npix;
    r->L = pm_ctx_nlambda(ctx);
    r->models = models;
    r->chi2 = chi2;
    r->iters = iters;
    r->accepted = accepted;
    r->fit = fit;
    r->bytes = req_bytes(npix, r->L, slight != nullptr, fit != nullptr);
    if (r->bytes > s->budget) { delete r; return PM_ENOMEM; }
    r->cost = (double)r->L * pm_ctx_nterms(ctx);
    // chunks of about the same work whatever the sampling
    r->chunk = std::max(4096, (int)((long long)s->chunk_px * 30 / std::max(r->L, 1)));
    r->t_submit = Clock::now();
    std::lock_guard<std::mutex> g(s->mu);
    r->id = s->next_id++;
    s->all[r->id] = r;
    if (npix == 0) {
        r->state = DONE;
        r->t_done = r->t_submit;
    } else {
        s->queued.push_back(r);
    }
    *ticket = r->id;
    s->cv.notify_all();
    return PM_OK;
}

extern "C" int pm_poll(pm_server *s, long long ticket)
{
    if (!s) return PM_EINVAL;
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->all.find(ticket);
    if (it == s->all.end()) return PM_EINVAL;
    return it->second->state == DONE ? 1 : 0;
}

extern "C" int pm_wait(pm_server *s, long long ticket, double *latency_s)
{
    if (!s) return PM_EINVAL;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->all.find(ticket);
    if (it == s->all.end()) return PM_EINVAL;
    Req *r = it->second;
    s->cv.wait(lk, [&] { return r->state == DONE; });
    if (latency_s)
        *latency_s = std::chrono::duration<double>(r->t_done - r->t_submit).count();
    int rc = r->rc;
    s->all.erase(it);
    delete r;
    return rc;
}
