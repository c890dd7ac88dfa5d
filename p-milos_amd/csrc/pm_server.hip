// pm_server.hip -- the serving loop of the C ABI (pm_server_*, pm_submit,
// pm_poll, pm_wait): scans of any size and wavelength sampling arrive at
// any time and share the GPU.
//
// Host side only.  Three threads per server:
//   dispatcher  splits every admitted scan into pixel chunks of about the
//               same estimated work and keeps each compute stream busy with
//               one chunk, always taking the next chunk from the admitted
//               scan with the earliest deadline.  A scan's deadline is its
//               arrival plus 1.5 times its estimated time alone on the GPU
//               (its work over the work rate the dispatcher measures on the
//               chunks it runs): a small scan that arrives behind a large
//               one waits for one chunk per stream, not for the large scan,
//               and a large scan is not held back indefinitely by a stream
//               of smaller ones.  A scan still running at its deadline gets
//               a new one, the same way, from the work it has left: a late
//               large scan does not hold every stream until it ends.  A chunk's Stokes samples are uploaded on
//               its stream just before its kernel.  Chunks of different
//               scans run at the same time: a persistent kernel's workgroups
//               retire as its work queue drains and the next chunk's take
//               their place.
//   admitter    admits queued scans while the device bytes of the admitted
//               ones stay within the HBM budget (the first queued scan that
//               fits; any scan when none is admitted).
//   copier      copies the results of finished chunks back into the
//               caller's buffers, the chunk of the earliest deadline first
//               (a copy into pageable memory can take far longer than the
//               kernel; it delays later copies, not admissions or
//               launches); a scan is done when its last chunk is back, which
//               releases its budget.
// Each compute stream has its own work-queue counters.
// With PM_SERVER_TRACE set to a file name, the server writes one line per
// finished scan there: its times of submission, admission, first chunk
// launch, last kernel done and results back, in ms from the server's start.  A pixel's result
// does not depend on which chunk, stream or workgroup computes it, so a scan
// gives the same bits alone and in the mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
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

constexpr size_t ALIGN = 256;

size_t up(size_t n, size_t a) { return (n + a - 1) / a * a; }

// Device blocks are kept for reuse until the server ends: freeing device
// memory waits for the whole device, which would hold up every other scan.
// get() runs on the admitter thread outside the server's lock (a hipMalloc
// can take milliseconds) and put() on the copier thread under it, so the
// block list has a lock of its own; a new block is allocated outside it.
class DevPool {
    struct Block { char *p; size_t n; bool used; };
    std::vector<Block> blocks_;
    std::mutex mu_;

public:
    char *get(size_t n)
    {
        n = up(n, size_t(2) << 20);
        {
            std::lock_guard<std::mutex> g(mu_);
            Block *best = nullptr;
            for (auto &b : blocks_)
                if (!b.used && b.n >= n && b.n <= 2 * n && (!best || b.n < best->n)) best = &b;
            if (best) { best->used = true; return best->p; }
        }
        char *p = nullptr;
        if (hipMalloc(&p, n) == hipSuccess) {
            std::lock_guard<std::mutex> g(mu_);
            blocks_.push_back({p, n, true});
            return p;
        }
        // give back idle blocks and try once more
        std::lock_guard<std::mutex> g(mu_);
        for (auto it = blocks_.begin(); it != blocks_.end();) {
            if (!it->used) { (void)hipFree(it->p); it = blocks_.erase(it); }
            else ++it;
        }
        if (hipMalloc(&p, n) != hipSuccess) return nullptr;
        blocks_.push_back({p, n, true});
        return p;
    }
    void put(char *p)
    {
        std::lock_guard<std::mutex> g(mu_);
        for (auto &b : blocks_)
            if (b.p == p) b.used = false;
    }
    void release()
    {
        std::lock_guard<std::mutex> g(mu_);
        for (auto &b : blocks_) (void)hipFree(b.p);
        blocks_.clear();
    }
};

enum State { QUEUED, READY, DONE };

struct Req {
    long long id = 0;
    pm_ctx *ctx = nullptr;
    const float *stokes = nullptr, *slight = nullptr;
    double *models = nullptr;
    float *chi2 = nullptr, *fit = nullptr;
    int *iters = nullptr, *accepted = nullptr;
    int npix = 0, L = 0, chunk = 0;
    size_t bytes = 0;
    double cost = 1;                     // estimated work per pixel
    Clock::time_point deadline;          // the dispatcher serves the earliest first
    // device buffers, carved from one pool block
    char *blk = nullptr;
    float *d_stokes = nullptr, *d_slight = nullptr, *d_chi2 = nullptr, *d_fit = nullptr;
    double *d_models = nullptr;
    int *d_iters = nullptr, *d_acc = nullptr;
    int next_px = 0;                     // first pixel not yet dispatched
    int back_px = 0;                     // pixels whose results are in host memory
    int rc = PM_OK;
    State state = QUEUED;
    Clock::time_point t_submit, t_done;
    Clock::time_point t_admit, t_launch, t_computed;   // for the trace
};

struct Chunk {
    Req *r;
    int p0, px;
    hipEvent_t ev;
};

// A scan's deadline: its arrival plus SLACK times its estimated time alone
// (at least MIN_SLACK_S); a late scan's: now plus SLACK times the estimated
// time of the work it has left.
constexpr double SLACK = 1.5, MIN_SLACK_S = 0.02;

static Clock::time_point deadline_after(Clock::time_point t, double work, double rate)
{
    return t + std::chrono::duration_cast<Clock::duration>(
                   std::chrono::duration<double>(std::max(MIN_SLACK_S, SLACK * work / rate)));
}

struct Slot {
    hipStream_t s = nullptr;
    int *counters = nullptr;
    unsigned long long *fallbacks = nullptr;
    bool busy = false;
    Chunk cur{};
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
    double ref_cost = 1;                 // cost of the 30-sample, 9-parameter workload
    // work (pm_ctx_cost units) the GPU completes per second while chunks
    // are in flight, measured over windows of at least 50 ms of busy time
    double rate = 0;
    double win_work = 0, win_busy_s = 0;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Req *> queued;
    std::vector<Req *> ready;            // admitted, chunks left to dispatch
    std::deque<Chunk> back;              // finished chunks to copy back
    std::map<long long, Req *> all;
    size_t in_use = 0;
    int inflight = 0;                    // chunks dispatched, not yet copied back
    bool stop = false;
    long long next_id = 1;
    hipStream_t xfer = nullptr;
    std::vector<Slot> slots;
    DevPool pool;
    std::thread admitter, copier, dispatcher;
    FILE *trace = nullptr;               // PM_SERVER_TRACE
    Clock::time_point t0;

    void admit_loop();
    void copy_loop();
    void dispatcher_loop();
    bool admit_one(std::unique_lock<std::mutex> &lk);
    bool copy_back(const Chunk &c);
    bool launch(Slot &sl, Req *r, std::unique_lock<std::mutex> &lk);
    void complete(Req *r, int px);
    bool idle() const { return queued.empty() && ready.empty() && back.empty() && inflight == 0; }
};

// Admits the first queued scan that fits the budget and gives it its
// device memory.  Called with the lock held; false when none fits.
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
    char *p = pool.get(r->bytes);
    lk.lock();
    if (!p) {
        r->rc = PM_ENOMEM;
        complete(r, r->npix);
        return true;
    }
    size_t n = r->npix, cube = n * 4 * r->L * sizeof(float);
    r->blk = p;
    r->d_stokes = (float *)p; p += up(cube, ALIGN);
    if (r->slight) { r->d_slight = (float *)p; p += up(cube, ALIGN); }
    r->d_models = (double *)p; p += up(n * 11 * sizeof(double), ALIGN);
    r->d_chi2 = (float *)p; p += up(n * sizeof(float), ALIGN);
    r->d_iters = (int *)p; p += up(n * sizeof(int), ALIGN);
    r->d_acc = (int *)p; p += up(n * sizeof(int), ALIGN);
    if (r->fit) r->d_fit = (float *)p;
    r->state = READY;
    r->t_admit = Clock::now();
    ready.push_back(r);
    cv.notify_all();
    return true;
}

// Counts px pixels of r as back in host memory; the scan is done when all
// are.  Called with the lock held.
void pm_server::complete(Req *r, int px)
{
    r->back_px += px;
    if (r->back_px < r->npix) return;
    if (r->blk) pool.put(r->blk);
    r->blk = nullptr;
    in_use -= r->bytes;
    r->t_done = Clock::now();
    r->state = DONE;
    if (trace) {
        auto ms = [&](Clock::time_point t) {
            return t.time_since_epoch().count() == 0
                       ? -1.0 : std::chrono::duration<double, std::milli>(t - t0).count();
        };
        std::fprintf(trace, "%lld %d %d %.3f %.3f %.3f %.3f %.3f\n", r->id, r->npix, r->L,
                     ms(r->t_submit), ms(r->t_admit), ms(r->t_launch), ms(r->t_computed),
                     ms(r->t_done));
    }
    cv.notify_all();
}

// Copies a finished chunk's results into the caller's buffers (copier
// thread, without the lock); false on a HIP error.
bool pm_server::copy_back(const Chunk &c)
{
    const Req *r = c.r;
    const size_t p0 = c.p0, n = c.px, o4 = p0 * 4 * r->L;
    bool ok = hipMemcpyAsync(r->models + p0 * 11, r->d_models + p0 * 11, n * 11 * sizeof(double),
                             hipMemcpyDeviceToHost, xfer) == hipSuccess &&
              hipMemcpyAsync(r->chi2 + p0, r->d_chi2 + p0, n * sizeof(float),
                             hipMemcpyDeviceToHost, xfer) == hipSuccess &&
              hipMemcpyAsync(r->iters + p0, r->d_iters + p0, n * sizeof(int),
                             hipMemcpyDeviceToHost, xfer) == hipSuccess &&
              (!r->accepted || hipMemcpyAsync(r->accepted + p0, r->d_acc + p0, n * sizeof(int),
                                              hipMemcpyDeviceToHost, xfer) == hipSuccess) &&
              (!r->fit || hipMemcpyAsync(r->fit + o4, r->d_fit + o4, n * 4 * r->L * sizeof(float),
                                         hipMemcpyDeviceToHost, xfer) == hipSuccess) &&
              hipStreamSynchronize(xfer) == hipSuccess;
    return ok;
}

void pm_server::admit_loop()
{
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    while (true) {
        if (!queued.empty() && admit_one(lk)) continue;
        if (stop && idle()) break;
        cv.wait(lk);
    }
    cv.notify_all();
}

void pm_server::copy_loop()
{
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    while (true) {
        if (!back.empty()) {
            auto it = std::min_element(back.begin(), back.end(), [](const Chunk &a, const Chunk &b) {
                return a.r->deadline < b.r->deadline;
            });
            Chunk c = *it;
            back.erase(it);
            const bool want = c.r->rc == PM_OK;
            lk.unlock();
            const bool ok = !want || copy_back(c);
            lk.lock();
            if (!ok) c.r->rc = PM_EDEVICE;
            inflight--;
            complete(c.r, c.px);
            continue;
        }
        if (stop && idle()) break;
        cv.wait(lk);
    }
    cv.notify_all();
}

// Uploads and launches the next chunk of r on sl (dispatcher thread; the
// lock is released for the upload and the launch: the scan's device buffers
// stay put until all its chunks are back, and the slot is the dispatcher's).
bool pm_server::launch(Slot &sl, Req *r, std::unique_lock<std::mutex> &lk)
{
    const int p0 = r->next_px, px = std::min(r->chunk, r->npix - p0);
    const size_t o4 = (size_t)p0 * 4 * r->L, cube = (size_t)px * 4 * r->L * sizeof(float);
    if (p0 == 0) r->t_launch = Clock::now();
    r->next_px += px;
    if (r->next_px >= r->npix) ready.erase(std::find(ready.begin(), ready.end(), r));
    inflight++;
    const bool ok_before = r->rc == PM_OK;
    lk.unlock();
    hipEvent_t ev = nullptr;
    if (!sl.spare.empty()) { ev = sl.spare.back(); sl.spare.pop_back(); }
    else if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) ev = nullptr;
    int rc = ev ? PM_OK : PM_EDEVICE;
    if (rc == PM_OK && ok_before) {
        if (hipMemcpyAsync(r->d_stokes + o4, r->stokes + o4, cube, hipMemcpyHostToDevice, sl.s) !=
                hipSuccess ||
            (r->slight && hipMemcpyAsync(r->d_slight + o4, r->slight + o4, cube,
                                         hipMemcpyHostToDevice, sl.s) != hipSuccess))
            rc = PM_EDEVICE;
        if (rc == PM_OK)
            rc = pm_invert_launch(r->ctx, r->d_stokes + o4, r->d_slight ? r->d_slight + o4 : nullptr,
                                  px, r->d_models + (size_t)p0 * 11, r->d_chi2 + p0,
                                  r->d_iters + p0, r->d_acc + p0, r->d_fit ? r->d_fit + o4 : nullptr,
                                  sl.counters, sl.fallbacks, sl.s, 1.0);
        if (rc == PM_OK && hipEventRecord(ev, sl.s) != hipSuccess) rc = PM_EDEVICE;
    }
    lk.lock();
    if (rc != PM_OK || r->rc != PM_OK) {
        if (rc != PM_OK) r->rc = rc;
        if (ev) sl.spare.push_back(ev);
        back.push_back({r, p0, px, nullptr});      // completes without a copy
        cv.notify_all();
        return false;
    }
    sl.cur = {r, p0, px, ev};
    sl.busy = true;
    return true;
}

void pm_server::dispatcher_loop()
{
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    Clock::time_point last = Clock::now();
    bool was_busy = false;
    while (true) {
        // the work rate: completed chunk work over the time chunks were in
        // flight
        const Clock::time_point now = Clock::now();
        if (was_busy) win_busy_s += std::chrono::duration<double>(now - last).count();
        last = now;
        if (win_busy_s >= 0.05) {
            const double r = win_work / win_busy_s;
            rate = rate > 0 ? 0.75 * rate + 0.25 * r : r;
            win_work = win_busy_s = 0;
        }
        bool busy = false;
        for (auto &sl : slots) {
            if (sl.busy) {
                hipError_t e = hipEventQuery(sl.cur.ev);
                if (e == hipErrorNotReady) { busy = true; continue; }
                if (e != hipSuccess) sl.cur.r->rc = PM_EDEVICE;
                sl.cur.r->t_computed = std::max(sl.cur.r->t_computed, now);
                win_work += sl.cur.px * sl.cur.r->cost;
                sl.spare.push_back(sl.cur.ev);
                back.push_back(sl.cur);
                sl.busy = false;
                cv.notify_all();
            }
            // next chunk: from the admitted scan with the earliest deadline
            // (the least work left among equal deadlines)
            Req *best = nullptr;
            double best_w = 0;
            for (Req *r : ready) {
                const double w = (r->npix - r->next_px) * r->cost;
                if (r->deadline < now) r->deadline = deadline_after(now, w, rate);
                if (!best || r->deadline < best->deadline ||
                    (r->deadline == best->deadline && w < best_w)) {
                    best = r;
                    best_w = w;
                }
            }
            if (best && launch(sl, best, lk)) busy = true;
        }
        was_busy = busy;
        if (stop && idle()) break;
        if (busy) {
            lk.unlock();
            std::this_thread::sleep_for(std::chrono::microseconds(50));
            lk.lock();
        } else if (ready.empty()) {
            cv.wait(lk);
        }
    }
    cv.notify_all();
}

extern "C" int pm_server_create(const pm_server_opts *opts, pm_server **out)
{
    if (!out) return PM_EINVAL;
    *out = nullptr;
    pm_server_opts o = {};
    if (opts) o = *opts;
    if (o.streams <= 0) o.streams = 3;
    if (o.streams > 8 || o.chunk_px < 0) return PM_EINVAL;
    if (o.chunk_px == 0) o.chunk_px = 65536;
    pm_server *s = new pm_server();
    if (hipGetDevice(&s->device) != hipSuccess) { delete s; return PM_EDEVICE; }
    if (o.hbm_budget == 0) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) { delete s; return PM_EDEVICE; }
        o.hbm_budget = fr / 2;
    }
    s->budget = o.hbm_budget;
    s->chunk_px = o.chunk_px;
    s->t0 = Clock::now();
    if (const char *f = std::getenv("PM_SERVER_TRACE")) {
        s->trace = std::fopen(f, "w");
        if (s->trace)
            std::fprintf(s->trace, "# id npix L submit admit first_launch computed done (ms)\n");
    }
    // per-pixel cost of the reference workload: 30 samples, 9 parameters,
    // a PSF of about 23 taps
    s->ref_cost = 30 * (400.0 + 10.0 * 9 * 23) + 10.0 * 729;
    // the work rate until the first measurement: 5M pixels/s of it
    s->rate = 5e6 * s->ref_cost;
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
    s->admitter = std::thread(&pm_server::admit_loop, s);
    s->copier = std::thread(&pm_server::copy_loop, s);
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
    if (s->admitter.joinable()) s->admitter.join();
    if (s->copier.joinable()) s->copier.join();
    if (s->dispatcher.joinable()) s->dispatcher.join();
    (void)hipSetDevice(s->device);
    for (auto &sl : s->slots) {
        if (sl.s) (void)hipStreamSynchronize(sl.s);
        for (auto ev : sl.spare) (void)hipEventDestroy(ev);
        if (sl.s) (void)hipStreamDestroy(sl.s);
        if (sl.counters) (void)hipFree(sl.counters);
        if (sl.fallbacks) (void)hipFree(sl.fallbacks);
    }
    if (s->xfer) (void)hipStreamDestroy(s->xfer);
    s->pool.release();
    if (s->trace) std::fclose(s->trace);
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
    r->cost = pm_ctx_cost(ctx);
    // chunks of about the work of chunk_px pixels of the reference workload
    r->chunk = (int)std::max(1024.0, std::min(1e9, s->chunk_px * s->ref_cost / r->cost));
    r->t_submit = Clock::now();
    std::lock_guard<std::mutex> g(s->mu);
    r->deadline = deadline_after(r->t_submit, npix * r->cost, s->rate);
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
    // the ticket is looked up again on every wake-up: another thread's
    // pm_wait on the same ticket may have collected and freed it meanwhile
    auto it = s->all.find(ticket);
    s->cv.wait(lk, [&] {
        it = s->all.find(ticket);
        return it == s->all.end() || it->second->state == DONE;
    });
    if (it == s->all.end()) return PM_EINVAL;
    Req *r = it->second;
    if (latency_s)
        *latency_s = std::chrono::duration<double>(r->t_done - r->t_submit).count();
    int rc = r->rc;
    s->all.erase(it);
    delete r;
    return rc;
}
