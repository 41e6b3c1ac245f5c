// Serving: continuous batching of extraction jobs (motex_submit / motex_poll / motex_wait).
//
// A dispatcher thread per context owns kSlots HIP streams.  Jobs wait in FIFO order until
// a stream is free and their estimated HBM footprint fits beside the jobs already
// resident (the first job always fits so the queue cannot stall).  An admitted job's
// sequences are packed, planned and uploaded on its stream through the stream-ordered
// allocator (no device-wide synchronisation) by a set-up thread of its own, so that the
// dispatcher keeps detecting completions and admitting jobs while a large job's host
// set-up runs; its persistent kernel is enqueued with a grid capped to half the device, so
// jobs that arrive while others run overlap on the GPU.
// Completion is detected by the job's end event; u and v are then accumulated into the
// caller's arrays and the job's device memory is returned to the pool.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>

#include "runtime.h"

namespace motex {

static const int kSlots = 4;

static double now_ms()
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Job {
    uint64_t id = 0;
    std::vector<const char *> seqs;
    std::vector<uint32_t> lens;
    motex_params params{};
    std::vector<uint32_t> gmin, gmax, blen, berr;
    std::vector<uint32_t *> u, v;
    size_t footprint = 0;
    int state = 0;  // 0 pending, 1 running, 2 done, 3 failed
    int rc = MOTEX_OK;
    std::string err;
    double t_submit = 0, t_done = 0, device_ms = 0, cells = 0;
    motex_batch *batch = nullptr;
    int slot = -1;
};

struct Serving {
    hipStream_t streams[kSlots] = {};
    bool busy[kSlots] = {};
    // guarded by the context's mutex:
    std::vector<std::shared_ptr<Job>> running;  // launched jobs, in launch order
    int setups = 0;                             // set-up threads still working
};

static Serving *serving_of(motex_ctx *c);

// Device bytes a job will hold while resident (packed text, NE tables, u/v, work queue,
// kernel scratch), with headroom.
static size_t footprint(const Job &j, const motex_ctx *c)
{
    uint64_t T = 0, maxlen = 0;
    for (uint32_t l : j.lens) { T += l; maxlen = std::max<uint64_t>(maxlen, l); }
    uint64_t ns = 1;
    for (uint32_t g = 0; g < j.params.nb_gaps; ++g) ns *= j.params.gap_max[g] - j.params.gap_min[g] + 1;
    uint64_t grid = (uint64_t)c->num_cu * 24 / 2;
    uint64_t bytes = T + 8 * T * ns + 64 * 1024;
    if (j.params.nb_gaps) {
        bytes += grid * (((maxlen + 31) / 32 * maxlen) * (j.params.nb_gaps + 1) + 4 * (maxlen + 128) + ns * maxlen / 32) * 4;
        if (j.params.distance == 0)  // range-batched plan: unit slots, 64-bit accumulators,
                                     // overflow list, NE tables
            bytes += std::min<uint64_t>((uint64_t)j.lens.size() * (T / 4096 + 1) * 32768, 2ull << 30) + 8 * T * ns +
                     std::min<uint64_t>((uint64_t)j.lens.size() * j.lens.size(), 1u << 22) * 24 +
                     32 * (T + 2 * maxlen + 16384) / 8;
    } else if (j.params.distance == 0) {
        bytes += 32 * (T + kPadBits + 16384) / 8;                // NE tables (up to 32 symbols)
        bytes += (T / 256 + 1) * (T / 7000 + 1) * sizeof(HdItem); // work items (upper bound)
        bytes += grid * 2 * hd_hitq_cap(j.params.k, j.params.e, 4) * 5 * 4;  // hit queues (uncapped grid)
        // pair-symmetric text-side lists: per range, its positions x min(1, 3 x rate x rows)
        // summed over the sequences before it (rate < kSymMaxRate), 2 B each, and a bitmap's
        // worth per range
        bytes += (uint64_t)(2.0 * (double)T * std::min<double>((double)j.lens.size(), 3.0 * kSymMaxRate * (double)T)) +
                 4 * T;
    } else {
        // boundary streams (text-parallel kernel: up to 32 words per row, grid of 12 waves/CU)
        bytes += std::max<uint64_t>(grid * 2 * (maxlen + 128) * 8, (uint64_t)c->num_cu * 12 * 2 * maxlen * 32 * 4);
        bytes += (uint64_t)j.lens.size() * j.lens.size() * sizeof(EdItemDev) + T * 24;
    }
    return (size_t)(bytes * 1.25);
}

// Waits for a running job, accumulates its u/v into the caller's arrays and frees its
// device memory -- all without the context lock -- then publishes the outcome under the lock
// held by `lk`, so that motex_poll/motex_wait/motex_job_times (which read these fields under
// the same lock) see the caller's u/v written before the state becomes 2.
static void finish_job(motex_ctx *c, Job &j, std::unique_lock<std::mutex> &lk)
{
    Serving *sv = serving_of(c);
    lk.unlock();
    int rc = batch_finish(j.batch);
    if (!rc) rc = motex_batch_download(j.batch, j.u.data(), j.v.data());
    const double device_ms = j.batch->stats.kernel_ms, cells = j.batch->stats.cells;
    std::string err = rc ? motex_last_error() : "";
    motex_batch_destroy(j.batch);
    lk.lock();
    j.batch = nullptr;
    sv->busy[j.slot] = false;
    j.device_ms = device_ms;
    j.cells = cells;
    j.rc = rc;
    j.err = err;
    j.t_done = now_ms();
    j.state = rc ? 3 : 2;
}

// Host set-up of an admitted job (packing, planning, upload) and its launch, on a thread of
// its own so that the dispatcher keeps detecting completions and admitting jobs meanwhile.
// Hands the launched job to the dispatcher's running list (or fails it) under the lock.
static void setup_job(motex_ctx *c, std::shared_ptr<Job> jp)
{
    Serving *sv = serving_of(c);
    (void)hipSetDevice(c->device);  // the context's creation selected this device already
    motex_batch *b = nullptr;
    int rc = batch_create(c, jp->seqs.data(), jp->lens.data(), (uint32_t)jp->seqs.size(), &jp->params,
                          sv->streams[jp->slot], &b, 2);
    if (!rc) {
        b->grid_cap = std::max<uint32_t>(c->num_cu, (uint32_t)c->num_cu * b->tile.waves_per_cu / 2);
        rc = batch_launch(b);
    }
    std::string err = rc ? motex_last_error() : "";
    if (rc && b) {
        motex_batch_destroy(b);
        b = nullptr;
    }
    std::lock_guard<std::mutex> lk(c->mu);
    if (rc) {
        jp->rc = rc;
        jp->err = err;
        jp->state = 3;
        jp->t_done = now_ms();
        sv->busy[jp->slot] = false;
        c->resident_bytes -= jp->footprint;
    } else {
        jp->batch = b;
        sv->running.push_back(jp);
    }
    sv->setups--;
    c->cv.notify_all();
}

static void dispatcher(motex_ctx *c)
{
    (void)hipSetDevice(c->device);  // the context's creation selected this device already
    Serving *sv = serving_of(c);
    std::unique_lock<std::mutex> lk(c->mu);
    for (;;) {
        if (c->stop) break;
        // completions: taken off the running list first, since finishing a job drops the
        // lock and set-up threads append to the list meanwhile
        std::vector<std::shared_ptr<Job>> done;
        for (size_t t = 0; t < sv->running.size();) {
            if (hipEventQuery(sv->running[t]->batch->ev1) == hipSuccess) {
                done.push_back(sv->running[t]);
                sv->running.erase(sv->running.begin() + t);
            } else {
                ++t;
            }
        }
        for (auto &jp : done) {
            finish_job(c, *jp, lk);
            c->resident_bytes -= jp->footprint;
        }
        if (!done.empty()) c->cv.notify_all();
        // admissions, FIFO
        while (!c->pending.empty()) {
            std::shared_ptr<Job> jp = c->pending.front();
            int slot = -1;
            for (int s = 0; s < kSlots; ++s) if (!sv->busy[s]) { slot = s; break; }
            if (slot < 0) break;
            const bool resident = !sv->running.empty() || sv->setups > 0;
            if (resident && c->resident_bytes + jp->footprint > c->budget) break;
            c->pending.pop_front();
            sv->busy[slot] = true;
            jp->slot = slot;
            jp->state = 1;
            c->resident_bytes += jp->footprint;
            sv->setups++;
            std::thread(setup_job, c, jp).detach();
        }
        if (sv->running.empty() && sv->setups == 0)
            c->cv.wait(lk, [&] { return c->stop || !c->pending.empty(); });
        else
            c->cv.wait_for(lk, std::chrono::microseconds(200));
    }
    // shut down: let set-ups end, wait for running jobs, fail pending ones
    c->cv.wait(lk, [&] { return sv->setups == 0; });
    std::vector<std::shared_ptr<Job>> left;
    left.swap(sv->running);
    for (auto &jp : left) finish_job(c, *jp, lk);
    for (auto &jp : c->pending) {
        jp->state = 3;
        jp->rc = MOTEX_EDEVICE;
        jp->err = "context destroyed before the job ran";
    }
    c->pending.clear();
    c->cv.notify_all();
}

static std::mutex g_serving_mu;
static std::map<motex_ctx *, Serving *> g_serving;

static Serving *serving_of(motex_ctx *c)
{
    std::lock_guard<std::mutex> g(g_serving_mu);
    return g_serving[c];
}

static std::mutex g_start_mu;  // serialises the first submits of a context

static int ensure_worker(motex_ctx *c)
{
    std::lock_guard<std::mutex> g0(g_start_mu);
    if (c->worker_started) return MOTEX_OK;
    Serving *sv = new Serving();
    for (int s = 0; s < kSlots; ++s) {
        hipError_t e = hipStreamCreateWithFlags(&sv->streams[s], hipStreamNonBlocking);
        if (e != hipSuccess) {
            for (int t = 0; t < s; ++t) (void)hipStreamDestroy(sv->streams[t]);
            delete sv;
            return hip_fail(e, "serving stream");
        }
    }
    {
        std::lock_guard<std::mutex> g(g_serving_mu);
        g_serving[c] = sv;
    }
    c->stop = false;
    c->worker = std::thread(dispatcher, c);
    c->worker_started = true;
    return MOTEX_OK;
}

void serving_shutdown(motex_ctx *c)
{
    std::lock_guard<std::mutex> g0(g_start_mu);
    if (!c->worker_started) return;
    {
        std::lock_guard<std::mutex> lk(c->mu);
        c->stop = true;
    }
    c->cv.notify_all();
    c->worker.join();
    c->worker_started = false;
    Serving *sv = nullptr;
    {
        std::lock_guard<std::mutex> g(g_serving_mu);
        sv = g_serving[c];
        g_serving.erase(c);
    }
    if (sv) {
        for (int s = 0; s < kSlots; ++s) (void)hipStreamDestroy(sv->streams[s]);
        delete sv;
    }
}

}  // namespace motex

using namespace motex;

extern "C" int motex_submit(motex_ctx *c, const motex_job *job, motex_job_id *id)
{
    if (!c || !job || !id) return set_error(MOTEX_EINVAL, "NULL argument");
    int rc = validate_params(&job->params);
    if (rc) return rc;
    if (job->num_seqs && (!job->seqs || !job->lens || !job->u || !job->v))
        return set_error(MOTEX_EINVAL, "job arrays are NULL");
    auto j = std::make_shared<Job>();
    j->seqs.assign(job->seqs, job->seqs + job->num_seqs);
    j->lens.assign(job->lens, job->lens + job->num_seqs);
    j->u.assign(job->u, job->u + job->num_seqs);
    j->v.assign(job->v, job->v + job->num_seqs);
    j->params = job->params;
    const motex_params &p = job->params;
    j->gmin.assign(p.gap_min, p.gap_min + p.nb_gaps);
    j->gmax.assign(p.gap_max, p.gap_max + p.nb_gaps);
    j->blen.assign(p.box_len, p.box_len + p.nb_gaps);
    j->berr.assign(p.box_err, p.box_err + p.nb_gaps);
    j->params.gap_min = j->gmin.data();
    j->params.gap_max = j->gmax.data();
    j->params.box_len = j->blen.data();
    j->params.box_err = j->berr.data();
    j->footprint = footprint(*j, c);
    if (j->footprint > c->budget)
        return set_error(MOTEX_ENOMEM, "job needs about %zu bytes, over the context budget of %zu",
                         j->footprint, c->budget);
    if ((rc = ensure_worker(c))) return rc;
    std::lock_guard<std::mutex> lk(c->mu);
    j->id = c->next_id++;
    j->t_submit = now_ms();
    c->jobs[j->id] = j;
    c->pending.push_back(j);
    *id = j->id;
    c->cv.notify_all();
    return MOTEX_OK;
}

extern "C" int motex_poll(motex_ctx *c, motex_job_id id, int *done)
{
    if (!c || !done) return set_error(MOTEX_EINVAL, "NULL argument");
    std::lock_guard<std::mutex> lk(c->mu);
    auto it = c->jobs.find(id);
    if (it == c->jobs.end()) return set_error(MOTEX_ENOTFOUND, "unknown job %llu", (unsigned long long)id);
    Job &j = *it->second;
    *done = j.state >= 2;
    if (j.state == 3) return set_error(j.rc, "%s", j.err.c_str());
    return MOTEX_OK;
}

extern "C" int motex_wait(motex_ctx *c, motex_job_id id)
{
    if (!c) return set_error(MOTEX_EINVAL, "NULL argument");
    std::unique_lock<std::mutex> lk(c->mu);
    auto it = c->jobs.find(id);
    if (it == c->jobs.end()) return set_error(MOTEX_ENOTFOUND, "unknown job %llu", (unsigned long long)id);
    std::shared_ptr<Job> j = it->second;
    c->cv.wait(lk, [&] { return j->state >= 2; });
    if (j->state == 3) return set_error(j->rc, "%s", j->err.c_str());
    return MOTEX_OK;
}

extern "C" int motex_job_release(motex_ctx *c, motex_job_id id)
{
    if (!c) return set_error(MOTEX_EINVAL, "NULL argument");
    std::lock_guard<std::mutex> lk(c->mu);
    auto it = c->jobs.find(id);
    if (it == c->jobs.end()) return set_error(MOTEX_ENOTFOUND, "unknown job %llu", (unsigned long long)id);
    if (it->second->state < 2) return set_error(MOTEX_EBUSY, "job %llu has not finished", (unsigned long long)id);
    c->jobs.erase(it);
    return MOTEX_OK;
}

extern "C" int motex_job_times(motex_ctx *c, motex_job_id id, double *latency_ms, double *device_ms,
                               double *cells)
{
    if (!c) return set_error(MOTEX_EINVAL, "NULL argument");
    std::lock_guard<std::mutex> lk(c->mu);
    auto it = c->jobs.find(id);
    if (it == c->jobs.end()) return set_error(MOTEX_ENOTFOUND, "unknown job %llu", (unsigned long long)id);
    Job &j = *it->second;
    if (j.state < 2) return set_error(MOTEX_EBUSY, "job %llu has not finished", (unsigned long long)id);
    if (latency_ms) *latency_ms = j.t_done - j.t_submit;
    if (device_ms) *device_ms = j.device_ms;
    if (cells) *cells = j.cells;
    return MOTEX_OK;
}
