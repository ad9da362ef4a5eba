// Served form: a request loop inside the library (no counterpart in the
// reference, which partitions one sample per process).
//
// Jobs (one single-end FASTQ sample each) are queued as they arrive.  E
// engines, each an mp_ctx with its own stream and its own share of the HBM
// budget, run on host threads.  An engine that goes idle takes, in arrival
// order, every queued job whose estimated device footprint still fits its
// share (at most max_jobs_per_batch and 256 MiB of FASTQ) and partitions
// them as one batch
// (mp_batch_load / mp_batch_partition); jobs that arrive meanwhile wait for the
// next idle engine, so the work of two batches overlaps on the device and new
// jobs are admitted as engines and memory free up.  A batch that fails is
// re-run one job at a time, so one bad job only fails itself.  Every job's
// result is that of a plain run of its sample alone (the batch guarantee).
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "../../include/metapart_amd.h"

namespace {

struct Job {
    uint64_t id = 0;
    const char *fq = nullptr;
    uint64_t nbytes = 0;
    bool want_parent = false;
    uint64_t est = 0;
    mp_job_result res{};
    std::vector<uint32_t> parent;
    bool done = false;
    bool claimed = false;  // a thread waits for it (a job is waited for once)
};

// Device bytes of one job in a batch, from its FASTQ size: the FASTQ itself,
// per-read state (~100 B per record at >= 100 B per record), and up to one
// k-mer tuple per base (bases <= bytes / 2) at 26 B for the ping-pong tuple
// buffers plus sort, filter and writer state.
uint64_t estimate_bytes(uint64_t nbytes) { return nbytes + nbytes + (nbytes / 2) * 32 + (8u << 20); }

// FASTQ bytes of one batch (a larger job runs alone): about 800 k reads of
// 150 bp, a few ms of device work, so the jobs of a batch, which all finish
// with it, do not wait on a long one; batching pays off for the small jobs,
// whose fixed per-sample costs it shares.  Under a backlog, batches that grew
// with the queue (the queued bytes shared among the engines) moved fewer reads
// per second than this cap: a large batch leaves the other engines idle.
constexpr uint64_t kBatchBytes = 256ull << 20;

}  // namespace

struct mp_server {
    mp_server_cfg cfg{};
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Job>> queue;
    std::map<uint64_t, std::shared_ptr<Job>> jobs;
    uint64_t next_id = 1;
    bool stop = false;
    std::vector<std::thread> engines;
    std::vector<mp_ctx *> ctxs;
    uint64_t share = 0;  // HBM bytes per engine
    std::chrono::steady_clock::time_point t0;
    std::string err;

    double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
    void engine_loop(int e);
    void run_batch(mp_ctx *ctx, std::vector<std::shared_ptr<Job>> &batch);
};

static void finish_job(Job &j, int status, const mp_sample_result *r, int batch_jobs, double t_start, double t_done)
{
    j.res.status = status;
    if (r) {
        j.res.reads = r->reads;
        j.res.tuples = r->tuples;
        j.res.components = r->components;
        j.res.largest = r->largest;
        memcpy(j.res.W, r->W, sizeof(j.res.W));
    }
    j.res.batch_jobs = batch_jobs;
    j.res.t_start = t_start;
    j.res.t_done = t_done;
}

void mp_server::run_batch(mp_ctx *ctx, std::vector<std::shared_ptr<Job>> &batch)
{
    const int n = (int)batch.size();
    const double ts = now();
    std::vector<const char *> ptr(n);
    std::vector<uint64_t> sz(n);
    for (int i = 0; i < n; i++) {
        ptr[i] = batch[i]->fq;
        sz[i] = batch[i]->nbytes;
    }
    std::vector<mp_sample_result> res(n);
    int rc = mp_batch_load(ctx, n, ptr.data(), sz.data());
    if (!rc) rc = mp_batch_partition(ctx, res.data(), nullptr, nullptr);
    for (int i = 0; !rc && i < n; i++)
        if (batch[i]->want_parent) {
            batch[i]->parent.resize(res[i].reads);
            rc = mp_batch_get_parent(ctx, i, batch[i]->parent.data());
        }
    if (rc && n > 1) {  // isolate the failing job: one at a time
        for (auto &j : batch) {
            std::vector<std::shared_ptr<Job>> one{j};
            run_batch(ctx, one);
        }
        return;
    }
    const double td = now();
    std::lock_guard<std::mutex> lk(mu);
    for (int i = 0; i < n; i++) {
        finish_job(*batch[i], rc, rc ? nullptr : &res[i], n, ts, td);
        if (rc) batch[i]->parent.clear();
        batch[i]->done = true;
    }
    cv_done.notify_all();
}

void mp_server::engine_loop(int e)
{
    mp_ctx *ctx = ctxs[e];
    (void)hipSetDevice(cfg.device);
    for (;;) {
        std::vector<std::shared_ptr<Job>> batch;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_work.wait(lk, [&] { return stop || !queue.empty(); });
            if (queue.empty()) return;  // stopping, nothing left
            uint64_t used = 0, bytes = 0;
            while (!queue.empty() && (int)batch.size() < cfg.max_jobs_per_batch) {
                const Job &j = *queue.front();
                if (!batch.empty() && (used + j.est > share || bytes + j.nbytes > kBatchBytes)) break;
                used += j.est;
                bytes += j.nbytes;
                batch.push_back(queue.front());
                queue.pop_front();
            }
        }
        run_batch(ctx, batch);
    }
}

extern "C" {

int mp_server_create(const mp_server_cfg *cfg, mp_server **out)
{
    if (!cfg || !out || cfg->k < 11 || cfg->k > 32 || (cfg->min3 != 0 && cfg->min3 != 1)) return MP_EINVAL;
    auto *s = new mp_server();
    s->cfg = *cfg;
    if (s->cfg.engines <= 0) s->cfg.engines = 3;
    if (s->cfg.max_jobs_per_batch <= 0 || s->cfg.max_jobs_per_batch > MP_MAX_SAMPLES)
        s->cfg.max_jobs_per_batch = 64;
    if (hipSetDevice(s->cfg.device) != hipSuccess) {
        delete s;
        return MP_EDEVICE;
    }
    uint64_t budget = s->cfg.hbm_budget;
    if (budget == 0) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
            delete s;
            return MP_EDEVICE;
        }
        budget = (uint64_t)fr / 4 * 3;
    }
    s->cfg.hbm_budget = budget;
    s->share = budget / (uint64_t)s->cfg.engines;
    for (int e = 0; e < s->cfg.engines; e++) {
        mp_cfg c{};
        c.device = s->cfg.device;
        c.k = s->cfg.k;
        c.min3 = s->cfg.min3;
        c.hbm_budget = s->share;
        mp_ctx *ctx = nullptr;
        const int rc = mp_ctx_create(&c, &ctx);
        if (rc) {
            for (mp_ctx *x : s->ctxs) mp_ctx_destroy(x);
            delete s;
            return rc;
        }
        s->ctxs.push_back(ctx);
    }
    s->t0 = std::chrono::steady_clock::now();
    for (int e = 0; e < s->cfg.engines; e++) s->engines.emplace_back([s, e] { s->engine_loop(e); });
    *out = s;
    return 0;
}

int mp_server_submit(mp_server *s, const char *fq, uint64_t nbytes, int want_parent, uint64_t *job)
{
    if (!s || !fq || !job || nbytes == 0) return MP_EINVAL;
    auto j = std::make_shared<Job>();
    j->fq = fq;
    j->nbytes = nbytes;
    j->want_parent = want_parent != 0;
    j->est = estimate_bytes(nbytes);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) return MP_ESTATE;
        j->id = s->next_id++;
        j->res.job = j->id;
        j->res.t_submit = s->now();
        s->jobs[j->id] = j;
        s->queue.push_back(j);
    }
    s->cv_work.notify_one();
    *job = j->id;
    return 0;
}

int mp_server_wait(mp_server *s, uint64_t job, mp_job_result *res, uint32_t *parent, uint64_t cap)
{
    if (!s || !res) return MP_EINVAL;
    std::shared_ptr<Job> j;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->jobs.find(job);
        if (it == s->jobs.end()) return MP_EINVAL;
        j = it->second;
        if (j->claimed) return MP_ESTATE;  // another thread waits for it
        j->claimed = true;
        s->cv_done.wait(lk, [&] { return j->done; });
        s->jobs.erase(job);  // by id: the map may have changed while this thread waited
    }
    *res = j->res;
    if (parent && !j->parent.empty()) {
        if (cap < j->parent.size()) return MP_EINVAL;
        memcpy(parent, j->parent.data(), j->parent.size() * sizeof(uint32_t));
    }
    return 0;
}

double mp_server_clock(const mp_server *s) { return s ? s->now() : 0.0; }

int mp_server_destroy(mp_server *s)
{
    if (!s) return 0;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto &t : s->engines) t.join();
    for (mp_ctx *c : s->ctxs) mp_ctx_destroy(c);
    delete s;
    return 0;
}

}  // extern "C"
