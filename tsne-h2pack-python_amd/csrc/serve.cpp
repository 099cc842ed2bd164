// serve.cpp -- job server: continuous batching of independent t-SNE jobs on one GPU.
//
// Jobs (a joint P, an initial Y, run parameters) are queued in arrival order and
// admitted into `slots` worker slots while their estimated device footprint fits the
// HBM budget.  Each slot is a host thread that runs its job through the ordinary
// entry points on the job's own HIP stream (every workspace owns one), so the kernels
// and host round trips of different jobs overlap on the device; a slot takes the next
// queued job as soon as its job ends, and a job arriving while others run is admitted
// as soon as a slot and memory are free.  Per-job device state (P, Y, gains, momentum,
// the H2 structure) stays resident for the whole job.
//
// The reference runs one job per process (its library keeps process-global state,
// src/tsne_h2pack.c:109-111); this server has no counterpart there.
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "api_internal.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Job {
    int id = 0;
    int n = 0, dim = 0;
    int precision = -1;  // H2 product precision for the job (-1: the workspace's default)
    tsne_params_t params{};
    sparse_matrix_t *P = nullptr;
    std::vector<double> init_Y;
    long long footprint = 0;
    int status = TSNE_JOB_QUEUED;
    double t_submit = 0, t_start = 0, t_end = 0;
    double kl = 0;
    std::vector<double> Y;
    bool claimed = false;  // a tsne_server_wait owns the job (later waiters get -1)
};

}  // namespace

struct tsne_server {
    int slots = 1;
    long long budget = 0;
    long long in_use = 0;
    int running = 0;
    bool stop = false;
    int waiters = 0;  // tsne_server_wait calls inside the server (destroy waits for them)
    int precision = -1;  // for jobs submitted from now on (tsne_server_set_h2_precision)
    std::mutex mu;
    std::condition_variable cv_work, cv_done, cv_idle;
    std::deque<Job *> queue;
    std::map<int, Job *> jobs;
    int next_id = 0;
    std::vector<std::thread> workers;
    Clock::time_point t0;

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

void fill_info(const Job *j, tsne_job_info_t *info)
{
    if (!info) return;
    info->job_id = j->id;
    info->status = j->status;
    info->n_point = j->n;
    info->low_dim = j->dim;
    info->kl = j->kl;
    info->submit_s = j->t_submit;
    info->start_s = j->t_start;
    info->end_s = j->t_end;
    info->footprint_bytes = j->footprint;
}

// Runs the job outside the server lock; returns its final status, which the caller
// publishes under the lock (with the end time) before any waiter may see it.
int run_job(Job *j)
{
    tsne_params_t p = j->params;
    p.profiling_output_file = nullptr;
    p.enable_profiling = 0;
    if (!j->init_Y.empty()) {
        p.init_method = 1;
        p.init_Y = j->init_Y.data();
    } else {
        p.init_Y = nullptr;
    }
    // launch shapes of served jobs: the tuning table's tree-level batch and ID classes,
    // with force-reusing iterations replayed from graphs (fewer host launches, which the
    // slots' host threads share) and no second stream for the attractive pass (the jobs
    // running together already fill the GPU's queues: with it, 128 jobs ran at 5.61
    // jobs/s instead of 7.96); TSNE_GRAPH / TSNE_OVERLAP, when set, still decide
    const tsne_tuning_t served = {0, 0, 1, -1, 0.0};
    bool taken = false;
    tsne_workspace_t *ws = tsne::init_with_P(j->n, 0, j->dim, &p, j->P, &taken, &served);
    if (!taken) sparse_matrix_destroy(j->P);
    j->P = nullptr;
    if (ws && j->precision >= 0) tsne_set_h2_precision(ws, j->precision);
    bool ok = ws != nullptr && tsne_run(ws) == 0;
    if (ok) {
        j->Y.resize((size_t)j->n * j->dim);
        tsne_get_embedding(ws, j->Y.data());
        j->kl = compute_kl_divergence(ws);
    }
    if (ws) tsne_destroy(ws);
    std::vector<double>().swap(j->init_Y);
    return ok ? TSNE_JOB_DONE : TSNE_JOB_FAILED;
}

void worker(tsne_server *s)
{
    for (;;) {
        Job *j = nullptr;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            // FIFO admission: the oldest queued job starts when a slot is free and its
            // footprint fits next to the running jobs (alone, it always starts)
            s->cv_work.wait(lk, [&] {
                if (s->stop) return true;
                if (s->queue.empty()) return false;
                const Job *h = s->queue.front();
                return s->running == 0 || s->in_use + h->footprint <= s->budget;
            });
            if (s->stop) return;
            j = s->queue.front();
            s->queue.pop_front();
            s->in_use += j->footprint;
            s->running++;
            j->status = TSNE_JOB_RUNNING;
            j->t_start = s->now();
        }
        const int status = run_job(j);
        {
            std::lock_guard<std::mutex> lk(s->mu);
            j->status = status;
            j->t_end = s->now();
            s->in_use -= j->footprint;
            s->running--;
        }
        s->cv_work.notify_all();
        s->cv_done.notify_all();
    }
}

}  // namespace

extern "C" {

long long tsne_job_footprint(int n_point, int low_dim, long long nnz)
{
    // P on the device (int32 columns, fp32 + fp64 values, 16-bit tile indices) and, until
    // the last renumbering, a second copy with its sort scratch and the tile build's;
    // per-point state, gather copies and forces; the H2 structure (node arrays, pair
    // lists and the U pool, whose regions hold each node's FP64 R entries during its ID:
    // ~1.1 KB per point in 2D, ~2.4 KB in 3D at leaf size 100); fixed buffers (proxy
    // tables, sort scratch)
    const long long h2 = (low_dim == 2) ? 1100 : 2400;
    return 64LL * nnz + (long long)n_point * (32LL * low_dim + h2) + (48LL << 20);
}

tsne_server_t *tsne_server_create(int slots, long long hbm_budget_bytes)
{
    if (slots < 1) slots = 1;
    if (slots > 64) slots = 64;
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
        fprintf(stderr, "tsne_server_create: no HIP device available\n");
        return nullptr;
    }
    tsne_server *s = new tsne_server();
    s->slots = slots;
    s->budget = hbm_budget_bytes > 0 ? hbm_budget_bytes : (long long)(0.9 * (double)free_b);
    {
        // jobs allocate their device arrays from the device's default stream-ordered pool;
        // by default it returns freed memory to the system at every synchronisation, so
        // each new job would allocate afresh.  Keep up to the budget cached instead.
        int dev = 0;
        hipMemPool_t pool = nullptr;
        uint64_t keep = (uint64_t)s->budget;
        if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
            uint64_t cur = 0;
            if (hipMemPoolGetAttribute(pool, hipMemPoolAttrReleaseThreshold, &cur) == hipSuccess &&
                cur < keep)
                (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
        }
    }
    s->t0 = Clock::now();
    for (int k = 0; k < slots; k++) s->workers.emplace_back(worker, s);
    return s;
}

int tsne_server_submit(tsne_server_t *s, int n_point, int low_dim, const tsne_params_t *params,
                       sparse_matrix_t *P, const double *init_Y)
{
    if (!s || !params || !P || (low_dim != 2 && low_dim != 3) || n_point < 2 ||
        P->n_rows != n_point || P->n_cols != n_point) {
        fprintf(stderr, "tsne_server_submit: invalid job\n");
        return -1;
    }
    Job *j = new Job();
    j->n = n_point;
    j->dim = low_dim;
    j->params = *params;
    j->P = P;
    if (init_Y) j->init_Y.assign(init_Y, init_Y + (size_t)n_point * low_dim);
    j->footprint = tsne_job_footprint(n_point, low_dim, P->row_ptr ? P->row_ptr[n_point] : P->nnz);
    int id;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) {
            delete j;
            return -1;
        }
        id = j->id = s->next_id++;
        j->precision = s->precision;
        j->t_submit = s->now();
        s->jobs[j->id] = j;
        s->queue.push_back(j);
    }
    s->cv_work.notify_all();
    return id;
}

int tsne_server_set_h2_precision(tsne_server_t *s, int mode)
{
    if (!s || mode < 0 || mode > 1) return -1;
    std::lock_guard<std::mutex> lk(s->mu);
    s->precision = mode;
    return 0;
}

int tsne_server_poll(tsne_server_t *s, int job_id, tsne_job_info_t *info)
{
    if (!s) return -2;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(job_id);
    if (it == s->jobs.end()) return -2;
    fill_info(it->second, info);
    return it->second->status;
}

int tsne_server_wait(tsne_server_t *s, int job_id, double *Y_out, tsne_job_info_t *info)
{
    if (!s) return -1;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->jobs.find(job_id);
    // one waiter per job: the first claims it and releases it; any other gets -1 at once
    if (it == s->jobs.end() || it->second->claimed) return -1;
    Job *j = it->second;
    j->claimed = true;
    s->waiters++;
    s->cv_done.wait(lk, [&] { return j->status == TSNE_JOB_DONE || j->status == TSNE_JOB_FAILED; });
    fill_info(j, info);
    const int rc = j->status == TSNE_JOB_DONE ? 0 : -1;
    if (rc == 0 && Y_out) memcpy(Y_out, j->Y.data(), j->Y.size() * sizeof(double));
    s->jobs.erase(j->id);
    delete j;
    if (--s->waiters == 0) s->cv_idle.notify_all();
    return rc;
}

int tsne_server_status(tsne_server_t *s, int *queued, int *running, long long *hbm_in_use)
{
    if (!s) return -1;
    std::lock_guard<std::mutex> lk(s->mu);
    if (queued) *queued = (int)s->queue.size();
    if (running) *running = s->running;
    if (hbm_in_use) *hbm_in_use = s->in_use;
    return 0;
}

void tsne_server_destroy(tsne_server_t *s)
{
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        for (Job *j : s->queue) {
            sparse_matrix_destroy(j->P);
            j->P = nullptr;
            j->status = TSNE_JOB_FAILED;
        }
        s->queue.clear();
    }
    s->cv_work.notify_all();
    for (auto &t : s->workers) t.join();
    {
        // every job has ended (run or dropped): wake the waiters and let them leave
        // before their jobs and the server are freed
        std::unique_lock<std::mutex> lk(s->mu);
        s->cv_done.notify_all();
        s->cv_idle.wait(lk, [&] { return s->waiters == 0; });
    }
    for (auto &kv : s->jobs) {
        if (kv.second->P) sparse_matrix_destroy(kv.second->P);
        delete kv.second;
    }
    delete s;
}

}  // extern "C"
