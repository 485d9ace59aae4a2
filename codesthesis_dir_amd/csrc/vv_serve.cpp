// vv_serve.cpp -- the request loop of include/vvserve.h: admission by slot
// and device footprint, one worker thread per slot, each admitted run on its
// own stream, results posted to a completion queue.  Runs stay resident after
// their request: a later request of the same grid, geometry and solver takes
// an idle run (vv_reset keeps its tables, buffers and stream) instead of
// creating one; idle runs are destroyed when the budget needs their memory
// or more than two per slot wait.

#include "vv_internal.h"

#include "../../include/vvserve.h"

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

namespace {

struct Job {
    long long id;
    vv_request req;
    std::vector<double> th0;   // copy of the initial temperature
    double need;               // admitted footprint
    double t_submit;
};

struct Idle {
    vv_run *r;
    vv_config cfg;
    double need;
};

}  // namespace

struct vv_server {
    int device = 0;
    double budget = 0.0;
    size_t max_idle = 0;
    std::deque<Idle> idle;     // resident runs between requests, oldest first
    double mem_idle = 0.0;     // their footprint (part of mem)
    int reused = 0, created = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Job> pending;
    std::deque<vv_result> done;
    std::vector<std::thread> workers;
    bool stop = false;
    long long next_id = 0;
    int running = 0, ndone = 0, peak_running = 0;
    double mem = 0.0, peak_mem = 0.0;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

    double now() const
    {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }

    // an idle run that can take the job, or -1 (under mu)
    int idle_match(const Job &j) const
    {
        for (size_t q = 0; q < idle.size(); ++q)
            if (vv::same_shape(idle[q].cfg, j.req.cfg)) return (int)q;
        return -1;
    }
    // the head of the queue can start: a resident run takes it, or the
    // running runs leave room for a new one (idle runs can be evicted)
    bool fits(const Job &j) const { return idle_match(j) >= 0 || mem - mem_idle + j.need <= budget; }

    // runs one admitted job start to finish on this thread, on the resident
    // run r (or a new one); r returns the run to keep (null: none)
    vv_result run(Job &j, vv_run *&r)
    {
        vv_result res;
        std::memset(&res, 0, sizeof(res));
        res.id = j.id;
        res.user = j.req.user;
        res.t_submit = j.t_submit;
        int nlev = 0;
        int st;
        if (r) {
            st = vv_reset(r, &j.req.cfg);
        } else {
            st = vv_create(&j.req.cfg, device, &r, &nlev);
            if (st) r = nullptr;
        }
        if (!st && !j.th0.empty()) st = vv_put(r, VV_TH, j.th0.data(), j.req.ld);
        if (!st) st = vv_start(r);
        vv_step_info info;
        std::memset(&info, 0, sizeof(info));
        if (!st || st == VV_NOT_CONVERGED) {
            const int s2 = vv_step(r, j.req.nsteps, &info);
            st = s2;
        }
        if (st >= 0) {
            res.steps = j.req.nsteps;
            res.trans1 = info.trans1;
            res.trans2 = info.trans2;
            int s3 = VV_OK;
            if (j.req.th_out) s3 = vv_get(r, VV_TH, j.req.th_out, j.req.ld);
            if (!s3 && j.req.vel_out) s3 = vv_get(r, VV_VEL, j.req.vel_out, j.req.ld);
            if (s3) st = s3;
        }
        if (st < 0) std::strncpy(res.error, vv_last_error(), sizeof(res.error) - 1);
        res.status = st;
        return res;
    }

    void worker()
    {
        for (;;) {
            Job j;
            vv_run *r = nullptr;
            double need = 0.0;          // the footprint this worker's run is counted with
            std::vector<vv_run *> evict;
            {
                std::unique_lock<std::mutex> lk(mu);
                // first come, first served: the head waits until it fits;
                // after stop the queue is still drained
                cv_work.wait(lk, [&] { return (!pending.empty() && fits(pending.front())) || (stop && pending.empty()); });
                if (pending.empty()) return;
                j = std::move(pending.front());
                pending.pop_front();
                const int q = idle_match(j);
                if (q >= 0) {           // a resident run of the same shape
                    r = idle[q].r;
                    need = idle[q].need;
                    mem_idle -= need;
                    idle.erase(idle.begin() + q);
                    reused += 1;
                } else {                // room for a new run: evict idle runs, oldest first
                    while (mem + j.need > budget && !idle.empty()) {
                        evict.push_back(idle.front().r);
                        mem -= idle.front().need;
                        mem_idle -= idle.front().need;
                        idle.pop_front();
                    }
                    need = j.need;
                    mem += need;
                    created += 1;
                }
                running += 1;
                if (running > peak_running) peak_running = running;
                if (mem > peak_mem) peak_mem = mem;
            }
            for (vv_run *e : evict) vv_destroy(e);
            evict.clear();
            const double t_admit = now();
            vv_result res = run(j, r);
            res.t_admit = t_admit;
            res.t_done = now();
            {
                std::lock_guard<std::mutex> lk(mu);
                running -= 1;
                ndone += 1;
                if (r) {                // stays resident
                    idle.push_back(Idle{r, j.req.cfg, need});
                    mem_idle += need;
                    while (idle.size() > max_idle) {
                        evict.push_back(idle.front().r);
                        mem -= idle.front().need;
                        mem_idle -= idle.front().need;
                        idle.pop_front();
                    }
                } else {
                    mem -= need;
                }
                done.push_back(res);
            }
            for (vv_run *e : evict) vv_destroy(e);
            cv_done.notify_all();
            cv_work.notify_all();
        }
    }
};

extern "C" {

double vv_run_footprint(const vv_config *cfg)
{
    if (!cfg || cfg->nr < 2 || cfg->np < 2) return 0.0;
    // 24 fields of ld*(np+1) doubles (ld = nr+2 rounded up to 8) and, per
    // Poisson problem (two), the levels of the multigrid hierarchy: three
    // arrays of n x ncol and a reduction table of at most 2 + 23 n words per
    // smoothed level; the direct level also holds one such table per Fourier
    // mode and its FFT twiddles (below 8 columns: two ncol x ncol transform
    // matrices)
    const double ld = (double)(((cfg->nr + 2) + 7) / 8 * 8);
    const double n = (double)cfg->nr;
    const int dc = cfg->direct_cols > 0 ? cfg->direct_cols : vv::default_direct_cols(cfg->np);
    double mg = 0.0;
    for (int c = cfg->np;; c /= 2) {
        const double tab = 2.0 + 23.0 * n;
        if (c <= dc || c < 2 || (c & 1)) {   // phi coarsening ends at an odd count
            // FFT twiddles (no dense transform weights): 2^bits radix-2 words
            // plus 2c for the mixed-radix combination
            const bool fft = vv::fft_cols(c);
            const double twd = (double)(c & -c) + ((c & -c) == c ? 0.0 : 2.0 * c);
            mg += 3.0 * n * c + (c / 2 + 1) * tab + (fft ? twd : 2.0 * (double)c * c);
            break;
        }
        mg += 3.5 * n * c + tab;
    }
    return 8.0 * (24.0 * ld * (cfg->np + 1) + 2.0 * mg) + (1 << 20);
}

int vv_server_create(int device, int slots, double mem_budget, vv_server **out)
{
    if (!out || slots < 1 || slots > 64 || !(mem_budget > 0)) {
        vv::set_error("vv_server_create: need 1 <= slots <= 64 and a positive budget");
        return VV_EINVAL;
    }
    VV_HIP(hipSetDevice(device));
    vv_server *s = new vv_server();
    s->device = device;
    s->budget = mem_budget;
    s->max_idle = 2 * (size_t)slots;
    for (int q = 0; q < slots; ++q) s->workers.emplace_back([s, device] {
        (void)hipSetDevice(device);
        s->worker();
    });
    *out = s;
    return VV_OK;
}

long long vv_server_submit(vv_server *s, const vv_request *req)
{
    if (!s || !req || req->nsteps < 0 || req->cfg.nr < 2 || req->cfg.np < 2) {
        vv::set_error("vv_server_submit: bad request");
        return VV_EINVAL;
    }
    if (req->th0 && req->ld < req->cfg.nr + 2) {
        vv::set_error("vv_server_submit: ld < nr+2");
        return VV_EINVAL;
    }
    Job j;
    j.req = *req;
    j.need = vv_run_footprint(&req->cfg);
    if (req->th0) j.th0.assign(req->th0, req->th0 + (size_t)req->ld * (req->cfg.np + 1));
    long long id;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (j.need > s->budget) {
            vv::set_error("vv_server_submit: the run's footprint exceeds the memory budget");
            return VV_EINVAL;
        }
        id = j.id = s->next_id++;
        j.t_submit = s->now();
        s->pending.push_back(std::move(j));
    }
    s->cv_work.notify_all();
    return id;
}

int vv_server_wait(vv_server *s, vv_result *res, double timeout_s)
{
    if (!s || !res) {
        vv::set_error("vv_server_wait: null argument");
        return VV_EINVAL;
    }
    std::unique_lock<std::mutex> lk(s->mu);
    auto ready = [&] { return !s->done.empty(); };
    if (timeout_s < 0) s->cv_done.wait(lk, ready);
    else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready)) return 1;
    *res = s->done.front();
    s->done.pop_front();
    return VV_OK;
}

int vv_server_stats_read(vv_server *s, vv_server_stats *st)
{
    if (!s || !st) {
        vv::set_error("vv_server_stats_read: null argument");
        return VV_EINVAL;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    st->queued = (int)s->pending.size();
    st->running = s->running;
    st->done = s->ndone;
    st->peak_running = s->peak_running;
    st->mem_in_use = s->mem;
    st->peak_mem = s->peak_mem;
    st->resident = (int)s->idle.size();
    st->reused = s->reused;
    st->created = s->created;
    return VV_OK;
}

int vv_server_destroy(vv_server *s)
{
    if (!s) return VV_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto &t : s->workers) t.join();
    (void)hipSetDevice(s->device);
    for (Idle &e : s->idle) vv_destroy(e.r);
    delete s;
    return VV_OK;
}

}  // extern "C"
