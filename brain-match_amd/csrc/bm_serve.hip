// Serving loop: refinement jobs arrive at any time, are admitted against the
// HBM budget, run concurrently on the device (one context and stream per
// job, step chunks kept enqueued by one scheduler thread) and retire when
// their stop rule holds.
//
// Threads:
//   preparation workers (prep_threads): take the oldest queued job once an
//     active slot is free and its footprint fits what is left of the budget,
//     then upload, set the mapping, capture the step graph and start the run;
//     they also destroy retired contexts (hipFree may wait for the device,
//     which the scheduler must never do);
//   scheduler (one): enqueues the next chunk of every admitted job as soon as
//     its previous chunk finished, applies the stop rules to the counters
//     each chunk copies back, and collects results.
// A job's stop rules are checked after every steps_per_check steps, exactly
// as bm_refine_run checks them, so a job's result does not depend on what
// else shares the device.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "bm_internal.h"

using namespace bm;

namespace {

struct Job {
    int64_t id = 0;
    int32_t n = 0;
    std::vector<int32_t> g1s, g1d, g1w, g2s, g2d, g2w, map;
    int64_t move_budget = 0, max_steps = 0;
    int32_t patience = 1;
    bm_refine_params prm{};
    int64_t reserve = 0;
    bm_ctx *ctx = nullptr;
    bm_job_result res{};
    std::vector<int32_t> out_map;
    bool done = false;
};

double now_s(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

struct bm_server {
    int device = 0;
    bm_server_config cfg{};
    std::chrono::steady_clock::time_point t0;
    std::mutex mu;
    std::condition_variable cv_prep;   // workers: queue / slots / budget / retired changed
    std::condition_variable cv_sched;  // scheduler: a job became ready
    std::condition_variable cv_done;   // waiters
    std::deque<std::shared_ptr<Job>> queued;
    std::vector<std::shared_ptr<Job>> ready;    // prepared, not yet picked up by the scheduler
    std::vector<std::shared_ptr<Job>> retired;  // contexts to destroy
    std::map<int64_t, std::shared_ptr<Job>> jobs;
    int64_t next_id = 1;
    int64_t slots_used = 0;  // preparing + active + retiring
    int64_t reserved = 0, peak_reserved = 0;
    int64_t submitted = 0, ndone = 0, nfailed = 0, active = 0, peak_active = 0;
    bool stop = false;
    std::vector<std::thread> workers;
    std::thread sched;
};

extern "C" int64_t bm_job_footprint(int32_t n, int64_t e1, int64_t e2, const bm_refine_params *p) {
    const int64_t N1 = (int64_t)n + 1, ld = (N1 + 63) / 64 * 64, xrows = (N1 + 63) / 64 * 64;
    int64_t b = N1 * ld * 2 * 4     // W2, W2^T, P, P^T
                + xrows * ld * 4    // all-pairs gain matrix
                + 2 * e1 * 36       // comb lists (r, w, dw, (r,w)), de-duplicated view
                + e1 * 16           // flat out-edge list
                + e2 * 28           // G2 lists, translated lists, upload staging
                + N1 * 64 + (N1 + 8) * 8
                + (std::max<int64_t>(64, N1 / 8) + 512) * N1 * 4;  // column-recompute scratch of the step
    if (p) {
        const int64_t M = N1 + p->triples + p->sets;
        b += (int64_t)p->triples * 20 + (int64_t)p->sets * (4 * std::max(p->set_k, 1) + 8) + M * 16 +
             (int64_t)(p->rounds > 0 ? p->rounds : (n >= 8000 ? 64 : 32)) * N1 * 8 + N1 * 6 + 24 * 4096 + 4096;
        if (p->triples > 0 && p->pairs)  // dense-form lists and G1 cell matrix of the triple deltas
            b += N1 * ld * 2 + 2 * e1 * 16 + N1 * 24;
    }
    return b + (16 << 20);  // allocation granularity, graph executable
}

static void finish_job(bm_server *s, const std::shared_ptr<Job> &j, int status) {
    // caller holds s->mu
    j->res.status = status;
    std::vector<int32_t>().swap(j->map);
    j->res.t_done = now_s(s->t0);
    j->done = true;
    if (status == BM_OK) s->ndone++;
    else s->nfailed++;
    s->cv_done.notify_all();
}

static int next_chunk(Job &j, int64_t steps_done);

// Preparation: upload, mapping, capture, start, first chunk.  Runs without the lock.
static int prepare(bm_server *s, Job &j) {
    int rc = bm_create(s->device, &j.ctx);
    if (rc) return rc;
    if ((rc = bm_upload_graphs(j.ctx, j.n, j.g1s.data(), j.g1d.data(), j.g1w.data(), (int64_t)j.g1s.size(),
                               j.g2s.data(), j.g2d.data(), j.g2w.data(), (int64_t)j.g2s.size())))
        return rc;
    if ((rc = bm_set_mapping(j.ctx, j.map.data()))) return rc;
    if ((rc = refine_configure(j.ctx, &j.prm))) return rc;
    BM_HIP(hipSetDevice(s->device));
    if ((rc = refine_start(j.ctx))) return rc;
    if ((rc = next_chunk(j, 0))) return rc;
    // inputs live on the device now
    std::vector<int32_t>().swap(j.g1s), std::vector<int32_t>().swap(j.g1d), std::vector<int32_t>().swap(j.g1w);
    std::vector<int32_t>().swap(j.g2s), std::vector<int32_t>().swap(j.g2d), std::vector<int32_t>().swap(j.g2w);
    j.res.device_bytes = bm_device_bytes(j.ctx);
    return BM_OK;
}

// Server threads make allocation and copy calls while other threads capture
// step graphs; in HIP's default (global) mode such a call would invalidate
// those captures, so every server thread runs in relaxed mode.
static void relaxed_capture_mode() {
    hipStreamCaptureMode m = hipStreamCaptureModeRelaxed;
    (void)hipThreadExchangeStreamCaptureMode(&m);
}

static void worker_main(bm_server *s) {
    (void)hipSetDevice(s->device);
    relaxed_capture_mode();
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        // retired contexts first: they hold memory and slots
        if (!s->retired.empty()) {
            auto j = s->retired.back();
            s->retired.pop_back();
            lk.unlock();
            bm_destroy(j->ctx);
            lk.lock();
            j->ctx = nullptr;
            s->reserved -= j->reserve;
            s->slots_used--;
            s->cv_prep.notify_all();
            continue;
        }
        if (s->stop && s->queued.empty() && s->slots_used == 0) return;
        if (!s->queued.empty() && s->slots_used < s->cfg.max_active &&
            s->reserved + s->queued.front()->reserve <= s->cfg.hbm_budget) {
            auto j = s->queued.front();
            s->queued.pop_front();
            s->slots_used++;
            s->reserved += j->reserve;
            s->peak_reserved = std::max(s->peak_reserved, s->reserved);
            j->res.t_admit = now_s(s->t0);
            lk.unlock();
            const int rc = prepare(s, *j);
            if (rc) snprintf(j->res.error, sizeof j->res.error, "%s", bm_last_error());
            lk.lock();
            if (rc) {
                if (j->ctx) s->retired.push_back(j);
                else {
                    s->reserved -= j->reserve;
                    s->slots_used--;
                }
                finish_job(s, j, rc);
                continue;
            }
            j->res.t_start = now_s(s->t0);
            const int64_t actual = j->res.device_bytes + (16 << 20);
            if (actual < j->reserve) {  // keep only what the job holds
                s->reserved -= j->reserve - actual;
                j->reserve = actual;
                s->cv_prep.notify_all();
            }
            s->ready.push_back(j);
            s->cv_sched.notify_one();
            continue;
        }
        s->cv_prep.wait(lk);
    }
}

// Enqueue the next chunk of j: steps_per_check steps, fewer at the step limit.
static int next_chunk(Job &j, int64_t steps_done) {
    int64_t c = j.prm.steps_per_check;
    if (j.max_steps > 0) c = std::min<int64_t>(c, j.max_steps - steps_done);
    return refine_enqueue(j.ctx, (int)c);
}

// Polls job j once: 0 while its chunk runs, 1 when it retired (result set).
static int service(Job &j) {
    int rc = refine_poll(j.ctx, false);
    if (rc == 0) return 0;
    bm_refine_stats st;
    bool fin = false;
    if (rc > 0) {
        rc = BM_OK;
        refine_stats(j.ctx, &st);
        if (!st.verified)
            rc = set_error(BM_ERANGE, "job %lld: exact score check failed at step %lld", (long long)j.id,
                           (long long)st.bad_step);
        else if ((j.max_steps > 0 && st.steps >= j.max_steps) || (j.move_budget > 0 && st.moves >= j.move_budget) ||
                 st.idle_steps >= j.patience)
            fin = true;
        else if ((rc = next_chunk(j, st.steps)) == BM_OK)
            return 0;
        j.res.initial_score = st.initial_score;
        j.res.final_score = st.final_score;
        j.res.steps = st.steps;
        j.res.moves = st.moves;
    }
    if (fin && !rc && (rc = refine_finish(j.ctx)) == BM_OK) {
        j.out_map.resize((size_t)j.n + 1);
        rc = bm_get_mapping(j.ctx, j.out_map.data());
    }
    j.res.status = rc;
    if (rc) snprintf(j.res.error, sizeof j.res.error, "%s", bm_last_error());
    return 1;
}

static void sched_main(bm_server *s) {
    (void)hipSetDevice(s->device);
    relaxed_capture_mode();
    std::vector<std::shared_ptr<Job>> act;
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(s->mu);
            auto idle_done = [&] { return s->stop && s->queued.empty() && s->slots_used == 0; };
            if (act.empty()) s->cv_sched.wait(lk, [&] { return !s->ready.empty() || idle_done(); });
            if (act.empty() && s->ready.empty() && idle_done()) return;
            s->active += (int64_t)s->ready.size();
            s->peak_active = std::max(s->peak_active, s->active);
            act.insert(act.end(), s->ready.begin(), s->ready.end());
            s->ready.clear();
        }
        bool progressed = false;
        for (size_t i = 0; i < act.size();) {
            if (!service(*act[i])) {
                i++;
                continue;
            }
            progressed = true;
            std::lock_guard<std::mutex> lk(s->mu);
            finish_job(s, act[i], act[i]->res.status);
            s->retired.push_back(act[i]);
            s->active--;
            s->cv_prep.notify_all();
            act.erase(act.begin() + (long)i);
        }
        if (!progressed) std::this_thread::sleep_for(std::chrono::microseconds(50));
    }
}

extern "C" int bm_server_create(int device, const bm_server_config *cfg, bm_server **out) {
    if (!cfg || !out) return set_error(BM_EINVAL, "NULL argument");
    if (cfg->max_active < 1 || cfg->prep_threads < 1 || cfg->hbm_budget < 0)
        return set_error(BM_EINVAL, "bad server config");
    int count = 0;
    BM_HIP(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) return set_error(BM_ENODEV, "no HIP device %d", device);
    BM_HIP(hipSetDevice(device));
    auto *s = new bm_server();
    s->device = device;
    s->cfg = *cfg;
    if (s->cfg.hbm_budget == 0) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
            delete s;
            return set_error(BM_EHIP, "hipMemGetInfo failed");
        }
        s->cfg.hbm_budget = (int64_t)fr;
    }
    s->t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < cfg->prep_threads; i++) s->workers.emplace_back(worker_main, s);
    s->sched = std::thread(sched_main, s);
    *out = s;
    return BM_OK;
}

extern "C" int bm_server_submit(bm_server *s, const bm_job_spec *js, int64_t *job_id) {
    if (!s || !js || !job_id) return set_error(BM_EINVAL, "NULL argument");
    if (js->n < 3 || js->e1 < 0 || js->e2 < 0 || !js->mapping || (js->e1 && (!js->g1_src || !js->g1_dst || !js->g1_w)) ||
        (js->e2 && (!js->g2_src || !js->g2_dst || !js->g2_w)))
        return set_error(BM_EINVAL, "bad job arrays");
    if (js->patience < 1 || js->move_budget < 0 || js->max_steps < 0 || js->params.steps_per_check < 1)
        return set_error(BM_EINVAL, "bad stop rule");
    auto j = std::make_shared<Job>();
    j->n = js->n;
    j->g1s.assign(js->g1_src, js->g1_src + js->e1);
    j->g1d.assign(js->g1_dst, js->g1_dst + js->e1);
    j->g1w.assign(js->g1_w, js->g1_w + js->e1);
    j->g2s.assign(js->g2_src, js->g2_src + js->e2);
    j->g2d.assign(js->g2_dst, js->g2_dst + js->e2);
    j->g2w.assign(js->g2_w, js->g2_w + js->e2);
    j->map.assign(js->mapping, js->mapping + js->n + 1);
    j->move_budget = js->move_budget;
    j->max_steps = js->max_steps;
    j->patience = js->patience;
    j->prm = js->params;
    j->reserve = bm_job_footprint(js->n, js->e1, js->e2, &js->params);
    j->res.n = js->n;
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return set_error(BM_ESTATE, "server is shutting down");
    if (j->reserve > s->cfg.hbm_budget)
        return set_error(BM_ENOMEM, "job needs %lld bytes, budget %lld", (long long)j->reserve,
                         (long long)s->cfg.hbm_budget);
    j->id = s->next_id++;
    j->res.job_id = j->id;
    j->res.t_submit = now_s(s->t0);
    s->jobs[j->id] = j;
    s->queued.push_back(j);
    s->submitted++;
    *job_id = j->id;
    s->cv_prep.notify_all();
    return BM_OK;
}

extern "C" int bm_server_wait(bm_server *s, int64_t id, bm_job_result *res, int32_t *mapping) {
    if (!s || !res) return set_error(BM_EINVAL, "NULL argument");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return set_error(BM_EINVAL, "unknown job %lld", (long long)id);
    auto j = it->second;
    s->cv_done.wait(lk, [&] { return j->done; });
    *res = j->res;
    if (mapping && !j->out_map.empty()) memcpy(mapping, j->out_map.data(), j->out_map.size() * 4);
    s->jobs.erase(it);
    return BM_OK;
}

extern "C" int bm_server_stats_get(bm_server *s, bm_server_stats *st) {
    if (!s || !st) return set_error(BM_EINVAL, "NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    st->submitted = s->submitted;
    st->done = s->ndone;
    st->failed = s->nfailed;
    st->queued = (int64_t)s->queued.size();
    st->active = s->active;
    st->peak_active = s->peak_active;
    st->reserved_bytes = s->reserved;
    st->peak_reserved_bytes = s->peak_reserved;
    st->hbm_budget = s->cfg.hbm_budget;
    return BM_OK;
}

extern "C" int bm_server_destroy(bm_server *s) {
    if (!s) return BM_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->cv_prep.notify_all();
        s->cv_sched.notify_all();
    }
    // workers leave once the queue is empty and nothing is left to destroy;
    // the scheduler once every slot has been released
    for (;;) {
        std::unique_lock<std::mutex> lk(s->mu);
        if (s->queued.empty() && s->slots_used == 0 && s->retired.empty()) break;
        lk.unlock();
        std::this_thread::sleep_for(std::chrono::milliseconds(1));
        s->cv_prep.notify_all();
        s->cv_sched.notify_all();
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->cv_prep.notify_all();
        s->cv_sched.notify_all();
    }
    for (auto &t : s->workers) t.join();
    s->sched.join();
    delete s;
    return BM_OK;
}
