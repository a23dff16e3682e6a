// lr_serve.hip - serving (include/lr_hip.h: lrh_server_*): a request loop over one engine.
//
// Requests from any host thread wait in arrival order; the worker thread admits the oldest
// while the resident scenes plus the admitted images fit the device-memory budget (and the
// engine has a job slot), then keeps the engine's graph of iterations running: a job admitted
// while others run starts at once in the slots their ended paths free (continuous batching).
// A job's latency runs from its submission to the device clock of its last sample. The
// reference has no serving loop; its one-view trace loop (engine/gpu/view_gpu.c:1081-1187) is
// what the engine's plain form keeps.
#include <string.h>

#include <algorithm>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "lr_engine.h"

namespace {

struct SrvJob {
    int state = 0;              // 0 waiting, LRH_JOB_RUNNING, LRH_JOB_DONE, -1 failed
    LrhJob job{};
    int64_t eng_id = -1;
    size_t bytes = 0;
    double t_arrive = 0.0;
    LrhJobInfo info{};
    int err = 0;
    std::string msg;
};

}  // namespace

struct LrhServer {
    LrhEngine *eng = nullptr;
    LrhContext *ctx = nullptr;
    // emu: the engine (the worker holds it while a graph runs); mu: the requests and counters
    // (held briefly: a submission never waits for the device). Order: emu, then mu.
    std::mutex emu, mu;
    std::condition_variable cv_work, cv_done;
    std::thread worker;
    bool stop = false;
    int err = 0;
    std::string errmsg;
    std::deque<int64_t> waiting;
    std::vector<int64_t> running;
    std::unordered_map<int64_t, SrvJob> jobs;
    int64_t next_id = 1;
    uint64_t changes = 0;       // submissions and releases: what can let a waiting job in
    size_t budget = 0, scene_bytes = 0, image_bytes = 0;
    std::vector<const LrhPtScene *> scenes;
    int64_t submitted = 0, admitted = 0, done = 0;
    std::vector<double> lat;
    double busy_ms = 0.0;
    float *d_tmp = nullptr;
    size_t tmp_bytes = 0;
};

namespace {

size_t image_bytes_of(const LrhJob &j) { return lr_engine_job_bytes(j); }

void fail_job(LrhServer *s, SrvJob &j, int rc)
{
    j.state = -1;
    j.err = rc;
    j.msg = lrh_last_error();
    s->cv_done.notify_all();
}

// under the lock: the oldest waiting requests that fit, into the engine
void admit(LrhServer *s)
{
    while (!s->waiting.empty()) {
        const int64_t id = s->waiting.front();
        SrvJob &j = s->jobs[id];
        if (s->scene_bytes + s->image_bytes + j.bytes > s->budget) {
            if (s->scene_bytes + j.bytes > s->budget) {   // alone it would not fit: it never will
                s->waiting.pop_front();
                lr_fail(LRH_EREFUSED, "server: job %lld needs %zu bytes over a budget of %zu", (long long)id, j.bytes,
                        s->budget - s->scene_bytes);
                fail_job(s, j, LRH_EREFUSED);
                continue;
            }
            return;   // until running jobs end and finished ones are released
        }
        int64_t eid = -1;
        int rc = lrh_engine_submit(s->eng, &j.job, &eid);
        if (rc == LRH_EAGAIN) {
            // the image may not fit the arena's chunks as they lie: free those holding no image
            // (what the cap then leaves can be made into one chunk) and try once more
            lr_engine_arena_trim(s->eng);
            rc = lrh_engine_submit(s->eng, &j.job, &eid);
            if (rc == LRH_EAGAIN) return;
        }
        s->waiting.pop_front();
        if (rc) {
            fail_job(s, j, rc);
            continue;
        }
        j.eng_id = eid;
        j.state = LRH_JOB_RUNNING;
        s->image_bytes += j.bytes;
        s->admitted++;
        s->running.push_back(id);
    }
}

// under the lock: running jobs whose last sample has ended
void collect(LrhServer *s)
{
    for (size_t k = 0; k < s->running.size();) {
        SrvJob &j = s->jobs[s->running[k]];
        LrhJobInfo info;
        if (lrh_engine_job(s->eng, j.eng_id, &info) == 0 && info.state == LRH_JOB_DONE) {
            j.state = LRH_JOB_DONE;
            j.info = info;
            s->lat.push_back(info.t_done_ms - j.t_arrive);
            s->done++;
            s->running[k] = s->running.back();
            s->running.pop_back();
            s->cv_done.notify_all();
        } else k++;
    }
}

void work(LrhServer *s)
{
    (void)hipSetDevice(s->ctx->device);
    for (;;) {
        std::unique_lock<std::mutex> ek(s->emu);
        std::unique_lock<std::mutex> lk(s->mu);
        if (s->stop) break;
        admit(s);
        if (s->running.empty()) {
            ek.unlock();
            const uint64_t seen = s->changes;
            s->cv_work.wait(lk, [&] { return s->stop || s->changes != seen; });
            continue;
        }
        lk.unlock();
        const double t0 = lrh_now_ms();
        const int rc = lrh_engine_pump(s->eng, 1);   // the engine's lock only: requests keep coming
        lk.lock();
        s->busy_ms += lrh_now_ms() - t0;
        if (rc < 0) {
            // the device failed: every job still open fails with it
            s->err = rc;
            s->errmsg = lrh_last_error();
            for (auto &kv : s->jobs)
                if (kv.second.state == 0 || kv.second.state == LRH_JOB_RUNNING) {
                    kv.second.state = -1;
                    kv.second.err = rc;
                    kv.second.msg = s->errmsg;
                }
            s->waiting.clear();
            s->running.clear();
            s->cv_done.notify_all();
            ek.unlock();
            s->cv_work.wait(lk, [&] { return s->stop; });
            break;
        }
        collect(s);
        lk.unlock();
        ek.unlock();
        std::this_thread::yield();   // readers and releasers get the engine between graphs
    }
}

}  // namespace

extern "C" {

int lrh_server_create(LrhContext *c, int slots_log2, int max_tree_depth, size_t budget_bytes, const LrhTune *tune,
                      LrhServer **out)
{
    if (!c || !out || budget_bytes < 4096) return lr_fail(LRH_EINVAL, "lrh_server_create: bad arguments");
    *out = nullptr;
    // the image arena is made chunk by chunk as admitted images need it, within what the budget
    // leaves beside the resident scenes: scenes plus chunks never pass the budget
    LrhEngine *e = nullptr;
    int rc = lr_engine_create(c, slots_log2, max_tree_depth, budget_bytes, true, &e);
    if (rc) return rc;
    if (tune && (rc = lrh_engine_tune_set(e, tune)) != 0) {
        lrh_engine_destroy(e);
        return rc;
    }
    LrhServer *s = new LrhServer();
    s->eng = e;
    s->ctx = c;
    s->budget = budget_bytes;
    s->worker = std::thread(work, s);
    *out = s;
    return 0;
}

int lrh_server_add_scene(LrhServer *s, const LrhPtScene *scene)
{
    if (!s || !scene) return lr_fail(LRH_EINVAL, "lrh_server_add_scene: bad arguments");
    std::lock_guard<std::mutex> eg(s->emu);
    std::lock_guard<std::mutex> g(s->mu);
    if (std::find(s->scenes.begin(), s->scenes.end(), scene) != s->scenes.end()) return 0;
    // the arena's chunks are device memory too: those holding no image go first if need be
    if (s->scene_bytes + scene->bytes + lr_engine_arena_held(s->eng) > s->budget) lr_engine_arena_trim(s->eng);
    if (s->scene_bytes + scene->bytes + lr_engine_arena_held(s->eng) > s->budget)
        return lr_fail(LRH_EREFUSED, "lrh_server_add_scene: %zu bytes do not fit the budget", scene->bytes);
    s->scenes.push_back(scene);
    s->scene_bytes += scene->bytes;
    lr_engine_arena_cap(s->eng, s->budget - s->scene_bytes);
    return 0;
}

int lrh_server_submit(LrhServer *s, const LrhJob *job, int64_t *id)
{
    if (!s || !job || !id || !job->scene) return lr_fail(LRH_EINVAL, "lrh_server_submit: bad arguments");
    std::lock_guard<std::mutex> g(s->mu);
    if (s->err) return lr_fail(s->err, "lrh_server_submit: %s", s->errmsg.c_str());
    if (std::find(s->scenes.begin(), s->scenes.end(), job->scene) == s->scenes.end())
        return lr_fail(LRH_EINVAL, "lrh_server_submit: the job's scene is not resident (lrh_server_add_scene)");
    SrvJob j;
    j.job = *job;
    j.bytes = image_bytes_of(*job);
    j.t_arrive = lrh_now_ms();
    const int64_t k = s->next_id++;
    s->jobs.emplace(k, j);
    s->waiting.push_back(k);
    s->submitted++;
    s->changes++;
    *id = k;
    s->cv_work.notify_one();
    return 0;
}

int lrh_server_wait(LrhServer *s, int64_t id, double timeout_ms, LrhJobInfo *info)
{
    if (!s) return lr_fail(LRH_EINVAL, "lrh_server_wait: no server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return lr_fail(LRH_EINVAL, "lrh_server_wait: no job %lld", (long long)id);
    auto ready = [&] { return it->second.state == LRH_JOB_DONE || it->second.state < 0; };
    if (timeout_ms < 0) s->cv_done.wait(lk, ready);
    else if (!s->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), ready))
        return lr_fail(LRH_EAGAIN, "lrh_server_wait: job %lld is not done", (long long)id);
    if (it->second.state < 0) return lr_fail(it->second.err, "%s", it->second.msg.c_str());
    if (info) {
        *info = it->second.info;
        info->t_submit_ms = it->second.t_arrive;
    }
    return 0;
}

int lrh_server_image(LrhServer *s, int64_t id, float *h_out)
{
    if (!s || !h_out) return lr_fail(LRH_EINVAL, "lrh_server_image: bad arguments");
    std::lock_guard<std::mutex> eg(s->emu);
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return lr_fail(LRH_EINVAL, "lrh_server_image: no job %lld", (long long)id);
    if (it->second.state != LRH_JOB_DONE) return lr_fail(LRH_EAGAIN, "lrh_server_image: job %lld is not done", (long long)id);
    const size_t bytes = (size_t)it->second.job.w * it->second.job.h * 16;
    if (bytes > s->tmp_bytes) {
        if (s->d_tmp) (void)hipFree(s->d_tmp);
        s->d_tmp = nullptr;
        s->tmp_bytes = 0;
        LR_HIP(hipMalloc((void **)&s->d_tmp, bytes));
        s->tmp_bytes = bytes;
    }
    int rc = lrh_engine_image(s->eng, it->second.eng_id, s->d_tmp, 0);
    if (rc) return rc;
    LR_HIP(hipMemcpy(h_out, s->d_tmp, bytes, hipMemcpyDeviceToHost));
    return 0;
}

int lrh_server_release(LrhServer *s, int64_t id)
{
    if (!s) return lr_fail(LRH_EINVAL, "lrh_server_release: no server");
    std::lock_guard<std::mutex> eg(s->emu);
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return lr_fail(LRH_EINVAL, "lrh_server_release: no job %lld", (long long)id);
    if (it->second.state == 0 || it->second.state == LRH_JOB_RUNNING)
        return lr_fail(LRH_EAGAIN, "lrh_server_release: job %lld is not done", (long long)id);
    if (it->second.state == LRH_JOB_DONE) {
        int rc = lrh_engine_release(s->eng, it->second.eng_id);
        if (rc) return rc;
        s->image_bytes -= it->second.bytes;
    }
    s->jobs.erase(it);
    s->changes++;
    s->cv_work.notify_one();
    return 0;
}

int lrh_server_stats(LrhServer *s, LrhServerStats *st)
{
    if (!s || !st) return lr_fail(LRH_EINVAL, "lrh_server_stats: bad arguments");
    std::lock_guard<std::mutex> eg(s->emu);
    std::lock_guard<std::mutex> g(s->mu);
    LrhEngineStats es;
    int rc = lrh_engine_stats(s->eng, &es);
    if (rc) return rc;
    memset(st, 0, sizeof(*st));
    st->submitted = s->submitted;
    st->admitted = s->admitted;
    st->done = s->done;
    st->waiting = (int64_t)s->waiting.size();
    st->rays = es.rays_ext + es.rays_shadow;
    st->busy_ms = s->busy_ms;
    st->budget = s->budget;
    st->in_use = s->scene_bytes + lr_engine_arena_held(s->eng);
    if (!s->lat.empty()) {
        std::vector<double> v = s->lat;
        std::sort(v.begin(), v.end());
        auto pct = [&](double p) {
            const double x = p * (double)(v.size() - 1);
            const size_t i = (size_t)x;
            const double f = x - (double)i;
            return i + 1 < v.size() ? v[i] + f * (v[i + 1] - v[i]) : v[i];
        };
        st->p50_ms = pct(0.50);
        st->p99_ms = pct(0.99);
    }
    return 0;
}

void lrh_server_destroy(LrhServer *s)
{
    if (!s) return;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    (void)hipSetDevice(s->ctx->device);
    if (s->d_tmp) (void)hipFree(s->d_tmp);
    lrh_engine_destroy(s->eng);
    delete s;
}

}  // extern "C"
