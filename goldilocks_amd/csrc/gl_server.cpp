// Continuous-batching request server for NTT and commit jobs.
//
// One scheduler thread owns `slots` non-blocking HIP streams.  Jobs queue in
// FIFO order; the head job is admitted as soon as a stream is idle and its
// device footprint fits in the unused HBM budget.  Admitted jobs are
// enqueued whole on their stream (copies in, passes, copies out, an event),
// so jobs that arrive while others run overlap with them on the device.
// Completion is observed by polling the job's event; the job's buffers then
// return to the device pool and its budget is released.
//
// Plans and twiddle tables are built at submission, in the submitting
// thread (the job keeps references to them), so the scheduler never builds
// a table or waits for the device; it enqueues a job without holding the
// server lock, so submit/poll/wait/stats are never stalled by an enqueue.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "gl_runtime.hpp"

using namespace gl;

namespace gl {
const char *last_error_cstr();
}

namespace {

double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

struct Job {
    gl_job spec;
    uint64_t ticket = 0;
    size_t footprint = 0;
    double t_submit = 0, t_done = 0;
    int status = GL_OK;
    std::string msg;  // error message of a failed job
    bool done = false;
    PlanRef plan;  // NTT jobs
    Keep keep;     // tables the job's passes read
    // in flight
    int slot = -1;
    std::vector<void *> bufs;
    hipEvent_t event = nullptr;
};

struct Slot {
    hipStream_t stream = nullptr;
    std::shared_ptr<Job> job;
};

size_t ntt_footprint(const gl_job &j, const Plan *plan) {
    size_t words = j.n * j.count;
    size_t f = plan && plan->passes.size() > 1 ? words * 8 : 0;
    if (plan) f += plan_extra_words(*plan, j.count, j.n) * 8;
    if (!j.device) f += words * 8;
    return f;
}

CommitTree tree_of(const gl_job &j) { return j.kind == GL_JOB_COMMIT_ROWS ? TREE_ROWS : TREE_BYTES; }

size_t commit_footprint(const gl_job &j) {
    size_t rows = j.n, cols = j.count, ext = rows * j.blowup;
    size_t f = rows * cols * 8 + ext * cols * 8;  // coef + tmp
    if (j.kind == GL_JOB_COMMIT_ROWS) f += row_tree_cv_words(ext, (u32)cols) * 8;  // K12 tree-mode chaining values
    if (!j.device) f += rows * cols * 8 + ext * cols * 8 + commit_tree_size(rows, cols, j.blowup, tree_of(j));
    return f;
}

// Latencies of the most recent completions (a bounded reservoir).
constexpr size_t kLatencyWindow = 1 << 16;

}  // namespace

struct gl_server {
    size_t budget = 0, in_use = 0, peak = 0;
    std::vector<Slot> slots;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Job>> queue;
    std::map<uint64_t, std::shared_ptr<Job>> jobs;  // until wait/poll reports them done
    uint64_t next_ticket = 1;
    bool stop = false;
    std::thread worker;
    // stats
    uint64_t submitted = 0, completed = 0, failed = 0, rejected = 0;
    std::vector<double> latencies;  // ring of kLatencyWindow entries
    size_t lat_next = 0;
    double elements = 0, busy_ms = 0, busy_since = -1;
    int inflight = 0;

    void loop();
    int start_job(Slot &slot, const std::shared_ptr<Job> &job);
    void finish_job(Slot &slot, int status);
};

// Enqueue a job on its slot's stream (no server lock held).
int gl_server::start_job(Slot &slot, const std::shared_ptr<Job> &job) {
    const gl_job &j = job->spec;
    hipStream_t s = slot.stream;
    auto alloc = [&](size_t bytes) -> void * {
        void *p = bytes ? pool_alloc(bytes, s) : nullptr;
        if (p) job->bufs.push_back(p);
        return p;
    };
    int rc = GL_OK;
    if (j.kind == GL_JOB_NTT || j.kind == GL_JOB_INTT) {
        const bool inverse = j.kind == GL_JOB_INTT;
        if (!job->plan) return GL_OK;  // intt of n <= 1
        const Plan &plan = *job->plan;
        size_t words = j.n * j.count;
        u64 *data = j.data;
        if (!j.device) {
            data = (u64 *)alloc(words * 8);
            if (!data) return set_error(GL_E_NOMEM, "server: allocation failed");
            GL_HIP(hipMemcpyAsync(data, j.data, words * 8, hipMemcpyHostToDevice, s));
        }
        u64 *tmp = nullptr;
        if (plan.passes.size() > 1) {
            tmp = (u64 *)alloc(words * 8);
            if (!tmp) return set_error(GL_E_NOMEM, "server: allocation failed");
        }
        if ((rc = run_plan(plan, data, tmp, 1, j.count, j.n, inverse, s))) return rc;
        if (!j.device) GL_HIP(hipMemcpyAsync(j.data, data, words * 8, hipMemcpyDeviceToHost, s));
    } else {
        size_t rows = j.n, cols = j.count, ext = rows * j.blowup;
        size_t tree = commit_tree_size(rows, cols, j.blowup, tree_of(j));
        const u64 *m = j.data;
        u64 *lde = j.lde_out;
        uint8_t *layers = j.layers_out;
        if (!j.device) {
            u64 *dm = (u64 *)alloc(rows * cols * 8);
            lde = (u64 *)alloc(ext * cols * 8);
            layers = (uint8_t *)alloc(tree + 256);
            if (!dm || !lde || !layers) return set_error(GL_E_NOMEM, "server: allocation failed");
            GL_HIP(hipMemcpyAsync(dm, j.data, rows * cols * 8, hipMemcpyHostToDevice, s));
            m = dm;
        }
        u64 *coef = (u64 *)alloc(rows * cols * 8), *tmp = (u64 *)alloc(ext * cols * 8);
        if (!coef || !tmp) return set_error(GL_E_NOMEM, "server: allocation failed");
        if ((rc = commit_enqueue(m, rows, cols, j.blowup, lde, layers, coef, tmp, s, tree_of(j)))) return rc;
        if (!j.device) {
            GL_HIP(hipMemcpyAsync(j.lde_out, lde, ext * cols * 8, hipMemcpyDeviceToHost, s));
            if (tree) GL_HIP(hipMemcpyAsync(j.layers_out, layers, tree, hipMemcpyDeviceToHost, s));
        }
    }
    return GL_OK;
}

// Caller holds mu.
void gl_server::finish_job(Slot &slot, int status) {
    std::shared_ptr<Job> job = slot.job;
    slot.job.reset();
    // a finished job's event has completed: its blocks go back idle (no event
    // to record); after a failure they stay ordered on the slot's stream
    for (void *p : job->bufs) {
        if (status == GL_OK) pool_free_idle(p);
        else pool_free(p, slot.stream);
    }
    job->bufs.clear();
    job->keep.clear();
    job->plan.reset();
    if (job->event) (void)hipEventDestroy(job->event);
    job->event = nullptr;
    job->t_done = now_ms();
    job->status = status;
    if (status != GL_OK) job->msg = last_error_cstr();
    job->done = true;
    in_use -= job->footprint;
    inflight--;
    if (inflight == 0 && busy_since >= 0) {
        busy_ms += job->t_done - busy_since;
        busy_since = -1;
    }
    if (status == GL_OK) {
        completed++;
        const double lat = job->t_done - job->t_submit;
        if (latencies.size() < kLatencyWindow) latencies.push_back(lat);
        else latencies[lat_next] = lat;
        lat_next = (lat_next + 1) % kLatencyWindow;
        const gl_job &j = job->spec;
        elements += j.kind == GL_JOB_NTT || j.kind == GL_JOB_INTT ? (double)j.n * j.count
                                                                  : (double)j.n * j.blowup * j.count;
    } else {
        failed++;
    }
}

void gl_server::loop() {
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        // completions
        bool progressed = false;
        for (Slot &sl : slots) {
            if (!sl.job || !sl.job->event) continue;  // idle, or being enqueued
            hipError_t q = hipEventQuery(sl.job->event);
            if (q == hipErrorNotReady) continue;
            finish_job(sl, q == hipSuccess ? GL_OK : hip_error(q, "server job"));
            progressed = true;
        }
        if (progressed) cv_done.notify_all();
        // admissions (FIFO head of line)
        while (!queue.empty()) {
            std::shared_ptr<Job> job = queue.front();
            if (budget && in_use + job->footprint > budget) break;
            Slot *free_slot = nullptr;
            for (Slot &sl : slots)
                if (!sl.job) {
                    free_slot = &sl;
                    break;
                }
            if (!free_slot) break;
            queue.pop_front();
            free_slot->job = job;
            in_use += job->footprint;
            peak = std::max(peak, in_use);
            if (inflight++ == 0) busy_since = now_ms();
            // enqueue without the lock; the slot is ours (job set, no event yet)
            lk.unlock();
            int rc = start_job(*free_slot, job);
            hipEvent_t ev = nullptr;
            hipError_t e = hipSuccess;
            if (rc == GL_OK) e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
            if (rc == GL_OK && e == hipSuccess) e = hipEventRecord(ev, free_slot->stream);
            if (rc == GL_OK && e != hipSuccess) rc = hip_error(e, "server event");
            if (rc != GL_OK) (void)hipStreamSynchronize(free_slot->stream);
            lk.lock();
            job->event = ev;
            if (rc != GL_OK) {
                finish_job(*free_slot, rc);
                cv_done.notify_all();
            }
        }
        if (stop && queue.empty() && inflight == 0) break;
        // Sleep briefly: new submissions wake us; in-flight work is polled.
        cv_work.wait_for(lk, std::chrono::microseconds(inflight ? 20 : 2000));
    }
}

extern "C" {

gl_server *gl_server_create(size_t hbm_budget_bytes, int slots) {
    if (ensure_init() != GL_OK) return nullptr;
    if (slots < 1 || slots > 16) {
        set_error(GL_E_ARG, "server slots must be in [1, 16]");
        return nullptr;
    }
    gl_server *srv = new gl_server();
    srv->budget = hbm_budget_bytes;
    srv->slots.resize(slots);
    for (Slot &sl : srv->slots) {
        if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess) {
            (void)hipGetLastError();
            for (Slot &t : srv->slots)
                if (t.stream) (void)hipStreamDestroy(t.stream);
            delete srv;
            set_error(GL_E_DEVICE, "server: stream creation failed");
            return nullptr;
        }
    }
    srv->worker = std::thread([srv] { srv->loop(); });
    return srv;
}

int gl_server_submit(gl_server *srv, const gl_job *job, uint64_t *ticket) {
    if (!srv || !job) return set_error(GL_E_ARG, "null pointer");
    auto j = std::make_shared<Job>();
    j->spec = *job;
    const gl_job &s = j->spec;
    // plans and tables are built here, in the submitting thread
    if (s.kind == GL_JOB_NTT || s.kind == GL_JOB_INTT) {
        if (!(s.kind == GL_JOB_INTT && s.n <= 1)) {
            int rc = get_plan(s.n, &j->plan);
            if (rc) return rc;
            if (j->plan->passes.empty()) j->plan.reset();
        }
        if (s.count == 0 || (!s.data && s.n)) return set_error(GL_E_ARG, "server: empty or null NTT job");
        if (words_overflow(s.n, s.count)) return set_error(GL_E_ARG, "server: n * count words overflow the address space");
        if (j->plan) {
            int rc = warm_plan(*j->plan, 1, &j->keep);
            if (rc) return rc;
        }
        j->footprint = ntt_footprint(s, j->plan.get());
    } else if (s.kind == GL_JOB_COMMIT || s.kind == GL_JOB_COMMIT_ROWS) {
        int rc = commit_check(s.n, s.count, s.blowup, tree_of(s));
        if (rc) return rc;
        if (!s.data || !s.lde_out || !s.layers_out) return set_error(GL_E_ARG, "server: null commit buffer");
        if (s.device && s.kind == GL_JOB_COMMIT_ROWS && (uintptr_t)s.layers_out % 8)
            return set_error(GL_E_ARG, "server: row tree output must be 8-byte aligned");
        if ((rc = commit_prepare(s.n, s.count, s.blowup, &j->keep))) return rc;
        j->footprint = commit_footprint(s);
    } else {
        return set_error(GL_E_ARG, "server: unknown job kind");
    }
    std::lock_guard<std::mutex> lk(srv->mu);
    if (srv->budget && j->footprint > srv->budget) {
        srv->rejected++;
        return set_error(GL_E_NOMEM, "server: job footprint exceeds the HBM budget");
    }
    j->ticket = srv->next_ticket++;
    j->t_submit = now_ms();
    srv->jobs[j->ticket] = j;
    srv->queue.push_back(j);
    srv->submitted++;
    if (ticket) *ticket = j->ticket;
    srv->cv_work.notify_one();
    return GL_OK;
}

int gl_server_poll(gl_server *srv, uint64_t ticket, int *done) {
    if (!srv) return set_error(GL_E_ARG, "null server");
    std::lock_guard<std::mutex> lk(srv->mu);
    auto it = srv->jobs.find(ticket);
    if (it == srv->jobs.end()) return set_error(GL_E_ARG, "unknown ticket");
    const bool fin = it->second->done;
    const int status = fin ? it->second->status : GL_OK;
    const std::string it_msg = it->second->msg;
    if (done) *done = fin;
    if (fin) srv->jobs.erase(it);  // reported: the ticket is released
    if (status != GL_OK) return set_error(status, "server job: " + it_msg);
    return status;
}

int gl_server_wait(gl_server *srv, uint64_t ticket, double *latency_ms) {
    if (!srv) return set_error(GL_E_ARG, "null server");
    std::unique_lock<std::mutex> lk(srv->mu);
    auto it = srv->jobs.find(ticket);
    if (it == srv->jobs.end()) return set_error(GL_E_ARG, "unknown ticket");
    std::shared_ptr<Job> j = it->second;
    srv->cv_done.wait(lk, [&] { return j->done; });
    if (latency_ms) *latency_ms = j->t_done - j->t_submit;
    srv->jobs.erase(ticket);  // reported: the ticket is released
    if (j->status != GL_OK) return set_error(j->status, "server job: " + j->msg);
    return GL_OK;
}

int gl_server_drain(gl_server *srv) {
    if (!srv) return set_error(GL_E_ARG, "null server");
    std::unique_lock<std::mutex> lk(srv->mu);
    srv->cv_done.wait(lk, [&] { return srv->queue.empty() && srv->inflight == 0; });
    return GL_OK;
}

int gl_server_stats_get(gl_server *srv, gl_server_stats *out) {
    if (!srv || !out) return set_error(GL_E_ARG, "null pointer");
    gl_server_stats st{};
    std::vector<double> l;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        st.submitted = srv->submitted;
        st.completed = srv->completed;
        st.failed = srv->failed;
        st.rejected = srv->rejected;
        st.elements = srv->elements;
        st.busy_seconds = srv->busy_ms / 1e3;
        st.peak_bytes = srv->peak;
        l = srv->latencies;  // at most kLatencyWindow entries
    }
    std::sort(l.begin(), l.end());  // outside the lock
    if (!l.empty()) {
        auto pct = [&](double q) { return l[std::min(l.size() - 1, (size_t)(q * (l.size() - 1) + 0.5))]; };
        st.p50_ms = pct(0.50);
        st.p99_ms = pct(0.99);
        double sum = 0;
        for (double v : l) sum += v;
        st.mean_ms = sum / l.size();
        st.max_ms = l.back();
    }
    *out = st;
    return GL_OK;
}

void gl_server_destroy(gl_server *srv) {
    if (!srv) return;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stop = true;
        srv->cv_work.notify_all();
    }
    if (srv->worker.joinable()) srv->worker.join();
    for (Slot &sl : srv->slots)
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
    delete srv;
}

}  // extern "C"
