// Serving runtime for the hip engine: a request loop with continuous batching.
//
// Callers submit homomorphic requests (multiply + relinearize [+ rescale], or
// automorphism/rotation + key-switch) from any thread; a worker thread admits
// them by memory footprint, groups requests of one class (operation, ring and
// rank; level, key, Galois element and rescale mask may differ) into one
// batched engine call, and
// runs batches on several engine lanes (HIP streams of default priority) so
// that work of different classes overlaps on the device.  A batch launched on
// a free lane does not wait for batches on other lanes; requests that arrive
// while batches run are admitted as soon as a lane and memory are free.  Lanes
// beyond the first four take a batch only while the queue is short (light load:
// more small batches in flight shorten the wait); under heavy load the queue
// grows past that and batches form on the first four lanes, where each batch
// grows instead of more batches competing for the device.
// Completion is observed with a HIP event per batch; a request's latency is
// submit -> completion observed.  Outputs are allocated by the runtime and
// handed to the caller, who frees them (free_mlwe_RNS_sample).
//
// Inputs may still be in flight on the submitting thread's lane: submit records
// an event on that lane's stream, and the batch's lane waits for the events of
// all its requests before its first launch, so a ciphertext produced
// asynchronously by the caller is read only once it is complete.
//
// The runtime calls only the engine's public C ABI (include/vfhe_hip.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/vfhe_hip.h"

namespace {

using clk = std::chrono::steady_clock;

double now_s()
{
    return std::chrono::duration<double>(clk::now().time_since_epoch()).count();
}

void srv_check(hipError_t e, const char *what)
{
    if (e != hipSuccess) {
        fprintf(stderr, "vfhe_hip serve: %s: %s\n", what, hipGetErrorString(e));
        abort();
    }
}

struct Req {
    uint64_t id;
    int op;                 // VFHE_HIP_REQ_MULTIPLY or VFHE_HIP_REQ_AUTOMORPHISM
    RNS_MLWE in1, in2;
    RNS_MLWE **ksk;
    uint64_t gen, divide_mask;
    uint64_t bytes;         // footprint reserved while in flight
    hipEvent_t ready;       // recorded on the submitter's stream: the inputs are complete
    double t_submit, t_admit, t_launched, t_done;
    RNS_MLWE out;
};

struct Batch {
    std::vector<Req *> reqs;
    hipEvent_t ev0, ev;   // timing events around the batch on its lane
    int lane;
    uint64_t bytes;
};

// Requests batch together when they share the operation, the ring and the rank;
// level, key, Galois element and rescale mask may differ per request.
bool same_class(const Req *a, const Req *b)
{
    return a->op == b->op && a->in1->b->ntt == b->in1->b->ntt && a->in1->r == b->in1->r;
}

uint64_t popcount(uint64_t x) { return (uint64_t)__builtin_popcountll(x); }

// Device bytes a request holds while in flight: its output, the engine's
// temporaries (tensor / permuted input) and the ModUp buffer of its digits.
// Lanes 0..kBaseLanes-1 always take work; the others only while at most
// kShortQueue requests wait.
constexpr unsigned kBaseLanes = 4, kShortQueue = 8;

uint64_t footprint(const Req *r)
{
    const incNTT t = r->in1->b->ntt;
    const uint64_t row = t->N * 8, l = t->l, rr = r->in1->r;
    uint64_t ext = 0;
    for (uint64_t i = 0; i < rr + 1 && r->ksk; i++)
        if (r->ksk[i]) { ext = popcount(r->ksk[i][0]->b->rns_mask); break; }
    const uint64_t digits = popcount(r->in1->b->rns_mask);
    const uint64_t out = (rr + 1) * l * row;
    if (r->op == VFHE_HIP_REQ_MULTIPLY) {
        const uint64_t R = rr * (rr + 3) / 2;
        return out + (R + 1) * l * row + rr * (rr + 1) / 2 * digits * ext * row;
    }
    return out + (rr + 1) * l * row + rr * digits * ext * row;
}

struct Server {
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Req *> incoming, waiting;
    std::vector<Batch> inflight;
    std::vector<Req *> done;
    uint64_t budget, in_use = 0, peak = 0, next_id = 1, pending = 0;
    unsigned lanes, max_batch;
    std::vector<int> lane_busy;
    bool stop = false;
    std::thread worker;
    uint64_t n_batches = 0, n_reqs = 0, max_busy = 0;
    double host_s = 0, device_s = 0, t_start = now_s(), max_launch_s = 0;
    std::vector<hipEvent_t> ev_free;

    std::vector<hipEvent_t> ready_free;   // guarded by mu (submitters and the worker)

    hipEvent_t event()
    {
        if (!ev_free.empty()) { hipEvent_t e = ev_free.back(); ev_free.pop_back(); return e; }
        hipEvent_t e;
        srv_check(hipEventCreate(&e), "hipEventCreate");
        return e;
    }

    // with mu held
    hipEvent_t ready_event()
    {
        if (!ready_free.empty()) { hipEvent_t e = ready_free.back(); ready_free.pop_back(); return e; }
        hipEvent_t e;
        srv_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreateWithFlags");
        return e;
    }

    void launch(std::vector<Req *> &reqs, int lane)
    {
        const double th0 = now_s();
        const int prev = vfhe_hip_set_lane(lane + 1);   // lanes 1..lanes; lane 0 stays the caller's
        hipStream_t st = (hipStream_t)vfhe_hip_current_stream();
        for (Req *r : reqs) srv_check(hipStreamWaitEvent(st, r->ready, 0), "hipStreamWaitEvent");
        {
            // the waits above captured the events' state: they can be recorded again
            std::lock_guard<std::mutex> g(mu);
            for (Req *r : reqs) { ready_free.push_back(r->ready); r->ready = nullptr; }
        }
        hipEvent_t ev0 = event();
        srv_check(hipEventRecord(ev0, st), "hipEventRecord");
        const size_t B = reqs.size();
        std::vector<RNS_MLWE> out(B), in1(B), in2(B);
        uint64_t bytes = 0;
        for (size_t i = 0; i < B; i++) {
            Req *r = reqs[i];
            const incNTT t = r->in1->b->ntt;
            r->out = vfhe_hip_mlwe_alloc_RNS_sample_uninit(t->N, r->in1->r, r->in1->b->rns_mask, t);
            out[i] = r->out;
            in1[i] = r->in1;
            in2[i] = r->in2;
            bytes += r->bytes;
        }
        std::vector<RNS_MLWE **> ksks(B);
        std::vector<uint64_t> arg(B);
        for (size_t i = 0; i < B; i++) {
            ksks[i] = reqs[i]->ksk;
            arg[i] = reqs[i]->op == VFHE_HIP_REQ_MULTIPLY ? reqs[i]->divide_mask : reqs[i]->gen;
        }
        if (reqs[0]->op == VFHE_HIP_REQ_MULTIPLY)
            vfhe_hip_mlwe_multiply_batch_multi(out.data(), in1.data(), in2.data(), B, ksks.data(), arg.data());
        else
            vfhe_hip_mlwe_automorphism_batch_multi(out.data(), in1.data(), B, arg.data(), ksks.data());
        Batch b;
        b.reqs = reqs;
        b.ev0 = ev0;
        b.ev = event();
        b.lane = lane;
        b.bytes = bytes;
        srv_check(hipEventRecord(b.ev, st), "hipEventRecord");
        vfhe_hip_set_lane(prev);
        lane_busy[lane] = 1;
        inflight.push_back(b);
        n_batches++;
        n_reqs += B;
        const double th1 = now_s();
        for (Req *r : reqs) r->t_launched = th1;
        host_s += th1 - th0;
        max_launch_s = std::max(max_launch_s, th1 - th0);
    }

    // One worker step; returns true when it made progress.
    bool step()
    {
        bool progress = false;
        // completions
        for (size_t k = 0; k < inflight.size();) {
            hipError_t q = hipEventQuery(inflight[k].ev);
            if (q == hipErrorNotReady) { k++; continue; }
            srv_check(q, "batch completion");
            const double t = now_s();
            Batch b = inflight[k];
            inflight.erase(inflight.begin() + (long)k);
            float ms = 0;
            if (hipEventElapsedTime(&ms, b.ev0, b.ev) == hipSuccess) device_s += ms * 1e-3;
            ev_free.push_back(b.ev0);
            ev_free.push_back(b.ev);
            lane_busy[b.lane] = 0;
            {
                std::lock_guard<std::mutex> g(mu);
                for (Req *r : b.reqs) { r->t_done = t; done.push_back(r); }
                in_use -= b.bytes;
                pending -= b.reqs.size();
            }
            cv_done.notify_all();
            progress = true;
        }
        // admission and dispatch
        for (;;) {
            int lane = -1;
            for (unsigned k = 0; k < lanes; k++) if (!lane_busy[k]) { lane = (int)k; break; }
            if (lane < 0) break;
            std::vector<Req *> batch;
            {
                std::lock_guard<std::mutex> g(mu);
                while (!incoming.empty()) { waiting.push_back(incoming.front()); incoming.pop_front(); }
                if (waiting.empty()) break;
                if ((unsigned)lane >= kBaseLanes && waiting.size() > kShortQueue) break;
                Req *head = waiting.front();
                uint64_t room = budget > in_use ? budget - in_use : 0;
                for (auto it = waiting.begin(); it != waiting.end() && batch.size() < max_batch;) {
                    Req *r = *it;
                    if (!same_class(head, r)) { ++it; continue; }
                    // The oldest request is admitted alone when nothing else is in flight,
                    // so a request larger than the budget cannot stall the loop.
                    const bool fits = r->bytes <= room || (batch.empty() && inflight.empty());
                    if (!fits) break;
                    room = r->bytes <= room ? room - r->bytes : 0;
                    batch.push_back(r);
                    it = waiting.erase(it);
                }
                if (batch.empty()) break;
                const double ta = now_s();
                for (Req *r : batch) { in_use += r->bytes; r->t_admit = ta; }
                peak = std::max(peak, in_use);
            }
            launch(batch, lane);
            uint64_t busy = 0;
            for (int x : lane_busy) busy += (uint64_t)x;
            max_busy = std::max(max_busy, busy);
            progress = true;
        }
        return progress;
    }

    void run()
    {
        for (;;) {
            const bool progress = step();
            std::unique_lock<std::mutex> g(mu);
            if (stop && incoming.empty() && waiting.empty() && inflight.empty()) break;
            if (!progress) {
                if (inflight.empty() && incoming.empty() && waiting.empty())
                    cv_work.wait(g, [&] { return stop || !incoming.empty(); });
                else
                    cv_work.wait_for(g, std::chrono::microseconds(20));
            }
        }
    }
};

}  // namespace

extern "C" void *vfhe_hip_server_new(uint64_t memory_budget, uint32_t lanes, uint32_t max_batch)
{
    if (lanes < 1 || lanes > 7 || max_batch < 1) return nullptr;
    vfhe_hip_synchronize();   // initialise the engine on the caller's thread first
    Server *s = new Server();
    s->budget = memory_budget;
    s->lanes = lanes;
    s->max_batch = max_batch;
    s->lane_busy.assign(lanes, 0);
    for (unsigned k = 0; k < lanes; k++) {
        const int prev = vfhe_hip_set_lane((int)k + 1);
        vfhe_hip_set_lane(prev);
    }
    s->worker = std::thread([s] { s->run(); });
    return s;
}

extern "C" uint64_t vfhe_hip_server_submit(void *srv, int op, RNS_MLWE in1, RNS_MLWE in2, RNS_MLWE **ksk,
                                           uint64_t gen, uint64_t divide_mask)
{
    Server *s = (Server *)srv;
    if (!s || !in1 || !ksk || (op != VFHE_HIP_REQ_MULTIPLY && op != VFHE_HIP_REQ_AUTOMORPHISM)) return 0;
    if (op == VFHE_HIP_REQ_MULTIPLY && (!in2 || in2->r != in1->r || in2->b->ntt != in1->b->ntt)) return 0;
    Req *r = new Req();
    r->op = op;
    r->in1 = in1;
    r->in2 = op == VFHE_HIP_REQ_MULTIPLY ? in2 : nullptr;
    r->ksk = ksk;
    r->gen = op == VFHE_HIP_REQ_AUTOMORPHISM ? gen : 0;
    r->divide_mask = op == VFHE_HIP_REQ_MULTIPLY ? divide_mask : 0;
    r->bytes = footprint(r);
    {
        std::lock_guard<std::mutex> g(s->mu);
        r->ready = s->ready_event();
    }
    srv_check(hipEventRecord(r->ready, (hipStream_t)vfhe_hip_current_stream()), "hipEventRecord");
    {
        std::lock_guard<std::mutex> g(s->mu);
        r->id = s->next_id++;
        r->t_submit = now_s();
        s->incoming.push_back(r);
        s->pending++;
    }
    s->cv_work.notify_one();
    return r->id;
}

static uint64_t poll_impl(Server *s, uint64_t *ids, RNS_MLWE *outs, double *latency_s, double *parts, uint64_t max,
                          int wait)
{
    std::unique_lock<std::mutex> g(s->mu);
    if (wait) s->cv_done.wait(g, [&] { return !s->done.empty() || s->pending == 0; });
    const uint64_t n = std::min<uint64_t>(max, s->done.size());
    for (uint64_t k = 0; k < n; k++) {
        Req *r = s->done[k];
        ids[k] = r->id;
        outs[k] = r->out;
        latency_s[k] = r->t_done - r->t_submit;
        if (parts) {
            parts[3 * k] = r->t_admit - r->t_submit;        // waiting for a lane and memory
            parts[3 * k + 1] = r->t_launched - r->t_admit;  // host enqueue of its batch
            parts[3 * k + 2] = r->t_done - r->t_launched;   // device time + completion observed
        }
        delete r;
    }
    s->done.erase(s->done.begin(), s->done.begin() + (long)n);
    return n;
}

extern "C" uint64_t vfhe_hip_server_poll(void *srv, uint64_t *ids, RNS_MLWE *outs, double *latency_s, uint64_t max,
                                         int wait)
{
    return poll_impl((Server *)srv, ids, outs, latency_s, nullptr, max, wait);
}

extern "C" uint64_t vfhe_hip_server_poll_timed(void *srv, uint64_t *ids, RNS_MLWE *outs, double *latency_s,
                                               double *parts, uint64_t max, int wait)
{
    return poll_impl((Server *)srv, ids, outs, latency_s, parts, max, wait);
}

extern "C" void vfhe_hip_server_stats(void *srv, double *stats)
{
    Server *s = (Server *)srv;
    std::lock_guard<std::mutex> g(s->mu);
    stats[0] = (double)s->n_batches;
    stats[1] = (double)s->n_reqs;
    stats[2] = s->n_batches ? (double)s->n_reqs / (double)s->n_batches : 0.0;
    stats[3] = (double)s->peak;
    stats[4] = (double)s->max_busy;
    stats[5] = (double)s->pending;
    stats[6] = s->host_s;
    stats[7] = s->device_s;
    stats[8] = now_s() - s->t_start;
    stats[9] = s->max_launch_s;
}

extern "C" void vfhe_hip_server_free(void *srv)
{
    Server *s = (Server *)srv;
    if (!s) return;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    s->worker.join();
    for (hipEvent_t e : s->ev_free) (void)hipEventDestroy(e);
    for (hipEvent_t e : s->ready_free) (void)hipEventDestroy(e);
    for (Req *r : s->done) delete r;   // outputs of unpolled requests stay with the engine's allocator
    delete s;
}
