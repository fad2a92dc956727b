// Decode server: continuous batching over resident pipeline slots (include/tc_server.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../../include/tc_server.h"
#include "../tc_internal.hpp"
#include "codestream.hpp"
#include "pipeline.hpp"
#include "plan_util.hpp"

using namespace tc;

namespace {

struct Request {
    uint64_t id = 0;
    std::vector<uint8_t> data;
    Image img;
    uint64_t footprint = 0, out_bytes = 0;
    tc_request_info info;
    std::vector<uint8_t> out;
    bool done = false;
};

struct Slot {
    std::unique_ptr<Pipeline> pipe;
    std::vector<Request*> batch;
    bool busy = false;
    tc_event ev = nullptr;
    void* pinned = nullptr;
    size_t pinned_cap = 0;
    std::vector<size_t> offs;
    uint64_t footprint = 0;
};

// Device bytes a request needs in a pipeline: coefficient planes and work
// buffers (4 B per element), the packed output, codewords and descriptors.
uint64_t footprint_of(const Image& im) {
    uint64_t f = 0;
    for (const auto& t : im.tiles)
        for (const auto& c : t.comps) {
            uint64_t sz[2];
            work_sizes(c, sz);
            f += ((uint64_t)c.width() * c.height() + sz[0] + sz[1] + 192) * 4;
        }
    f += (uint64_t)im.width() * im.height() * im.ncomp * (im.max_prec() > 8 ? 2 : 1);
    f += im.cs_len + im.cblks.size() * 40 + 4096;
    return f;
}

}  // namespace

struct tc_server_s {
    tc_device dev;
    tc_server_config cfg;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> queue;
    std::unordered_map<uint64_t, std::unique_ptr<Request>> reqs;
    std::vector<Slot> slots;
    std::thread worker;
    bool stop = false;
    uint64_t next_id = 1, batch_seq = 0;
    std::chrono::steady_clock::time_point t0;
    tc_server_stats st;
    double busy_since = -1;
    int busy_slots = 0;

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }

    void finish_slot(Slot& s) {
        double t = now();
        for (size_t i = 0; i < s.batch.size(); i++) {
            Request* r = s.batch[i];
            r->out.assign((uint8_t*)s.pinned + s.offs[i], (uint8_t*)s.pinned + s.offs[i] + r->out_bytes);
            r->info.done_s = t;
            r->info.status = TC_SUCCESS;
            r->done = true;
            st.completed++;
            st.pixels += (uint64_t)r->img.width() * r->img.height();
        }
        s.batch.clear();
        s.busy = false;
        s.footprint = 0;
        if (--busy_slots == 0 && busy_since >= 0) {
            st.busy_s += t - busy_since;
            busy_since = -1;
        }
    }

    void fail_batch(Slot& s, tc_status code) {
        double t = now();
        for (Request* r : s.batch) {
            r->info.status = code;
            r->info.done_s = t;
            r->done = true;
            st.failed++;
        }
        s.batch.clear();
    }

    // Launches a batch on slot s (called without the lock held).
    tc_status launch(Slot& s) {
        std::vector<const Image*> imgs;
        size_t total = 0;
        s.offs.clear();
        for (Request* r : s.batch) {
            imgs.push_back(&r->img);
            s.offs.push_back(total);
            total += (r->out_bytes + 255) & ~(size_t)255;
        }
        std::string err;
        if (s.pipe->prepare(imgs, err) != Status::Ok) {
            set_error(err);
            return TC_ERR_DEVICE;
        }
        if (!s.pinned || s.pinned_cap < total) {
            if (s.pinned) tc_host_free_pinned(s.pinned);
            s.pinned = nullptr;
            s.pinned_cap = 0;
            tc_status e = tc_host_alloc_pinned(std::max<size_t>(total, 1 << 20), &s.pinned);
            if (e) return e;
            s.pinned_cap = std::max<size_t>(total, 1 << 20);
        }
        if (s.pipe->run(false) != Status::Ok) return TC_ERR_DEVICE;
        hipStream_t hs = stream_handle(s.pipe->stream());
        for (size_t i = 0; i < s.batch.size(); i++) {
            hipError_t e = hipMemcpyAsync((uint8_t*)s.pinned + s.offs[i], s.pipe->output_ptr((int)i),
                                          s.batch[i]->out_bytes, hipMemcpyDeviceToHost, hs);
            if (e != hipSuccess) return hip_fail(e, "result copy");
        }
        return tc_event_record(s.ev, s.pipe->stream());
    }

    void loop() {
        std::unique_lock<std::mutex> lk(mu);
        const uint64_t slot_budget = cfg.memory_budget / (uint64_t)std::max(1, cfg.slots);
        for (;;) {
            bool progress = false;
            for (Slot& s : slots)
                if (s.busy && tc_event_query(s.ev) == TC_SUCCESS) {
                    finish_slot(s);
                    progress = true;
                }
            if (progress) cv_done.notify_all();
            if (stop && queue.empty() && busy_slots == 0) break;
            for (Slot& s : slots) {
                if (s.busy || queue.empty()) continue;
                // admission: FIFO while the batch fits the slot's memory share
                uint64_t fp = 0;
                while (!queue.empty() && (int)s.batch.size() < cfg.max_batch) {
                    Request* r = queue.front();
                    if (!s.batch.empty() && fp + r->footprint > slot_budget) break;
                    fp += r->footprint;
                    s.batch.push_back(r);
                    queue.pop_front();
                }
                uint64_t seq = ++batch_seq;
                double t = now();
                for (Request* r : s.batch) {
                    r->info.start_s = t;
                    r->info.batch = seq;
                    r->info.batch_size = (int32_t)s.batch.size();
                }
                s.footprint = fp;
                s.busy = true;
                if (busy_slots++ == 0) busy_since = t;
                st.batches++;
                uint64_t inflight = 0;
                for (Slot& o : slots) inflight += o.footprint;
                st.peak_bytes = std::max(st.peak_bytes, inflight);
                lk.unlock();
                tc_status e = launch(s);
                lk.lock();
                if (e != TC_SUCCESS) {
                    tc_stream_sync(s.pipe->stream());
                    fail_batch(s, e);
                    s.busy = false;
                    s.footprint = 0;
                    if (--busy_slots == 0 && busy_since >= 0) {
                        st.busy_s += now() - busy_since;
                        busy_since = -1;
                    }
                    cv_done.notify_all();
                }
                progress = true;
            }
            if (!progress) cv_work.wait_for(lk, std::chrono::microseconds(busy_slots ? 50 : 2000));
        }
    }
};

extern "C" {

tc_status tc_server_create(tc_device dev, const tc_server_config* cfg, tc_server* out) {
    if (!dev || !out) return fail(TC_ERR_INVALID_ARG, "tc_server_create: null");
    auto s = new tc_server_s();
    s->dev = dev;
    s->cfg.slots = cfg && cfg->slots > 0 ? cfg->slots : 3;
    s->cfg.max_batch = cfg && cfg->max_batch > 0 ? cfg->max_batch : 256;
    s->cfg.memory_budget = cfg && cfg->memory_budget ? cfg->memory_budget : (32ull << 30);
    s->cfg.bpw_small = cfg ? cfg->bpw_small : 0;
    s->cfg.bpw_large = cfg ? cfg->bpw_large : 0;
    s->cfg.dwt_shape = cfg ? cfg->dwt_shape : 0;
    std::memset(&s->st, 0, sizeof s->st);
    s->t0 = std::chrono::steady_clock::now();
    s->slots.resize(s->cfg.slots);
    for (Slot& sl : s->slots) {
        sl.pipe.reset(new Pipeline(dev));
        if (!sl.pipe->stream() || tc_event_create(&sl.ev) != TC_SUCCESS) {
            delete s;
            return fail(TC_ERR_DEVICE, "tc_server_create: stream/event creation failed");
        }
        Tuning t;
        t.bpw32 = s->cfg.bpw_small;
        t.bpw64 = s->cfg.bpw_large;
        t.dwt_shape = s->cfg.dwt_shape;
        sl.pipe->set_tuning(t);
    }
    s->worker = std::thread([s]() { s->loop(); });
    *out = s;
    return TC_SUCCESS;
}

tc_status tc_server_destroy(tc_server s) {
    if (!s) return TC_SUCCESS;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    for (Slot& sl : s->slots) {
        if (sl.pipe) tc_stream_sync(sl.pipe->stream());
        tc_event_destroy(sl.ev);
        if (sl.pinned) tc_host_free_pinned(sl.pinned);
    }
    delete s;
    return TC_SUCCESS;
}

tc_status tc_server_submit(tc_server s, const uint8_t* cs, size_t len, uint64_t* id) {
    if (!s || !cs || !id) return fail(TC_ERR_INVALID_ARG, "tc_server_submit: null");
    std::unique_ptr<Request> r(new Request());
    std::memset(&r->info, 0, sizeof r->info);
    r->data.assign(cs, cs + len);
    std::string err;
    Status st = parse_codestream(r->data.data(), r->data.size(), r->img, err);
    std::lock_guard<std::mutex> lk(s->mu);
    r->id = s->next_id++;
    r->info.submit_s = s->now();
    *id = r->id;
    s->st.submitted++;
    if (st != Status::Ok) {
        r->info.status = TC_ERR_UNSUPPORTED;
        r->info.done_s = r->info.submit_s;
        r->done = true;
        s->st.failed++;
        set_error(err);
        s->reqs[r->id] = std::move(r);
        return TC_SUCCESS;
    }
    r->footprint = footprint_of(r->img);
    r->info.width = r->img.width();
    r->info.height = r->img.height();
    r->info.ncomp = r->img.ncomp;
    r->info.bytes_per_sample = r->img.max_prec() > 8 ? 2 : 1;
    r->info.footprint = r->footprint;
    r->out_bytes = (uint64_t)r->info.width * r->info.height * r->info.ncomp * r->info.bytes_per_sample;
    s->queue.push_back(r.get());
    s->reqs[r->id] = std::move(r);
    s->cv_work.notify_one();
    return TC_SUCCESS;
}

tc_status tc_server_poll(tc_server s, uint64_t id, int* done) {
    if (!s || !done) return fail(TC_ERR_INVALID_ARG, "tc_server_poll: null");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return fail(TC_ERR_NOT_FOUND, "tc_server_poll: unknown request");
    *done = it->second->done ? 1 : 0;
    return TC_SUCCESS;
}

tc_status tc_server_wait(tc_server s, uint64_t id, double timeout_s) {
    if (!s) return fail(TC_ERR_INVALID_ARG, "tc_server_wait: null");
    std::unique_lock<std::mutex> lk(s->mu);
    auto deadline = std::chrono::steady_clock::now() + std::chrono::duration<double>(timeout_s);
    for (;;) {
        auto it = s->reqs.find(id);
        if (it == s->reqs.end()) return fail(TC_ERR_NOT_FOUND, "tc_server_wait: unknown request");
        if (it->second->done) return TC_SUCCESS;
        if (s->cv_done.wait_until(lk, deadline) == std::cv_status::timeout) {
            it = s->reqs.find(id);
            if (it != s->reqs.end() && it->second->done) return TC_SUCCESS;
            return TC_ERR_BUSY;
        }
    }
}

tc_status tc_server_result(tc_server s, uint64_t id, void* out, size_t cap, tc_request_info* info) {
    if (!s) return fail(TC_ERR_INVALID_ARG, "tc_server_result: null");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return fail(TC_ERR_NOT_FOUND, "tc_server_result: unknown request");
    Request* r = it->second.get();
    if (!r->done) return fail(TC_ERR_BUSY, "tc_server_result: not finished");
    if (info) *info = r->info;
    if (!out) return TC_SUCCESS;
    if (r->info.status != TC_SUCCESS) {
        tc_status st = r->info.status;
        s->reqs.erase(it);
        return fail(st, "request failed");
    }
    if (cap < r->out.size()) return fail(TC_ERR_CAPACITY, "tc_server_result: output buffer too small");
    std::memcpy(out, r->out.data(), r->out.size());
    s->reqs.erase(it);
    return TC_SUCCESS;
}

tc_status tc_server_release(tc_server s, uint64_t id) {
    if (!s) return fail(TC_ERR_INVALID_ARG, "tc_server_release: null");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return fail(TC_ERR_NOT_FOUND, "tc_server_release: unknown request");
    if (!it->second->done) return fail(TC_ERR_BUSY, "tc_server_release: not finished");
    s->reqs.erase(it);
    return TC_SUCCESS;
}

tc_status tc_server_get_stats(tc_server s, tc_server_stats* out) {
    if (!s || !out) return fail(TC_ERR_INVALID_ARG, "tc_server_get_stats: null");
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->st;
    if (s->busy_since >= 0) out->busy_s += s->now() - s->busy_since;
    out->queued = (int32_t)s->queue.size();
    out->in_flight = s->busy_slots;
    return TC_SUCCESS;
}

double tc_server_now(tc_server s) { return s ? s->now() : 0.0; }

}  // extern "C"
