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
    std::vector<uint8_t> data;     // owned copy of the codestream (tc_server_submit)
    const uint8_t* cs = nullptr;   // the codestream (the copy, or the caller's bytes)
    size_t cs_len = 0;
    Image img;                     // main header at submit; fully parsed by the worker
    uint64_t footprint = 0, out_bytes = 0;
    tc_request_info info;
    std::vector<uint8_t> out;     // the result, once moved out of pinned memory
    const uint8_t* pin = nullptr;  // the result in a slot's pinned buffer
    bool done = false;
    bool moving = false;           // a worker is copying the result out of pinned memory
};

// A batch of one slot: its requests, where their results go in the slot's
// pinned result buffer, its plan and the host copy of its upload region.
struct Batch {
    std::vector<Request*> reqs;
    std::vector<Request*> rejected;  // failed to parse
    std::vector<size_t> offs;
    uint64_t footprint = 0;
    int buf = 0;                     // pinned result buffer (0 or 1)
    size_t out_total = 0;            // bytes of the results in it
    Plan plan;
    tc_status status = TC_SUCCESS;
    double t_plan = 0, t_launch = 0;
};

struct Slot {
    std::unique_ptr<Pipeline> pipe;
    tc_event ev = nullptr;     // the running batch's results are in pinned memory
    tc_event kdone = nullptr;  // the running batch's kernels have finished
    // Results land in one of two pinned buffers and stay there until read; a
    // buffer is reused every other batch, after moving out what is unread.
    void* pinned[2] = {nullptr, nullptr};
    size_t pinned_cap[2] = {0, 0};
    std::vector<uint64_t> holders[2];  // request ids with results in each buffer
    int cur = 0;
    // Host copy of the upload region.  The next batch is staged into it once
    // the running batch's upload has left (up_done; long before its kernels
    // end).
    void* staging = nullptr;
    size_t staging_cap = 0;
    tc_event up_done = nullptr;
    bool up_pending = false;
    std::thread worker;
};

// Upper estimate of the device bytes a request needs in a pipeline, from its
// main header:
// - coefficient planes: 4 B per sample of every component over the whole
//   image (the tiles partition it), plus one 64-element alignment per
//   tile-component;
// - inverse-DWT work buffers: per tile-component at most two quarter-tiles of
//   the nominal tile size (the level-1 low band, ping-pong), 4 B each;
// - tier-1 state of a lockstep packing, whichever the planner picks: per
//   code-block (4 + passes / 3) rows of the block height, 8 B per row for
//   64-wide blocks and 4 B for narrower ones, with passes bounded by the
//   bit-planes (precision + 2 guard bits + levels + 2), blocks counted over
//   the image with partial blocks at every band edge;
// - the packed output, the codestream (codewords) and 40 B of descriptor per
//   16 codeword bytes.
uint64_t footprint_of(const Image& im) {
    if (im.tiles.empty() || im.tiles[0].comps.empty()) return (uint64_t)im.cs_len * 4 + (1u << 20);
    const uint64_t ntiles = (uint64_t)std::max(1, im.ntx * im.nty);
    uint64_t f = 0;
    for (int c = 0; c < im.ncomp; c++) {
        int x0, y0, x1, y1;
        im.comp_rect(c, x0, y0, x1, y1);
        const uint64_t samples = (uint64_t)std::max(0, x1 - x0) * std::max(0, y1 - y0);
        const TileComp& tc0 = im.tiles[0].comps[c];
        const uint64_t tw = (uint64_t)(im.tw + im.dx[c] - 1) / im.dx[c], th = (uint64_t)(im.th + im.dy[c] - 1) / im.dy[c];
        const uint64_t work = tc0.nlev > 1 ? 2 * ((tw + 1) / 2) * ((th + 1) / 2) : 0;
        f += (samples + ntiles * (64 + work + 128)) * 4;
        const Band& B = tc0.res[tc0.nlev].band[0];
        const uint64_t cbw = (uint64_t)1 << B.xcb, cbh = (uint64_t)1 << B.ycb;
        const uint64_t blocks = samples / (cbw * cbh) + ntiles * (uint64_t)(3 * tc0.nlev + 1) * (tw / cbw + th / cbh + 2);
        const uint64_t passes = 3 * (uint64_t)(im.prec[c] + 2 + tc0.nlev + 2);
        f += blocks * (4 + passes / 3) * cbh * (cbw > 32 ? 8 : 4);
    }
    f += im.out_bytes();
    f += im.cs_len + im.cs_len / 16 * 40 + 4096;
    return f;
}

// Resident buffers of a slot's pipeline grow by this factor when they must
// (admission charges it).
constexpr double kSlotGrowth = 1.25;
}  // namespace

// One worker thread per slot: it takes the next batch from the queue (FIFO,
// within the slot's share of the memory budget), plans it while the slot's
// previous batch runs, then launches it on the slot's stream once that one
// has finished.  Host work (parsing, planning, staging) overlaps the slot's
// own device work and that of the other slots, and batches of different
// slots run concurrently on the device.
// Threads one slot uses to parse a large batch's codestreams.
constexpr int kParseThreads = 4;

struct tc_server_s {
    tc_device dev;
    tc_server_config cfg;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> queue;
    std::unordered_map<uint64_t, std::unique_ptr<Request>> reqs;
    std::vector<Slot> slots;
    bool stop = false;
    uint64_t next_id = 1, batch_seq = 0;
    std::chrono::steady_clock::time_point t0;
    tc_server_stats st;
    double busy_since = -1;
    int busy_slots = 0;
    uint64_t inflight_bytes = 0;
    // Kernels of different slots run one batch after the other: each batch's
    // first kernel waits for the last kernel of the batch launched before it
    // (on whichever slot), while uploads, result copies and host work of the
    // slots overlap them.  Two batches whose kernels shared the device took
    // longer together than one after the other: each is bound by its longest
    // tier-1 chains, which slow down when another batch's waves share the
    // compute units.
    std::mutex compute_mu;
    tc_event last_compute = nullptr;  // kdone of the slot that launched kernels last
    int idle_workers = 0;

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }

    // caller holds mu
    void batch_done(Batch& b) {
        inflight_bytes -= b.footprint;
        b.footprint = 0;
        if (--busy_slots == 0 && busy_since >= 0) {
            st.busy_s += now() - busy_since;
            busy_since = -1;
        }
    }

    // Host half of a batch, without the lock and while the slot's previous
    // batch runs on the device: parse (on a few threads when the batch is
    // large; a request whose packets do not parse fails alone), plan, and
    // gather the upload region into the batch's staging buffer.
    tc_status plan_host(Slot& s, Batch& b) {
        const double tp = now();
        std::vector<char> good(b.reqs.size(), 0);
        size_t bytes = 0;
        for (Request* r : b.reqs) bytes += r->cs_len;
        const int nt = (int)std::min<size_t>({(size_t)kParseThreads, b.reqs.size(), 1 + bytes / (4u << 20)});
        auto parse_part = [&](int k) {
            std::string perr;
            for (size_t i = (size_t)k; i < b.reqs.size(); i += (size_t)nt)
                // (sub-sampled images, whose output is planar, are not served)
                good[i] = parse_codestream(b.reqs[i]->cs, b.reqs[i]->cs_len, b.reqs[i]->img, perr) == Status::Ok &&
                          !b.reqs[i]->img.planar();
        };
        if (nt <= 1) {
            parse_part(0);
        } else {
            std::vector<std::thread> th;
            for (int k = 1; k < nt; k++) th.emplace_back(parse_part, k);
            parse_part(0);
            for (auto& t : th) t.join();
        }
        std::vector<Request*> ok;
        for (size_t i = 0; i < b.reqs.size(); i++) (good[i] ? ok : b.rejected).push_back(b.reqs[i]);
        b.reqs.swap(ok);
        b.offs.clear();
        if (b.reqs.empty()) return TC_SUCCESS;
        std::vector<const Image*> imgs;
        size_t total = 0;
        for (Request* r : b.reqs) {
            imgs.push_back(&r->img);
            b.offs.push_back(total);
            total += (r->out_bytes + 255) & ~(size_t)255;
        }
        std::string err;
        if (s.pipe->plan_batch(imgs, b.plan, err) != Status::Ok) {
            set_error(err);
            return TC_ERR_DEVICE;
        }
        b.out_total = total;
        // Pinned memory for the upload.  Allocating and freeing pinned memory
        // stalls the whole device, so this buffer and the result buffers
        // start large (the upload at the batch byte cap), grow by doubling
        // and are freed only with the server.  The staging buffer is replaced
        // or rewritten only once the running batch's upload has left.
        const size_t up = b.plan.upload_bytes();
        if (s.up_pending) {
            tc_status e = tc_event_sync(s.up_done);
            if (e) return e;
            s.up_pending = false;
        }
        if (!s.staging || s.staging_cap < up) {
            if (s.staging) tc_host_free_pinned(s.staging);
            s.staging = nullptr;
            const size_t want = std::max<size_t>({2 * up, 2 * s.staging_cap, (size_t)cfg.max_batch_bytes + (64u << 20)});
            s.staging_cap = 0;
            tc_status e = tc_host_alloc_pinned(want, &s.staging);
            if (e) return e;
            s.staging_cap = want;
        }
        stage_upload(b.plan, (uint8_t*)s.staging);
        std::lock_guard<std::mutex> lk(mu);
        st.plan_s += now() - tp;
        return TC_SUCCESS;
    }

    // Device half, once the slot's previous batch has finished: fit the
    // buffers, upload, launch, copy the results to pinned memory and record
    // the slot's event (nothing waits here).
    tc_status launch(Slot& s, Batch& b) {
        b.t_launch = now();
        if (b.reqs.empty()) return TC_SUCCESS;
        vacate(s, b.buf);
        void*& pin = s.pinned[b.buf];
        if (!pin || s.pinned_cap[b.buf] < b.out_total) {
            if (pin) tc_host_free_pinned(pin);
            pin = nullptr;
            const size_t want = std::max<size_t>({2 * b.out_total, 2 * s.pinned_cap[b.buf], (size_t)256 << 20});
            s.pinned_cap[b.buf] = 0;
            tc_status e = tc_host_alloc_pinned(want, &pin);
            if (e) return e;
            s.pinned_cap[b.buf] = want;
        }
        std::string err;
        if (s.pipe->adopt(std::move(b.plan), s.staging, err) != Status::Ok) {
            set_error(err);
            return TC_ERR_DEVICE;
        }
        if (tc_event_record(s.up_done, s.pipe->stream()) == TC_SUCCESS) s.up_pending = true;
        {
            std::lock_guard<std::mutex> g(compute_mu);
            if (!cfg.overlap_kernels && last_compute &&
                tc_stream_wait_event(s.pipe->stream(), last_compute) != TC_SUCCESS)
                return TC_ERR_DEVICE;
            if (s.pipe->run(false) != Status::Ok) return TC_ERR_DEVICE;
            tc_status e = tc_event_record(s.kdone, s.pipe->stream());
            if (e != TC_SUCCESS) return e;
            last_compute = s.kdone;
        }
        hipStream_t hs = stream_handle(s.pipe->stream());
        for (size_t i = 0; i < b.reqs.size(); i++) {
            hipError_t e = hipMemcpyAsync((uint8_t*)s.pinned[b.buf] + b.offs[i], s.pipe->output_ptr((int)i),
                                          b.reqs[i]->out_bytes, hipMemcpyDeviceToHost, hs);
            if (e != hipSuccess) return hip_fail(e, "result copy");
        }
        return tc_event_record(s.ev, s.pipe->stream());
    }

    // Waits for a launched batch and publishes its results.
    void finish(Slot& s, Batch& b) {
        tc_status e = b.status;
        if (e == TC_SUCCESS && !b.reqs.empty()) e = tc_event_sync(s.ev);
        const double tc0 = now();
        if (e != TC_SUCCESS) tc_stream_sync(s.pipe->stream());
        // a slot keeps its buffers for the next batch unless they outgrew its
        // share of the budget (a request larger than the share runs alone and
        // its buffers go with it)
        const uint64_t held = s.pipe->device_bytes();
        if (held > slot_budget()) s.pipe->release_buffers();
        std::lock_guard<std::mutex> lk(mu);
        if (!b.reqs.empty()) st.device_s += tc0 - b.t_launch;
        st.slot_device_peak = std::max<uint64_t>(st.slot_device_peak, held);
        if (e == TC_SUCCESS)
            for (size_t i = 0; i < b.reqs.size(); i++) {
                Request* r = b.reqs[i];
                r->pin = (const uint8_t*)s.pinned[b.buf] + b.offs[i];
                s.holders[b.buf].push_back(r->id);
            }
        const double td = now();
        st.copy_s += td - tc0;
        for (Request* r : b.rejected) {
            r->info.done_s = td;
            r->info.status = TC_ERR_UNSUPPORTED;
            r->done = true;
            st.failed++;
        }
        for (Request* r : b.reqs) {
            r->info.done_s = td;
            r->info.status = e;
            r->done = true;
            if (e == TC_SUCCESS) {
                st.completed++;
                st.pixels += (uint64_t)r->img.width() * r->img.height();
            } else {
                st.failed++;
            }
        }
        batch_done(b);
        cv_done.notify_all();
    }

    uint64_t slot_budget() const { return cfg.memory_budget / (uint64_t)std::max(1, cfg.slots); }

    // Takes the next batch from the queue into b (caller holds mu, queue not
    // empty): FIFO while the batch fits the slot's memory share and the
    // per-batch byte cap (which also bounds the staging buffers, so that they
    // stop growing once warm), and at most this worker's share of the queued
    // codestream bytes when other slots are idle (so that large requests
    // spread over slots and one slot's host work overlaps another's device
    // work).
    void admit(Slot& s, Batch& b) {
        const uint64_t batch_budget = (uint64_t)((double)slot_budget() / kSlotGrowth);
        uint64_t queued = 0;
        for (const Request* r : queue) queued += r->cs_len;
        const uint64_t share = std::min<uint64_t>(queued / (uint64_t)(idle_workers + 1), cfg.max_batch_bytes);
        uint64_t fp = 0, taken = 0;
        while (!queue.empty() && (int)b.reqs.size() < cfg.max_batch) {
            Request* r = queue.front();
            if (!b.reqs.empty() &&
                (fp + r->footprint > batch_budget || taken >= share || taken + r->cs_len > cfg.max_batch_bytes))
                break;
            fp += r->footprint;
            taken += r->cs_len;
            b.reqs.push_back(r);
            queue.pop_front();
        }
        const uint64_t seq = ++batch_seq;
        const double t = now();
        for (Request* r : b.reqs) {
            r->info.start_s = t;
            r->info.batch = seq;
            r->info.batch_size = (int32_t)b.reqs.size();
        }
        b.footprint = fp;  // charged to inflight_bytes while the batch holds the slot's device buffers
        if (busy_slots++ == 0) busy_since = t;
        st.batches++;
        if (!queue.empty()) cv_work.notify_one();  // another slot can start on the rest
        s.cur ^= 1;
        b.buf = s.cur;
    }

    // Results still unread in pinned buffer k leave it before the buffer is
    // written again (caller does not hold mu): they are copied out without
    // the lock, and readers of these requests wait for the copy.  This runs
    // when the batch that reuses the buffer launches, one whole batch after
    // the results were published, so readers have usually taken them.
    void vacate(Slot& s, int k) {
        std::unique_lock<std::mutex> lk(mu);
        std::vector<Request*> unread;
        for (uint64_t id : s.holders[k]) {
            auto it = reqs.find(id);
            if (it != reqs.end() && it->second->pin) unread.push_back(it->second.get());
        }
        s.holders[k].clear();
        if (unread.empty()) return;
        for (Request* r : unread) r->moving = true;
        lk.unlock();
        for (Request* r : unread) r->out.assign(r->pin, r->pin + r->out_bytes);
        lk.lock();
        for (Request* r : unread) {
            r->pin = nullptr;
            r->moving = false;
        }
        cv_done.notify_all();
    }

    // One worker per slot.  Its batches are pipelined: while batch n runs on
    // the slot's stream, the worker admits batch n+1 and does its host work
    // (parse, plan, gather), then waits for batch n, publishes it and
    // launches n+1.  So the slot's device stream is fed without a gap for
    // host work, and several slots still overlap with each other.
    void work(Slot& s) {
        std::unique_ptr<Batch> running;  // launched, not yet finished
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            if (!running) {
                idle_workers++;
                cv_work.wait(lk, [&] { return stop || !queue.empty(); });
                idle_workers--;
                if (queue.empty()) break;  // stopping and drained
            }
            std::unique_ptr<Batch> next;
            if (!queue.empty()) {
                next.reset(new Batch());
                admit(s, *next);
            }
            lk.unlock();
            if (next) next->status = plan_host(s, *next);
            if (running) {
                finish(s, *running);
                running.reset();
            }
            lk.lock();
            if (next) {
                // a slot's batches take its device buffers one after the other
                inflight_bytes += next->footprint;
                st.peak_bytes = std::max(st.peak_bytes, inflight_bytes);
                lk.unlock();
                if (next->status == TC_SUCCESS) next->status = launch(s, *next);
                running = std::move(next);
                lk.lock();
            }
        }
        lk.unlock();
        if (running) finish(s, *running);
    }
};

extern "C" {

tc_status tc_server_create(tc_device dev, const tc_server_config* cfg, tc_server* out) {
    if (!dev || !out) return fail(TC_ERR_INVALID_ARG, "tc_server_create: null");
    auto s = new tc_server_s();
    s->dev = dev;
    s->cfg.slots = cfg && cfg->slots > 0 ? cfg->slots : 2;
    s->cfg.max_batch = cfg && cfg->max_batch > 0 ? cfg->max_batch : 256;
    s->cfg.memory_budget = cfg && cfg->memory_budget ? cfg->memory_budget : (32ull << 30);
    s->cfg.max_batch_bytes = cfg && cfg->max_batch_bytes ? cfg->max_batch_bytes : (1ull << 30);
    s->cfg.bpw_small = cfg ? cfg->bpw_small : 0;
    s->cfg.bpw_large = cfg ? cfg->bpw_large : 0;
    s->cfg.dwt_shape = cfg ? cfg->dwt_shape : 0;
    s->cfg.overlap_kernels = cfg ? cfg->overlap_kernels : 0;
    std::memset(&s->st, 0, sizeof s->st);
    s->t0 = std::chrono::steady_clock::now();
    s->slots.resize(s->cfg.slots);
    for (Slot& sl : s->slots) {
        sl.pipe.reset(new Pipeline(dev));
        if (!sl.pipe->stream() || tc_event_create(&sl.ev) != TC_SUCCESS || tc_event_create(&sl.kdone) != TC_SUCCESS ||
            tc_event_create(&sl.up_done) != TC_SUCCESS) {
            delete s;
            return fail(TC_ERR_DEVICE, "tc_server_create: stream/event creation failed");
        }
        sl.pipe->set_growth(kSlotGrowth);
        Tuning t;
        t.bpw32 = s->cfg.bpw_small;
        t.bpw64 = s->cfg.bpw_large;
        t.dwt_shape = s->cfg.dwt_shape;
        sl.pipe->set_tuning(t);
    }
    for (Slot& sl : s->slots) sl.worker = std::thread([s, &sl]() { s->work(sl); });
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
    for (Slot& sl : s->slots)
        if (sl.worker.joinable()) sl.worker.join();
    for (Slot& sl : s->slots) {
        if (sl.pipe) tc_stream_sync(sl.pipe->stream());
        tc_event_destroy(sl.ev);
        tc_event_destroy(sl.kdone);
        for (void* p : sl.pinned)
            if (p) tc_host_free_pinned(p);
        if (sl.staging) tc_host_free_pinned(sl.staging);
        tc_event_destroy(sl.up_done);
    }
    delete s;
    return TC_SUCCESS;
}

static tc_status submit(tc_server s, const uint8_t* cs, size_t len, uint64_t* id, bool copy) {
    if (!s || !cs || !id) return fail(TC_ERR_INVALID_ARG, "tc_server_submit: null");
    std::unique_ptr<Request> r(new Request());
    std::memset(&r->info, 0, sizeof r->info);
    if (copy) {
        r->data.assign(cs, cs + len);
        r->cs = r->data.data();
    } else {
        r->cs = cs;
    }
    r->cs_len = len;
    std::string err;
    Status st = probe_codestream(r->cs, r->cs_len, r->img, err);
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

tc_status tc_server_submit(tc_server s, const uint8_t* cs, size_t len, uint64_t* id) {
    return submit(s, cs, len, id, true);
}

tc_status tc_server_submit_ref(tc_server s, const uint8_t* cs, size_t len, uint64_t* id) {
    return submit(s, cs, len, id, false);
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
    std::unique_ptr<Request> dead;  // destroyed after the lock is released
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return fail(TC_ERR_NOT_FOUND, "tc_server_result: unknown request");
    if (out) {
        while (it->second->moving) {
            s->cv_done.wait(lk);
            it = s->reqs.find(id);
            if (it == s->reqs.end()) return fail(TC_ERR_NOT_FOUND, "tc_server_result: unknown request");
        }
    }
    Request* r = it->second.get();
    if (!r->done) return fail(TC_ERR_BUSY, "tc_server_result: not finished");
    if (info) *info = r->info;
    if (!out) return TC_SUCCESS;
    if (r->info.status != TC_SUCCESS) {
        tc_status st = r->info.status;
        dead = std::move(it->second);
        s->reqs.erase(it);
        lk.unlock();
        return fail(st, "request failed");
    }
    if (cap < r->out_bytes) return fail(TC_ERR_CAPACITY, "tc_server_result: output buffer too small");
    // a result still in pinned memory is copied under the lock: the slot may
    // reuse that buffer as soon as the request is gone
    std::memcpy(out, r->pin ? r->pin : r->out.data(), r->out_bytes);
    dead = std::move(it->second);
    s->reqs.erase(it);
    lk.unlock();
    return TC_SUCCESS;
}

tc_status tc_server_release(tc_server s, uint64_t id) {
    if (!s) return fail(TC_ERR_INVALID_ARG, "tc_server_release: null");
    std::unique_ptr<Request> dead;  // its codestream copy, parse and result are freed after the lock
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return fail(TC_ERR_NOT_FOUND, "tc_server_release: unknown request");
    if (!it->second->done) return fail(TC_ERR_BUSY, "tc_server_release: not finished");
    while (it->second->moving) {
        s->cv_done.wait(lk);
        it = s->reqs.find(id);
        if (it == s->reqs.end()) return fail(TC_ERR_NOT_FOUND, "tc_server_release: unknown request");
    }
    dead = std::move(it->second);
    s->reqs.erase(it);
    lk.unlock();
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

tc_status tc_server_reset_stats(tc_server s) {
    if (!s) return fail(TC_ERR_INVALID_ARG, "tc_server_reset_stats: null");
    std::lock_guard<std::mutex> lk(s->mu);
    std::memset(&s->st, 0, sizeof s->st);
    s->st.peak_bytes = s->inflight_bytes;
    if (s->busy_since >= 0) s->busy_since = s->now();
    return TC_SUCCESS;
}

double tc_server_now(tc_server s) { return s ? s->now() : 0.0; }

}  // extern "C"
