// C ABI of the decode path (include/tc_decode.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../../include/tc_decode.h"
#include "../tc_internal.hpp"
#include "codestream.hpp"
#include "pipeline.hpp"

using namespace tc;

struct tc_decoder_s {
    tc_device dev;
    std::unique_ptr<Pipeline> pipe;
    std::vector<std::vector<uint8_t>> data;  // owned copies of the prepared codestreams
    std::vector<std::unique_ptr<Image>> imgs;
    int reduce = 0;  // finest DWT levels not reconstructed
    int region[4] = {0, 0, 0, 0};  // x, y, w, h of the output window (w or h 0: whole image)
};

namespace {

tc_status to_status(Status s) {
    switch (s) {
    case Status::Ok: return TC_SUCCESS;
    case Status::InvalidArgument: return TC_ERR_INVALID_ARG;
    case Status::Truncated: return TC_ERR_TRUNCATED;
    case Status::BadMarker: return TC_ERR_BAD_MARKER;
    case Status::Unsupported: return TC_ERR_UNSUPPORTED;
    case Status::BadGeometry: return TC_ERR_GEOMETRY;
    case Status::OutOfMemory: return TC_ERR_OUT_OF_MEMORY;
    case Status::Device: return TC_ERR_DEVICE;
    case Status::Capacity: return TC_ERR_CAPACITY;
    }
    return TC_ERR_DEVICE;
}

tc_status report(Status s, const std::string& err) {
    if (s == Status::Ok) return TC_SUCCESS;
    set_error(err.empty() ? status_name(s) : err);
    return to_status(s);
}

void fill_info(const Image& im, tc_image_info* info) {
    std::memset(info, 0, sizeof *info);
    info->x0 = im.wx0;
    info->y0 = im.wy0;
    info->width = im.out_width();
    info->height = im.out_height();
    info->ncomp = im.ncomp;
    info->max_prec = im.max_prec();
    info->bytes_per_sample = info->max_prec > 8 ? 2 : 1;
    info->ntx = im.ntx;
    info->nty = im.nty;
    info->nlayers = im.nlayers;
    info->mct = im.mct;
    info->progression = im.progression;
    if (!im.tiles.empty() && !im.tiles[0].comps.empty()) {
        const TileComp& tc = im.tiles[0].comps[0];
        info->reversible = tc.reversible;
        info->nlevels = tc.nlev;
        const Band& b = tc.res[tc.nlev].band[0];
        info->cbw = 1 << b.xcb;
        info->cbh = 1 << b.ycb;
    }
    for (int c = 0; c < 4 && c < im.ncomp; c++) {
        info->prec[c] = im.prec[c];
        info->sgnd[c] = im.sgnd[c];
        int x0, y0, x1, y1;
        if (im.planar()) {
            im.comp_rect(c, x0, y0, x1, y1);
        } else {
            x0 = im.wx0, y0 = im.wy0, x1 = im.wx1, y1 = im.wy1;
        }
        info->comp_width[c] = x1 - x0;
        info->comp_height[c] = y1 - y0;
    }
    info->planar = im.planar() ? 1 : 0;
    info->jp2 = im.jp2;
    info->colourspace = im.colourspace;
}

uint32_t fnv1a(const uint8_t* p, size_t n, uint32_t h) {
    for (size_t i = 0; i < n; i++) {
        h ^= p[i];
        h *= 16777619u;
    }
    return h;
}

// ---- working memory of tc_decode_noalloc ----
//
// One static device array serves every no-allocation decode on a GPU.  Host
// code must not pass the array's name to a kernel or a copy: on the host it is
// only a stand-in for the symbol.  Its device address, which differs per GPU,
// comes from hipGetSymbolAddress once per device (with that device current)
// and is the only handle used afterwards.
constexpr size_t kNoallocBytes = TC_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

}  // namespace

static __device__ __align__(256) uint8_t scratch[kNoallocBytes];

namespace {

struct NoallocDevice {
    std::once_flag once;
    hipError_t err = hipSuccess;
    uint8_t* ptr = nullptr;  // this GPU's copy of `scratch`
    std::mutex use;          // held for a whole call: calls share the array
};
NoallocDevice g_noalloc[kMaxDevices];

uint64_t align256(uint64_t v) { return (v + 255) / 256 * 256; }

// Makes `ordinal` the current device for the lifetime of the guard.
struct DeviceGuard {
    int prev = -1;
    hipError_t err = hipSuccess;
    explicit DeviceGuard(int ordinal) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        err = hipSetDevice(ordinal);
    }
    ~DeviceGuard() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

}  // namespace

extern "C" {

tc_status tc_image_probe(const uint8_t* cs, size_t len, tc_image_info* info) {
    if (!cs || !info) return report(Status::InvalidArgument, "tc_image_probe: null");
    Image im;
    std::string err;
    Status s = probe_codestream(cs, len, im, err);
    if (s != Status::Ok) return report(s, err);
    fill_info(im, info);
    return TC_SUCCESS;
}

tc_status tc_image_probe_reduced(const uint8_t* cs, size_t len, int reduce, tc_image_info* info) {
    if (!cs || !info) return report(Status::InvalidArgument, "tc_image_probe_reduced: null");
    Image im;
    std::string err;
    Status s = probe_codestream(cs, len, im, err);
    if (s == Status::Ok) s = reduce_image(im, reduce, err);
    if (s != Status::Ok) return report(s, err);
    fill_info(im, info);
    return TC_SUCCESS;
}

tc_status tc_image_list_cblks(const uint8_t* cs, size_t len, tc_cblk_info* out, int max, int* count) {
    if (!cs || !count) return report(Status::InvalidArgument, "tc_image_list_cblks: null");
    Image im;
    std::string err;
    Status s = parse_codestream(cs, len, im, err, 8);
    if (s != Status::Ok) return report(s, err);
    // crc over each block's segments in order
    std::vector<uint32_t> crc(im.cblks.size(), 2166136261u);
    for (const Segment& g : im.segs) crc[g.cblk] = fnv1a(im.cs + g.off, g.len, crc[g.cblk]);
    *count = (int)im.cblks.size();
    for (size_t k = 0; out && k < im.cblks.size() && (int)k < max; k++) {
        const CodeBlock& cb = im.cblks[k];
        const Band& B = im.tiles[cb.tile].comps[cb.comp].res[cb.res].band[cb.band];
        tc_cblk_info& o = out[k];
        o.tile = cb.tile;
        o.comp = cb.comp;
        o.res = cb.res;
        o.band = B.orient;
        o.x0 = cb.x0;
        o.y0 = cb.y0;
        o.x1 = cb.x1;
        o.y1 = cb.y1;
        o.npasses = cb.npasses;
        o.zero_bp = cb.zbp;
        o.mb = B.mb;
        o.nbytes = (int32_t)cb.nbytes;
        o.crc = crc[k];
    }
    return TC_SUCCESS;
}

tc_status tc_image_cblk_data(const uint8_t* cs, size_t len, int index, uint8_t* out, size_t cap, size_t* nbytes) {
    if (!cs || !nbytes || index < 0) return report(Status::InvalidArgument, "tc_image_cblk_data: args");
    Image im;
    std::string err;
    Status s = parse_codestream(cs, len, im, err);
    if (s != Status::Ok) return report(s, err);
    if ((size_t)index >= im.cblks.size()) return report(Status::InvalidArgument, "tc_image_cblk_data: index");
    size_t n = 0;
    for (const Segment& g : im.segs)
        if (g.cblk == (uint32_t)index) {
            if (out && n + g.len <= cap) std::memcpy(out + n, im.cs + g.off, g.len);
            n += g.len;
        }
    *nbytes = n;
    if (out && n > cap) return report(Status::InvalidArgument, "tc_image_cblk_data: buffer too small");
    return TC_SUCCESS;
}

tc_status tc_decoder_create(tc_device dev, tc_decoder* out) {
    if (!dev || !out) return report(Status::InvalidArgument, "tc_decoder_create: null");
    auto d = new tc_decoder_s();
    d->dev = dev;
    d->pipe.reset(new Pipeline(dev));
    if (!d->pipe->stream()) {
        delete d;
        return report(Status::Device, std::string("stream creation failed: ") + tc_last_error());
    }
    *out = d;
    return TC_SUCCESS;
}

tc_status tc_decoder_destroy(tc_decoder d) {
    delete d;
    return TC_SUCCESS;
}

static bool valid_bpw(int b) {
    return b == 0 || b == 1 || b == TC_BPW_LOCKSTEP || b == TC_BPW_LOCKSTEP32 || b == TC_BPW_LOCKSTEP16;
}

tc_status tc_decoder_set_blocks_per_wave(tc_decoder d, int small, int bpw) {
    if (!d || !valid_bpw(bpw))
        return report(Status::InvalidArgument,
                      "blocks per wave must be 0 (auto), 1 (one wave per block), TC_BPW_LOCKSTEP, TC_BPW_LOCKSTEP32 "
                      "or TC_BPW_LOCKSTEP16");
    Tuning t = d->pipe->tuning();
    (small ? t.bpw32 : t.bpw64) = bpw;
    d->pipe->set_tuning(t);
    return TC_SUCCESS;
}

tc_status tc_decoder_set_reduce(tc_decoder d, int reduce) {
    if (!d || reduce < 0 || reduce > 32) return report(Status::InvalidArgument, "tc_decoder_set_reduce: 0..32");
    d->reduce = reduce;
    return TC_SUCCESS;
}

tc_status tc_decoder_set_region(tc_decoder d, int x, int y, int w, int h) {
    if (!d || x < 0 || y < 0 || w < 0 || h < 0) return report(Status::InvalidArgument, "tc_decoder_set_region");
    d->region[0] = x;
    d->region[1] = y;
    d->region[2] = w;
    d->region[3] = h;
    return TC_SUCCESS;
}

tc_status tc_decoder_set_host_output(tc_decoder d, int enable) {
    if (!d) return report(Status::InvalidArgument, "tc_decoder_set_host_output: null");
    d->pipe->set_host_output(enable != 0);
    return TC_SUCCESS;
}

tc_status tc_decoder_set_t1_kernel(tc_decoder d, int kind) {
    if (!d || kind < 0 || kind > 1) return report(Status::InvalidArgument, "t1 kernel must be 0 (stripe) or 1 (row)");
    Tuning t = d->pipe->tuning();
    t.t1_kernel = kind;
    d->pipe->set_tuning(t);
    return TC_SUCCESS;
}

namespace {

Status reprepare(tc_decoder d, std::string& err) {
    std::vector<const Image*> v;
    for (auto& im : d->imgs) v.push_back(im.get());
    return d->pipe->prepare(v, err);
}

void to_tuning(const tc_tuning& in, Tuning& t) {
    t.bpw32 = in.bpw_small;
    t.bpw64 = in.bpw_large;
    t.t1_kernel = in.t1_kernel;
    t.dwt_shape = in.dwt_shape;
    t.heavy32 = in.heavy_small;
    t.heavy64 = in.heavy_large;
    t.heavy_waves_per_cu = in.heavy_waves > 0 ? in.heavy_waves : 4;
}

}  // namespace

tc_status tc_decoder_get_tuning(tc_decoder d, tc_tuning* out) {
    if (!d || !out) return report(Status::InvalidArgument, "tc_decoder_get_tuning: null");
    const Tuning& t = d->pipe->tuning();
    *out = tc_tuning{};
    out->bpw_small = t.bpw32;
    out->bpw_large = t.bpw64;
    out->t1_kernel = t.t1_kernel;
    out->dwt_shape = t.dwt_shape;
    out->heavy_small = t.heavy32;
    out->heavy_large = t.heavy64;
    out->heavy_waves = t.heavy_waves_per_cu;
    return TC_SUCCESS;
}

tc_status tc_decoder_set_tuning(tc_decoder d, const tc_tuning* in) {
    int nshape = 0;
    tc_idwt_tile_shape(0, nullptr, nullptr, &nshape);
    if (!d || !in || !valid_bpw(in->bpw_small) || !valid_bpw(in->bpw_large) || in->t1_kernel < 0 || in->t1_kernel > 1 ||
        in->dwt_shape < 0 || in->dwt_shape >= nshape || in->heavy_small < -1 || in->heavy_large < -1 ||
        in->heavy_waves < 0 || in->heavy_waves > 32)
        return report(Status::InvalidArgument, "tc_decoder_set_tuning: value out of range");
    Tuning t = d->pipe->tuning();
    to_tuning(*in, t);
    d->pipe->set_tuning(t);
    return TC_SUCCESS;
}

// Coordinate search over the prepared batch: blocks-per-wave of each block
// class (the other class held), then the inverse-DWT tile shape, each
// candidate timed with the pipeline's stage events (median of reps runs).
tc_status tc_decoder_autotune(tc_decoder d, int reps, tc_tuning* best) {
    if (!d || reps <= 0) return report(Status::InvalidArgument, "tc_decoder_autotune: args");
    if (d->imgs.empty()) return report(Status::InvalidArgument, "tc_decoder_autotune: nothing prepared");
    std::string err;
    auto measure = [&](const Tuning& t, float* t1, float* dwt) -> Status {
        d->pipe->set_tuning(t);
        Status s = reprepare(d, err);
        if (s != Status::Ok) return s;
        std::vector<float> a, b;
        for (int r = 0; r <= reps; r++) {
            if (d->pipe->run(true) != Status::Ok || d->pipe->sync() != Status::Ok) return Status::Device;
            StageTimes st = d->pipe->times();
            if (r == 0) continue;  // warm-up
            a.push_back(st.t1_ms);
            b.push_back(st.dwt_ms + st.final_ms);
        }
        std::nth_element(a.begin(), a.begin() + a.size() / 2, a.end());
        std::nth_element(b.begin(), b.begin() + b.size() / 2, b.end());
        *t1 = a[a.size() / 2];
        *dwt = b[b.size() / 2];
        return Status::Ok;
    };
    // start from what the prepared plan uses; a class without blocks is not tuned
    const Tuning start = d->pipe->tuning();
    Tuning cur = start;
    const bool has[2] = {d->pipe->blocks_per_wave(0) != 0, d->pipe->blocks_per_wave(1) != 0};
    const bool styled[2] = {d->pipe->styled(0), d->pipe->styled(1)};
    cur.bpw32 = has[0] ? d->pipe->blocks_per_wave(0) : 1;
    cur.bpw64 = has[1] ? d->pipe->blocks_per_wave(1) : 1;
    // on any failure the batch is prepared again with the tuning it came with,
    // so the decoder is left as it was (or reports why it cannot be)
    auto fail = [&](Status s) -> tc_status {
        const std::string why = err.empty() ? "autotune run failed" : err;
        d->pipe->set_tuning(start);
        std::string e2;
        Status r = reprepare(d, e2);
        return report(s, r == Status::Ok ? why : why + "; preparing again with the starting tuning failed too: " + e2);
    };
    float t1 = 0, dw = 0, best_t1, best_dwt;
    // A candidate replaces the incumbent only when it is faster by more than
    // the run-to-run spread of one measurement (about 1 %): near-equal
    // packings otherwise win by noise, and a heavy-block split chosen that
    // way costs its extra launch for nothing.
    constexpr float kBetter = 0.99f;
    Status s = measure(cur, &best_t1, &best_dwt);
    if (s != Status::Ok) return fail(s);
    const int cand[] = {1, TC_BPW_LOCKSTEP, TC_BPW_LOCKSTEP32, TC_BPW_LOCKSTEP16};
    // The two classes' launches run side by side and the stage lasts as long
    // as the longer one, so the packings are searched jointly: tuned one
    // class at a time, a class whose launch is the shorter one sees only
    // noise, and the search stopped at pairs like (16, 32 blocks per wave)
    // on the mixed batch, 11 % slower than (32, 16).
    std::vector<int> c32, c64;
    for (int b : cand) {
        // styled code-blocks decode only under a lockstep packing; the planner
        // would run any other request as lockstep anyway
        if (has[0] && !(styled[0] && b > 0)) c32.push_back(b);
        if (has[1] && !(styled[1] && b > 0)) c64.push_back(b);
    }
    if (c32.empty()) c32.push_back(cur.bpw32);
    if (c64.empty()) c64.push_back(cur.bpw64);
    struct Cand {
        Tuning t;
        float t1;
    };
    std::vector<Cand> tried;
    for (int b32 : c32)
        for (int b64 : c64) {
            Tuning t = cur;
            t.bpw32 = b32;
            t.bpw64 = b64;
            if (b32 == cur.bpw32 && b64 == cur.bpw64) {
                tried.push_back({t, best_t1});
                continue;
            }
            s = measure(t, &t1, &dw);
            if (s != Status::Ok) return fail(s);
            tried.push_back({t, t1});
        }
    // the three fastest measured again, and the incumbent kept unless beaten
    // by more than the spread
    std::sort(tried.begin(), tried.end(), [](const Cand& x, const Cand& y) { return x.t1 < y.t1; });
    for (size_t k = 0; k < tried.size() && k < 3; k++) {
        s = measure(tried[k].t, &t1, &dw);
        if (s != Status::Ok) return fail(s);
        tried[k].t1 = std::min(tried[k].t1, t1);
    }
    std::sort(tried.begin(), tried.begin() + std::min<size_t>(3, tried.size()),
              [](const Cand& x, const Cand& y) { return x.t1 < y.t1; });
    if (!tried.empty() && tried[0].t1 < best_t1 * kBetter) {
        cur = tried[0].t;
        best_t1 = tried[0].t1;
    }
    // heavy-block thresholds of a lockstep class (estimated symbols; 0: no split)
    for (int cls = 0; cls < 2; cls++) {
        if (!has[cls] || (cls == 0 ? cur.bpw32 : cur.bpw64) >= 0) continue;
        static const int cand_small[] = {0, 3000, 4500, 6000, 8000};
        static const int cand_large[] = {0, 10000, 13000, 16000, 20000, 25000};
        const int* c = cls == 0 ? cand_small : cand_large;
        const int nc = cls == 0 ? 5 : 6;
        int& slot_cur = cls == 0 ? cur.heavy32 : cur.heavy64;
        const int now = slot_cur < 0 ? default_heavy(cls == 0) : slot_cur;
        for (int q = 0; q < nc; q++) {
            if (c[q] == now) continue;
            Tuning t = cur;
            (cls == 0 ? t.heavy32 : t.heavy64) = c[q];
            s = measure(t, &t1, &dw);
            if (s != Status::Ok) return fail(s);
            if (t1 < best_t1 * kBetter) {
                // one of several candidates beating the incumbent once is
                // often noise: it must do so again
                float again = 0, dw2 = 0;
                s = measure(t, &again, &dw2);
                if (s != Status::Ok) return fail(s);
                t1 = std::max(t1, again);
            }
            if (t1 < best_t1 * kBetter) {
                best_t1 = t1;
                cur = t;
            }
        }
    }
    int nshape = 0;
    tc_idwt_tile_shape(0, nullptr, nullptr, &nshape);
    for (int sh = 0; sh < nshape; sh++) {
        if (sh == cur.dwt_shape) continue;
        Tuning t = cur;
        t.dwt_shape = sh;
        s = measure(t, &t1, &dw);
        if (s != Status::Ok) return fail(s);
        if (dw < best_dwt * kBetter) {
            float t1b = 0, again = 0;
            s = measure(t, &t1b, &again);
            if (s != Status::Ok) return fail(s);
            dw = std::max(dw, again);
        }
        if (dw < best_dwt * kBetter) {
            best_dwt = dw;
            cur = t;
        }
    }
    d->pipe->set_tuning(cur);
    s = reprepare(d, err);
    if (s != Status::Ok) return fail(s);
    if (best) {
        tc_decoder_get_tuning(d, best);
        // the packings the plan runs (a class without blocks keeps its setting)
        if (has[0]) best->bpw_small = d->pipe->blocks_per_wave(0);
        if (has[1]) best->bpw_large = d->pipe->blocks_per_wave(1);
        best->heavy_small = d->pipe->plan().heavy[0];
        best->heavy_large = d->pipe->plan().heavy[1];
        best->t1_ms = best_t1;
        best->dwt_ms = best_dwt;
    }
    return TC_SUCCESS;
}

tc_status tc_decoder_prepare(tc_decoder d, int n, const uint8_t* const* cs, const size_t* lens) {
    if (!d || n < 0 || (n > 0 && (!cs || !lens))) return report(Status::InvalidArgument, "tc_decoder_prepare: args");
    d->pipe->sync();
    // the old plan goes first: a prepare that fails leaves nothing to run
    d->pipe->invalidate();
    d->data.clear();
    d->imgs.clear();
    for (int i = 0; i < n; i++)
        if (!cs[i]) return report(Status::InvalidArgument, "tc_decoder_prepare: null codestream");
    d->data.resize(n);
    d->imgs.resize(n);
    std::vector<Status> st(n, Status::Ok);
    std::vector<std::string> errs(n);
    // copy and parse the codestreams on up to eight threads (tier-2 of each
    // image is independent); the resolution / region rewrites follow
    auto work = [&](int k, int nt) {
        // the threads not needed across images read the tiles of each image
        const int tile_threads = std::max(1, 8 / nt);
        for (int i = k; i < n; i += nt) {
            d->data[i].assign(cs[i], cs[i] + lens[i]);
            d->imgs[i].reset(new Image());
            Image& im = *d->imgs[i];
            std::string& err = errs[i];
            Status s = parse_codestream(d->data[i].data(), d->data[i].size(), im, err, tile_threads);
            if (s == Status::Ok) s = reduce_image(im, d->reduce, err);
            if (s == Status::Ok && d->region[2] > 0 && d->region[3] > 0)
                s = window_image(im, d->region[0], d->region[1], d->region[2], d->region[3], err);
            if (s == Status::Ok && im.planar() && im.ncomp > 4) {
                s = Status::Unsupported;
                err = "sub-sampled images of more than four components are not supported";
            }
            st[i] = s;
        }
    };
    size_t bytes = 0;
    for (int i = 0; i < n; i++) bytes += lens[i];
    const int nt = (int)std::min<size_t>({(size_t)8, (size_t)std::max(n, 1), 1 + bytes / (2u << 20)});
    if (nt <= 1) {
        work(0, 1);
    } else {
        std::vector<std::thread> th;
        for (int k = 1; k < nt; k++) th.emplace_back(work, k, nt);
        work(0, nt);
        for (auto& t : th) t.join();
    }
    std::vector<const Image*> v;
    for (int i = 0; i < n; i++) {
        if (st[i] != Status::Ok) return report(st[i], "image " + std::to_string(i) + ": " + errs[i]);
        v.push_back(d->imgs[i].get());
    }
    std::string err;
    return report(d->pipe->prepare(v, err), err);
}

tc_status tc_decoder_run(tc_decoder d, int timed) {
    if (!d) return report(Status::InvalidArgument, "tc_decoder_run: null");
    if (!d->pipe->ready())
        return report(Status::InvalidArgument, "tc_decoder_run: nothing prepared (the last prepare failed or none ran)");
    Status s = d->pipe->run(timed != 0);
    return report(s, s == Status::Ok ? "" : std::string("launch failed: ") + tc_last_error());
}

tc_status tc_decoder_sync(tc_decoder d) {
    if (!d) return report(Status::InvalidArgument, "tc_decoder_sync: null");
    Status s = d->pipe->sync();
    return report(s, s == Status::Ok ? "" : std::string("stream sync failed: ") + tc_last_error());
}

tc_status tc_decoder_fetch(tc_decoder d, int i, void* out, size_t cap) {
    if (!d || !out) return report(Status::InvalidArgument, "tc_decoder_fetch: null");
    std::string err;
    return report(d->pipe->fetch(i, out, cap, err), err);
}

tc_status tc_decoder_output(tc_decoder d, int i, const void** ptr, size_t* bytes) {
    if (!d || !ptr || !bytes || i < 0 || i >= d->pipe->num_images())
        return report(Status::InvalidArgument, "tc_decoder_output: args");
    *ptr = d->pipe->output_ptr(i);
    *bytes = d->pipe->output_bytes(i);
    return TC_SUCCESS;
}

tc_status tc_decoder_decode(tc_decoder d, const uint8_t* cs, size_t len, void* out, size_t cap, tc_image_info* info) {
    if (!d || !cs || !out) return report(Status::InvalidArgument, "tc_decoder_decode: null");
    tc_status s = tc_decoder_prepare(d, 1, &cs, &len);
    if (s) return s;
    if (info) fill_info(*d->imgs[0], info);
    s = tc_decoder_run(d, 0);
    if (s) return s;
    return tc_decoder_fetch(d, 0, out, cap);
}

tc_status tc_decode_noalloc(tc_stream stream, const uint8_t* cs, size_t len, void* out_device, size_t out_cap,
                            tc_image_info* info) {
    if (!stream || !cs || !out_device) return report(Status::InvalidArgument, "tc_decode_noalloc: null argument");
    tc_device dev = tc_stream_device(stream);
    const int ord = device_ordinal(dev);
    if (ord < 0 || ord >= kMaxDevices) return report(Status::InvalidArgument, "tc_decode_noalloc: device ordinal");
    DeviceGuard guard(ord);
    if (guard.err != hipSuccess) return hip_fail(guard.err, "hipSetDevice");
    NoallocDevice& D = g_noalloc[ord];
    std::call_once(D.once, [&D] { D.err = hipGetSymbolAddress((void**)&D.ptr, HIP_SYMBOL(scratch)); });
    if (D.err != hipSuccess || !D.ptr)
        return hip_fail(D.err != hipSuccess ? D.err : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");

    Image im;
    std::string err;
    Status s = parse_codestream(cs, len, im, err, 4);
    if (s == Status::Ok && im.planar() && im.ncomp > 4) {
        s = Status::Unsupported;
        err = "sub-sampled images of more than four components are not supported";
    }
    if (s != Status::Ok) return report(s, err);
    Plan P;
    s = build_plan({&im}, Tuning(), device_cu_count(dev), P, err);
    if (s != Status::Ok) return report(s, err);
    // the array holds the upload region, then the sample planes, then the
    // tier-1 state, each 256-byte aligned
    const uint64_t smp_off = align256(P.upload_bytes());
    const uint64_t scr_off = align256(smp_off + std::max<uint64_t>(P.sample_elems, 1) * 4);
    const uint64_t need = scr_off + P.scratch_bytes;
    if (need > kNoallocBytes)
        return report(Status::Capacity, "tc_decode_noalloc: the image needs " + std::to_string(need) +
                                            " bytes of device working memory; the static array holds " +
                                            std::to_string(kNoallocBytes));
    if (out_cap < P.out_bytes)
        return report(Status::Capacity, "tc_decode_noalloc: output needs " + std::to_string(P.out_bytes) + " bytes");
    std::vector<uint8_t> stage(P.upload_bytes());
    stage_upload(P, stage.data());
    PlanRegions R;
    R.upload = D.ptr;
    R.samples = D.ptr + smp_off;
    R.out = out_device;
    R.scratch = P.scratch_bytes ? D.ptr + scr_off : nullptr;
    R.scratch_bytes = P.scratch_bytes;
    hipStream_t hs = stream_handle(stream);
    std::lock_guard<std::mutex> lock(D.use);
    hipError_t e = hipMemcpyAsync(R.upload, stage.data(), stage.size(), hipMemcpyHostToDevice, hs);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync (upload)");
    tc_status ts = launch_plan(P, R, device_cu_count(dev), stream, nullptr, nullptr);
    // the staging vector and the array stay in use until the stream drains
    e = hipStreamSynchronize(hs);
    if (ts != TC_SUCCESS) return ts == TC_ERR_INVALID_ARG ? report(Status::InvalidArgument, "launch_plan") : ts;
    if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    if (info) fill_info(im, info);
    return TC_SUCCESS;
}

tc_status tc_decoder_capture_graph(tc_decoder d) {
    if (!d) return report(Status::InvalidArgument, "tc_decoder_capture_graph: null");
    std::string err;
    return report(d->pipe->capture_graph(err), err);
}

tc_status tc_decoder_launch_graph(tc_decoder d) {
    if (!d) return report(Status::InvalidArgument, "tc_decoder_launch_graph: null");
    if (!d->pipe->ready())
        return report(Status::InvalidArgument, "tc_decoder_launch_graph: nothing prepared (the last prepare failed)");
    Status s = d->pipe->launch_graph();
    return report(s, s == Status::Ok ? "" : std::string("graph launch failed: ") + tc_last_error());
}

tc_status tc_decoder_get_stats(tc_decoder d, tc_decoder_stats* st) {
    if (!d || !st) return report(Status::InvalidArgument, "tc_decoder_get_stats: null");
    std::memset(st, 0, sizeof *st);
    StageTimes t = d->pipe->times();
    st->fill_ms = t.fill_ms;
    st->t1_ms = t.t1_ms;
    st->dwt_ms = t.dwt_ms;
    st->final_ms = t.final_ms;
    st->total_ms = t.total_ms;
    st->cblks = d->pipe->num_cblks();
    st->codeword_bytes = d->pipe->codeword_bytes();
    st->samples = d->pipe->sample_count();
    st->t1_samples = d->pipe->t1_sample_count();
    st->pixels = d->pipe->pixel_count();
    st->dwt_bytes = d->pipe->dwt_bytes();
    st->final_bytes = d->pipe->final_bytes();
    st->bpw_small = d->pipe->blocks_per_wave(0);
    st->bpw_large = d->pipe->blocks_per_wave(1);
    st->images = d->pipe->num_images();
    st->device_bytes = (int64_t)d->pipe->device_bytes();
    const Plan& P = d->pipe->plan();
    st->t1_launches = (int32_t)std::min<size_t>(P.groups.size(), kMaxT1Launches);
    for (int g = 0; g < st->t1_launches; g++) {
        st->t1_launch_class[g] = P.groups[g].small ? 0 : 1;
        st->t1_launch_bpw[g] = P.groups[g].bpw;
        st->t1_launch_blocks[g] = P.groups[g].n;
        st->t1_launch_bytes[g] = P.groups[g].bytes;
        st->t1_launch_ms[g] = t.t1_launch_ms[g];
    }
    st->heavy_small = P.heavy[0];
    st->heavy_large = P.heavy[1];
    return TC_SUCCESS;
}

tc_stream tc_decoder_stream(tc_decoder d) { return d ? d->pipe->stream() : nullptr; }

}  // extern "C"
