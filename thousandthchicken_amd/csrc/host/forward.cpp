// GPU forward transform (encoder front end): plans the per-level job lists of
// a single-tile image and drives the fdwt launches (include/tc_forward.h).
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../../include/tc_forward.h"
#include "../tc_internal.hpp"
#include "../tc_kernels.h"
#include "forward.hpp"

namespace tc {
int idwt_tile_w(int shape);
int idwt_tile_h(int shape);
int idwt_shape_count();

namespace {
inline int cdiv(int v, int k) { return (int)(((int64_t)v + ((int64_t)1 << k) - 1) >> k); }
inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }
}  // namespace

// Level n >= first of a forward DWT of one w x h plane at origin 0.  Level
// `first` reads the raster at (in_off, in_pitch); each later level reads the
// LL its predecessor left in a work buffer (work[0] and work[1] alternate);
// the last level's LL and every level's HL/LH/HH go to the plane in Mallat
// layout (pitch w).
void plan_forward_levels(int w, int h, int levels, int first, uint64_t in_off, uint32_t in_pitch, uint64_t plane,
                         const uint64_t work[2], int TX, int TY, std::vector<std::vector<tc_fdwt_job>>& out) {
    out.resize(levels + 1);
    uint64_t src = in_off;
    uint32_t src_pitch = in_pitch;
    for (int n = first; n <= levels; n++) {
        const int rw = cdiv(w, n - 1), rh = cdiv(h, n - 1);
        const int lw = (rw + 1) / 2, lh = (rh + 1) / 2;
        tc_fdwt_job j{};
        j.in_off = src;
        j.in_pitch = src_pitch;
        j.u0 = 0;
        j.u1 = rw;
        j.v0 = 0;
        j.v1 = rh;
        j.hl_off = plane + (uint64_t)lw;
        j.lh_off = plane + (uint64_t)lh * w;
        j.hh_off = plane + (uint64_t)lh * w + lw;
        j.band_pitch = (uint32_t)w;
        if (n == levels) {
            j.ll_off = plane;
            j.ll_pitch = (uint32_t)w;
        } else {
            j.ll_off = work[(n - first) & 1];
            j.ll_pitch = (uint32_t)lw;
        }
        j.tiles_x = (rw + TX - 1) / TX;
        out[n].push_back(j);
        src = j.ll_off;
        src_pitch = j.ll_pitch;
    }
}

}  // namespace tc

using namespace tc;

struct tc_forward_s {
    tc_device dev = nullptr;
    tc_stream s = nullptr;
    int w = 0, h = 0, nc = 0, bits = 0, levels = 0, irr = 0, mct = 0, shape = 0, prepared = 0;
    tc_buf img = nullptr, planes = nullptr, work = nullptr, plan = nullptr;
    struct Launch {
        bool first;
        size_t jobs_off, map_off;
        int nwg;
    };
    std::vector<Launch> launches;
    void* base = nullptr;
};

namespace {
uint64_t eoff(const void* base, const void* p) { return (uint64_t)((const char*)p - (const char*)base) / 4; }

tc_status realloc_buf(tc_device dev, tc_buf& b, size_t bytes) {
    if (b && tc_buf_size(b) >= bytes) return TC_SUCCESS;
    tc_buf_free(b);
    b = nullptr;
    return tc_buf_alloc(dev, std::max<size_t>(bytes, 256), &b);
}
}  // namespace

extern "C" {

tc_status tc_forward_create(tc_device dev, tc_forward* out) {
    if (!dev || !out) return fail(TC_ERR_INVALID_ARG, "tc_forward_create");
    tc_forward f = new tc_forward_s();
    f->dev = dev;
    tc_status s = tc_stream_create(dev, &f->s);
    if (s) {
        delete f;
        return s;
    }
    *out = f;
    return TC_SUCCESS;
}

tc_status tc_forward_destroy(tc_forward f) {
    if (!f) return TC_SUCCESS;
    if (f->s) tc_stream_sync(f->s);
    tc_buf_free(f->img);
    tc_buf_free(f->planes);
    tc_buf_free(f->work);
    tc_buf_free(f->plan);
    if (f->s) tc_stream_destroy(f->s);
    delete f;
    return TC_SUCCESS;
}

tc_status tc_forward_set_shape(tc_forward f, int shape) {
    if (!f || shape < 0 || shape >= idwt_shape_count()) return fail(TC_ERR_INVALID_ARG, "tc_forward_set_shape");
    f->shape = shape;
    return TC_SUCCESS;
}

tc_status tc_forward_prepare(tc_forward f, int w, int h, int nc, int bits, int levels, int irr, int mct) {
    if (!f || w <= 0 || h <= 0 || (nc != 1 && nc != 3) || bits < 1 || bits > 16 || levels < 0 || levels > 32 ||
        (mct && nc != 3) || (int64_t)w * h > ((int64_t)1 << 31))
        return fail(TC_ERR_INVALID_ARG, "tc_forward_prepare: bad geometry");
    f->prepared = 0;
    f->w = w;
    f->h = h;
    f->nc = nc;
    f->bits = bits;
    f->levels = levels;
    f->irr = irr ? 1 : 0;
    f->mct = mct ? 1 : 0;
    const size_t px = (size_t)w * h;
    const size_t in_bytes = bits > 8 ? 2 : 1;
    const size_t half = (size_t)((w + 1) / 2) * ((h + 1) / 2);
    tc_status s = realloc_buf(f->dev, f->img, px * nc * in_bytes);
    if (!s) s = realloc_buf(f->dev, f->planes, px * nc * 4);
    if (!s) s = realloc_buf(f->dev, f->work, 2 * (align_up(half * nc, 64) * 4) + 256);
    if (s) return s;
    // every offset is in 4-byte elements from the lower of planes and work
    void* pp = tc_buf_device_ptr(f->planes);
    void* wp = tc_buf_device_ptr(f->work);
    f->base = std::min(pp, wp);
    const int TX = idwt_tile_w(f->shape), TY = idwt_tile_h(f->shape);
    const uint64_t work0 = eoff(f->base, wp);
    const uint64_t work1 = work0 + align_up(half * nc, 64);
    // first launch: image -> level-1 subbands of every component (or the
    // shifted samples themselves when levels == 0)
    tc_fdwt_first_job fj{};
    const int lw = (w + 1) / 2, lh = (h + 1) / 2;
    for (int c = 0; c < nc; c++) {
        const uint64_t plane = eoff(f->base, pp) + (uint64_t)c * px;
        fj.shift[c] = 1 << (bits - 1);
        if (levels == 0) {
            fj.ll_off[c] = plane;
            fj.ll_pitch[c] = (uint32_t)w;
            continue;
        }
        fj.hl_off[c] = plane + lw;
        fj.lh_off[c] = plane + (uint64_t)lh * w;
        fj.hh_off[c] = plane + (uint64_t)lh * w + lw;
        fj.band_pitch[c] = (uint32_t)w;
        if (levels == 1) {
            fj.ll_off[c] = plane;
            fj.ll_pitch[c] = (uint32_t)w;
        } else {
            fj.ll_off[c] = work0 + (uint64_t)c * half;
            fj.ll_pitch[c] = (uint32_t)lw;
        }
    }
    fj.u0 = 0;
    fj.u1 = w;
    fj.v0 = 0;
    fj.v1 = h;
    fj.in_off = 0;
    fj.in_pitch = (uint32_t)(w * nc);
    fj.in_cstride = (uint32_t)nc;
    fj.tiles_x = (w + TX - 1) / TX;
    fj.wg_base = 0;
    fj.levels = levels;
    fj.mct = f->mct;
    // later levels: per component, from the LL in work buffer 0/1
    std::vector<std::vector<tc_fdwt_job>> per_level(levels + 1);
    for (int c = 0; c < nc; c++) {
        std::vector<std::vector<tc_fdwt_job>> lv;
        const uint64_t plane = eoff(f->base, pp) + (uint64_t)c * px;
        const uint64_t wk[2] = {work1 + (uint64_t)c * half, work0 + (uint64_t)c * half};
        // level 2 reads the first launch's LL in work0; its LL goes to work1, and so on
        if (levels >= 2)
            plan_forward_levels(w, h, levels, 2, work0 + (uint64_t)c * half, (uint32_t)lw, plane, wk, TX, TY, lv);
        for (int n = 2; n <= levels; n++)
            for (auto& j : lv[n]) per_level[n].push_back(j);
    }
    // one blob: first job + its map, then each level's jobs + map
    std::vector<uint8_t> blob;
    auto put = [&](const void* p, size_t n) {
        size_t off = align_up(blob.size(), 256);
        blob.resize(off + n);
        std::memcpy(blob.data() + off, p, n);
        return off;
    };
    f->launches.clear();
    {
        const int nwg = fj.tiles_x * ((h + TY - 1) / TY);
        std::vector<uint32_t> map(nwg, 0u);
        tc_forward_s::Launch L{true, put(&fj, sizeof fj), 0, nwg};
        L.map_off = put(map.data(), map.size() * 4);
        f->launches.push_back(L);
    }
    for (int n = 2; n <= levels; n++) {
        std::vector<tc_fdwt_job>& js = per_level[n];
        std::vector<uint32_t> map;
        for (size_t k = 0; k < js.size(); k++) {
            js[k].wg_base = (int32_t)map.size();
            const int ntiles = js[k].tiles_x * ((js[k].v1 - js[k].v0 + TY - 1) / TY);
            map.insert(map.end(), (size_t)ntiles, (uint32_t)k);
        }
        tc_forward_s::Launch L{false, put(js.data(), js.size() * sizeof(tc_fdwt_job)), 0, (int)map.size()};
        L.map_off = put(map.data(), map.size() * 4);
        f->launches.push_back(L);
    }
    s = realloc_buf(f->dev, f->plan, blob.size());
    if (!s) s = tc_buf_copy_h2d(f->s, f->plan, 0, blob.data(), blob.size());
    if (!s) s = tc_stream_sync(f->s);
    if (s) return s;
    f->prepared = 1;
    return TC_SUCCESS;
}

tc_status tc_forward_upload(tc_forward f, const void* pixels) {
    if (!f || !f->prepared || !pixels) return fail(TC_ERR_INVALID_ARG, "tc_forward_upload");
    size_t bytes = (size_t)f->w * f->h * f->nc * (f->bits > 8 ? 2 : 1);
    tc_status s = tc_buf_copy_h2d(f->s, f->img, 0, pixels, bytes);
    return s ? s : tc_stream_sync(f->s);
}

static tc_status run_impl(tc_forward f, tc_event after_first) {
    const uint8_t* plan = (const uint8_t*)tc_buf_device_ptr(f->plan);
    tc_status s = TC_SUCCESS;
    for (const auto& L : f->launches) {
        if (s) break;
        if (after_first && &L == &f->launches[1]) s = tc_event_record(after_first, f->s);
        if (s) break;
        if (L.first)
            s = tc_launch_fdwt_first(f->s, f->irr, f->shape, f->nc, f->bits > 8 ? 2 : 1, plan + L.jobs_off,
                                     (const uint32_t*)(plan + L.map_off), L.nwg, tc_buf_device_ptr(f->img), f->base);
        else
            s = tc_launch_fdwt_level(f->s, f->irr, f->shape, plan + L.jobs_off, (const uint32_t*)(plan + L.map_off),
                                     L.nwg, f->base);
    }
    if (!s && after_first && f->launches.size() == 1) s = tc_event_record(after_first, f->s);
    return s;
}

tc_status tc_forward_run(tc_forward f) {
    if (!f || !f->prepared) return fail(TC_ERR_INVALID_ARG, "tc_forward_run: not prepared");
    return run_impl(f, nullptr);
}

tc_status tc_forward_run_timed(tc_forward f, float* first_ms, float* total_ms) {
    if (!f || !f->prepared) return fail(TC_ERR_INVALID_ARG, "tc_forward_run_timed: not prepared");
    tc_event e[3] = {nullptr, nullptr, nullptr};
    tc_status s = TC_SUCCESS;
    for (int i = 0; i < 3 && !s; i++) s = tc_event_create(&e[i]);
    if (!s) s = tc_event_record(e[0], f->s);
    if (!s) s = run_impl(f, e[1]);
    if (!s) s = tc_event_record(e[2], f->s);
    if (!s) s = tc_event_sync(e[2]);
    float a = 0.f, b = 0.f;
    if (!s) s = tc_event_elapsed_ms(e[0], e[1], &a);
    if (!s) s = tc_event_elapsed_ms(e[0], e[2], &b);
    for (int i = 0; i < 3; i++)
        if (e[i]) tc_event_destroy(e[i]);
    if (first_ms) *first_ms = a;
    if (total_ms) *total_ms = b;
    return s;
}

tc_status tc_forward_sync(tc_forward f) {
    if (!f) return fail(TC_ERR_INVALID_ARG, "tc_forward_sync");
    return tc_stream_sync(f->s);
}

tc_status tc_forward_download(tc_forward f, void* planes) {
    if (!f || !f->prepared || !planes) return fail(TC_ERR_INVALID_ARG, "tc_forward_download");
    tc_status s = tc_buf_copy_d2h(f->s, planes, f->planes, 0, (size_t)f->w * f->h * f->nc * 4);
    return s ? s : tc_stream_sync(f->s);
}

tc_stream tc_forward_stream(tc_forward f) { return f ? f->s : nullptr; }
void* tc_forward_image_ptr(tc_forward f) { return f && f->img ? tc_buf_device_ptr(f->img) : nullptr; }
void* tc_forward_planes_ptr(tc_forward f) { return f && f->planes ? tc_buf_device_ptr(f->planes) : nullptr; }

}  // extern "C"
