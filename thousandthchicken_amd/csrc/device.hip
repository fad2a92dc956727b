// C ABI device layer over the HIP runtime (see include/tc_device.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/tc_device.h"
#include "tc_kernels.h"
#include "tc_internal.hpp"

namespace tc {
hipError_t launch_t1_wave(hipStream_t s, bool narrow, int grid, const tc_cblk_desc* descs, const uint32_t* order,
                          int n, const uint8_t* arena, uint32_t arena_bytes, uint32_t* coef);
hipError_t launch_t1_stripe(hipStream_t s, bool small, int grid, const tc_cblk_desc* descs, const uint32_t* order,
                            int n, const uint8_t* arena, uint32_t arena_bytes, uint32_t* coef);
size_t t1_lockstep_scratch_bytes(bool cls32, int n, int max_passes, int bpw);
hipError_t launch_t1_lockstep(hipStream_t s, bool cls32, int bpw, const tc_cblk_desc* descs, const uint32_t* order,
                              int n, const uint8_t* arena, uint32_t arena_bytes, uint32_t* coef, void* scratch,
                              int max_passes);
int idwt_shape_count();
int idwt_tile_w(int shape);
int idwt_tile_h(int shape);
hipError_t launch_idwt_level(hipStream_t s, int shape, bool is97, const tc_dwt_job* jobs, const uint32_t* wg_job, int nwg,
                             void* base);
hipError_t launch_idwt_final(hipStream_t s, int shape, bool is97, int nc, int out_bytes, const tc_final_job* jobs,
                             const uint32_t* wg_job, int nwg, const void* base, void* out);
hipError_t launch_fdwt_level(hipStream_t s, int shape, bool is97, const tc_fdwt_job* jobs, const uint32_t* wg_job,
                             int nwg, void* base);
hipError_t launch_fdwt_first(hipStream_t s, int shape, bool is97, int nc, int in_bytes,
                             const tc_fdwt_first_job* jobs, const uint32_t* wg_job, int nwg, const void* img,
                             void* base);
hipError_t launch_fill32(hipStream_t s, uint32_t* p, uint32_t v, uint64_t n);
hipError_t launch_colour(hipStream_t s, int kind, void* c0, void* c1, void* c2, size_t n, int shift, int minv,
                         int maxv);

static thread_local std::string g_last_error;

void* host_device_ptr(void* host) {
    void* d = nullptr;
    if (hipHostGetDevicePointer(&d, host, 0) != hipSuccess) return nullptr;
    return d;
}

void set_error(const std::string& m) { g_last_error = m; }

tc_status fail(tc_status code, const std::string& m) {
    g_last_error = m;
    return code;
}

tc_status hip_fail(hipError_t e, const char* what) {
    if (e == hipSuccess) return TC_SUCCESS;
    char b[256];
    snprintf(b, sizeof b, "%s: %s", what, hipGetErrorString(e));
    g_last_error = b;
    return e == hipErrorOutOfMemory ? TC_ERR_OUT_OF_MEMORY : TC_ERR_DEVICE;
}

}  // namespace tc

using namespace tc;

struct tc_device_s {
    int ordinal;
    hipDeviceProp_t prop;
};
struct tc_stream_s {
    tc_device dev;
    hipStream_t s;
    hipGraph_t capturing;
};
struct tc_buf_s {
    tc_device dev;
    void* p;
    size_t n;
};
struct tc_event_s {
    hipEvent_t e;
};
struct tc_graph_s {
    hipGraph_t g;
    hipGraphExec_t x;
};

extern "C" {

const char* tc_last_error(void) { return g_last_error.c_str(); }

const char* tc_status_string(tc_status s) {
    switch (s) {
    case TC_SUCCESS: return "success";
    case TC_ERR_INVALID_ARG: return "invalid argument";
    case TC_ERR_TRUNCATED: return "truncated codestream";
    case TC_ERR_BAD_MARKER: return "bad marker";
    case TC_ERR_UNSUPPORTED: return "unsupported feature";
    case TC_ERR_GEOMETRY: return "bad geometry";
    case TC_ERR_OUT_OF_MEMORY: return "out of memory";
    case TC_ERR_DEVICE: return "device error";
    case TC_ERR_CAPACITY: return "capacity exceeded";
    case TC_ERR_BUSY: return "busy";
    case TC_ERR_NOT_FOUND: return "not found";
    }
    return "unknown status";
}

tc_status tc_device_count(int* count) {
    if (!count) return fail(TC_ERR_INVALID_ARG, "tc_device_count: null");
    return hip_fail(hipGetDeviceCount(count), "hipGetDeviceCount");
}

tc_status tc_device_open(int ordinal, tc_device* out) {
    if (!out) return fail(TC_ERR_INVALID_ARG, "tc_device_open: null out");
    int n = 0;
    tc_status st = hip_fail(hipGetDeviceCount(&n), "hipGetDeviceCount");
    if (st) return st;
    if (ordinal < 0 || ordinal >= n) return fail(TC_ERR_NOT_FOUND, "tc_device_open: no such device");
    tc_device d = new tc_device_s();
    d->ordinal = ordinal;
    st = hip_fail(hipGetDeviceProperties(&d->prop, ordinal), "hipGetDeviceProperties");
    if (st) {
        delete d;
        return st;
    }
    if (std::strncmp(d->prop.gcnArchName, "gfx950", 6) != 0) {
        std::string a = d->prop.gcnArchName;
        delete d;
        return fail(TC_ERR_UNSUPPORTED, "kernels are built for gfx950 only; device is " + a);
    }
    *out = d;
    return TC_SUCCESS;
}

tc_status tc_device_close(tc_device dev) {
    delete dev;
    return TC_SUCCESS;
}

tc_status tc_device_get_props(tc_device dev, tc_device_props* p) {
    if (!dev || !p) return fail(TC_ERR_INVALID_ARG, "tc_device_get_props: null");
    std::memset(p, 0, sizeof *p);
    std::snprintf(p->name, sizeof p->name, "%s", dev->prop.name);
    std::snprintf(p->arch, sizeof p->arch, "%s", dev->prop.gcnArchName);
    p->compute_units = dev->prop.multiProcessorCount;
    p->wavefront_size = dev->prop.warpSize;
    p->xcds = 8;
    p->lds_bytes_per_workgroup = dev->prop.sharedMemPerBlock;
    p->global_mem_bytes = dev->prop.totalGlobalMem;
    p->mem_base_align = 256;
    p->clock_khz = dev->prop.clockRate;
    p->memory_clock_khz = dev->prop.memoryClockRate;
    p->memory_bus_bits = dev->prop.memoryBusWidth;
    return TC_SUCCESS;
}

tc_status tc_stream_create(tc_device dev, tc_stream* out) {
    if (!dev || !out) return fail(TC_ERR_INVALID_ARG, "tc_stream_create: null");
    tc_status st = hip_fail(hipSetDevice(dev->ordinal), "hipSetDevice");
    if (st) return st;
    hipStream_t s;
    st = hip_fail(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
    if (st) return st;
    *out = new tc_stream_s{dev, s, nullptr};
    return TC_SUCCESS;
}

tc_status tc_stream_destroy(tc_stream s) {
    if (!s) return TC_SUCCESS;
    tc_status st = hip_fail(hipStreamDestroy(s->s), "hipStreamDestroy");
    delete s;
    return st;
}

tc_status tc_stream_sync(tc_stream s) {
    if (!s) return fail(TC_ERR_INVALID_ARG, "tc_stream_sync: null");
    return hip_fail(hipStreamSynchronize(s->s), "hipStreamSynchronize");
}

tc_status tc_stream_flush(tc_stream s) { return s ? TC_SUCCESS : fail(TC_ERR_INVALID_ARG, "null stream"); }

tc_status tc_buf_alloc(tc_device dev, size_t bytes, tc_buf* out) {
    if (!dev || !out) return fail(TC_ERR_INVALID_ARG, "tc_buf_alloc: null");
    tc_status st = hip_fail(hipSetDevice(dev->ordinal), "hipSetDevice");
    if (st) return st;
    void* p = nullptr;
    st = hip_fail(hipMalloc(&p, bytes ? bytes : 1), "hipMalloc");
    if (st) return st;
    *out = new tc_buf_s{dev, p, bytes};
    return TC_SUCCESS;
}

tc_status tc_buf_free(tc_buf b) {
    if (!b) return TC_SUCCESS;
    tc_status st = hip_fail(hipFree(b->p), "hipFree");
    delete b;
    return st;
}

void* tc_buf_device_ptr(tc_buf b) { return b ? b->p : nullptr; }
tc_device tc_buf_device(tc_buf b) { return b ? b->dev : nullptr; }
tc_device tc_stream_device(tc_stream s) { return s ? s->dev : nullptr; }
size_t tc_buf_size(tc_buf b) { return b ? b->n : 0; }

static bool in_range(tc_buf b, size_t off, size_t n) { return b && off <= b->n && n <= b->n - off; }

tc_status tc_buf_fill32(tc_stream s, tc_buf b, size_t off, uint32_t v, size_t count) {
    if (!s || !in_range(b, off, count * 4) || (off & 3)) return fail(TC_ERR_INVALID_ARG, "tc_buf_fill32: range");
    return hip_fail(launch_fill32(s->s, (uint32_t*)((char*)b->p + off), v, count), "fill32");
}

tc_status tc_buf_copy_h2d(tc_stream s, tc_buf dst, size_t off, const void* src, size_t n) {
    if (!s || !src || !in_range(dst, off, n)) return fail(TC_ERR_INVALID_ARG, "tc_buf_copy_h2d: range");
    return hip_fail(hipMemcpyAsync((char*)dst->p + off, src, n, hipMemcpyHostToDevice, s->s), "hipMemcpyAsync H2D");
}

tc_status tc_buf_copy_d2h(tc_stream s, void* dst, tc_buf src, size_t off, size_t n) {
    if (!s || !dst || !in_range(src, off, n)) return fail(TC_ERR_INVALID_ARG, "tc_buf_copy_d2h: range");
    return hip_fail(hipMemcpyAsync(dst, (char*)src->p + off, n, hipMemcpyDeviceToHost, s->s), "hipMemcpyAsync D2H");
}

tc_status tc_buf_copy_d2d(tc_stream s, tc_buf dst, size_t doff, tc_buf src, size_t soff, size_t n) {
    if (!s || !in_range(dst, doff, n) || !in_range(src, soff, n)) return fail(TC_ERR_INVALID_ARG, "tc_buf_copy_d2d: range");
    return hip_fail(hipMemcpyAsync((char*)dst->p + doff, (char*)src->p + soff, n, hipMemcpyDeviceToDevice, s->s),
                    "hipMemcpyAsync D2D");
}

tc_status tc_buf_copy_rect(tc_stream s, tc_buf dst, tc_buf src, const size_t so[3], const size_t doo[3],
                           const size_t region[3], size_t srp, size_t ssp, size_t drp, size_t dsp) {
    if (!s || !dst || !src || !so || !doo || !region) return fail(TC_ERR_INVALID_ARG, "tc_buf_copy_rect: null");
    if (!srp) srp = region[0];
    if (!ssp) ssp = region[1] * srp;
    if (!drp) drp = region[0];
    if (!dsp) dsp = region[1] * drp;
    if (region[0] > srp || region[0] > drp) return fail(TC_ERR_INVALID_ARG, "tc_buf_copy_rect: row wider than pitch");
    // every slice is checked before the first is copied: a bad region copies nothing
    const size_t last = (region[1] ? region[1] - 1 : 0);
    for (size_t z = 0; z < region[2]; z++) {
        size_t sb = (so[2] + z) * ssp + so[1] * srp + so[0];
        size_t db = (doo[2] + z) * dsp + doo[1] * drp + doo[0];
        if (!in_range(src, sb, last * srp + region[0]) || !in_range(dst, db, last * drp + region[0]))
            return fail(TC_ERR_INVALID_ARG, "tc_buf_copy_rect: region outside buffer");
    }
    for (size_t z = 0; z < region[2] && region[0] && region[1]; z++) {
        size_t sb = (so[2] + z) * ssp + so[1] * srp + so[0];
        size_t db = (doo[2] + z) * dsp + doo[1] * drp + doo[0];
        tc_status st = hip_fail(hipMemcpy2DAsync((char*)dst->p + db, drp, (char*)src->p + sb, srp, region[0],
                                                 region[1], hipMemcpyDeviceToDevice, s->s),
                                "hipMemcpy2DAsync");
        if (st) return st;
    }
    return TC_SUCCESS;
}

tc_status tc_buf_map_read(tc_stream s, tc_buf b, size_t off, size_t n, void** host) {
    if (!host || !in_range(b, off, n)) return fail(TC_ERR_INVALID_ARG, "tc_buf_map_read: range");
    void* h = nullptr;
    tc_status st = hip_fail(hipHostMalloc(&h, n ? n : 1, hipHostMallocDefault), "hipHostMalloc");
    if (st) return st;
    st = tc_buf_copy_d2h(s, h, b, off, n);
    if (!st) st = tc_stream_sync(s);
    if (st) {
        (void)hipHostFree(h);
        return st;
    }
    *host = h;
    return TC_SUCCESS;
}

tc_status tc_buf_unmap(tc_buf b, void* host) {
    (void)b;
    return hip_fail(hipHostFree(host), "hipHostFree");
}

// Fresh pinned pages are faulted in on first touch (about 3 GB/s from one
// thread on the MI355X hosts), so the allocation touches them itself, on a
// few threads, instead of leaving that cost to the first copy into them.
tc_status tc_host_alloc_pinned(size_t n, void** out) {
    if (!out) return fail(TC_ERR_INVALID_ARG, "tc_host_alloc_pinned: null");
    tc_status st = hip_fail(hipHostMalloc(out, n ? n : 1, hipHostMallocDefault), "hipHostMalloc");
    if (st != TC_SUCCESS) return st;
    const int nt = (int)std::min<size_t>(8, 1 + n / (32u << 20));
    char* p = static_cast<char*>(*out);
    auto touch = [=](int k) {
        const size_t a = n * (size_t)k / (size_t)nt, b = n * (size_t)(k + 1) / (size_t)nt;
        std::memset(p + a, 0, b - a);
    };
    if (nt == 1) {
        touch(0);
    } else {
        std::vector<std::thread> th;
        for (int k = 1; k < nt; k++) th.emplace_back(touch, k);
        touch(0);
        for (auto& t : th) t.join();
    }
    return TC_SUCCESS;
}

tc_status tc_host_free_pinned(void* p) { return hip_fail(hipHostFree(p), "hipHostFree"); }

tc_status tc_event_create(tc_event* out) {
    if (!out) return fail(TC_ERR_INVALID_ARG, "tc_event_create: null");
    hipEvent_t e;
    tc_status st = hip_fail(hipEventCreate(&e), "hipEventCreate");
    if (st) return st;
    *out = new tc_event_s{e};
    return TC_SUCCESS;
}

tc_status tc_event_destroy(tc_event e) {
    if (!e) return TC_SUCCESS;
    tc_status st = hip_fail(hipEventDestroy(e->e), "hipEventDestroy");
    delete e;
    return st;
}

tc_status tc_event_record(tc_event e, tc_stream s) {
    if (!e || !s) return fail(TC_ERR_INVALID_ARG, "tc_event_record: null");
    return hip_fail(hipEventRecord(e->e, s->s), "hipEventRecord");
}

tc_status tc_stream_wait_event(tc_stream s, tc_event e) {
    if (!s || !e) return fail(TC_ERR_INVALID_ARG, "tc_stream_wait_event: null");
    return hip_fail(hipStreamWaitEvent(s->s, e->e, 0), "hipStreamWaitEvent");
}

tc_status tc_event_sync(tc_event e) {
    if (!e) return fail(TC_ERR_INVALID_ARG, "tc_event_sync: null");
    return hip_fail(hipEventSynchronize(e->e), "hipEventSynchronize");
}

tc_status tc_event_query(tc_event e) {
    if (!e) return fail(TC_ERR_INVALID_ARG, "tc_event_query: null");
    hipError_t r = hipEventQuery(e->e);
    if (r == hipErrorNotReady) return TC_ERR_BUSY;
    return hip_fail(r, "hipEventQuery");
}

tc_status tc_event_elapsed_ms(tc_event a, tc_event b, float* ms) {
    if (!a || !b || !ms) return fail(TC_ERR_INVALID_ARG, "tc_event_elapsed_ms: null");
    return hip_fail(hipEventElapsedTime(ms, a->e, b->e), "hipEventElapsedTime");
}

tc_status tc_graph_begin(tc_stream s) {
    if (!s) return fail(TC_ERR_INVALID_ARG, "tc_graph_begin: null");
    return hip_fail(hipStreamBeginCapture(s->s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
}

tc_status tc_graph_end(tc_stream s, tc_graph* out) {
    if (!s || !out) return fail(TC_ERR_INVALID_ARG, "tc_graph_end: null");
    hipGraph_t g;
    tc_status st = hip_fail(hipStreamEndCapture(s->s, &g), "hipStreamEndCapture");
    if (st) return st;
    hipGraphExec_t x;
    st = hip_fail(hipGraphInstantiate(&x, g, nullptr, nullptr, 0), "hipGraphInstantiate");
    if (st) {
        (void)hipGraphDestroy(g);
        return st;
    }
    *out = new tc_graph_s{g, x};
    return TC_SUCCESS;
}

tc_status tc_graph_launch(tc_graph g, tc_stream s) {
    if (!g || !s) return fail(TC_ERR_INVALID_ARG, "tc_graph_launch: null");
    return hip_fail(hipGraphLaunch(g->x, s->s), "hipGraphLaunch");
}

tc_status tc_graph_destroy(tc_graph g) {
    if (!g) return TC_SUCCESS;
    (void)hipGraphExecDestroy(g->x);
    (void)hipGraphDestroy(g->g);
    delete g;
    return TC_SUCCESS;
}

tc_status tc_launch_t1_decode_wave(tc_stream s, int small_blocks, int grid, const void* descs, const uint32_t* order,
                                   int n, const uint8_t* arena, size_t arena_bytes, void* coef) {
    if (!s || n < 0 || grid <= 0 || (n > 0 && (!descs || !arena || !coef)) || (arena_bytes & 3) ||
        arena_bytes > 0xFFFFFFFFull || ((uintptr_t)arena & 3))
        return fail(TC_ERR_INVALID_ARG, "tc_launch_t1_decode_wave");
    return hip_fail(launch_t1_wave(s->s, small_blocks != 0, grid, (const tc_cblk_desc*)descs, order, n, arena,
                                   (uint32_t)arena_bytes,
                                   (uint32_t*)coef),
                    "t1_decode_wave launch");
}

tc_status tc_launch_t1_decode_stripe(tc_stream s, int small_blocks, int grid, const void* descs, const uint32_t* order,
                                     int n, const uint8_t* arena, size_t arena_bytes, void* coef) {
    if (!s || n < 0 || grid <= 0 || (n > 0 && (!descs || !arena || !coef)) || (arena_bytes & 3) ||
        arena_bytes > 0xFFFFFFFFull || ((uintptr_t)arena & 3))
        return fail(TC_ERR_INVALID_ARG, "tc_launch_t1_decode_stripe");
    return hip_fail(launch_t1_stripe(s->s, small_blocks != 0, grid, (const tc_cblk_desc*)descs, order, n, arena,
                                     (uint32_t)arena_bytes, (uint32_t*)coef),
                    "t1_decode_stripe launch");
}

static bool lockstep_bpw_ok(int b) { return b == 64 || b == 32 || b == 16; }

size_t tc_t1_lockstep_scratch_bytes(int small_blocks, int blocks_per_wave, int n, int max_passes) {
    return n > 0 && max_passes >= 0 && lockstep_bpw_ok(blocks_per_wave)
               ? t1_lockstep_scratch_bytes(small_blocks != 0, n, max_passes, blocks_per_wave)
               : 0;
}

tc_status tc_launch_t1_decode_lockstep(tc_stream s, int small_blocks, int blocks_per_wave, const void* descs,
                                       const uint32_t* order, int n, int max_passes, const uint8_t* arena,
                                       size_t arena_bytes, void* scratch, size_t scratch_bytes, void* coef) {
    if (!s || n < 0 || (n > 0 && (!descs || !arena || !coef || !scratch)) || ((uintptr_t)arena & 15) ||
        ((uintptr_t)scratch & 7) || (arena_bytes & 15) || arena_bytes < 64 || arena_bytes > 0xFFFFFFF0ull ||
        max_passes < 0 || max_passes > 255 || !lockstep_bpw_ok(blocks_per_wave))
        return fail(TC_ERR_INVALID_ARG, "tc_launch_t1_decode_lockstep");
    if (n > 0 && scratch_bytes < t1_lockstep_scratch_bytes(small_blocks != 0, n, max_passes, blocks_per_wave))
        return fail(TC_ERR_INVALID_ARG, "tc_launch_t1_decode_lockstep: scratch smaller than tc_t1_lockstep_scratch_bytes");
    return hip_fail(launch_t1_lockstep(s->s, small_blocks != 0, blocks_per_wave, (const tc_cblk_desc*)descs, order, n,
                                       arena, (uint32_t)arena_bytes, (uint32_t*)coef, scratch, max_passes),
                    "t1_decode_lockstep launch");
}

tc_status tc_idwt_tile_shape(int shape, int* w, int* h, int* count) {
    if (count) *count = idwt_shape_count();
    if (shape < 0 || shape >= idwt_shape_count()) return fail(TC_ERR_INVALID_ARG, "tc_idwt_tile_shape: shape");
    if (w) *w = idwt_tile_w(shape);
    if (h) *h = idwt_tile_h(shape);
    return TC_SUCCESS;
}

tc_status tc_launch_idwt_level(tc_stream s, int irr, int shape, const void* jobs, const uint32_t* wg_job, int nwg,
                               void* base) {
    if (!s || nwg < 0 || (nwg > 0 && (!jobs || !wg_job || !base)) || shape < 0 || shape >= idwt_shape_count())
        return fail(TC_ERR_INVALID_ARG, "tc_launch_idwt_level");
    return hip_fail(launch_idwt_level(s->s, shape, irr != 0, (const tc_dwt_job*)jobs, wg_job, nwg, base),
                    "idwt_level launch");
}

tc_status tc_launch_idwt_final(tc_stream s, int irr, int shape, int nc, int ob, const void* jobs, const uint32_t* wg_job,
                               int nwg, const void* base, void* out) {
    if (!s || nwg < 0 || (nwg > 0 && (!jobs || !wg_job || !base || !out)) || (nc != 1 && nc != 3) ||
        (ob != 1 && ob != 2) || shape < 0 || shape >= idwt_shape_count())
        return fail(TC_ERR_INVALID_ARG, "tc_launch_idwt_final");
    return hip_fail(launch_idwt_final(s->s, shape, irr != 0, nc, ob, (const tc_final_job*)jobs, wg_job, nwg, base,
                                      out),
                    "idwt_final launch");
}

tc_status tc_launch_fdwt_level(tc_stream s, int irr, int shape, const void* jobs, const uint32_t* wg_job, int nwg,
                               void* base) {
    if (!s || nwg < 0 || (nwg > 0 && (!jobs || !wg_job || !base)) || shape < 0 || shape >= idwt_shape_count())
        return fail(TC_ERR_INVALID_ARG, "tc_launch_fdwt_level");
    return hip_fail(launch_fdwt_level(s->s, shape, irr != 0, (const tc_fdwt_job*)jobs, wg_job, nwg, base),
                    "fdwt_level launch");
}

tc_status tc_launch_fdwt_first(tc_stream s, int irr, int shape, int nc, int in_bytes, const void* jobs,
                               const uint32_t* wg_job, int nwg, const void* img, void* base) {
    if (!s || nwg < 0 || (nwg > 0 && (!jobs || !wg_job || !base || !img)) || (nc != 1 && nc != 3) ||
        (in_bytes != 1 && in_bytes != 2) || shape < 0 || shape >= idwt_shape_count())
        return fail(TC_ERR_INVALID_ARG, "tc_launch_fdwt_first");
    return hip_fail(launch_fdwt_first(s->s, shape, irr != 0, nc, in_bytes, (const tc_fdwt_first_job*)jobs, wg_job,
                                      nwg, img, base),
                    "fdwt_first launch");
}

tc_status tc_launch_colour(tc_stream s, int kind, void* c0, void* c1, void* c2, size_t n, int32_t shift,
                           int32_t minv, int32_t maxv) {
    if (!s || kind < 0 || kind > 3 || !c0 || ((kind == 1 || kind == 2) && (!c1 || !c2)))
        return fail(TC_ERR_INVALID_ARG, "tc_launch_colour");
    return hip_fail(launch_colour(s->s, kind, c0, c1, c2, n, shift, minv, maxv), "colour launch");
}

tc_status tc_launch_fill32(tc_stream s, void* p, uint32_t v, uint64_t n) {
    if (!s || (n && !p)) return fail(TC_ERR_INVALID_ARG, "tc_launch_fill32");
    return hip_fail(launch_fill32(s->s, (uint32_t*)p, v, n), "fill32 launch");
}

}  // extern "C"

// Internal accessors used by the host pipeline.
namespace tc {
hipStream_t stream_handle(tc_stream s) { return s->s; }
int device_ordinal(tc_device d) { return d->ordinal; }
int device_cu_count(tc_device d) { return d->prop.multiProcessorCount; }
}  // namespace tc
