// Drop-in stage classes (include/tc/*.h) over the C ABI device layer.
//
// Each class keeps the reference's signature; its body builds the same work
// descriptors the fused pipeline uses and launches the typed kernels on the
// queue it was given (KernelInitInfoBase::cmd_queue), in stream order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>
#include <memory>

#include "../../../include/tc/Decoder.h"
#include "../../../include/tc_decode.h"
#include "../tc_internal.hpp"
#include "../../../include/tc/DWTForward53.h"
#include "../../../include/tc/DWTForward97.h"
#include "codestream.hpp"
#include "forward.hpp"
#include "pipeline.hpp"
#include "plan_util.hpp"

namespace tc {
int idwt_tile_w(int shape);
int idwt_tile_h(int shape);
}  // namespace tc

using namespace tc;

namespace {

void check(tc_status s, const char* what) {
    if (s != TC_SUCCESS) throw Error(std::string(what) + ": " + tc_last_error());
}

uint64_t align_up(uint64_t v, uint64_t a) { return (v + a - 1) / a * a; }

// Grows a device buffer (synchronising the queue first so no launch still uses it).
void ensure(tc_stream q, tc_device dev, tc_buf& b, size_t bytes) {
    if (b && tc_buf_size(b) >= bytes) return;
    if (b) {
        tc_stream_sync(q);
        tc_buf_free(b);
        b = nullptr;
    }
    check(tc_buf_alloc(dev, std::max<size_t>(bytes, 256), &b), "tc_buf_alloc");
}

// Element offset of p from base (both 4-byte aligned device addresses).
uint64_t eoff(const void* base, const void* p) { return (uint64_t)((const char*)p - (const char*)base) / 4; }

struct DeviceOf {
    tc_device dev;
};

// Per-image library state hung off type_image::impl.
struct ImageState {
    std::vector<uint8_t> data;
    Image img;
    tc_device dev = nullptr;
};

}  // namespace

// ---------------------------------------------------------------- CoefficientCoder

CoefficientCoder::CoefficientCoder(KernelInitInfoBase initInfo)
    : DeviceKernel(KernelInitInfo(initInfo, "t1_lockstep.hip", "t1_decode")) {}

CoefficientCoder::~CoefficientCoder(void) {
    tc_buf_free(d_upload_);
    tc_buf_free(d_scratch_);
}

void CoefficientCoder::decode_tile(type_tile* tile) {
    if (!tile || !tile->parent_img || !tile->parent_img->impl) throw Error("decode_tile: tile without parsed image");
    ImageState* S = (ImageState*)tile->parent_img->impl;
    const Image& im = S->img;
    const Tile& T = *(const Tile*)tile->impl;
    int t = (int)tile->tile_no;
    int nc = im.ncomp;
    // planes: one device buffer per tile-component (allocated here if missing)
    for (int c = 0; c < nc; c++) {
        type_tile_comp& tcp = tile->tile_comp[c];
        size_t bytes = align_up((uint64_t)tcp.width * tcp.height * 4, 256);
        if (!tcp.img_data_d || tc_buf_size(tcp.img_data_d) < bytes) {
            tc_buf_free(tcp.img_data_d);
            tcp.img_data_d = nullptr;
            check(tc_buf_alloc(S->dev, bytes, &tcp.img_data_d), "tile-component plane");
        }
        check(tc_buf_fill32(queue_, tcp.img_data_d, 0, 0u, (size_t)tcp.width * tcp.height), "zero plane");
    }
    // all planes addressed from the lowest one
    void* base = tc_buf_device_ptr(tile->tile_comp[0].img_data_d);
    for (int c = 1; c < nc; c++) base = std::min(base, tc_buf_device_ptr(tile->tile_comp[c].img_data_d));
    std::vector<uint64_t> planes(nc);
    for (int c = 0; c < nc; c++) planes[c] = eoff(base, tc_buf_device_ptr(tile->tile_comp[c].img_data_d));
    std::vector<uint8_t> cw;
    std::vector<tc_cblk_desc> descs;
    std::vector<uint32_t> cls[2];
    std::string err;
    Status st = append_tile_cblks(im, t, planes.data(), cw, descs, cls, err);
    if (st != Status::Ok) throw Error("decode_tile: " + err);
    (void)T;
    if (descs.empty()) {
        for (int c = 0; c < nc; c++) tile->tile_comp[c].stage = 1;
        return;
    }
    cw.resize(((cw.size() + 15) & ~(size_t)15) + 64, 0xFF);
    size_t cw_bytes = cw.size();
    size_t off_desc = align_up(cw_bytes, 256);
    size_t off_o0 = align_up(off_desc + descs.size() * sizeof(tc_cblk_desc), 256);
    size_t off_o1 = align_up(off_o0 + cls[0].size() * 4, 256);
    size_t total = off_o1 + cls[1].size() * 4 + 16;
    std::vector<uint8_t> blob(total, 0);
    std::memcpy(blob.data(), cw.data(), cw.size());
    std::memcpy(blob.data() + off_desc, descs.data(), descs.size() * sizeof(tc_cblk_desc));
    for (int k = 0; k < 2; k++) sort_for_packing(cls[k], descs, k == 0);
    if (!cls[0].empty()) std::memcpy(blob.data() + off_o0, cls[0].data(), cls[0].size() * 4);
    if (!cls[1].empty()) std::memcpy(blob.data() + off_o1, cls[1].data(), cls[1].size() * 4);
    ensure(queue_, S->dev, d_upload_, total);
    // the upload is synchronous from pageable memory: the blob may go away after
    check(tc_buf_copy_h2d(queue_, d_upload_, 0, blob.data(), total), "upload");
    check(tc_stream_sync(queue_), "upload sync");
    uint8_t* up = (uint8_t*)tc_buf_device_ptr(d_upload_);
    int cus = device_cu_count(S->dev);
    tc_event e0, e1;
    tc_event_create(&e0);
    tc_event_create(&e1);
    tc_event_record(e0, queue_);
    for (int k = 0; k < 2; k++) {
        if (cls[k].empty()) continue;
        int n = (int)cls[k].size();
        // the fused path's rule: lockstep from 64 blocks per CU up, else one wave per block
        Tuning tn;
        tn.bpw32 = bpw32_;
        tn.bpw64 = bpw64_;
        const int g = pick_blocks_per_wave(n, k == 0, cus, tn);
        const uint32_t* order = (const uint32_t*)(up + (k == 0 ? off_o0 : off_o1));
        int style = 0, max_passes = 0;
        for (uint32_t i : cls[k]) {
            style |= descs[i].mode & TC_MODE_STYLE;
            max_passes = std::max<int>(max_passes, descs[i].npasses);
        }
        if (style || g < 0) {
            // lockstep (only it decodes the code-block styles), -g blocks per wave
            const int bpw = g < 0 ? -g : 32;
            const size_t sb = tc_t1_lockstep_scratch_bytes(k == 0, bpw, n, max_passes);
            ensure(queue_, S->dev, d_scratch_, sb);
            check(tc_launch_t1_decode_lockstep(queue_, k == 0, bpw, up + off_desc, order, n, max_passes, up, cw_bytes,
                                               tc_buf_device_ptr(d_scratch_), tc_buf_size(d_scratch_), base),
                  "t1 launch");
        } else {
            check(tc_launch_t1_decode_wave(queue_, k == 0, std::min(n, cus * 16), up + off_desc, order, n, up, cw_bytes,
                                           base),
                  "t1 launch");
        }
    }
    tc_event_record(e1, queue_);
    tc_event_sync(e1);
    tc_event_elapsed_ms(e0, e1, &last_ms_);
    tc_event_destroy(e0);
    tc_event_destroy(e1);
    for (int c = 0; c < nc; c++) tile->tile_comp[c].stage = 1;
}

// ---------------------------------------------------------------- Quantizer

void Quantizer::dequantize_tile(type_tile* tile) {
    if (!tile) throw Error("dequantize_tile: null tile");
    for (unsigned c = 0; c < tile->parent_img->num_components; c++)
        if (tile->tile_comp[c].stage < 1) throw Error("dequantize_tile: tile not decoded by CoefficientCoder");
}

// ---------------------------------------------------------------- DWTKernel

template <typename T>
DWTKernel<T>::DWTKernel(int diameter, KernelInitInfo initInfo) : DeviceKernel(initInfo), waveletImpulseDiameter(diameter) {}

template <typename T>
DWTKernel<T>::~DWTKernel(void) {
    tc_buf_free(work_);
    tc_buf_free(jobs_);
    if (ownsMemory) {
        tc_buf_free(srcMem);
        tc_buf_free(dstMem);
    }
}

template <typename T>
tDeviceRC DWTKernel<T>::run(tDeviceMem in, tDeviceMem out, int sizeX, int sizeY, int levels) {
    return run_at(in, out, 0, 0, sizeX, sizeY, levels);
}

template <typename T>
tDeviceRC DWTKernel<T>::run_forward(tDeviceMem in, tDeviceMem out, int sizeX, int sizeY, int levels) {
    dimX = sizeX;
    dimY = sizeY;
    srcMem = in;
    dstMem = out;
    if (levels == 0) return tc_buf_copy_d2d(queue_, out, 0, in, 0, (size_t)sizeX * sizeY * 4);
    const size_t half = (size_t)((sizeX + 1) / 2) * ((sizeY + 1) / 2);
    const size_t wbytes = (align_up(half, 64) * 2) * 4 + 256;
    if (!work_ || tc_buf_size(work_) < wbytes) {
        tc_stream_sync(queue_);
        tc_buf_free(work_);
        work_ = nullptr;
        tc_status s = tc_buf_alloc(tc_buf_device(in), wbytes, &work_);
        if (s) return s;
    }
    void* base = std::min(std::min(tc_buf_device_ptr(in), tc_buf_device_ptr(out)), tc_buf_device_ptr(work_));
    const uint64_t w0 = eoff(base, tc_buf_device_ptr(work_));
    const uint64_t wk[2] = {w0, w0 + align_up(half, 64)};
    const int TX = idwt_tile_w(0), TY = idwt_tile_h(0);
    std::vector<std::vector<tc_fdwt_job>> lv;
    plan_forward_levels(sizeX, sizeY, levels, 1, eoff(base, tc_buf_device_ptr(in)), (uint32_t)sizeX,
                        eoff(base, tc_buf_device_ptr(out)), wk, TX, TY, lv);
    std::vector<uint8_t> blob;
    std::vector<size_t> joff(levels + 1), moff(levels + 1);
    std::vector<int> nwg(levels + 1, 0);
    for (int n = 1; n <= levels; n++) {
        tc_fdwt_job& j = lv[n][0];
        j.wg_base = 0;
        nwg[n] = j.tiles_x * ((j.v1 - j.v0 + TY - 1) / TY);
        joff[n] = align_up(blob.size(), 256);
        blob.resize(joff[n] + sizeof(tc_fdwt_job));
        std::memcpy(blob.data() + joff[n], &j, sizeof j);
        moff[n] = align_up(blob.size(), 256);
        blob.resize(moff[n] + (size_t)nwg[n] * 4, 0);
    }
    if (!jobs_ || tc_buf_size(jobs_) < blob.size()) {
        tc_stream_sync(queue_);
        tc_buf_free(jobs_);
        jobs_ = nullptr;
        tc_status s = tc_buf_alloc(tc_buf_device(in), blob.size() + 256, &jobs_);
        if (s) return s;
    }
    tc_status s = tc_buf_copy_h2d(queue_, jobs_, 0, blob.data(), blob.size());
    if (!s) s = tc_stream_sync(queue_);
    uint8_t* jp = (uint8_t*)tc_buf_device_ptr(jobs_);
    for (int n = 1; n <= levels && !s; n++)
        s = tc_launch_fdwt_level(queue_, irreversible() ? 1 : 0, 0, jp + joff[n], (const uint32_t*)(jp + moff[n]),
                                 nwg[n], base);
    return s;
}

template <typename T>
tDeviceRC DWTKernel<T>::run_at(tDeviceMem in, tDeviceMem out, int x0, int y0, int sizeX, int sizeY, int levels) {
    if (!in || !out || sizeX <= 0 || sizeY <= 0 || levels < 0 || levels > 32) return TC_ERR_INVALID_ARG;
    if (tc_buf_size(in) < (size_t)sizeX * sizeY * 4 || tc_buf_size(out) < (size_t)sizeX * sizeY * 4) return TC_ERR_INVALID_ARG;
    if (forward()) {
        if (x0 || y0) return TC_ERR_UNSUPPORTED;  // forward classes take origin (0, 0)
        return run_forward(in, out, sizeX, sizeY, levels);
    }
    dimX = sizeX;
    dimY = sizeY;
    srcMem = in;
    dstMem = out;
    if (levels == 0) return tc_buf_copy_d2d(queue_, out, 0, in, 0, (size_t)sizeX * sizeY * 4);
    // geometry of a single tile-component
    TileComp tc;
    tc.x0 = x0;
    tc.y0 = y0;
    tc.x1 = x0 + sizeX;
    tc.y1 = y0 + sizeY;
    tc.nlev = levels;
    tc.reversible = irreversible() ? 0 : 1;
    tc.res.assign(levels + 1, Resolution());
    for (int r = 0; r <= levels; r++) {
        int64_t s = (int64_t)1 << (levels - r);
        Resolution& R = tc.res[r];
        R.x0 = (int)((tc.x0 + s - 1) / s);
        R.y0 = (int)((tc.y0 + s - 1) / s);
        R.x1 = (int)((tc.x1 + s - 1) / s);
        R.y1 = (int)((tc.y1 + s - 1) / s);
    }
    uint64_t sz[2];
    work_sizes(tc, sz);
    size_t wbytes = (align_up(sz[0], 64) + align_up(sz[1], 64)) * 4 + 256;
    tc_device dev = nullptr;
    (void)dev;
    // work buffers and job arrays live with this object
    const int TX = idwt_tile_w(0), TY = idwt_tile_h(0);
    std::vector<std::vector<tc_dwt_job>> jobs(levels);
    std::vector<std::vector<uint32_t>> maps(levels);
    // addresses: everything relative to the lowest of in/out/work
    if (!work_ || tc_buf_size(work_) < wbytes) {
        tc_stream_sync(queue_);
        tc_buf_free(work_);
        work_ = nullptr;
        tc_status s = tc_buf_alloc(tc_buf_device(in), wbytes, &work_);
        if (s) return s;
    }
    void* base = tc_buf_device_ptr(in);
    base = std::min(base, tc_buf_device_ptr(out));
    base = std::min(base, tc_buf_device_ptr(work_));
    uint64_t plane = eoff(base, tc_buf_device_ptr(in));
    uint64_t wk[2] = {eoff(base, tc_buf_device_ptr(work_)), 0};
    wk[1] = wk[0] + align_up(sz[0], 64);
    uint64_t raster = eoff(base, tc_buf_device_ptr(out));
    size_t total = 0;
    std::vector<size_t> joff(levels), moff(levels);
    for (int n = levels, k = 0; n >= 1; n--, k++) {
        const Resolution& R = tc.res[levels - n + 1];
        if (R.width() <= 0 || R.height() <= 0) continue;
        tc_dwt_job j = level_job(tc, n, plane, wk, true, raster, (uint32_t)sizeX, TX, TY);
        j.wg_base = 0;
        int ntiles = j.tiles_x * ((R.height() + TY - 1) / TY);
        jobs[k].push_back(j);
        maps[k].assign(ntiles, 0u);
        joff[k] = align_up(total, 256);
        moff[k] = align_up(joff[k] + sizeof(tc_dwt_job), 256);
        total = moff[k] + maps[k].size() * 4;
    }
    if (!jobs_ || tc_buf_size(jobs_) < total + 256) {
        tc_stream_sync(queue_);
        tc_buf_free(jobs_);
        jobs_ = nullptr;
        tc_status s = tc_buf_alloc(tc_buf_device(in), total + 256, &jobs_);
        if (s) return s;
    }
    std::vector<uint8_t> blob(total + 256, 0);
    for (int k = 0; k < levels; k++) {
        if (jobs[k].empty()) continue;
        std::memcpy(blob.data() + joff[k], jobs[k].data(), sizeof(tc_dwt_job));
        std::memcpy(blob.data() + moff[k], maps[k].data(), maps[k].size() * 4);
    }
    tc_status s = tc_buf_copy_h2d(queue_, jobs_, 0, blob.data(), blob.size());
    if (!s) s = tc_stream_sync(queue_);
    uint8_t* jp = (uint8_t*)tc_buf_device_ptr(jobs_);
    for (int k = 0; k < levels && !s; k++) {
        if (jobs[k].empty()) continue;
        s = tc_launch_idwt_level(queue_, irreversible() ? 1 : 0, 0, jp + joff[k], (const uint32_t*)(jp + moff[k]),
                                 (int)maps[k].size(), base);
    }
    return s;
}

template <typename T>
tDeviceRC DWTKernel<T>::run(T* in, int sizeX, int sizeY, int levels) {
    if (!in || sizeX <= 0 || sizeY <= 0) return TC_ERR_INVALID_ARG;
    tc_device dev = tc_stream_device(queue_);
    size_t bytes = (size_t)sizeX * sizeY * 4;
    if (ownsMemory) {
        tc_buf_free(srcMem);
        tc_buf_free(dstMem);
    }
    srcMem = dstMem = nullptr;
    tc_status s = tc_buf_alloc(dev, bytes, &srcMem);
    if (!s) s = tc_buf_alloc(dev, bytes, &dstMem);
    ownsMemory = true;
    if (!s) s = tc_buf_copy_h2d(queue_, srcMem, 0, in, bytes);
    if (!s) s = tc_stream_sync(queue_);
    if (!s) s = run(srcMem, dstMem, sizeX, sizeY, levels);
    return s;
}

template <typename T>
T* DWTKernel<T>::mapOutputBufferToHost() {
    if (!dstMem) return nullptr;
    host_out_.resize((size_t)dimX * dimY);
    if (tc_buf_copy_d2h(queue_, host_out_.data(), dstMem, 0, host_out_.size() * 4) || tc_stream_sync(queue_))
        return nullptr;
    return host_out_.data();
}

template class DWTKernel<int>;
template class DWTKernel<float>;

DWTForward53::DWTForward53(KernelInitInfoBase initInfo) : DWTKernel<int>(3, KernelInitInfo(initInfo, "fdwt.hip", "fdwt_level_kernel<false>")) {}
DWTForward97::DWTForward97(KernelInitInfoBase initInfo) : DWTKernel<float>(9, KernelInitInfo(initInfo, "fdwt.hip", "fdwt_level_kernel<true>")) {}
DWTReverse53::DWTReverse53(KernelInitInfoBase initInfo) : DWTKernel<int>(5, KernelInitInfo(initInfo, "idwt.hip", "idwt_level_kernel<false>")) {}
DWTReverse97::DWTReverse97(KernelInitInfoBase initInfo) : DWTKernel<float>(9, KernelInitInfo(initInfo, "idwt.hip", "idwt_level_kernel<true>")) {}

// ---------------------------------------------------------------- DWT

DWT::DWT(KernelInitInfoBase info) : initInfo(info), r53(new DWTReverse53(info)), r97(new DWTReverse97(info)) {}

DWT::~DWT(void) {
    delete r53;
    delete r97;
}

void DWT::iwt(type_tile* tile) {
    ImageState* S = (ImageState*)tile->parent_img->impl;
    for (unsigned c = 0; c < tile->parent_img->num_components; c++) {
        type_tile_comp& tcp = tile->tile_comp[c];
        if (tcp.stage != 1) throw Error("iwt: tile-component has no coefficients");
        tc_buf out = nullptr;
        check(tc_buf_alloc(S->dev, align_up((uint64_t)tcp.width * tcp.height * 4, 256), &out), "iwt output");
        tc_status s = tcp.reversible ? r53->run_at(tcp.img_data_d, out, tcp.tlx, tcp.tly, tcp.width, tcp.height, tcp.num_dlvls)
                                     : r97->run_at(tcp.img_data_d, out, tcp.tlx, tcp.tly, tcp.width, tcp.height, tcp.num_dlvls);
        if (s) {
            tc_buf_free(out);
            check(s, "iwt");
        }
        check(tc_stream_sync(initInfo.cmd_queue), "iwt sync");
        // the output replaces the coefficient plane (TC/DWT.cpp:94)
        tc_buf_free(tcp.img_data_d);
        tcp.img_data_d = out;
        tcp.stage = 2;
    }
}

// ---------------------------------------------------------------- Preprocessor

int Preprocessor::color_trans_gpu(type_image* img, color_trans_type type) {
    if (img->num_components < 3) return -1;
    for (unsigned t = 0; t < img->num_tiles; t++) {
        type_tile* tile = &img->tile[t];
        type_tile_comp* c = tile->tile_comp;
        if (c[0].stage != 2 || c[1].stage != 2 || c[2].stage != 2) return -1;
        if (c[0].prec != c[1].prec || c[0].prec != c[2].prec || c[0].sgnd || c[1].sgnd || c[2].sgnd) return -1;
        int prec = c[0].prec;
        int shift = 1 << (prec - 1), lo = 0, hi = (1 << prec) - 1;
        int kind = type == TCR ? 1 : 2;
        size_t n = (size_t)c[0].width * c[0].height;
        check(tc_launch_colour(initInfo.cmd_queue, kind, tc_buf_device_ptr(c[0].img_data_d),
                               tc_buf_device_ptr(c[1].img_data_d), tc_buf_device_ptr(c[2].img_data_d), n, shift, lo, hi),
              "colour");
        for (int k = 0; k < 3; k++) c[k].stage = 3;
    }
    return 0;
}

int Preprocessor::color_decoder_lossless(type_image* img) { return color_trans_gpu(img, TCR); }
int Preprocessor::color_decoder_lossy(type_image* img) { return color_trans_gpu(img, TCI); }

void Preprocessor::idc_level_shifting(type_image* img) {
    for (unsigned t = 0; t < img->num_tiles; t++)
        for (unsigned k = 0; k < img->num_components; k++) {
            type_tile_comp& c = img->tile[t].tile_comp[k];
            if (c.stage != 2) continue;
            int prec = c.prec;
            int shift = c.sgnd ? 0 : 1 << (prec - 1);
            int lo = c.sgnd ? -(1 << (prec - 1)) : 0, hi = c.sgnd ? (1 << (prec - 1)) - 1 : (1 << prec) - 1;
            check(tc_launch_colour(initInfo.cmd_queue, c.reversible ? 0 : 3, tc_buf_device_ptr(c.img_data_d), nullptr,
                                   nullptr, (size_t)c.width * c.height, shift, lo, hi),
                  "dc shift");
            c.stage = 3;
        }
}

// ---------------------------------------------------------------- Decoder

Decoder::Decoder(ocl_args_d_t* ocl) : _ocl(ocl), coder(nullptr), quantizer(nullptr), dwt(nullptr), preprocessor(nullptr), fused_(nullptr) {
    if (!ocl || !ocl->device || !ocl->commandQueue) throw Error("Decoder: no device/queue");
    KernelInitInfoBase info(ocl->commandQueue, "");
    coder = new CoefficientCoder(info);
    quantizer = new Quantizer(info);
    dwt = new DWT(info);
    preprocessor = new Preprocessor(info);
    tc_decoder d = nullptr;
    check(tc_decoder_create(ocl->device, &d), "tc_decoder_create");
    fused_ = d;
}

Decoder::~Decoder(void) {
    delete coder;
    delete quantizer;
    delete dwt;
    delete preprocessor;
    tc_decoder_destroy((tc_decoder)fused_);
}

int Decoder::decode(std::string fileName) {
    std::ifstream f(fileName, std::ios::binary);
    if (!f) {
        printf("File not found\n");
        return -2;
    }
    std::vector<uint8_t> data((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    return decode_memory(data.data(), data.size());
}

int Decoder::decode_memory(const uint8_t* data, size_t len) {
    tc_image_info info;
    tc_status s = tc_image_probe(data, len, &info);
    if (s) return -1;
    width_ = info.width;
    height_ = info.height;
    ncomp_ = info.ncomp;
    bps_ = info.bytes_per_sample;
    out_.resize((size_t)width_ * height_ * ncomp_ * bps_);
    tc_decoder d = (tc_decoder)fused_;
    tc_event e0, e1;
    tc_event_create(&e0);
    tc_event_create(&e1);
    s = tc_decoder_prepare(d, 1, &data, &len);
    if (!s) {
        tc_event_record(e0, tc_decoder_stream(d));
        s = tc_decoder_run(d, 0);
        tc_event_record(e1, tc_decoder_stream(d));
    }
    if (!s) s = tc_decoder_fetch(d, 0, out_.data(), out_.size());
    if (!s) tc_event_elapsed_ms(e0, e1, &last_ms_);
    tc_event_destroy(e0);
    tc_event_destroy(e1);
    return s ? -1 : 0;
}

tDeviceRC Decoder::mapComponentToHost(type_tile_comp* tile_comp) {
    void* host = nullptr;
    tc_status s = tc_buf_map_read(_ocl->commandQueue, tile_comp->img_data_d, 0,
                                  (size_t)tile_comp->width * tile_comp->height * 4, &host);
    if (!s) tile_comp->img_data_h = host;
    return s;
}

int Decoder::decode_staged(const uint8_t* data, size_t len) {
    std::unique_ptr<ImageState> S(new ImageState());
    S->data.assign(data, data + len);
    S->dev = _ocl->device;
    std::string err;
    if (parse_codestream(S->data.data(), S->data.size(), S->img, err) != Status::Ok) {
        set_error(err);
        return -1;
    }
    if (S->img.planar()) {
        // the reference's type_image has one size for every component
        set_error("staged decode of sub-sampled components is not supported (use tc_decoder_decode)");
        return -1;
    }
    const Image& im = S->img;
    type_image img;
    std::memset(&img, 0, sizeof img);
    img.width = (unsigned short)im.width();
    img.height = (unsigned short)im.height();
    img.num_components = (unsigned short)im.ncomp;
    img.depth = (unsigned short)im.max_prec();
    img.sign = im.sgnd.empty() ? 0 : (unsigned char)im.sgnd[0];
    img.num_tiles = (unsigned)im.tiles.size();
    img.use_mct = (unsigned char)im.mct;
    img.num_layers = (unsigned short)im.nlayers;
    img.impl = S.get();
    std::vector<type_tile> tiles(im.tiles.size());
    std::vector<std::vector<type_tile_comp>> comps(im.tiles.size());
    img.tile = tiles.data();
    for (size_t t = 0; t < im.tiles.size(); t++) {
        const Tile& T = im.tiles[t];
        type_tile& tt = tiles[t];
        std::memset(&tt, 0, sizeof tt);
        tt.tile_no = (unsigned)t;
        tt.tlx = T.x0;
        tt.tly = T.y0;
        tt.brx = T.x1;
        tt.bry = T.y1;
        tt.width = (unsigned short)(T.x1 - T.x0);
        tt.height = (unsigned short)(T.y1 - T.y0);
        tt.parent_img = &img;
        tt.use_mct = T.mct;
        tt.impl = &T;
        comps[t].resize(im.ncomp);
        tt.tile_comp = comps[t].data();
        for (int c = 0; c < im.ncomp; c++) {
            const TileComp& tc = T.comps[c];
            type_tile_comp& k = comps[t][c];
            std::memset(&k, 0, sizeof k);
            k.tile_comp_no = (unsigned)c;
            k.tlx = tc.x0;
            k.tly = tc.y0;
            k.brx = tc.x1;
            k.bry = tc.y1;
            k.width = (unsigned short)tc.width();
            k.height = (unsigned short)tc.height();
            k.num_dlvls = (unsigned char)tc.nlev;
            k.num_rlvls = (unsigned char)(tc.nlev + 1);
            k.reversible = tc.reversible;
            k.prec = im.prec[c];
            k.sgnd = im.sgnd[c];
            k.parent_tile = &tt;
        }
    }
    img.wavelet_type = im.tiles[0].comps[0].reversible ? DWT_53 : DWT_97;
    img.num_dlvls = (unsigned char)im.tiles[0].comps[0].nlev;
    int rc = 0;
    try {
        // TC/Decoder.cpp:153-180: per tile T1 -> dequant -> iDWT, then colour
        for (unsigned i = 0; i < img.num_tiles; i++) {
            coder->decode_tile(&img.tile[i]);
            quantizer->dequantize_tile(&img.tile[i]);
            dwt->iwt(&img.tile[i]);
        }
        if (img.use_mct == 1 && img.num_components >= 3) {
            if (img.wavelet_type == DWT_53)
                rc = preprocessor->color_decoder_lossless(&img);
            else
                rc = preprocessor->color_decoder_lossy(&img);
            // components beyond the first three keep their own DC shift
            if (img.num_components > 3) preprocessor->idc_level_shifting(&img);
        } else {
            preprocessor->idc_level_shifting(&img);
        }
        check(tc_stream_sync(_ocl->commandQueue), "finish");
        // map every tile-component and pack the interleaved output
        width_ = im.width();
        height_ = im.height();
        ncomp_ = im.ncomp;
        bps_ = im.max_prec() > 8 ? 2 : 1;
        out_.assign((size_t)width_ * height_ * ncomp_ * bps_, 0);
        for (unsigned i = 0; i < img.num_tiles && rc == 0; i++)
            for (unsigned c = 0; c < img.num_components; c++) {
                type_tile_comp* k = &img.tile[i].tile_comp[c];
                check(mapComponentToHost(k), "map");
                const int32_t* src = (const int32_t*)k->img_data_h;
                for (int y = 0; y < k->height; y++)
                    for (int x = 0; x < k->width; x++) {
                        size_t o = ((size_t)(k->tly - im.y0 + y) * width_ + (k->tlx - im.x0 + x)) * ncomp_ + c;
                        int32_t v = src[(size_t)y * k->width + x];
                        if (bps_ == 1)
                            out_[o] = (uint8_t)v;
                        else {
                            out_[2 * o] = (uint8_t)(v & 0xFF);
                            out_[2 * o + 1] = (uint8_t)((v >> 8) & 0xFF);
                        }
                    }
                tc_buf_unmap(k->img_data_d, k->img_data_h);
                k->img_data_h = nullptr;
            }
    } catch (const Error& e) {
        set_error(e.what());
        rc = -1;
    }
    for (auto& v : comps)
        for (auto& k : v) tc_buf_free(k.img_data_d);
    return rc;
}

// ---------------------------------------------------------------- C ABI

extern "C" tc_status tc_decode_staged(tc_device dev, const uint8_t* cs, size_t len, void* out, size_t cap) {
    if (!dev || !cs || !out) return fail(TC_ERR_INVALID_ARG, "tc_decode_staged: null");
    tc_stream q = nullptr;
    tc_status s = tc_stream_create(dev, &q);
    if (s) return s;
    ocl_args_d_t ocl;
    ocl.device = dev;
    ocl.commandQueue = q;
    int rc = -1;
    try {
        Decoder d(&ocl);
        rc = d.decode_staged(cs, len);
        if (rc == 0) {
            if (cap < d.output().size())
                rc = -2;
            else
                std::memcpy(out, d.output().data(), d.output().size());
        }
    } catch (const Error& e) {
        set_error(e.what());
        rc = -1;
    }
    tc_stream_destroy(q);
    if (rc == -2) return fail(TC_ERR_CAPACITY, "tc_decode_staged: output buffer too small");
    return rc == 0 ? TC_SUCCESS : TC_ERR_DEVICE;
}

// Runs one DWT drop-in class over a host plane (DWTKernel::run(T*, ...) then
// mapOutputBufferToHost, TC/DWTKernel.cpp:45-78,180-191).
extern "C" tc_status tc_dwt_plane(tc_device dev, int fwd, int irr, const void* in, void* out, int w, int h,
                                  int levels) {
    if (!dev || !in || !out || w <= 0 || h <= 0 || levels < 0 || levels > 32)
        return fail(TC_ERR_INVALID_ARG, "tc_dwt_plane");
    tc_stream q = nullptr;
    tc_status s = tc_stream_create(dev, &q);
    if (s) return s;
    KernelInitInfoBase info(q, "");
    const size_t bytes = (size_t)w * h * 4;
    auto go = [&](auto& k, auto* src) -> tc_status {
        tc_status r = k.run(src, w, h, levels);
        if (r) return r;
        const void* res = k.mapOutputBufferToHost();
        if (!res) return fail(TC_ERR_DEVICE, "tc_dwt_plane: map");
        std::memcpy(out, res, bytes);
        return TC_SUCCESS;
    };
    std::vector<uint8_t> copy((const uint8_t*)in, (const uint8_t*)in + bytes);
    try {
        if (fwd && !irr) {
            DWTForward53 k(info);
            s = go(k, (int*)copy.data());
        } else if (fwd) {
            DWTForward97 k(info);
            s = go(k, (float*)copy.data());
        } else if (!irr) {
            DWTReverse53 k(info);
            s = go(k, (int*)copy.data());
        } else {
            DWTReverse97 k(info);
            s = go(k, (float*)copy.data());
        }
    } catch (const Error& e) {
        s = fail(TC_ERR_DEVICE, e.what());
    }
    tc_stream_destroy(q);
    return s;
}
