// Device work planning and execution for one image or a batch of images.
//
// The reference runs, per tile, CoefficientCoder::decode_tile ->
// Quantizer::dequantize_tile -> DWT::iwt, then the Preprocessor over the image
// (TC/Decoder.cpp:153-180), allocating device buffers inside every stage
// (TC/CoefficientCoder.cpp:185-205, TC/Quantizer.cpp:80, TC/DWT.cpp:79-84).
// Here a Pipeline turns any number of parsed images into flat device work
// lists and runs one launch per stage for all of them:
//   fill (zero the coefficient planes) -> tier-1 (one launch per code-block
//   size class) -> one inverse-DWT launch per level step -> the fused final
//   level (colour transform, DC shift, clamp, pack).
// Every buffer is kept resident and only grows, so repeated decodes allocate
// nothing; all descriptors travel in one host-to-device copy.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "../../../include/tc_decode.h"
#include "../../../include/tc_device.h"
#include "../tc_kernels.h"
#include "codestream.hpp"

namespace tc {

struct Tuning {
    int bpw64 = 0;  // blocks per wave for blocks up to 64x64 (0 = pick automatically)
    int bpw32 = 0;  // blocks per wave for blocks up to 32x32 (0 = pick automatically)
    int t1_kernel = 0;  // one-wave-per-block kernel: 0 stripe-column form, 1 row-mask form
    int dwt_shape = 0;  // inverse-DWT output tile shape (tc_idwt_tile_shape)
};

// Chooses blocks-per-wave from the block count and the machine size, by a rule
// taken from the autotuner's choices; a fixed value in Tuning overrides it.
int pick_blocks_per_wave(int nblocks, bool small_blocks, int cu_count, const Tuning& t);

struct StageTimes {
    float fill_ms = 0, t1_ms = 0, dwt_ms = 0, final_ms = 0, total_ms = 0;
};

class Pipeline {
  public:
    explicit Pipeline(tc_device dev);
    ~Pipeline();
    Pipeline(const Pipeline&) = delete;
    Pipeline& operator=(const Pipeline&) = delete;

    // Builds the work lists for `imgs` and uploads them (stream ordered).  A
    // failed prepare leaves no plan: run, capture_graph, fetch and the output
    // queries then fail with InvalidArgument until a prepare succeeds.
    Status prepare(const std::vector<const Image*>& imgs, std::string& err);
    // Drops the current plan (as a failed prepare does).
    void invalidate();
    bool ready() const { return ready_; }
    // Enqueues every stage on the resident data.  With events, per-stage
    // times become available after sync().
    Status run(bool timed = false);
    Status sync();
    // Copies image i's interleaved output to host memory.
    Status fetch(int i, void* dst, size_t cap, std::string& err);
    // Device address of image i's output and its size in bytes.
    const void* output_ptr(int i) const;
    size_t output_bytes(int i) const;
    int output_sample_bytes(int i) const { return images_[i].bps; }
    int num_images() const { return ready_ ? (int)images_.size() : 0; }
    StageTimes times() const;
    tc_stream stream() const { return stream_; }
    void set_tuning(const Tuning& t) { tuning_ = t; }
    // output images in pinned host memory, written by the final-level kernel
    // directly (no device buffer, no copy); from the next prepare
    void set_host_output(bool on) { host_out_ = on; }
    const Tuning& tuning() const { return tuning_; }
    // Statistics of the last prepare().
    int64_t num_cblks() const { return ncblk_; }
    int64_t codeword_bytes() const { return cw_bytes_; }
    int64_t sample_count() const { return nsamples_; }
    int64_t t1_sample_count() const { return t1_samples_; }
    int64_t pixel_count() const { return npixels_; }
    // packing the prepared plan runs for class cls (0: blocks up to 32x32), 0 if empty
    int blocks_per_wave(int cls) const { return bpw_[cls]; }
    // the class holds code-blocks with a style (only lockstep packings decode them)
    bool styled(int cls) const { return styled_[cls]; }
    // bytes moved by the HBM-bound stages (algorithmic): dwt levels and final
    int64_t dwt_bytes() const { return dwt_bytes_; }
    int64_t final_bytes() const { return final_bytes_; }
    // Capture run() as a graph and replay it (batched form).
    Status capture_graph(std::string& err);
    Status launch_graph();
    void drop_graph();
    size_t device_bytes() const;

  private:
    struct Group {  // one t1 launch
        bool small = false;
        int bpw = 1;
        int max_planes = 0;  // most bit-planes of a block in the class
        int max_passes = 0;  // most coding passes of a block in the class
        uint64_t order_off = 0;
        uint64_t scratch_off = 0;  // lockstep packing: its state area in scratch_
        int n = 0;
    };
    struct LevelLaunch {
        bool irr = false;
        uint64_t jobs_off = 0, map_off = 0;
        int nwg = 0;
    };
    struct FinalLaunch {
        bool irr = false;
        int nc = 1, ob = 1;
        uint64_t jobs_off = 0, map_off = 0;
        int nwg = 0;
    };
    struct ImageOut {
        uint64_t off = 0;  // bytes into the output arena
        uint64_t bytes = 0;
        int bps = 1;
    };

    Status prepare_plan(const std::vector<const Image*>& imgs, std::string& err);
    Status ensure(tc_buf& b, size_t bytes, const char* what, std::string& err);
    Status ensure_pinned(size_t bytes, std::string& err);

    tc_device dev_;
    tc_stream stream_ = nullptr;
    tc_buf upload_ = nullptr, samples_ = nullptr, out_ = nullptr, scratch_ = nullptr;
    bool host_out_ = false;
    void* host_buf_ = nullptr;  // pinned output (host_out_)
    void* host_dev_ = nullptr;  // its device address
    size_t host_cap_ = 0;
    void* pinned_ = nullptr;
    size_t pinned_cap_ = 0;
    std::vector<Group> groups_;
    std::vector<LevelLaunch> levels_;
    std::vector<FinalLaunch> finals_;
    std::vector<ImageOut> images_;
    uint64_t descs_off_ = 0, coef_elems_ = 0, cw_size_ = 0;
    int64_t t1_samples_ = 0;
    int64_t ncblk_ = 0, cw_bytes_ = 0, nsamples_ = 0, npixels_ = 0, dwt_bytes_ = 0, final_bytes_ = 0;
    int bpw_[2] = {0, 0};
    bool styled_[2] = {false, false};
    Tuning tuning_;
    int shape_ = 0;  // tile shape the prepared job lists were built for
    tc_event ev_[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    tc_stream side_ = nullptr;           // second tier-1 class runs here, beside the first
    tc_event fork_ = nullptr, join_ = nullptr;
    bool timed_ = false;
    tc_graph graph_ = nullptr;
    bool ready_ = false;  // a prepare succeeded and nothing has invalidated it since
};

}  // namespace tc
