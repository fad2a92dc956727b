// Device work planning and execution for one image or a batch of images.
//
// The reference runs, per tile, CoefficientCoder::decode_tile ->
// Quantizer::dequantize_tile -> DWT::iwt, then the Preprocessor over the image
// (TC/Decoder.cpp:153-180), allocating device buffers inside every stage
// (TC/CoefficientCoder.cpp:185-205, TC/Quantizer.cpp:80, TC/DWT.cpp:79-84).
// Here a Plan turns any number of parsed images into flat device work lists,
// and launch_plan runs one launch per stage for all of them:
//   fill (zero the coefficient planes) -> tier-1 (one launch per code-block
//   size class) -> one inverse-DWT launch per level step -> the fused final
//   level (colour transform, DC shift, clamp, pack).
// A Plan is host data only: it says how large each device region must be and
// where everything sits in it.  A Pipeline owns resident device buffers that
// only grow, so repeated decodes allocate nothing, and uploads all
// descriptors in one host-to-device copy; the no-allocation entry point runs
// the same plan in a static device array instead.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "../../../include/tc_decode.h"
#include "../../../include/tc_device.h"
#include "../tc_kernels.h"
#include "codestream.hpp"
#include "plan_util.hpp"

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
    // each tier-1 launch on its own stream, by block class (0: blocks up to 32x32)
    float t1_class_ms[2] = {0, 0};
};

// The device work of a batch: launches, their descriptor offsets in the upload
// region, and the sizes of the four device regions it needs.
struct Plan {
    struct Group {  // one t1 launch
        bool small = false;
        int bpw = 1;
        int max_planes = 0;  // most bit-planes of a block in the class
        int max_passes = 0;  // most coding passes of a block in the class
        uint64_t order_off = 0;
        uint64_t scratch_off = 0;  // lockstep packing: its state area in the tier-1 scratch region
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
        uint64_t off = 0;  // bytes into the output region
        uint64_t bytes = 0;
        int bps = 1;
    };

    std::vector<Group> groups;
    std::vector<LevelLaunch> levels;
    std::vector<FinalLaunch> finals;
    std::vector<ImageOut> images;
    int shape = 0;   // inverse-DWT tile shape the job lists were built for
    int t1_kernel = 0;
    // upload region: [0, cw_end) codewords, 0xFF up to blob_base, then blob
    uint64_t cw_end = 0, cw_size = 0, blob_base = 0;
    std::vector<uint8_t> blob;     // descriptors, packing orders, job lists
    std::vector<CwCopy> copies;    // codeword segments to gather into [0, cw_end)
    uint64_t descs_off = 0;
    uint64_t coef_elems = 0;       // coefficient planes (zeroed by the fill) ...
    uint64_t sample_elems = 0;     // ... followed by the DWT work buffers: 4 B each
    uint64_t scratch_bytes = 0;    // tier-1 state of the lockstep packings
    uint64_t out_bytes = 0;        // all output images
    int bpw[2] = {0, 0};           // packing per class (0: no blocks)
    // tier-1 algorithmic bytes per class: codeword bytes read + 4 B per
    // coefficient of the blocks that have coding passes
    int64_t t1_class_bytes[2] = {0, 0};
    bool styled[2] = {false, false};
    int64_t ncblk = 0, cw_bytes = 0, nsamples = 0, npixels = 0, dwt_bytes = 0, final_bytes = 0, t1_samples = 0;

    uint64_t upload_bytes() const { return blob_base + blob.size(); }
};

// Builds the plan of `imgs` for a device of `cus` compute units.
Status build_plan(const std::vector<const Image*>& imgs, const Tuning& tuning, int cus, Plan& P, std::string& err);

// Writes the upload region (upload_bytes() bytes) into host memory at dst.
void stage_upload(const Plan& P, uint8_t* dst);

// Device addresses of the regions a plan runs in.
struct PlanRegions {
    uint8_t* upload = nullptr;
    void* samples = nullptr;
    void* out = nullptr;
    uint8_t* scratch = nullptr;
    size_t scratch_bytes = 0;
};

// Enqueues the plan on `stream`.  With a side stream and fork/join events the
// second tier-1 class runs beside the first; with ev (nine events) the stage
// boundaries are recorded on `stream` (ev[0..4]) and each tier-1 launch is
// bracketed on the stream it runs on (ev[5..8], two per block class).
tc_status launch_plan(const Plan& P, const PlanRegions& R, int cus, tc_stream stream, tc_stream side, tc_event fork,
                      tc_event join, tc_event* ev);

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
    int output_sample_bytes(int i) const { return plan_.images[i].bps; }
    int num_images() const { return ready_ ? (int)plan_.images.size() : 0; }
    StageTimes times() const;
    tc_stream stream() const { return stream_; }
    void set_tuning(const Tuning& t) { tuning_ = t; }
    // output images in pinned host memory, written by the final-level kernel
    // directly (no device buffer, no copy); from the next prepare
    void set_host_output(bool on) { host_out_ = on; }
    const Tuning& tuning() const { return tuning_; }
    // Statistics of the last prepare().
    int64_t num_cblks() const { return plan_.ncblk; }
    int64_t codeword_bytes() const { return plan_.cw_bytes; }
    int64_t sample_count() const { return plan_.nsamples; }
    int64_t t1_sample_count() const { return plan_.t1_samples; }
    int64_t pixel_count() const { return plan_.npixels; }
    // packing the prepared plan runs for class cls (0: blocks up to 32x32), 0 if empty
    int blocks_per_wave(int cls) const { return plan_.bpw[cls]; }
    // the class holds code-blocks with a style (only lockstep packings decode them)
    bool styled(int cls) const { return plan_.styled[cls]; }
    // bytes moved by the HBM-bound stages (algorithmic): dwt levels and final
    int64_t dwt_bytes() const { return plan_.dwt_bytes; }
    int64_t t1_class_bytes(int cls) const { return plan_.t1_class_bytes[cls]; }
    int64_t final_bytes() const { return plan_.final_bytes; }
    // Capture run() as a graph and replay it (batched form).
    Status capture_graph(std::string& err);
    Status launch_graph();
    void drop_graph();
    // device memory the pipeline holds (its four buffers, as allocated)
    size_t device_bytes() const;
    // the most device memory a pipeline holds after preparing P, counting the
    // geometric growth of its buffers (each grows to twice what it is asked
    // for) and the 1 MiB floor of each
    static size_t device_bytes_bound(const Plan& P, bool host_out);

  private:
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
    Plan plan_;
    Tuning tuning_;
    tc_event ev_[9] = {};
    tc_stream side_ = nullptr;           // second tier-1 class runs here, beside the first
    tc_event fork_ = nullptr, join_ = nullptr;
    bool timed_ = false;
    tc_graph graph_ = nullptr;
    bool ready_ = false;  // a prepare succeeded and nothing has invalidated it since
};

}  // namespace tc
