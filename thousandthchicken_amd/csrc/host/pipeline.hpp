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
    // A class packed in lockstep sends its blocks of more estimated MQ
    // symbols than this to a one-wave-per-block launch of their own, whose
    // waves decode a long symbol chain several times faster than a lockstep
    // lane does (0: no split; -1: the default rule, default_heavy).
    int heavy64 = -1, heavy32 = -1;
    // waves per compute unit of the heavy blocks' launch (each wave decodes
    // its share of them one after the other), so that they leave room on
    // every compute unit for the lockstep waves
    int heavy_waves_per_cu = 4;
};

// Estimated MQ symbols of a code-block from its descriptor: a least-squares
// fit over the 250 000 blocks of the 256-image mixed batch (correlation 0.996
// with the oracle's counts).
inline int64_t estimated_symbols(const tc_cblk_desc& d) {
    return (int64_t)(6.3 * d.data_len + 7.5 * d.npasses * ((double)d.w * d.h / 64.0)) + 200;
}

// Heavy-block threshold used when Tuning says -1.
int default_heavy(bool small_blocks);

// Tier-1 launches of a plan: at most two per class (the heavy blocks' one
// wave per block launch, then the class's main launch).
constexpr int kMaxT1Launches = 4;

// Chooses blocks-per-wave from the block count and the machine size, by a rule
// taken from the autotuner's choices; a fixed value in Tuning overrides it.
int pick_blocks_per_wave(int nblocks, bool small_blocks, int cu_count, const Tuning& t);

struct StageTimes {
    float fill_ms = 0, t1_ms = 0, dwt_ms = 0, final_ms = 0, total_ms = 0;
    // each tier-1 launch, timed on the stream it runs on (order of Plan::groups)
    float t1_launch_ms[kMaxT1Launches] = {0, 0, 0, 0};
};

// The device work of a batch: launches, their descriptor offsets in the upload
// region, and the sizes of the four device regions it needs.
struct Plan {
    int heavy_waves_per_cu = 4;
    struct Group {  // one t1 launch
        bool small = false;
        int bpw = 1;
        int max_planes = 0;  // most bit-planes of a block in the class
        int max_passes = 0;  // most coding passes of a block in the class
        uint64_t order_off = 0;
        uint64_t scratch_off = 0;  // lockstep packing: its state area in the tier-1 scratch region
        int n = 0;
        bool heavy = false;  // the class's heaviest blocks, split off its lockstep launch
        int64_t bytes = 0;   // algorithmic bytes: codeword bytes + 4 B per coefficient
        int64_t max_symbols = 0;  // largest estimated_symbols of a block in the launch
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
    int heavy[2] = {0, 0};         // heavy-block threshold applied per class (0: none)
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

// Streams and events for running a plan's tier-1 launches side by side.
struct LaunchStreams {
    tc_stream side[kMaxT1Launches - 1] = {};  // launch g > 0 runs on side[g - 1]
    tc_event fork = nullptr;
    tc_event join[kMaxT1Launches - 1] = {};
    tc_event dwt_fork = nullptr, dwt_join = nullptr;  // the 9/7 DWT chain on side[0]
};

// Enqueues the plan on `stream`.  With side streams (S, may be null) the
// tier-1 launches run side by side, the heavy blocks' launches last, and the
// 9/7 inverse-DWT chain runs beside the 5/3 one; with
// ev (5 + 2 * kMaxT1Launches events) the stage boundaries are recorded on
// `stream` (ev[0..4]) and each tier-1 launch is bracketed on the stream it
// runs on (ev[5 + 2g], ev[6 + 2g]).
tc_status launch_plan(const Plan& P, const PlanRegions& R, int cus, tc_stream stream, const LaunchStreams* S,
                      tc_event* ev);

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
    // The two halves of prepare, for callers that plan the next batch while
    // the current one runs: plan_batch is host work only (it touches no device
    // state); adopt makes the buffers fit P and enqueues the upload of its
    // staged region (stage_upload's P.upload_bytes() bytes at `staged`, which
    // must stay unchanged until the stream has passed the copy).
    Status plan_batch(const std::vector<const Image*>& imgs, Plan& P, std::string& err) const;
    Status adopt(Plan&& P, const void* staged, std::string& err);
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
    const Plan& plan() const { return plan_; }
    int64_t final_bytes() const { return plan_.final_bytes; }
    // Capture run() as a graph and replay it (batched form).
    Status capture_graph(std::string& err);
    Status launch_graph();
    void drop_graph();
    // device memory the pipeline holds (its four buffers, as allocated)
    size_t device_bytes() const;
    // A buffer that must grow is allocated this many times the size asked
    // for (default 2: few reallocations; each one synchronises the stream).
    void set_growth(double factor) { growth_ = factor < 1.0 ? 1.0 : factor; }
    // Frees the device buffers (after the stream drains); the next prepare
    // allocates again.  The current plan is dropped.
    void release_buffers();


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
    tc_event ev_[5 + 2 * kMaxT1Launches] = {};
    LaunchStreams ls_;  // tier-1 launches after the first run on these, side by side
    bool timed_ = false;
    tc_graph graph_ = nullptr;
    bool ready_ = false;  // a prepare succeeded and nothing has invalidated it since
    double growth_ = 2.0;
};

}  // namespace tc
