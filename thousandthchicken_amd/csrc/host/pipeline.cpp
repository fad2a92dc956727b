// Device work planning and execution (see pipeline.hpp).
#include "pipeline.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cstring>
#include <thread>

#include "../tc_internal.hpp"
#include "plan_util.hpp"

namespace tc {

int idwt_tile_w(int shape);
int idwt_tile_h(int shape);

namespace {

uint64_t align_up(uint64_t v, uint64_t a) { return (v + a - 1) / a * a; }

// Upload blob: the codeword arena [0, base) is gathered straight into the
// staging buffer; v holds everything after it.
struct Blob {
    uint64_t base = 0;
    std::vector<uint8_t> v;
};

template <typename T>
uint64_t put(Blob& blob, const std::vector<T>& v) {
    uint64_t off = align_up(blob.v.size(), 256);
    blob.v.resize(off + v.size() * sizeof(T));
    if (!v.empty()) std::memcpy(blob.v.data() + off, v.data(), v.size() * sizeof(T));
    return blob.base + off;
}

// Gathers the codeword segments into the staging buffer, on up to six
// threads (one per 16 MB): one thread copies about 20 GB/s, six about 60.
void gather(uint8_t* dst, const std::vector<CwCopy>& copies, uint64_t bytes) {
    const int nt = (int)std::min<uint64_t>(6, 1 + bytes / (16u << 20));
    auto part = [&](int k) {
        const size_t a = copies.size() * k / nt, b = copies.size() * (k + 1) / nt;
        for (size_t i = a; i < b; i++) std::memcpy(dst + copies[i].dst, copies[i].src, copies[i].len);
    };
    if (nt == 1) {
        part(0);
        return;
    }
    std::vector<std::thread> th;
    for (int k = 1; k < nt; k++) th.emplace_back(part, k);
    part(0);
    for (auto& t : th) t.join();
}

Status from_tc(tc_status s) { return s == TC_SUCCESS ? Status::Ok : (s == TC_ERR_OUT_OF_MEMORY ? Status::OutOfMemory : Status::Device); }

}  // namespace

// Default packing when none is set, from timings on MI355X: lockstep from 64
// blocks per CU up, one wave per block (stripe kernel) below, where a few long
// lockstep waves would leave most SIMDs idle.  Lockstep waves hold 64 blocks
// from 8 such waves per CU up, else 32 (more, half-full waves run side by
// side).  Tier-1 of 7680x4320 16-bit lossless images (8227 blocks of 64x64
// each): one image 33 ms one wave per block, 74 ms in lockstep; four 134 ms
// one per wave, 67 ms lockstep-64, 58 ms lockstep-32; eight 52 ms
// lockstep-32 against 59 ms lockstep-64.
int pick_blocks_per_wave(int n, bool small, int cus, const Tuning& t) {
    int fixed = small ? t.bpw32 : t.bpw64;
    if (fixed) return fixed;
    const int c = std::max(cus, 1);
    if (n < 64 * c) return 1;
    return n >= 8 * 64 * c ? TC_BPW_LOCKSTEP : TC_BPW_LOCKSTEP32;
}

// Heavy-block threshold (estimated MQ symbols) when none is set: none.  Which
// split helps depends on the batch: on the 256-image mixed batch the blocks
// above 20 000 symbols gain 4 % (their launch last), those above 16 000 lose
// 40 %, and the one-wave-per-block waves compete with the lockstep waves for
// their compute units' scalar unit and LDS; the autotuner tries thresholds.
int default_heavy(bool) { return 0; }

Pipeline::Pipeline(tc_device dev) : dev_(dev) {
    if (tc_stream_create(dev, &stream_) != TC_SUCCESS) stream_ = nullptr;
    for (auto& sd : ls_.side)
        if (tc_stream_create(dev, &sd) != TC_SUCCESS) sd = nullptr;
    for (auto& e : ev_) tc_event_create(&e);
    tc_event_create(&ls_.fork);
    for (auto& e : ls_.join) tc_event_create(&e);
    tc_event_create(&ls_.dwt_fork);
    tc_event_create(&ls_.dwt_join);
}

Pipeline::~Pipeline() {
    if (stream_) tc_stream_sync(stream_);
    drop_graph();
    for (auto& sd : ls_.side)
        if (sd) tc_stream_sync(sd);
    for (auto& e : ev_) tc_event_destroy(e);
    tc_event_destroy(ls_.fork);
    for (auto& e : ls_.join) tc_event_destroy(e);
    tc_event_destroy(ls_.dwt_fork);
    tc_event_destroy(ls_.dwt_join);
    for (auto& sd : ls_.side) tc_stream_destroy(sd);
    tc_buf_free(upload_);
    tc_buf_free(samples_);
    tc_buf_free(out_);
    tc_buf_free(scratch_);
    if (pinned_) tc_host_free_pinned(pinned_);
    if (host_buf_) tc_host_free_pinned(host_buf_);
    tc_stream_destroy(stream_);
}

Status Pipeline::ensure(tc_buf& b, size_t bytes, const char* what, std::string& err) {
    if (b && tc_buf_size(b) >= bytes) return Status::Ok;
    if (b) {
        tc_stream_sync(stream_);
        tc_buf_free(b);
        b = nullptr;
    }
    // grow geometrically: reallocation syncs the device
    size_t want = std::max<size_t>((size_t)((double)bytes * growth_), 1 << 20);
    tc_status s = tc_buf_alloc(dev_, want, &b);
    if (s != TC_SUCCESS) {
        err = std::string("device allocation of ") + what + " failed: " + tc_last_error();
        b = nullptr;
        return from_tc(s);
    }
    return Status::Ok;
}

Status Pipeline::ensure_pinned(size_t bytes, std::string& err) {
    if (pinned_ && pinned_cap_ >= bytes) return Status::Ok;
    if (pinned_) tc_host_free_pinned(pinned_);
    pinned_ = nullptr;
    size_t want = std::max<size_t>(2 * bytes, 1 << 20);  // grow geometrically: reallocation syncs the device
    tc_status s = tc_host_alloc_pinned(want, &pinned_);
    if (s != TC_SUCCESS) {
        err = std::string("pinned allocation failed: ") + tc_last_error();
        pinned_ = nullptr;
        return from_tc(s);
    }
    pinned_cap_ = want;
    return Status::Ok;
}

Status build_plan(const std::vector<const Image*>& imgs, const Tuning& tuning, int cus, Plan& P, std::string& err) {
    P = Plan();
    P.shape = tuning.dwt_shape;
    P.t1_kernel = tuning.t1_kernel;
    P.heavy_waves_per_cu = std::max(1, tuning.heavy_waves_per_cu);
    const int TX = idwt_tile_w(P.shape), TY = idwt_tile_h(P.shape);

    // ---- sample arena layout: coefficient planes, then work buffers ----
    struct TCPos {
        uint64_t plane = 0, work[2] = {0, 0};
    };
    std::vector<std::vector<std::vector<TCPos>>> pos(imgs.size());
    uint64_t coef = 0;
    for (size_t i = 0; i < imgs.size(); i++) {
        const Image& im = *imgs[i];
        pos[i].resize(im.tiles.size());
        for (size_t t = 0; t < im.tiles.size(); t++) {
            pos[i][t].resize(im.ncomp);
            for (int c = 0; c < im.ncomp; c++) {
                const TileComp& tc = im.tiles[t].comps[c];
                pos[i][t][c].plane = coef;
                coef += align_up((uint64_t)tc.width() * tc.height(), 64);
            }
        }
    }
    uint64_t work = coef;
    for (size_t i = 0; i < imgs.size(); i++) {
        const Image& im = *imgs[i];
        for (size_t t = 0; t < im.tiles.size(); t++)
            for (int c = 0; c < im.ncomp; c++) {
                uint64_t sz[2];
                work_sizes(im.tiles[t].comps[c], sz);
                for (int k = 0; k < 2; k++) {
                    pos[i][t][c].work[k] = work;
                    work += align_up(sz[k], 64);
                }
            }
    }
    P.coef_elems = coef;

    // ---- codewords and tier-1 descriptors ----
    uint64_t cw_end = 0;
    std::vector<CwCopy> copies;
    std::vector<tc_cblk_desc> descs;
    std::vector<uint32_t> cls_idx[2];
    for (size_t i = 0; i < imgs.size(); i++) {
        const Image& im = *imgs[i];
        for (size_t t = 0; t < im.tiles.size(); t++) {
            std::vector<uint64_t> planes(im.ncomp);
            for (int c = 0; c < im.ncomp; c++) planes[c] = pos[i][t][c].plane;
            Status st = plan_tile_cblks(im, (int)t, planes.data(), cw_end, copies, descs, cls_idx, err);
            if (st != Status::Ok) return st;
        }
    }
    P.cw_bytes = (int64_t)cw_end;
    P.cw_size = ((cw_end + 15) & ~(uint64_t)15) + 64;  // slack of 0xFF for look-ahead reads
    P.ncblk = (int64_t)descs.size();
    P.t1_samples = 0;
    for (const auto& d : descs) P.t1_samples += (int64_t)d.w * d.h;

    Blob blob;
    blob.base = align_up(P.cw_size, 256);
    P.descs_off = put(blob, descs);
    uint64_t scratch_total = 0;
    auto group_of = [&](const std::vector<uint32_t>& v, bool small, int bpw, bool heavy) {
        Plan::Group g;
        g.small = small;
        g.bpw = bpw;
        g.heavy = heavy;
        g.n = (int)v.size();
        for (uint32_t i : v) {
            g.max_planes = std::max<int>(g.max_planes, descs[i].numbps);
            g.max_passes = std::max<int>(g.max_passes, descs[i].npasses);
            g.bytes += (int64_t)descs[i].data_len + 4 * (int64_t)descs[i].w * descs[i].h;
            g.max_symbols = std::max(g.max_symbols, estimated_symbols(descs[i]));
        }
        g.order_off = put(blob, v);
        if (g.bpw < 0) {  // lockstep, -bpw blocks per wave
            g.scratch_off = scratch_total;
            scratch_total += align_up(tc_t1_lockstep_scratch_bytes(g.small ? 1 : 0, -g.bpw, g.n, g.max_passes), 256);
        }
        return g;
    };
    std::vector<Plan::Group> heavy_groups, main_groups;
    for (int k = 0; k < 2; k++) {
        auto& v = cls_idx[k];
        P.bpw[k] = 0;
        if (v.empty()) continue;
        const bool small = k == 0;
        int bpw = pick_blocks_per_wave((int)v.size(), small, cus, tuning);
        int style = 0;
        for (uint32_t i : v) style |= descs[i].mode & TC_MODE_STYLE;
        // only the lockstep kernel decodes the code-block styles
        if (style && bpw >= 0) bpw = (int)v.size() >= 8 * 64 * std::max(cus, 1) ? TC_BPW_LOCKSTEP : TC_BPW_LOCKSTEP32;
        // a lockstep class sends its longest symbol chains to one wave per block
        int thr = small ? tuning.heavy32 : tuning.heavy64;
        if (thr < 0) thr = default_heavy(small);
        std::vector<uint32_t> heavy;
        if (bpw < 0 && thr > 0) {
            std::vector<uint32_t> light;
            for (uint32_t i : v) {
                const tc_cblk_desc& d = descs[i];
                const bool h = !(d.mode & TC_MODE_STYLE) && d.numbps <= TC_T1_STRIPE_MAX_PLANES(small) &&
                               estimated_symbols(d) > thr;
                (h ? heavy : light).push_back(i);
            }
            v.swap(light);
            // longest chains first: they start first and bound the launch
            std::stable_sort(heavy.begin(), heavy.end(), [&](uint32_t a, uint32_t b) {
                return estimated_symbols(descs[a]) > estimated_symbols(descs[b]);
            });
        }
        if (!heavy.empty()) heavy_groups.push_back(group_of(heavy, small, 1, true));
        if (!v.empty()) {
            sort_for_packing(v, descs, small);
            main_groups.push_back(group_of(v, small, bpw, false));
        }
        P.bpw[k] = v.empty() ? 1 : bpw;
        P.heavy[k] = heavy.empty() ? 0 : thr;
        P.styled[k] = style != 0;
    }
    // The class whose longest block chain is longest goes first: launches
    // dispatch their waves in submission order, and the waves resident first
    // are the ones that bound the tier-1 stage (on the 256-image mixed batch,
    // the 64x64 class first takes the stage from 35.3 to 32.8 ms; its waves
    // otherwise wait for the 32x32 class's to leave).  Heavy blocks' launches
    // go last: their one-wave-per-block waves are short beside the lockstep
    // waves and fill the compute units those leave (first, they slowed the
    // lockstep waves and made the stage vary by 30 %; last, blocks above 20 000
    // estimated symbols take it from 30.1 to 29.0 ms on that batch).
    std::stable_sort(main_groups.begin(), main_groups.end(),
                     [](const Plan::Group& a, const Plan::Group& b) { return a.max_symbols > b.max_symbols; });
    for (auto& g : main_groups) P.groups.push_back(g);
    for (auto& g : heavy_groups) P.groups.push_back(g);

    // ---- inverse DWT levels ----
    int maxlev = 0;
    for (auto* im : imgs)
        for (auto& t : im->tiles)
            for (auto& c : t.comps) maxlev = std::max(maxlev, c.nlev);
    for (int s = 0; s + 2 <= maxlev; s++) {
        std::vector<tc_dwt_job> jobs[2];
        std::vector<uint32_t> maps[2];
        for (size_t i = 0; i < imgs.size(); i++) {
            const Image& im = *imgs[i];
            for (size_t t = 0; t < im.tiles.size(); t++)
                for (int c = 0; c < im.ncomp; c++) {
                    const TileComp& tc = im.tiles[t].comps[c];
                    int n = tc.nlev - s;
                    if (n < 2) continue;
                    const Resolution& R = tc.res[tc.nlev - n + 1];
                    if (R.width() <= 0 || R.height() <= 0) continue;
                    const TCPos& Q = pos[i][t][c];
                    tc_dwt_job j = level_job(tc, n, Q.plane, Q.work, false, 0, 0, TX, TY);
                    int irr = tc.reversible ? 0 : 1;
                    int tiles_y = (R.height() + TY - 1) / TY;
                    j.wg_base = (int32_t)maps[irr].size();
                    uint32_t jid = (uint32_t)jobs[irr].size();
                    for (int q = 0; q < j.tiles_x * tiles_y; q++) maps[irr].push_back(jid);
                    jobs[irr].push_back(j);
                    P.dwt_bytes += (int64_t)R.width() * R.height() * 8;
                }
        }
        for (int irr = 0; irr < 2; irr++) {
            if (jobs[irr].empty()) continue;
            Plan::LevelLaunch L;
            L.irr = irr != 0;
            L.jobs_off = put(blob, jobs[irr]);
            L.map_off = put(blob, maps[irr]);
            L.nwg = (int)maps[irr].size();
            P.levels.push_back(L);
        }
    }

    // ---- final level + colour + pack, and output placement ----
    uint64_t out_bytes_total = 0;
    struct FKey {
        int irr, nc, ob;
    };
    std::vector<tc_final_job> fjobs[8];
    std::vector<uint32_t> fmaps[8];
    auto key_index = [](int irr, int nc, int ob) { return irr * 4 + (nc == 3 ? 2 : 0) + (ob == 2 ? 1 : 0); };
    for (size_t i = 0; i < imgs.size(); i++) {
        const Image& im = *imgs[i];
        Plan::ImageOut io;
        io.bps = im.max_prec() > 8 ? 2 : 1;
        io.off = align_up(out_bytes_total, 256);
        const int ow = im.out_width(), oh = im.out_height();
        const bool planar = im.planar();
        io.bytes = im.out_bytes();
        out_bytes_total = io.off + io.bytes;
        P.images.push_back(io);
        P.npixels += (int64_t)ow * oh;
        P.nsamples += (int64_t)(io.bytes / io.bps);
        uint64_t out_elem_base = io.off / io.bps;
        // planar output (sub-sampled components): each component's plane, in turn
        std::vector<uint64_t> cbase(im.ncomp, 0);
        std::vector<std::array<int, 4>> crect(im.ncomp);
        for (int c = 0, acc = 0; planar && c < im.ncomp; c++) {
            im.comp_rect(c, crect[c][0], crect[c][1], crect[c][2], crect[c][3]);
            cbase[c] = out_elem_base + (uint64_t)acc;
            acc += (crect[c][2] - crect[c][0]) * (crect[c][3] - crect[c][1]);
        }
        for (size_t t = 0; t < im.tiles.size(); t++) {
            const Tile& T = im.tiles[t];
            std::vector<std::vector<int>> sets;
            bool mct3 = !planar && T.mct && im.ncomp >= 3 && T.comps[0].reversible == T.comps[1].reversible &&
                        T.comps[0].reversible == T.comps[2].reversible && T.comps[0].nlev == T.comps[1].nlev &&
                        T.comps[0].nlev == T.comps[2].nlev;
            int c0 = 0;
            if (mct3) {
                sets.push_back({0, 1, 2});
                c0 = 3;
            }
            for (int c = c0; c < im.ncomp; c++) sets.push_back({c});
            for (auto& set : sets) {
                const TileComp& tc0 = T.comps[set[0]];
                const Resolution& R = tc0.res[tc0.nlev];
                if (R.width() <= 0 || R.height() <= 0) continue;
                // the part of the output window this tile covers (planar: the
                // whole tile-component)
                const int jx0 = planar ? R.x0 : std::max(R.x0, im.wx0), jx1 = planar ? R.x1 : std::min(R.x1, im.wx1);
                const int jy0 = planar ? R.y0 : std::max(R.y0, im.wy0), jy1 = planar ? R.y1 : std::min(R.y1, im.wy1);
                if (jx1 <= jx0 || jy1 <= jy0) continue;
                tc_final_job f;
                std::memset(&f, 0, sizeof f);
                f.u0 = R.x0;
                f.u1 = R.x1;
                f.v0 = R.y0;
                f.v1 = R.y1;
                f.wx0 = jx0;
                f.wx1 = jx1;
                f.wy0 = jy0;
                f.wy1 = jy1;
                f.levels = tc0.nlev;
                f.mct = set.size() == 3 ? 1 : 0;
                for (size_t k = 0; k < set.size(); k++) {
                    int c = set[k];
                    const TileComp& tc = T.comps[c];
                    const TCPos& Q = pos[i][t][c];
                    if (tc.nlev == 0) {
                        f.ll_off[k] = f.hl_off[k] = f.lh_off[k] = f.hh_off[k] = Q.plane;
                        f.ll_pitch[k] = f.band_pitch[k] = (uint32_t)tc.width();
                    } else {
                        const Resolution& Rl = tc.res[tc.nlev - 1];
                        int lw = Rl.width(), lh = Rl.height();
                        if (tc.nlev == 1) {
                            f.ll_off[k] = Q.plane;
                            f.ll_pitch[k] = (uint32_t)tc.width();
                        } else {
                            f.ll_off[k] = Q.work[(tc.nlev - 2) % 2];
                            f.ll_pitch[k] = (uint32_t)lw;
                        }
                        f.hl_off[k] = Q.plane + lw;
                        f.lh_off[k] = Q.plane + (uint64_t)lh * tc.width();
                        f.hh_off[k] = Q.plane + (uint64_t)lh * tc.width() + lw;
                        f.band_pitch[k] = (uint32_t)tc.width();
                    }
                    int prec = im.prec[c];
                    if (im.sgnd[c]) {
                        f.shift[k] = 0;
                        f.minv[k] = -(1 << (prec - 1));
                        f.maxv[k] = (1 << (prec - 1)) - 1;
                    } else {
                        f.shift[k] = 1 << (prec - 1);
                        f.minv[k] = 0;
                        f.maxv[k] = (1 << prec) - 1;
                    }
                }
                if (planar) {
                    const auto& cr = crect[set[0]];
                    const int cw = cr[2] - cr[0];
                    f.out_off = cbase[set[0]] + (uint64_t)(jy0 - cr[1]) * cw + (uint64_t)(jx0 - cr[0]);
                    f.out_pitch = (uint32_t)cw;
                    f.out_cstride = 1;
                } else {
                    f.out_off = out_elem_base + ((uint64_t)(jy0 - im.wy0) * ow + (uint64_t)(jx0 - im.wx0)) * im.ncomp +
                                (uint64_t)set[0];
                    f.out_pitch = (uint32_t)((uint64_t)ow * im.ncomp);
                    f.out_cstride = (uint32_t)im.ncomp;
                }
                int irr = tc0.reversible ? 0 : 1, nc = (int)set.size();
                f.tiles_x = (jx1 - jx0 + TX - 1) / TX;
                int tiles_y = (jy1 - jy0 + TY - 1) / TY;
                int key = key_index(irr, nc, io.bps);
                f.wg_base = (int32_t)fmaps[key].size();
                uint32_t jid = (uint32_t)fjobs[key].size();
                for (int q = 0; q < f.tiles_x * tiles_y; q++) fmaps[key].push_back(jid);
                fjobs[key].push_back(f);
                P.final_bytes += (int64_t)(jx1 - jx0) * (jy1 - jy0) * nc * (4 + io.bps);
            }
        }
    }
    for (int irr = 0; irr < 2; irr++)
        for (int nc : {1, 3})
            for (int ob : {1, 2}) {
                int key = key_index(irr, nc, ob);
                if (fjobs[key].empty()) continue;
                Plan::FinalLaunch F;
                F.irr = irr != 0;
                F.nc = nc;
                F.ob = ob;
                F.jobs_off = put(blob, fjobs[key]);
                F.map_off = put(blob, fmaps[key]);
                F.nwg = (int)fmaps[key].size();
                P.finals.push_back(F);
            }

    P.cw_end = cw_end;
    P.copies = std::move(copies);
    P.blob_base = blob.base;
    P.blob = std::move(blob.v);
    P.sample_elems = work;
    P.scratch_bytes = scratch_total;
    P.out_bytes = out_bytes_total;
    return Status::Ok;
}

void stage_upload(const Plan& P, uint8_t* stage) {
    // alignment gaps between codewords are never read as data; the slack after
    // the last one reads as 0xFF
    std::memset(stage + P.cw_end, 0xFF, P.blob_base - P.cw_end);
    gather(stage, P.copies, P.cw_end);
    if (!P.blob.empty()) std::memcpy(stage + P.blob_base, P.blob.data(), P.blob.size());
}

tc_status launch_plan(const Plan& P, const PlanRegions& R, int cus, tc_stream stream, const LaunchStreams* S,
                      tc_event* ev) {
    uint8_t* up = R.upload;
    void* base = R.samples;
    if (!up || !base || !R.out) return TC_ERR_INVALID_ARG;
    if (P.groups.size() > (size_t)kMaxT1Launches) return TC_ERR_INVALID_ARG;
    for (const Plan::Group& g : P.groups)
        if (g.bpw < 0 && (!R.scratch || R.scratch_bytes <= g.scratch_off)) return TC_ERR_INVALID_ARG;
    tc_status s = TC_SUCCESS;
    if (ev) tc_event_record(ev[0], stream);
    s = tc_launch_fill32(stream, base, 0u, P.coef_elems);
    if (ev) tc_event_record(ev[1], stream);
    // Tier-1: launch g > 0 runs on side stream g - 1, so that every launch's
    // waves share the machine with the others' (a short launch's tail does not
    // idle the GPU while a long one still runs)
    const int ng = (int)P.groups.size();
    bool fork = S && S->fork && ng > 1;
    for (int g = 1; fork && g < ng; g++) fork = S->side[g - 1] && S->join[g - 1];
    if (fork && s == TC_SUCCESS) {
        s = tc_event_record(S->fork, stream);
        for (int g = 1; s == TC_SUCCESS && g < ng; g++) s = tc_stream_wait_event(S->side[g - 1], S->fork);
    }
    for (int gi = 0; gi < ng; gi++) {
        if (s != TC_SUCCESS) break;
        const Plan::Group& g = P.groups[gi];
        tc_stream st = (fork && gi > 0) ? S->side[gi - 1] : stream;
        tc_event* gev = ev ? ev + 5 + 2 * gi : nullptr;
        if (gev) tc_event_record(gev[0], st);
        if (g.bpw < 0)
            s = tc_launch_t1_decode_lockstep(st, g.small ? 1 : 0, -g.bpw, up + P.descs_off,
                                             (const uint32_t*)(up + g.order_off), g.n, g.max_passes, up, P.cw_size,
                                             R.scratch + g.scratch_off, R.scratch_bytes - g.scratch_off, base);
        else if (g.bpw == 1 && P.t1_kernel == 0 && g.max_planes <= TC_T1_STRIPE_MAX_PLANES(g.small))
            s = tc_launch_t1_decode_stripe(st, g.small ? 1 : 0,
                                           std::min(g.n, cus * (g.heavy ? P.heavy_waves_per_cu : 32)),
                                           up + P.descs_off, (const uint32_t*)(up + g.order_off), g.n, up, P.cw_size,
                                           base);
        else
            s = tc_launch_t1_decode_wave(st, g.small ? 1 : 0, std::min(g.n, cus * 16), up + P.descs_off,
                                         (const uint32_t*)(up + g.order_off), g.n, up, P.cw_size, base);
        if (gev) tc_event_record(gev[1], st);
    }
    for (int g = 1; fork && s == TC_SUCCESS && g < ng; g++) {
        s = tc_event_record(S->join[g - 1], S->side[g - 1]);
        if (s == TC_SUCCESS) s = tc_stream_wait_event(stream, S->join[g - 1]);
    }
    if (ev) tc_event_record(ev[2], stream);
    // The 5/3 and 9/7 tile-components form two independent chains (their
    // levels, then their final levels): the 9/7 chain runs on a side stream
    // beside the 5/3 one, so that the small coarse-level launches and the
    // final launches' tails of one filter overlap the other's.
    bool irr_any[2] = {false, false};
    for (const Plan::LevelLaunch& L : P.levels) irr_any[L.irr ? 1 : 0] = true;
    for (const Plan::FinalLaunch& F : P.finals) irr_any[F.irr ? 1 : 0] = true;
    const bool split = S && S->side[0] && S->dwt_fork && S->dwt_join && irr_any[0] && irr_any[1];
    if (split && s == TC_SUCCESS) {
        s = tc_event_record(S->dwt_fork, stream);
        if (s == TC_SUCCESS) s = tc_stream_wait_event(S->side[0], S->dwt_fork);
    }
    for (const Plan::LevelLaunch& L : P.levels) {
        if (s != TC_SUCCESS) break;
        s = tc_launch_idwt_level(split && L.irr ? S->side[0] : stream, L.irr ? 1 : 0, P.shape, up + L.jobs_off,
                                 (const uint32_t*)(up + L.map_off), L.nwg, base);
    }
    if (ev) tc_event_record(ev[3], stream);  // (split: the 5/3 chain's levels)
    for (const Plan::FinalLaunch& F : P.finals) {
        if (s != TC_SUCCESS) break;
        s = tc_launch_idwt_final(split && F.irr ? S->side[0] : stream, F.irr ? 1 : 0, P.shape, F.nc, F.ob,
                                 up + F.jobs_off, (const uint32_t*)(up + F.map_off), F.nwg, base, R.out);
    }
    if (split && s == TC_SUCCESS) {
        s = tc_event_record(S->dwt_join, S->side[0]);
        if (s == TC_SUCCESS) s = tc_stream_wait_event(stream, S->dwt_join);
    }
    if (ev) tc_event_record(ev[4], stream);
    return s;
}

// ---- Pipeline: resident buffers around a plan ----

void Pipeline::invalidate() {
    ready_ = false;
    drop_graph();
    plan_ = Plan();
}

void Pipeline::release_buffers() {
    if (stream_) tc_stream_sync(stream_);
    invalidate();
    for (tc_buf* b : {&upload_, &samples_, &out_, &scratch_}) {
        tc_buf_free(*b);
        *b = nullptr;
    }
}

Status Pipeline::plan_batch(const std::vector<const Image*>& imgs, Plan& P, std::string& err) const {
    return build_plan(imgs, tuning_, device_cu_count(dev_), P, err);
}

Status Pipeline::prepare(const std::vector<const Image*>& imgs, std::string& err) {
    invalidate();
    if (!stream_) {
        err = "no stream";
        return Status::Device;
    }
    tc_stream_sync(stream_);
    Plan P;
    Status st = plan_batch(imgs, P, err);
    if (st == Status::Ok) st = ensure_pinned(P.upload_bytes(), err);
    if (st == Status::Ok) {
        stage_upload(P, (uint8_t*)pinned_);
        st = adopt(std::move(P), pinned_, err);
    }
    if (st != Status::Ok) invalidate();
    return st;
}

Status Pipeline::adopt(Plan&& P, const void* staged, std::string& err) {
    invalidate();
    if (!stream_) {
        err = "no stream";
        return Status::Device;
    }
    plan_ = std::move(P);
    const uint64_t total = plan_.upload_bytes();
    Status st = ensure(upload_, total, "descriptors", err);
    if (st == Status::Ok) st = ensure(samples_, std::max<uint64_t>(plan_.sample_elems, 1) * 4, "sample planes", err);
    if (st == Status::Ok && !host_out_) st = ensure(out_, std::max<uint64_t>(plan_.out_bytes, 1), "output images", err);
    if (st == Status::Ok && host_out_ && host_cap_ < plan_.out_bytes) {
        tc_stream_sync(stream_);
        if (host_buf_) tc_host_free_pinned(host_buf_);
        host_buf_ = host_dev_ = nullptr;
        host_cap_ = 0;
        const size_t want = std::max<size_t>(plan_.out_bytes, 4096);
        if (tc_host_alloc_pinned(want, &host_buf_) != TC_SUCCESS || !(host_dev_ = host_device_ptr(host_buf_))) {
            err = std::string("pinned output: ") + tc_last_error();
            st = Status::OutOfMemory;
        } else {
            host_cap_ = want;
        }
    }
    if (st == Status::Ok && plan_.scratch_bytes) st = ensure(scratch_, plan_.scratch_bytes, "tier-1 state", err);
    if (st == Status::Ok) {
        tc_status ts = tc_buf_copy_h2d(stream_, upload_, 0, staged, total);
        if (ts != TC_SUCCESS) {
            err = tc_last_error();
            st = from_tc(ts);
        }
    }
    if (st != Status::Ok) {
        invalidate();
        return st;
    }
    ready_ = true;
    return Status::Ok;
}

Status Pipeline::run(bool timed) {
    if (!ready_) return Status::InvalidArgument;
    PlanRegions R;
    R.upload = (uint8_t*)tc_buf_device_ptr(upload_);
    R.samples = tc_buf_device_ptr(samples_);
    R.out = host_out_ ? host_dev_ : tc_buf_device_ptr(out_);
    R.scratch = (uint8_t*)tc_buf_device_ptr(scratch_);
    R.scratch_bytes = tc_buf_size(scratch_);
    timed_ = timed;
    tc_status s = launch_plan(plan_, R, device_cu_count(dev_), stream_, &ls_, timed ? ev_ : nullptr);
    if (s == TC_ERR_INVALID_ARG) return Status::InvalidArgument;
    return s == TC_SUCCESS ? Status::Ok : Status::Device;
}
Status Pipeline::sync() { return tc_stream_sync(stream_) == TC_SUCCESS ? Status::Ok : Status::Device; }

StageTimes Pipeline::times() const {
    StageTimes t;
    if (!timed_) return t;
    tc_event_elapsed_ms(ev_[0], ev_[1], &t.fill_ms);
    tc_event_elapsed_ms(ev_[1], ev_[2], &t.t1_ms);
    tc_event_elapsed_ms(ev_[2], ev_[3], &t.dwt_ms);
    tc_event_elapsed_ms(ev_[3], ev_[4], &t.final_ms);
    tc_event_elapsed_ms(ev_[0], ev_[4], &t.total_ms);
    for (size_t g = 0; g < plan_.groups.size() && g < (size_t)kMaxT1Launches; g++)
        tc_event_elapsed_ms(ev_[5 + 2 * g], ev_[6 + 2 * g], &t.t1_launch_ms[g]);
    return t;
}

Status Pipeline::fetch(int i, void* dst, size_t cap, std::string& err) {
    if (!ready_) {
        err = "nothing prepared (the last prepare failed or none ran)";
        return Status::InvalidArgument;
    }
    if (i < 0 || i >= (int)plan_.images.size()) {
        err = "image index out of range";
        return Status::InvalidArgument;
    }
    const Plan::ImageOut& o = plan_.images[i];
    if (cap < o.bytes) {
        err = "output buffer too small";
        return Status::Capacity;
    }
    tc_status s;
    if (host_out_) {
        s = tc_stream_sync(stream_);
        if (s == TC_SUCCESS) std::memcpy(dst, (const uint8_t*)host_buf_ + o.off, o.bytes);
    } else {
        s = tc_buf_copy_d2h(stream_, dst, out_, o.off, o.bytes);
        if (s == TC_SUCCESS) s = tc_stream_sync(stream_);
    }
    if (s != TC_SUCCESS) {
        err = tc_last_error();
        return Status::Device;
    }
    return Status::Ok;
}

const void* Pipeline::output_ptr(int i) const {
    if (!ready_ || i < 0 || i >= (int)plan_.images.size()) return nullptr;
    if (host_out_) return (const uint8_t*)host_buf_ + plan_.images[i].off;
    return (const uint8_t*)tc_buf_device_ptr(out_) + plan_.images[i].off;
}

size_t Pipeline::output_bytes(int i) const {
    return ready_ && i >= 0 && i < (int)plan_.images.size() ? plan_.images[i].bytes : 0;
}

Status Pipeline::capture_graph(std::string& err) {
    drop_graph();
    if (!ready_) {
        err = "nothing prepared (the last prepare failed or none ran)";
        return Status::InvalidArgument;
    }
    if (tc_graph_begin(stream_) != TC_SUCCESS) {
        err = tc_last_error();
        return Status::Device;
    }
    Status s = run(false);
    tc_graph g = nullptr;
    tc_status e = tc_graph_end(stream_, &g);
    if (s != Status::Ok || e != TC_SUCCESS) {
        err = tc_last_error();
        if (g) tc_graph_destroy(g);
        return Status::Device;
    }
    graph_ = g;
    return Status::Ok;
}

Status Pipeline::launch_graph() {
    if (!ready_ || !graph_) return Status::InvalidArgument;
    return tc_graph_launch(graph_, stream_) == TC_SUCCESS ? Status::Ok : Status::Device;
}

void Pipeline::drop_graph() {
    if (graph_) tc_graph_destroy(graph_);
    graph_ = nullptr;
}

size_t Pipeline::device_bytes() const {
    return tc_buf_size(upload_) + tc_buf_size(samples_) + tc_buf_size(out_) + tc_buf_size(scratch_);
}

}  // namespace tc
