// Work-list builders (see plan_util.hpp).
#include "plan_util.hpp"

#include <algorithm>
#include <cstring>

namespace tc {

Status plan_tile_cblks(const Image& im, int t, const uint64_t* plane, uint64_t& cw_size, std::vector<CwCopy>& copies,
                       std::vector<tc_cblk_desc>& descs, std::vector<uint32_t> (&cls)[2], std::string& err) {
    // segment lists of this tile's blocks (segments are in codestream order)
    size_t first = im.cblks.size(), last = 0;
    for (size_t k = 0; k < im.cblks.size(); k++)
        if (im.cblks[k].tile == t) {
            first = std::min(first, k);
            last = k + 1;
        }
    if (first >= last) return Status::Ok;
    // first segment of each block, and the next segment of the same block
    std::vector<uint32_t> head(last - first, UINT32_MAX), next(im.segs.size(), UINT32_MAX);
    std::vector<uint32_t> tail(last - first, UINT32_MAX);
    for (uint32_t s = 0; s < im.segs.size(); s++) {
        const uint32_t c = im.segs[s].cblk;
        if (c < first || c >= last) continue;
        if (head[c - first] == UINT32_MAX)
            head[c - first] = s;
        else
            next[tail[c - first]] = s;
        tail[c - first] = s;
    }
    for (size_t k = first; k < last; k++) {
        const CodeBlock& cb = im.cblks[k];
        int w = cb.x1 - cb.x0, h = cb.y1 - cb.y0;
        if (cb.tile != t || !cb.included || cb.numbps <= 0 || cb.npasses <= 0 || w <= 0 || h <= 0) continue;
        if (w > 64 || h > 64) {
            err = "code-blocks larger than 64x64 are not supported";
            return Status::Unsupported;
        }
        const TileComp& tc = im.tiles[t].comps[cb.comp];
        const Band& B = tc.res[cb.res].band[cb.band];
        int ox = 0, oy = 0;
        if (cb.res > 0) {
            if (B.orient == 1 || B.orient == 3) ox = tc.res[cb.res - 1].width();
            if (B.orient == 2 || B.orient == 3) oy = tc.res[cb.res - 1].height();
        }
        tc_cblk_desc d;
        std::memset(&d, 0, sizeof d);
        cw_size = (cw_size + 3) & ~(uint64_t)3;
        if (cw_size + cb.nbytes > 0xFFFFFFF0ull) {
            err = "codeword arena beyond 4 GiB";
            return Status::Unsupported;
        }
        if (cb.term >= 0) {
            // BYPASS / TERMALL: segment lengths, then their count, before the codeword
            const std::vector<uint32_t>& tl = im.termlens[cb.term];
            copies.push_back(CwCopy{reinterpret_cast<const uint8_t*>(tl.data()), (uint32_t)(tl.size() * 4), cw_size});
            cw_size += tl.size() * 4;
        }
        d.data_off = (uint32_t)cw_size;
        d.data_len = cb.nbytes;
        for (uint32_t s = head[k - first]; s != UINT32_MAX; s = next[s]) {
            const Segment& sg = im.segs[s];
            copies.push_back(CwCopy{im.cs + sg.off, sg.len, cw_size});
            cw_size += sg.len;
        }
        d.out_off = plane[cb.comp] + (uint64_t)(oy + cb.y0 - B.y0) * tc.width() + (uint64_t)(ox + cb.x0 - B.x0);
        d.out_pitch = (uint32_t)tc.width();
        d.w = (uint16_t)w;
        d.h = (uint16_t)h;
        d.numbps = (uint8_t)cb.numbps;
        d.npasses = (uint8_t)std::min(cb.npasses, 255);
        d.cls = (uint8_t)(B.orient == 1 ? 1 : (B.orient == 3 ? 2 : 0));
        d.mode = (uint8_t)((tc.reversible ? 0 : TC_MODE_IRREVERSIBLE) | ((tc.cbsty & 2) ? TC_MODE_RESET : 0) |
                           ((tc.cbsty & 8) ? TC_MODE_CAUSAL : 0) | ((tc.cbsty & 32) ? TC_MODE_SEGSYM : 0) |
                           ((tc.cbsty & 1) ? TC_MODE_BYPASS : 0) | ((tc.cbsty & 4) ? TC_MODE_TERMALL : 0));
        d.delta = B.delta;
        cls[(w <= 32 && h <= 32) ? 0 : 1].push_back((uint32_t)descs.size());
        descs.push_back(d);
    }
    return Status::Ok;
}

Status append_tile_cblks(const Image& im, int t, const uint64_t* plane, std::vector<uint8_t>& cw,
                         std::vector<tc_cblk_desc>& descs, std::vector<uint32_t> (&cls)[2], std::string& err) {
    uint64_t size = cw.size();
    std::vector<CwCopy> copies;
    Status st = plan_tile_cblks(im, t, plane, size, copies, descs, cls, err);
    if (st != Status::Ok) return st;
    cw.resize(size, 0);
    for (const CwCopy& c : copies) std::memcpy(cw.data() + c.dst, c.src, c.len);
    return Status::Ok;
}

void sort_for_packing(std::vector<uint32_t>& v, const std::vector<tc_cblk_desc>& descs, bool small) {
    // A lockstep wave walks its blocks' passes and stripes together and every
    // stripe pass lasts as long as its busiest lane, so blocks of one wave
    // should share the pass count and the number of 4-row stripes; within
    // that, similar codeword lengths.  (On the 256-image mixed batch this cuts
    // the longest wave of the 64x64 class at 16 blocks per wave from 48.4 k to
    // 42.6 k steps, and the class's total steps by 10 %.)  Blocks up to 32x32
    // go by width before length: their total steps fall by 1.8 % at 64 blocks
    // per wave, the longest wave unchanged (for the 64x64 class that order
    // lengthens the longest wave by 10 %).
    std::stable_sort(v.begin(), v.end(), [&](uint32_t a, uint32_t b) {
        const tc_cblk_desc &x = descs[a], &y = descs[b];
        if (x.npasses != y.npasses) return x.npasses > y.npasses;
        const int sx = (x.h + 3) / 4, sy = (y.h + 3) / 4;
        if (sx != sy) return sx > sy;
        if (small && x.w != y.w) return x.w > y.w;
        if (x.data_len != y.data_len) return x.data_len > y.data_len;
        if (x.h != y.h) return x.h > y.h;
        return x.w > y.w;
    });
}

void work_sizes(const TileComp& tc, uint64_t sz[2]) {
    sz[0] = sz[1] = 0;
    for (int n = 2; n <= tc.nlev; n++) {
        const Resolution& R = tc.res[tc.nlev - n + 1];
        uint64_t e = (uint64_t)R.width() * R.height();
        sz[(tc.nlev - n) % 2] = std::max(sz[(tc.nlev - n) % 2], e);
    }
}

tc_dwt_job level_job(const TileComp& tc, int n, uint64_t plane, const uint64_t work[2], bool raster_out,
                     uint64_t raster_off, uint32_t raster_pitch, int TX, int TY) {
    int r = tc.nlev - n + 1;
    const Resolution &R = tc.res[r], &Rl = tc.res[r - 1];
    int lw = Rl.width(), lh = Rl.height();
    tc_dwt_job j;
    std::memset(&j, 0, sizeof j);
    if (n == tc.nlev) {
        j.ll_off = plane;
        j.ll_pitch = (uint32_t)tc.width();
    } else {
        j.ll_off = work[(tc.nlev - n - 1) % 2];
        j.ll_pitch = (uint32_t)lw;
    }
    j.hl_off = plane + lw;
    j.lh_off = plane + (uint64_t)lh * tc.width();
    j.hh_off = plane + (uint64_t)lh * tc.width() + lw;
    j.band_pitch = (uint32_t)tc.width();
    if (raster_out && n == 1) {
        j.out_off = raster_off;
        j.out_pitch = raster_pitch;
    } else {
        j.out_off = work[(tc.nlev - n) % 2];
        j.out_pitch = (uint32_t)R.width();
    }
    j.u0 = R.x0;
    j.u1 = R.x1;
    j.v0 = R.y0;
    j.v1 = R.y1;
    j.tiles_x = (R.width() + TX - 1) / TX;
    return j;
}

}  // namespace tc
