// Tier-2 codestream parsing and Annex B geometry for the decode path.
//
// Replaces the reference's host C parser TC/codestream.c (main header
// :233-258, COD :90-124, QCD :140-230, tile header :260-298, packet header
// :336-453, packet body :455-470), its geometry TC/codestream_image.c:14-277
// and tag trees TC/codestream_tag_tree_encode.c:23-132.  Differences: the parse
// returns a flat table of code-blocks with codeword segments (offsets into the
// codestream, no per-block malloc/memcpy as in TC/Decoder.cpp:71-79), reports
// errors instead of exit(0) (TC/io_buffered_stream.c:42-46), and covers what
// the reference lacks: expounded quantization, COC/QCC, tile-part headers,
// COM/TLM/PLM/PLT/CRG skipping, several layers and tile-parts, the five
// progression orders with one precinct per resolution, empty packets,
// SOP/EPH, edge tiles and per-component precision and signedness.
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <algorithm>
#include <string>
#include <vector>

namespace tc {

enum class Status : int32_t {
    Ok = 0,
    InvalidArgument = 1,
    Truncated = 2,
    BadMarker = 3,
    Unsupported = 4,
    BadGeometry = 5,
    OutOfMemory = 6,
    Device = 7,
    Capacity = 8,
};

struct Band {
    int orient = 0;  // 0 LL, 1 HL, 2 LH, 3 HH
    int x0 = 0, y0 = 0, x1 = 0, y1 = 0;
    int xcb = 6, ycb = 6;      // code-block exponents after the precinct limit
    int ncx = 0, ncy = 0;      // code-blocks across / down
    int eps = 0, mu = 0, mb = 0;
    float delta = 1.0f;        // quantization step (irreversible)
    int first_cblk = 0;        // index into Image::cblks
};

struct Resolution {
    int x0 = 0, y0 = 0, x1 = 0, y1 = 0;
    int nbands = 0;
    Band band[3];
    int width() const { return x1 - x0; }
    int height() const { return y1 - y0; }
};

struct TileComp {
    int x0 = 0, y0 = 0, x1 = 0, y1 = 0;
    int nlev = 0;
    int reversible = 1;  // 1: 5/3, 0: 9/7
    int cbsty = 0;
    std::vector<Resolution> res;  // nlev + 1
    int width() const { return x1 - x0; }
    int height() const { return y1 - y0; }
};

struct Tile {
    int x0 = 0, y0 = 0, x1 = 0, y1 = 0;
    int mct = 0;
    std::vector<TileComp> comps;
};

struct CodeBlock {
    int tile = 0, comp = 0, res = 0, band = 0;  // band index within the resolution
    int x0 = 0, y0 = 0, x1 = 0, y1 = 0;         // subband coordinates
    int npasses = 0;
    int zbp = 0;          // missing most-significant bit-planes
    int numbps = 0;       // bit-planes present: M_b - zbp
    int lblock = 3;
    bool included = false;
    uint32_t nbytes = 0;  // codeword bytes over all segments
    int nseg = 0;
    uint32_t seg0_off = 0;  // first segment offset in the codestream
    int term = -1;          // BYPASS / TERMALL: index into Image::termlens
};

struct Segment {
    uint32_t cblk;
    uint32_t off;  // offset in the codestream
    uint32_t len;
};

struct Image {
    // SIZ
    int x0 = 0, y0 = 0, x1 = 0, y1 = 0;
    int tx0 = 0, ty0 = 0, tw = 0, th = 0, ntx = 0, nty = 0;
    int ncomp = 0;
    std::vector<int> prec, sgnd;
    std::vector<int> dx, dy;  // component sub-sampling XRsiz, YRsiz
    // COD (main header)
    int progression = 0, nlayers = 1, mct = 0;
    std::vector<Tile> tiles;
    std::vector<CodeBlock> cblks;
    std::vector<Segment> segs;  // in codestream order; contiguous per block when nseg == 1
    // BYPASS / TERMALL code-blocks: bytes of each codeword segment (the
    // block's contributions concatenated in codestream order hold them in turn)
    std::vector<std::vector<uint32_t>> termlens;
    const uint8_t* cs = nullptr;  // codestream (inside a JP2 wrapper if any)
    size_t cs_len = 0;
    // JP2 file format (T.800 Annex I): 1 when the codestream came in a JP2
    // file; the colour space of its first usable colour specification box
    // (enumerated EnumCS: 16 sRGB, 17 greyscale, 18 sYCC; -1 an ICC profile;
    // 0 none)
    int jp2 = 0, colourspace = 0;
    // output window on the reference grid: the whole image unless
    // window_image narrowed it
    int wx0 = 0, wy0 = 0, wx1 = 0, wy1 = 0;
    int width() const { return x1 - x0; }
    int height() const { return y1 - y0; }
    int out_width() const { return wx1 - wx0; }
    int out_height() const { return wy1 - wy0; }
    // some component sub-sampled: the output holds each component as its own
    // plane (comp_rect), one after the other
    bool planar() const {
        for (size_t c = 0; c < dx.size(); c++)
            if (dx[c] != 1 || dy[c] != 1) return true;
        return false;
    }
    // full-resolution image area and the levels a reduce discarded
    int fx0 = 0, fy0 = 0, fx1 = 0, fy1 = 0, reduced = 0;
    // component c's area in its own coordinates: ceil(area / XRsiz) (B-12),
    // then ceil(/ 2^reduced)
    void comp_rect(int c, int& x0c, int& y0c, int& x1c, int& y1c) const {
        auto cd = [](int64_t v, int64_t d) { return v >= 0 ? (v + d - 1) / d : -((-v) / d); };
        const int64_t s = (int64_t)1 << reduced;
        x0c = (int)cd(cd(fx0, dx[c]), s);
        y0c = (int)cd(cd(fy0, dy[c]), s);
        x1c = (int)cd(cd(fx1, dx[c]), s);
        y1c = (int)cd(cd(fy1, dy[c]), s);
    }
    // bytes of the decoded output
    uint64_t out_bytes() const {
        const uint64_t bps = max_prec() > 8 ? 2 : 1;
        if (!planar()) return (uint64_t)out_width() * out_height() * ncomp * bps;
        uint64_t n = 0;
        for (int c = 0; c < ncomp; c++) {
            int a, b, e, f;
            comp_rect(c, a, b, e, f);
            n += (uint64_t)(e - a) * (f - b);
        }
        return n * bps;
    }
    int max_prec() const;
    bool all_reversible() const;
};

// Parses a raw codestream or a JP2 file.  The image keeps a pointer into
// `data`, which must outlive it.
// Parses the whole codestream.  threads > 1 reads the packets of up to that
// many tiles at once (one thread per megabyte of codestream at most).
Status parse_codestream(const uint8_t* data, size_t len, Image& img, std::string& err, int threads = 1);

// Main header only (SIZ/COD/QCD), no packets.
Status probe_codestream(const uint8_t* data, size_t len, Image& img, std::string& err);

// Resolution-reduced decode: turns a parsed image into the image k DWT levels
// down (its LL band k levels below full resolution, then the usual colour
// transform and level shift).  Each tile-component loses its k finest
// levels, code-blocks of those levels are dropped, and every rectangle moves
// to that resolution's coordinates, ceil(x / 2^k) (T.800 B.5).  Fails when a
// tile-component has fewer than k levels.
Status reduce_image(Image& img, int k, std::string& err);

// BYPASS (1) / TERMALL (4) codeword segments (T.800 D.6, Table D.9): pass k
// (0 = the first cleanup) ends a segment.
inline bool seg_end(int style, int k) {
    if (style & 4) return true;
    if (style & 1) return k == 9 || (k > 9 && (k - 1) % 3 != 0);
    return false;
}
inline int seg_index(int style, int k) {
    int n = 0;
    for (int j = 0; j < k; j++) n += seg_end(style, j) ? 1 : 0;
    return n;
}

// Region decode: narrows the output window to the w x h rectangle at (x, y)
// from the image origin (clipped to the image) and drops every code-block
// that cannot influence it.  Per decomposition level the needed band range is
// the next finer range mapped down with a margin of 4 samples on each side
// (the longest lifting support, 9/7), so every sample inside the window is
// reconstructed exactly as in the full decode.  Apply after reduce_image.
Status window_image(Image& img, int x, int y, int w, int h, std::string& err);

const char* status_name(Status s);

}  // namespace tc
