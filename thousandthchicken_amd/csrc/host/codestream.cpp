// Tier-2 codestream parsing and Annex B geometry (see codestream.hpp).
#include "codestream.hpp"

#include <algorithm>
#include <thread>
#include <array>
#include <tuple>
#include <cmath>
#include <cstring>

namespace tc {

namespace {

constexpr int kMaxLevels = 32;

struct CodingStyle {
    int nlev = 0, xcb = 6, ycb = 6, cbsty = 0, xform = 1;
    uint8_t ppx[kMaxLevels + 1], ppy[kMaxLevels + 1];
    CodingStyle() {
        std::fill(ppx, ppx + kMaxLevels + 1, 15);
        std::fill(ppy, ppy + kMaxLevels + 1, 15);
    }
};

struct Quant {
    int style = 0, guard = 2;
    std::vector<uint16_t> v;  // style 0: exponent only (eps); 1/2: eps<<11|mu
};

// One progression order change (POC, A.6.6): resolutions rs..re-1,
// components cs..ce-1, layers below le, in order prog.
struct Poc {
    int rs, cs, le, re, ce, prog;
};

// A PPM or PPT marker segment: index byte and payload range in the codestream.
struct PackedSeg {
    int z;
    size_t off, len;
};

struct Params {
    int prog = 0, nlayers = 1, mct = 0;
    bool sop = false, eph = false;
    std::vector<CodingStyle> cs;
    std::vector<Quant> q;
    std::vector<Poc> poc;
    bool poc_tile = false;  // poc came from the tile's own headers
    std::vector<PackedSeg> ppm;  // main header only (A.7.4)
};

struct Range {
    size_t off, len;
};

inline int64_t floor_div(int64_t a, int64_t b) {
    int64_t q = a / b;
    return (a % b != 0 && ((a < 0) != (b < 0))) ? q - 1 : q;
}
inline int64_t ceil_div(int64_t a, int64_t b) { return -floor_div(-a, b); }

class Reader {
  public:
    Reader(const uint8_t* p, size_t n, size_t pos = 0) : p_(p), n_(n), pos_(pos) {}
    bool ok() const { return ok_; }
    size_t pos() const { return pos_; }
    void seek(size_t p) { pos_ = p; }
    size_t size() const { return n_; }
    uint32_t u8() {
        if (pos_ >= n_) {
            ok_ = false;
            return 0;
        }
        return p_[pos_++];
    }
    uint32_t u16() {
        uint32_t a = u8();
        return (a << 8) | u8();
    }
    uint32_t u32() {
        uint32_t a = u16();
        return (a << 16) | u16();
    }

  private:
    const uint8_t* p_;
    size_t n_, pos_;
    bool ok_ = true;
};

// Packet-header bit reader with the stuffing rule of T.800 B.10.1: after an
// 0xFF byte only 7 bits of the next byte carry data.
class BitReader {
  public:
    BitReader(const uint8_t* p, size_t end, size_t pos) : p_(p), end_(end), pos_(pos) {}
    int bit() {
        if (avail_ == 0) {
            if (pos_ >= end_) {
                bad_ = true;
                return 0;
            }
            avail_ = (last_ == 0xFF) ? 7 : 8;
            last_ = p_[pos_++];
        }
        --avail_;
        return (last_ >> avail_) & 1;
    }
    uint32_t bits(int n) {
        uint32_t v = 0;
        while (n-- > 0) v = (v << 1) | (uint32_t)bit();
        return v;
    }
    // End of header: a final 0xFF owns the following stuffed byte.
    void align() {
        if (last_ == 0xFF) {
            if (pos_ < end_)
                ++pos_;
            else
                bad_ = true;
        }
        avail_ = 0;
        last_ = 0;
    }
    size_t pos() const { return pos_; }
    bool bad() const { return bad_; }

  private:
    const uint8_t* p_;
    size_t end_, pos_;
    uint32_t last_ = 0;
    int avail_ = 0;
    bool bad_ = false;
};

// Tag tree over a grid of leaves (T.800 B.10.2).  Each node keeps its lower
// bound and whether its value is known; parents are found by halving.
class TagTree {
  public:
    void init(int w, int h) {
        levels_.clear();
        int lw = w, lh = h;
        size_t total = 0;
        for (;;) {
            levels_.push_back({lw, lh, total});
            total += (size_t)lw * lh;
            if (lw == 1 && lh == 1) break;
            lw = (lw + 1) / 2;
            lh = (lh + 1) / 2;
        }
        low_.assign(total, 0);
        known_.assign(total, 0);
    }
    // true when the value of leaf (x,y) is below thr
    bool decode(BitReader& br, int x, int y, int thr) {
        int path[40];
        int nl = (int)levels_.size();
        for (int l = 0; l < nl; l++) {
            path[l] = (int)(levels_[l].base + (size_t)y * levels_[l].w + x);
            x >>= 1;
            y >>= 1;
        }
        int low = 0;
        for (int l = nl - 1; l >= 0; l--) {
            int q = path[l];
            if (low_[q] < low) low_[q] = low;
            while (!known_[q] && low_[q] < thr) {
                if (br.bit())
                    known_[q] = 1;
                else
                    ++low_[q];
                if (br.bad()) return false;
            }
            low = low_[q];
            if (!known_[q]) return false;
        }
        return low < thr;
    }
    // full value of leaf (x,y)
    int value(BitReader& br, int x, int y) {
        int t = 1;
        while (!decode(br, x, y, t)) {
            if (br.bad() || t > 64) return -1;
            ++t;
        }
        return t - 1;
    }

  private:
    struct Lvl {
        int w, h;
        size_t base;
    };
    std::vector<Lvl> levels_;
    std::vector<int> low_;
    std::vector<uint8_t> known_;
};

// The code-blocks of one band inside one precinct: columns i0..i0+w-1 and rows
// j0..j0+h-1 of the band's code-block grid, with their tag trees (B.10.2).
struct PrecBand {
    int i0 = 0, j0 = 0, w = 0, h = 0;
    TagTree incl, zbp;
};

// Precinct partition of one resolution, anchored at 0 (B.6), and each band's
// share of every precinct.
struct ResState {
    int ppx = 15, ppy = 15;
    int64_t px0 = 0, py0 = 0;  // precinct index of the resolution's origin
    int pw = 0, ph = 0;        // precincts across and down (0: empty resolution)
    std::vector<PrecBand> prc[3];
    std::vector<int> nextl;    // per precinct: the layer of its next packet
};
using TileState = std::vector<std::vector<ResState>>;  // [component][resolution]

Status read_cs_params(Reader& r, CodingStyle& c, bool user_prec) {
    c.nlev = (int)r.u8();
    c.xcb = (int)r.u8() + 2;
    c.ycb = (int)r.u8() + 2;
    c.cbsty = (int)r.u8();
    c.xform = (int)r.u8();
    if (!r.ok()) return Status::Truncated;
    if (c.nlev > kMaxLevels || c.xcb > 10 || c.ycb > 10 || c.xcb + c.ycb > 12 || c.xform > 1)
        return Status::Unsupported;
    for (int i = 0; i <= c.nlev; i++) {
        if (user_prec) {
            uint32_t b = r.u8();
            c.ppx[i] = (uint8_t)(b & 15);
            c.ppy[i] = (uint8_t)(b >> 4);
        } else {
            c.ppx[i] = c.ppy[i] = 15;
        }
    }
    return r.ok() ? Status::Ok : Status::Truncated;
}

Status read_quant(Reader& r, size_t end, Quant& q) {
    uint32_t s = r.u8();
    q.style = (int)(s & 31);
    q.guard = (int)(s >> 5);
    q.v.clear();
    if (q.style == 0) {
        while (r.pos() < end) q.v.push_back((uint16_t)((r.u8() >> 3) << 11));
    } else if (q.style == 1 || q.style == 2) {
        while (r.pos() + 2 <= end) q.v.push_back((uint16_t)r.u16());
    } else {
        return Status::Unsupported;
    }
    return r.ok() ? Status::Ok : Status::Truncated;
}

// One parameter marker segment (COD/COC/QCD/QCC) or a skippable one.
Status read_segment(Reader& r, uint32_t marker, int ncomp, Params& P, std::vector<uint8_t>& coc_set,
                    std::vector<uint8_t>& qcc_set, std::string& err, bool tile_hdr = false) {
    size_t lpos = r.pos();
    uint32_t L = r.u16();
    if (!r.ok() || L < 2 || lpos + L > r.size()) return Status::Truncated;
    size_t end = lpos + L;
    Status st = Status::Ok;
    switch (marker) {
    case 0xFF52: {  // COD
        uint32_t scod = r.u8();
        P.prog = (int)r.u8();
        P.nlayers = (int)r.u16();
        P.mct = (int)r.u8();
        P.sop = (scod & 2) != 0;
        P.eph = (scod & 4) != 0;
        CodingStyle c;
        st = read_cs_params(r, c, (scod & 1) != 0);
        if (st != Status::Ok) break;
        for (int i = 0; i < ncomp; i++)
            if (!coc_set[i]) P.cs[i] = c;
        if (P.prog > 4 || P.nlayers < 1) st = Status::Unsupported;
        break;
    }
    case 0xFF53: {  // COC
        uint32_t ci = ncomp < 257 ? r.u8() : r.u16();
        uint32_t scoc = r.u8();
        if ((int)ci >= ncomp) return Status::BadMarker;
        st = read_cs_params(r, P.cs[ci], (scoc & 1) != 0);
        coc_set[ci] = 1;
        break;
    }
    case 0xFF5C: {  // QCD
        Quant q;
        st = read_quant(r, end, q);
        if (st != Status::Ok) break;
        for (int i = 0; i < ncomp; i++)
            if (!qcc_set[i]) P.q[i] = q;
        break;
    }
    case 0xFF5D: {  // QCC
        uint32_t ci = ncomp < 257 ? r.u8() : r.u16();
        if ((int)ci >= ncomp) return Status::BadMarker;
        st = read_quant(r, end, P.q[ci]);
        qcc_set[ci] = 1;
        break;
    }
    case 0xFF5F: {  // POC: a tile's own entries replace the main header's
        const bool wide = ncomp >= 257;
        const uint32_t esz = wide ? 9 : 7;
        if (L - 2 < esz || (L - 2) % esz) return Status::BadMarker;
        if (tile_hdr && !P.poc_tile) {
            P.poc.clear();
            P.poc_tile = true;
        }
        for (uint32_t i = 0; i < (L - 2) / esz; i++) {
            Poc e;
            e.rs = (int)r.u8();
            e.cs = (int)(wide ? r.u16() : r.u8());
            e.le = (int)r.u16();
            e.re = (int)r.u8();
            e.ce = (int)(wide ? r.u16() : r.u8());
            e.prog = (int)r.u8();
            if (e.ce == 0) e.ce = wide ? 16384 : 256;
            if (e.prog > 4) {
                err = "bad progression order in POC";
                return Status::Unsupported;
            }
            if (P.poc.size() >= 256) {
                err = "more than 256 progression order changes";
                return Status::Unsupported;
            }
            P.poc.push_back(e);
        }
        break;
    }
    case 0xFF60:  // PPM: packed packet headers of every tile-part (A.7.4)
        if (tile_hdr || L < 3) {
            err = "PPM outside the main header";
            return Status::BadMarker;
        }
        {
            int z = (int)r.u8();
            P.ppm.push_back(PackedSeg{z, r.pos(), end - r.pos()});
        }
        break;
    case 0xFF61:  // PPT: the tile-part loop takes these
        err = "PPT outside a tile-part header";
        return Status::BadMarker;
    case 0xFF64:  // COM
    case 0xFF55:  // TLM
    case 0xFF57:  // PLM
    case 0xFF58:  // PLT
    case 0xFF63:  // CRG
        break;
    default: {
        char b[64];
        snprintf(b, sizeof b, "unsupported marker 0x%04X", marker);
        err = b;
        return Status::Unsupported;  // RGN, PPM, PPT and unknown markers
    }
    }
    if (st != Status::Ok) return st;
    r.seek(end);
    return Status::Ok;
}

// ---- JP2 file format (T.800 Annex I) ----
//
// The reference reads the signature, File Type, JP2 Header (Image Header
// first) and Contiguous Codestream boxes in TC/boxes.c:142-307: it rejects a
// brand other than 'jp2 ', takes the image size and component count from
// 'ihdr', and decodes the 'jp2c' payload.  Here the same boxes are read and
// checked against the codestream: 'ihdr' (and 'bpcc') must agree with SIZ,
// the compression type must be 7, and the colour specification is reported.
// Palettes ('pclr') and component mappings ('cmap') would map the decoded
// components to other channels, which this decoder does not do: such files
// are rejected rather than returned unmapped.

struct Jp2Header {
    bool have_ihdr = false;
    uint32_t height = 0, width = 0;
    int nc = 0, bpc = 0, compression = 0;
    std::vector<int> bpcc;  // bits per component box (BPC 255)
    int colourspace = 0;    // EnumCS of the first usable 'colr', -1 for an ICC profile, 0 none
};

uint32_t be32(const uint8_t* p) { return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3]; }

// One box at pos of [0, n): its type, payload [*off, *off + *len) and the
// position after it.  A length of 0 runs to the end of the file (last box).
bool next_box(const uint8_t* d, size_t n, size_t pos, uint32_t* type, size_t* off, size_t* len, size_t* end) {
    if (pos + 8 > n) return false;
    uint64_t L = be32(d + pos);
    size_t hdr = 8;
    if (L == 1) {
        if (pos + 16 > n) return false;
        L = ((uint64_t)be32(d + pos + 8) << 32) | be32(d + pos + 12);
        hdr = 16;
    } else if (L == 0) {
        L = n - pos;
    }
    if (L < hdr) return false;
    *type = be32(d + pos + 4);
    *off = pos + hdr;
    *len = (size_t)std::min<uint64_t>(L, n - pos) - hdr;  // a truncated last box keeps what is there
    *end = pos + (size_t)std::min<uint64_t>(L, n - pos);
    return true;
}

constexpr uint32_t box_id(const char (&t)[5]) {
    return ((uint32_t)(uint8_t)t[0] << 24) | ((uint32_t)(uint8_t)t[1] << 16) | ((uint32_t)(uint8_t)t[2] << 8) |
           (uint32_t)(uint8_t)t[3];
}

Status read_jp2_header(const uint8_t* d, size_t n, Jp2Header& H, std::string& err) {
    size_t pos = 0;
    bool first = true, have_colr = false;
    while (pos < n) {
        uint32_t t;
        size_t off, len, end;
        if (!next_box(d, n, pos, &t, &off, &len, &end)) {
            err = "JP2 header box: malformed sub-box";
            return Status::BadMarker;
        }
        const uint8_t* p = d + off;
        if (first && t != box_id("ihdr")) {
            err = "JP2 header box: the image header box must come first";
            return Status::BadMarker;
        }
        first = false;
        if (t == box_id("ihdr")) {
            if (len < 14) {
                err = "JP2 image header box too short";
                return Status::Truncated;
            }
            H.have_ihdr = true;
            H.height = be32(p);
            H.width = be32(p + 4);
            H.nc = (p[8] << 8) | p[9];
            H.bpc = p[10];
            H.compression = p[11];
        } else if (t == box_id("bpcc")) {
            H.bpcc.assign(p, p + len);
        } else if (t == box_id("colr")) {
            // the first colour specification of a method this reader knows is used (I.5.3.3)
            if (!have_colr && len >= 3 && (p[0] == 1 || p[0] == 2)) {
                if (p[0] == 1) {
                    if (len < 7) {
                        err = "JP2 colour specification box too short";
                        return Status::Truncated;
                    }
                    H.colourspace = (int)be32(p + 3);
                } else {
                    H.colourspace = -1;
                }
                have_colr = true;
            }
        } else if (t == box_id("pclr") || t == box_id("cmap")) {
            err = std::string("JP2 ") + (t == box_id("pclr") ? "palette (pclr)" : "component mapping (cmap)") +
                  " boxes are not supported: the decoded components would need mapping to other channels";
            return Status::Unsupported;
        }
        pos = end;
    }
    if (!H.have_ihdr) {
        err = "JP2 header box without an image header box";
        return Status::BadMarker;
    }
    if (H.compression != 7) {
        err = "JP2 image header: compression type " + std::to_string(H.compression) + " (only 7, T.800, is defined)";
        return Status::Unsupported;
    }
    return Status::Ok;
}

// Finds the codestream: inside a JP2 file (signature box, File Type box, JP2
// Header box, then 'jp2c'), or the input itself when it does not start with
// the JP2 signature.
Status locate_codestream(const uint8_t* d, size_t n, const uint8_t** cs, size_t* cl, bool* is_jp2, Jp2Header& H,
                         std::string& err) {
    static const uint8_t sig[12] = {0, 0, 0, 12, 'j', 'P', ' ', ' ', 0x0D, 0x0A, 0x87, 0x0A};
    *is_jp2 = n >= 12 && std::memcmp(d, sig, 12) == 0;
    if (!*is_jp2) {
        *cs = d;
        *cl = n;
        return Status::Ok;
    }
    size_t pos = 12;
    uint32_t t;
    size_t off, len, end;
    // File Type box right after the signature (I.5.2): brand or compatibility list 'jp2 '
    if (!next_box(d, n, pos, &t, &off, &len, &end) || t != box_id("ftyp") || len < 8) {
        err = "JP2 file: the File Type box must follow the signature box";
        return Status::BadMarker;
    }
    bool jp2 = be32(d + off) == box_id("jp2 ");
    for (size_t k = 8; k + 4 <= len; k += 4) jp2 = jp2 || be32(d + off + k) == box_id("jp2 ");
    if (!jp2) {
        err = "JP2 file: neither the brand nor the compatibility list names 'jp2 '";
        return Status::Unsupported;
    }
    pos = end;
    bool have_header = false;
    while (next_box(d, n, pos, &t, &off, &len, &end)) {
        if (t == box_id("jp2h")) {
            if (have_header) {
                err = "JP2 file: more than one JP2 header box";
                return Status::BadMarker;
            }
            Status s = read_jp2_header(d + off, len, H, err);
            if (s != Status::Ok) return s;
            have_header = true;
        } else if (t == box_id("jp2c")) {
            if (!have_header) {
                err = "JP2 file: the JP2 header box must precede the codestream box";
                return Status::BadMarker;
            }
            *cs = d + off;
            *cl = len;
            return Status::Ok;
        }
        pos = end;
    }
    err = "JP2 file without a contiguous codestream box";
    return Status::BadMarker;
}

// The JP2 image header must describe the codestream's image (I.5.3.1, I.5.3.2).
Status check_jp2_header(const Jp2Header& H, const Image& img, std::string& err) {
    auto bits_ok = [&](int bpc, int c) { return (bpc & 0x7F) + 1 == img.prec[c] && (bpc >> 7) == img.sgnd[c]; };
    std::string what;
    if ((int64_t)H.width != img.width() || (int64_t)H.height != img.height())
        what = "size " + std::to_string(H.width) + "x" + std::to_string(H.height) + " against " +
               std::to_string(img.width()) + "x" + std::to_string(img.height());
    else if (H.nc != img.ncomp)
        what = std::to_string(H.nc) + " components against " + std::to_string(img.ncomp);
    else if (H.bpc == 255 && (int)H.bpcc.size() != img.ncomp)
        what = "a bits-per-component box of " + std::to_string(H.bpcc.size()) + " entries";
    else
        for (int c = 0; c < img.ncomp && what.empty(); c++)
            if (!bits_ok(H.bpc == 255 ? H.bpcc[c] : H.bpc, c)) what = "bit depth of component " + std::to_string(c);
    if (what.empty()) return Status::Ok;
    err = "JP2 image header disagrees with the codestream (SIZ): " + what;
    return Status::BadGeometry;
}

Status parse_main_header(Reader& r, Image& img, Params& P, std::string& err) {
    if (r.u16() != 0xFF4F) {
        err = "missing SOC";
        return Status::BadMarker;
    }
    if (r.u16() != 0xFF51) {
        err = "missing SIZ";
        return Status::BadMarker;
    }
    size_t lpos = r.pos();
    uint32_t L = r.u16();
    r.u16();  // Rsiz
    img.x1 = (int)r.u32();
    img.y1 = (int)r.u32();
    img.x0 = (int)r.u32();
    img.y0 = (int)r.u32();
    img.tw = (int)r.u32();
    img.th = (int)r.u32();
    img.tx0 = (int)r.u32();
    img.ty0 = (int)r.u32();
    img.ncomp = (int)r.u16();
    if (!r.ok()) return Status::Truncated;
    if (img.ncomp < 1 || img.ncomp > 16384) return Status::Unsupported;
    img.prec.resize(img.ncomp);
    img.sgnd.resize(img.ncomp);
    img.dx.resize(img.ncomp);
    img.dy.resize(img.ncomp);
    for (int i = 0; i < img.ncomp; i++) {
        uint32_t s = r.u8();
        img.prec[i] = (int)(s & 0x7F) + 1;
        img.sgnd[i] = (int)(s >> 7);
        img.dx[i] = (int)r.u8();
        img.dy[i] = (int)r.u8();
        if (r.ok() && (img.dx[i] < 1 || img.dy[i] < 1)) {
            err = "component sub-sampling of 0";
            return Status::BadMarker;
        }
        if (img.prec[i] > 16) {
            err = "bit depth above 16";
            return Status::Unsupported;
        }
    }
    if (!r.ok()) return Status::Truncated;
    r.seek(lpos + L);
    if (img.x1 <= img.x0 || img.y1 <= img.y0 || img.tw <= 0 || img.th <= 0 || img.tx0 > img.x0 ||
        img.ty0 > img.y0 || (int64_t)img.tx0 + img.tw <= img.x0 || (int64_t)img.ty0 + img.th <= img.y0) {
        err = "bad image/tile geometry";
        return Status::BadGeometry;
    }
    img.ntx = (int)ceil_div(img.x1 - img.tx0, img.tw);
    img.nty = (int)ceil_div(img.y1 - img.ty0, img.th);
    if ((int64_t)img.ntx * img.nty > 65535) {  // Isot is 16 bits (A.4.2)
        err = "more than 65535 tiles";
        return Status::BadGeometry;
    }
    if ((int64_t)img.width() * img.height() > ((int64_t)1 << 34)) {
        err = "image area beyond 2^34 samples";
        return Status::Capacity;
    }
    P.cs.assign(img.ncomp, CodingStyle());
    P.q.assign(img.ncomp, Quant());
    std::vector<uint8_t> coc(img.ncomp, 0), qcc(img.ncomp, 0);
    bool cod = false, qcd = false;
    for (;;) {
        uint32_t m = r.u16();
        if (!r.ok()) return Status::Truncated;
        if (m == 0xFF90) {
            r.seek(r.pos() - 2);
            break;
        }
        cod |= m == 0xFF52;
        qcd |= m == 0xFF5C;
        Status s = read_segment(r, m, img.ncomp, P, coc, qcc, err);
        if (s != Status::Ok) return s;
    }
    if (!cod || !qcd) {
        err = "main header lacks COD or QCD";
        return Status::BadMarker;
    }
    img.progression = P.prog;
    img.nlayers = P.nlayers;
    img.mct = P.mct;
    return Status::Ok;
}

// Geometry of one tile and its code-blocks (Annex B.3-B.7).
Status build_tile(Image& img, int t, const Params& P, TileState& bstate, std::string& err) {
    Tile& T = img.tiles[t];
    int p = t % img.ntx, q = t / img.ntx;
    T.x0 = std::max(img.tx0 + p * img.tw, img.x0);
    T.y0 = std::max(img.ty0 + q * img.th, img.y0);
    T.x1 = std::min((int64_t)img.tx0 + (int64_t)(p + 1) * img.tw, (int64_t)img.x1);
    T.y1 = std::min((int64_t)img.ty0 + (int64_t)(q + 1) * img.th, (int64_t)img.y1);
    T.mct = P.mct;
    T.comps.resize(img.ncomp);
    bstate.resize(img.ncomp);
    for (int c = 0; c < img.ncomp; c++) {
        const CodingStyle& C = P.cs[c];
        const Quant& Q = P.q[c];
        TileComp& tc = T.comps[c];
        // tile-component rectangle, T.800 B-12
        tc.x0 = (int)ceil_div(T.x0, img.dx[c]);
        tc.y0 = (int)ceil_div(T.y0, img.dy[c]);
        tc.x1 = (int)ceil_div(T.x1, img.dx[c]);
        tc.y1 = (int)ceil_div(T.y1, img.dy[c]);
        tc.nlev = C.nlev;
        tc.reversible = C.xform;
        tc.cbsty = C.cbsty;
        // decodable styles: RESET (2), CAUSAL (8), PTERM (16, nothing to do for a
        // decoder) and SEGSYM (32); BYPASS (1) and TERMALL (4) split a codeword
        // into per-pass segments, which the tier-1 kernels do not take
        if (C.cbsty & ~63) {
            err = "reserved code-block style bits";
            return Status::Unsupported;
        }
        int nq = Q.style == 1 ? 1 : 3 * C.nlev + 1;
        if ((int)Q.v.size() < nq) {
            err = "quantization marker too short";
            return Status::BadMarker;
        }
        tc.res.assign(C.nlev + 1, Resolution());
        bstate[c].assign(C.nlev + 1, ResState());
        for (int r = 0; r <= C.nlev; r++) {
            Resolution& R = tc.res[r];
            int64_t s = (int64_t)1 << (C.nlev - r);
            R.x0 = (int)ceil_div(tc.x0, s);
            R.y0 = (int)ceil_div(tc.y0, s);
            R.x1 = (int)ceil_div(tc.x1, s);
            R.y1 = (int)ceil_div(tc.y1, s);
            ResState& RS = bstate[c][r];
            RS.ppx = C.ppx[r];
            RS.ppy = C.ppy[r];
            if (r > 0 && (RS.ppx == 0 || RS.ppy == 0)) {
                err = "precinct exponent 0 above the lowest resolution";
                return Status::BadMarker;
            }
            if (R.x1 > R.x0 && R.y1 > R.y0) {
                RS.px0 = floor_div(R.x0, (int64_t)1 << RS.ppx);
                RS.py0 = floor_div(R.y0, (int64_t)1 << RS.ppy);
                int64_t npx = ceil_div(R.x1, (int64_t)1 << RS.ppx) - RS.px0;
                int64_t npy = ceil_div(R.y1, (int64_t)1 << RS.ppy) - RS.py0;
                if (npx * npy > ((int64_t)1 << 22)) {
                    err = "more than 2^22 precincts in a resolution";
                    return Status::Capacity;
                }
                RS.pw = (int)npx;
                RS.ph = (int)npy;
                RS.nextl.assign((size_t)RS.pw * RS.ph, 0);
            }
            R.nbands = r == 0 ? 1 : 3;
            int nb = r == 0 ? C.nlev : C.nlev - r + 1;
            for (int k = 0; k < R.nbands; k++) {
                Band& B = R.band[k];
                B.orient = r == 0 ? 0 : k + 1;
                int xo = (B.orient == 1 || B.orient == 3), yo = (B.orient == 2 || B.orient == 3);
                if (nb == 0) {
                    B.x0 = tc.x0;
                    B.y0 = tc.y0;
                    B.x1 = tc.x1;
                    B.y1 = tc.y1;
                } else {
                    int64_t bs = (int64_t)1 << nb, hs = (int64_t)1 << (nb - 1);
                    B.x0 = (int)ceil_div(tc.x0 - hs * xo, bs);
                    B.y0 = (int)ceil_div(tc.y0 - hs * yo, bs);
                    B.x1 = (int)ceil_div(tc.x1 - hs * xo, bs);
                    B.y1 = (int)ceil_div(tc.y1 - hs * yo, bs);
                }
                B.xcb = std::min<int>(C.xcb, r == 0 ? C.ppx[r] : C.ppx[r] - 1);
                B.ycb = std::min<int>(C.ycb, r == 0 ? C.ppy[r] : C.ppy[r] - 1);
                // quantization (E.1): derived style scales the LL exponent per level
                int qi = r == 0 ? 0 : 3 * (r - 1) + k + 1;
                uint16_t qv = Q.style == 1 ? Q.v[0] : Q.v[qi];
                B.eps = qv >> 11;
                B.mu = qv & 0x7FF;
                if (Q.style == 1 && r > 0) B.eps = B.eps - C.nlev + nb;
                B.mb = Q.guard + B.eps - 1;
                int gain = B.orient == 0 ? 0 : (B.orient == 3 ? 2 : 1);
                B.delta = (float)std::ldexp(1.0 + B.mu / 2048.0, img.prec[c] + gain - B.eps);
                B.first_cblk = (int)img.cblks.size();
                if (B.x1 > B.x0 && B.y1 > B.y0) {
                    int64_t cw = (int64_t)1 << B.xcb, ch = (int64_t)1 << B.ycb;
                    int gx0 = (int)floor_div(B.x0, cw), gy0 = (int)floor_div(B.y0, ch);
                    B.ncx = (int)(ceil_div(B.x1, cw) - gx0);
                    B.ncy = (int)(ceil_div(B.y1, ch) - gy0);
                    if ((int64_t)img.cblks.size() + (int64_t)B.ncx * B.ncy > ((int64_t)1 << 24)) {
                        err = "more than 2^24 code-blocks";
                        return Status::Capacity;
                    }
                    for (int j = 0; j < B.ncy; j++)
                        for (int i = 0; i < B.ncx; i++) {
                            CodeBlock cb;
                            cb.tile = t;
                            cb.comp = c;
                            cb.res = r;
                            cb.band = k;
                            cb.x0 = std::max<int64_t>(B.x0, (gx0 + i) * cw);
                            cb.y0 = std::max<int64_t>(B.y0, (gy0 + j) * ch);
                            cb.x1 = std::min<int64_t>(B.x1, (gx0 + i + 1) * cw);
                            cb.y1 = std::min<int64_t>(B.y1, (gy0 + j + 1) * ch);
                            img.cblks.push_back(cb);
                        }
                    // each precinct's code-blocks: the band's precinct partition is
                    // the resolution's halved above r = 0, and no code-block
                    // straddles it (B.7: xcb' <= the band's precinct exponent)
                    const int sx = (r == 0 ? RS.ppx : RS.ppx - 1) - B.xcb;
                    const int sy = (r == 0 ? RS.ppy : RS.ppy - 1) - B.ycb;
                    RS.prc[k].resize((size_t)RS.pw * RS.ph);
                    for (int q = 0; q < RS.pw * RS.ph; q++) {
                        int64_t X = RS.px0 + q % RS.pw, Y = RS.py0 + q / RS.pw;
                        int64_t i0 = std::max<int64_t>(0, (X << sx) - gx0);
                        int64_t i1 = std::min<int64_t>(B.ncx, ((X + 1) << sx) - gx0);
                        int64_t j0 = std::max<int64_t>(0, (Y << sy) - gy0);
                        int64_t j1 = std::min<int64_t>(B.ncy, ((Y + 1) << sy) - gy0);
                        if (i1 <= i0 || j1 <= j0) continue;
                        PrecBand& PB = RS.prc[k][q];
                        PB.i0 = (int)i0;
                        PB.j0 = (int)j0;
                        PB.w = (int)(i1 - i0);
                        PB.h = (int)(j1 - j0);
                        PB.incl.init(PB.w, PB.h);
                        PB.zbp.init(PB.w, PB.h);
                    }
                } else {
                    B.ncx = B.ncy = 0;
                }
            }
        }
    }
    return Status::Ok;
}

int pass_count(BitReader& br) {  // B.10.6
    if (!br.bit()) return 1;
    if (!br.bit()) return 2;
    uint32_t v = br.bits(2);
    if (v != 3) return 3 + (int)v;
    v = br.bits(5);
    if (v != 31) return 6 + (int)v;
    return 37 + (int)br.bits(7);
}

int ilog2(int v) {
    int l = 0;
    while (v > 1) {
        v >>= 1;
        ++l;
    }
    return l;
}

// Reads the packet of (component c, resolution r, precinct q, layer l): its
// body at pos within the tile data ranges, its header there too or, for
// packed headers (PPM/PPT), in *ph at hpos.
// Where the packets of one tile record what they carry: the codeword segments
// (in codestream order) and the BYPASS / TERMALL segment lengths, whose index
// a code-block's `term` holds.  Tiles parse into their own and are merged in
// tile order.
struct PacketOut {
    std::vector<Segment> segs;
    std::vector<std::vector<uint32_t>> termlens;
};

Status read_packet(Image& img, int t, int c, int r, int q, int layer, const Params& P,
                   TileState& bstate, const std::vector<Range>& ranges, size_t& ri,
                   size_t& pos, const std::vector<uint8_t>* ph, size_t& hpos, PacketOut& out, std::string& err) {
    const TileComp& tc = img.tiles[t].comps[c];
    const Resolution& R = tc.res[r];
    ResState& RS = bstate[c][r];
    if (q >= RS.pw * RS.ph) return Status::Ok;  // empty resolution: no precinct
    while (ri < ranges.size() && pos >= ranges[ri].off + ranges[ri].len) {
        ++ri;
        if (ri < ranges.size()) pos = ranges[ri].off;
    }
    if (ph ? hpos >= ph->size() : ri >= ranges.size()) return Status::Ok;  // stream ends at a packet boundary
    const uint8_t* d = img.cs;
    size_t end = ri < ranges.size() ? ranges[ri].off + ranges[ri].len : pos;
    if (P.sop && pos + 6 <= end && d[pos] == 0xFF && d[pos + 1] == 0x91) pos += 6;
    BitReader br = ph ? BitReader(ph->data(), ph->size(), hpos) : BitReader(d, end, pos);
    struct Inc {
        uint32_t cb;
        uint32_t len;
        int seg;  // codeword segment (BYPASS / TERMALL), else 0
    };
    const int style = tc.cbsty;
    const bool multi = (style & (1 | 4)) != 0;
    Inc incs[512];
    std::vector<Inc> more;
    int ninc = 0;
    if (br.bit()) {
        for (int k = 0; k < R.nbands; k++) {
            const Band& B = R.band[k];
            if (RS.prc[k].empty()) continue;
            PrecBand& S = RS.prc[k][q];
            for (int j = 0; j < S.h; j++)
                for (int i = 0; i < S.w; i++) {
                    uint32_t id = (uint32_t)(B.first_cblk + (S.j0 + j) * B.ncx + S.i0 + i);
                    CodeBlock& cb = img.cblks[id];
                    bool inc = cb.included ? br.bit() != 0 : S.incl.decode(br, i, j, layer + 1);
                    if (br.bad()) break;
                    if (!inc) continue;
                    if (!cb.included) {
                        int z = S.zbp.value(br, i, j);
                        if (z < 0) {
                            err = "bad zero bit-plane tag tree";
                            return Status::Truncated;
                        }
                        cb.zbp = z;
                        cb.included = true;
                        cb.numbps = B.mb - z;
                    }
                    int np = pass_count(br);
                    while (br.bit()) {
                        if (br.bad() || ++cb.lblock > 32) break;
                    }
                    // one length per codeword segment the passes touch (B.10.7.2)
                    for (int k = cb.npasses, k1 = cb.npasses + np; k < k1 && !br.bad();) {
                        int e = k1 - 1;
                        if (multi)
                            for (e = k; e < k1 - 1 && !seg_end(style, e);) e++;
                        uint32_t L = br.bits(cb.lblock + ilog2(e - k + 1));
                        Inc inc{id, L, multi ? seg_index(style, k) : 0};
                        if (ninc < 512)
                            incs[ninc] = inc;
                        else
                            more.push_back(inc);
                        ++ninc;
                        k = e + 1;
                    }
                    cb.npasses += np;
                }
        }
    }
    br.align();
    if (br.bad()) {
        err = "truncated packet header";
        return Status::Truncated;
    }
    // the header ends here, with EPH when signalled, in whichever stream holds it
    const uint8_t* hd = ph ? ph->data() : d;
    const size_t hend = ph ? ph->size() : end;
    size_t& hp = ph ? hpos : pos;
    hp = br.pos();
    if (P.eph) {
        if (hp + 2 <= hend && hd[hp] == 0xFF && hd[hp + 1] == 0x92)
            hp += 2;
        else {
            err = "missing EPH";
            return Status::BadMarker;
        }
    }
    for (int s = 0; s < ninc; s++) {
        const Inc& e = s < 512 ? incs[s] : more[s - 512];
        if (pos + e.len > end) {
            err = "truncated packet body";
            return Status::Truncated;
        }
        CodeBlock& cb = img.cblks[e.cb];
        if (multi) {
            if (cb.term < 0) {
                cb.term = (int)out.termlens.size();
                out.termlens.emplace_back();
            }
            std::vector<uint32_t>& tl = out.termlens[cb.term];
            if ((int)tl.size() <= e.seg) tl.resize(e.seg + 1, 0);
            tl[e.seg] += e.len;
        }
        if (e.len) {
            if (cb.nseg == 0) cb.seg0_off = (uint32_t)pos;
            out.segs.push_back(Segment{e.cb, (uint32_t)pos, e.len});
            cb.nseg++;
            cb.nbytes += e.len;
        }
        pos += e.len;
    }
    return Status::Ok;
}

Status read_packets(Image& img, int t, const Params& P, TileState& bstate, const std::vector<Range>& ranges,
                    const std::vector<uint8_t>* ph, PacketOut& out, std::string& err) {
    if (ranges.empty() && !ph) return Status::Ok;
    size_t ri = 0, pos = ranges.empty() ? 0 : ranges[0].off, hpos = 0;
    const Tile& T = img.tiles[t];
    int C = img.ncomp, maxr = 0;
    for (int c = 0; c < C; c++) maxr = std::max(maxr, T.comps[c].nlev + 1);
    // packet (c, r, q, l) unless the precinct has none or its layer-l packet
    // came earlier (B.12.2: a packet appears once, in layer order per precinct)
    auto emit = [&](int c, int r, int q, int l) -> Status {
        if (r > T.comps[c].nlev) return Status::Ok;
        ResState& RS = bstate[c][r];
        if (q >= RS.pw * RS.ph || RS.nextl[q] != l) return Status::Ok;
        RS.nextl[q]++;
        return read_packet(img, t, c, r, q, l, P, bstate, ranges, ri, pos, ph, hpos, out, err);
    };
    // RPCL, PCRL and CPRL (B.12.1.3-B.12.1.5) step through reference-grid
    // positions and emit a precinct where it starts: at a multiple of its size
    // projected onto the grid, or at the tile origin for the first precinct of
    // a resolution whose origin is not such a multiple.  Every precinct starts
    // at exactly one position, so sorting the precincts by that position gives
    // the same sequence.
    struct Start {
        int64_t y, x;
        int c, r, q;
    };
    std::vector<Start> starts;
    bool positional = P.prog >= 2;
    for (const Poc& v : P.poc) positional |= v.prog >= 2;
    if (positional)
        for (int c = 0; c < C; c++) {
            const TileComp& tc = T.comps[c];
            for (int r = 0; r <= tc.nlev; r++) {
                const ResState& RS = bstate[c][r];
                const Resolution& R = tc.res[r];
                const int lv = tc.nlev - r;
                const int64_t sx = (int64_t)1 << (RS.ppx + lv), sy = (int64_t)1 << (RS.ppy + lv);
                const bool ux = (((int64_t)R.x0 << lv) % sx) != 0, uy = (((int64_t)R.y0 << lv) % sy) != 0;
                // on the reference grid a precinct spans XRsiz * 2^(PPx + levels) samples
                for (int q = 0; q < RS.pw * RS.ph; q++) {
                    int i = q % RS.pw, j = q / RS.pw;
                    int64_t x = (i == 0 && ux) ? T.x0 : (RS.px0 + i) * sx * img.dx[c];
                    int64_t y = (j == 0 && uy) ? T.y0 : (RS.py0 + j) * sy * img.dy[c];
                    starts.push_back(Start{y, x, c, r, q});
                }
            }
        }
    // one progression volume (B.12.1) in order v.prog
    auto volume = [&](const Poc& v) -> Status {
        const int ce = std::min(v.ce, C), le = std::min(v.le, P.nlayers);
        Status s = Status::Ok;
        if (v.prog <= 1) {
            const bool lrcp = v.prog == 0;
            const int n0 = lrcp ? le : v.re, n1 = lrcp ? v.re : le;
            for (int a = lrcp ? 0 : v.rs; a < n0; a++)
                for (int b = lrcp ? v.rs : 0; b < n1; b++) {
                    const int l = lrcp ? a : b, r = lrcp ? b : a;
                    if (r >= maxr) continue;
                    for (int c = v.cs; c < ce; c++) {
                        if (r > T.comps[c].nlev) continue;
                        const ResState& RS = bstate[c][r];
                        for (int q = 0; q < RS.pw * RS.ph; q++)
                            if ((s = emit(c, r, q, l)) != Status::Ok) return s;
                    }
                }
            return Status::Ok;
        }
        std::vector<Start> seq;
        for (const Start& e : starts)
            if (e.c >= v.cs && e.c < ce && e.r >= v.rs && e.r < v.re) seq.push_back(e);
        auto key = [&](const Start& a) {
            if (v.prog == 2) return std::make_tuple((int64_t)a.r, a.y, a.x, (int64_t)a.c);  // RPCL
            if (v.prog == 3) return std::make_tuple(a.y, a.x, (int64_t)a.c, (int64_t)a.r);  // PCRL
            return std::make_tuple((int64_t)a.c, a.y, a.x, (int64_t)a.r);                   // CPRL
        };
        std::sort(seq.begin(), seq.end(), [&](const Start& a, const Start& b) { return key(a) < key(b); });
        for (const Start& e : seq)
            for (int l = 0; l < le; l++)
                if ((s = emit(e.c, e.r, e.q, l)) != Status::Ok) return s;
        return Status::Ok;
    };
    // the POC volumes in order, then the COD progression over what they left (B.12)
    Status s = Status::Ok;
    for (const Poc& v : P.poc)
        if ((s = volume(v)) != Status::Ok) return s;
    return volume(Poc{0, 0, P.nlayers, maxr, C, P.prog});
}

Status parse_impl(const uint8_t* data, size_t len, Image& img, std::string& err, bool packets, int threads) {
    img = Image();
    const uint8_t* cs = nullptr;
    size_t cl = 0;
    if (!data) {
        err = "no codestream";
        return Status::InvalidArgument;
    }
    bool is_jp2 = false;
    Jp2Header jh;
    Status s = locate_codestream(data, len, &cs, &cl, &is_jp2, jh, err);
    if (s != Status::Ok) return s;
    img.cs = cs;
    img.cs_len = cl;
    img.jp2 = is_jp2 ? 1 : 0;
    img.colourspace = jh.colourspace;
    Reader r(cs, cl);
    Params main;
    s = parse_main_header(r, img, main, err);
    if (s == Status::Ok && is_jp2) s = check_jp2_header(jh, img, err);
    img.wx0 = img.x0;
    img.wy0 = img.y0;
    img.wx1 = img.x1;
    img.wy1 = img.y1;
    img.fx0 = img.x0;
    img.fy0 = img.y0;
    img.fx1 = img.x1;
    img.fy1 = img.y1;
    if (s != Status::Ok || !packets) {
        if (s == Status::Ok) {
            // geometry of tile 0 only, for probing
            img.tiles.resize((size_t)img.ntx * img.nty);
            TileState bs;
            s = build_tile(img, 0, main, bs, err);
            img.cblks.clear();
        }
        return s;
    }
    int nt = img.ntx * img.nty;
    img.tiles.resize(nt);
    std::vector<Params> tp(nt);
    std::vector<uint8_t> seen(nt, 0), packed(nt, 0);
    std::vector<std::vector<Range>> ranges(nt);
    std::vector<std::vector<uint8_t>> hdrs(nt);  // packed packet headers per tile
    // payloads of marker segments in the order of their index bytes
    auto concat = [&](std::vector<PackedSeg> segs, std::vector<uint8_t>& dst) {
        std::stable_sort(segs.begin(), segs.end(), [](const PackedSeg& a, const PackedSeg& b) { return a.z < b.z; });
        for (const PackedSeg& g : segs) dst.insert(dst.end(), cs + g.off, cs + g.off + g.len);
    };
    std::vector<uint8_t> ppm;  // Nppm + Ippm per tile-part, in codestream order
    size_t ppm_pos = 0;
    concat(main.ppm, ppm);
    // tile-parts (A.4.2)
    for (;;) {
        if (r.pos() + 2 > cl) break;  // tolerate a missing EOC
        size_t sot = r.pos();
        uint32_t m = r.u16();
        if (m == 0xFFD9) break;
        if (m != 0xFF90) {
            err = "expected SOT";
            return Status::BadMarker;
        }
        uint32_t Lsot = r.u16();
        uint32_t I = r.u16();
        uint32_t Psot = r.u32();
        uint32_t TP = r.u8();
        r.u8();
        if (!r.ok()) return Status::Truncated;
        if (Lsot != 10 || (int)I >= nt) {
            err = "bad SOT";
            return Status::BadMarker;
        }
        size_t end = Psot ? sot + Psot : (cl >= 2 ? cl - 2 : cl);
        if (end > cl || end < r.pos()) {
            err = "tile-part runs past the codestream";
            return Status::Truncated;
        }
        if (!seen[I]) {
            tp[I] = main;
            seen[I] = 1;
        }
        std::vector<uint8_t> coc(img.ncomp, 0), qcc(img.ncomp, 0);
        std::vector<PackedSeg> ppt;
        for (;;) {
            uint32_t mk = r.u16();
            if (!r.ok()) return Status::Truncated;
            if (mk == 0xFF93) break;
            if (mk == 0xFF61) {  // PPT (A.7.5)
                size_t lpos = r.pos();
                uint32_t L = r.u16();
                if (!r.ok() || L < 3 || lpos + L > end) {
                    err = "bad PPT";
                    return Status::BadMarker;
                }
                int z = (int)r.u8();
                ppt.push_back(PackedSeg{z, r.pos(), lpos + L - r.pos()});
                r.seek(lpos + L);
                continue;
            }
            if (TP != 0 && (mk == 0xFF52 || mk == 0xFF53 || mk == 0xFF5C || mk == 0xFF5D)) {
                err = "coding parameters in a later tile-part";
                return Status::BadMarker;
            }
            s = read_segment(r, mk, img.ncomp, tp[I], coc, qcc, err, true);
            if (s != Status::Ok) return s;
        }
        if (!ppt.empty() && !main.ppm.empty()) {
            err = "PPM and PPT together";
            return Status::BadMarker;
        }
        if (!ppt.empty()) {
            concat(ppt, hdrs[I]);
            packed[I] = 1;
        }
        if (!main.ppm.empty()) {
            if (ppm_pos + 4 > ppm.size()) {
                err = "PPM holds fewer tile-parts than the codestream";
                return Status::Truncated;
            }
            size_t n = ((size_t)ppm[ppm_pos] << 24) | ((size_t)ppm[ppm_pos + 1] << 16) |
                       ((size_t)ppm[ppm_pos + 2] << 8) | ppm[ppm_pos + 3];
            ppm_pos += 4;
            if (ppm_pos + n > ppm.size()) {
                err = "truncated PPM";
                return Status::Truncated;
            }
            hdrs[I].insert(hdrs[I].end(), ppm.begin() + ppm_pos, ppm.begin() + ppm_pos + n);
            ppm_pos += n;
            packed[I] = 1;
        }
        if (end > r.pos()) ranges[I].push_back(Range{r.pos(), end - r.pos()});
        r.seek(end);
    }
    // geometry of every tile (its code-blocks take consecutive indices), then
    // the packets of each tile -- independent of the other tiles' (B.9), so
    // several tiles parse at once -- merged in tile order
    std::vector<TileState> bs(nt);
    std::vector<size_t> cb0(nt + 1, 0);
    for (int t = 0; t < nt; t++) {
        cb0[t] = img.cblks.size();
        s = build_tile(img, t, seen[t] ? tp[t] : main, bs[t], err);
        if (s != Status::Ok) return s;
    }
    cb0[nt] = img.cblks.size();
    std::vector<PacketOut> outs(nt);
    std::vector<Status> st(nt, Status::Ok);
    std::vector<std::string> errs(nt);
    auto work = [&](int k, int n) {
        for (int t = k; t < nt; t += n)
            st[t] = read_packets(img, t, seen[t] ? tp[t] : main, bs[t], ranges[t], packed[t] ? &hdrs[t] : nullptr,
                                 outs[t], errs[t]);
    };
    // a thread is worth starting for about a megabyte of packet data
    const int nthr = std::max(1, std::min<int>({threads, nt, (int)(1 + cl / (1u << 20))}));
    if (nthr <= 1) {
        work(0, 1);
    } else {
        std::vector<std::thread> th;
        for (int k = 1; k < nthr; k++) th.emplace_back(work, k, nthr);
        work(0, nthr);
        for (auto& x : th) x.join();
    }
    for (int t = 0; t < nt; t++)
        if (st[t] != Status::Ok) {
            err = errs[t];
            return st[t];
        }
    for (int t = 0; t < nt; t++) {
        const int off = (int)img.termlens.size();
        for (size_t k = cb0[t]; k < cb0[t + 1]; k++)
            if (img.cblks[k].term >= 0) img.cblks[k].term += off;
        for (auto& tl : outs[t].termlens) img.termlens.push_back(std::move(tl));
        img.segs.insert(img.segs.end(), outs[t].segs.begin(), outs[t].segs.end());
    }
    // each BYPASS / TERMALL block's segment lengths end with their count (the
    // table the tier-1 kernel reads before the codeword)
    for (auto& tl : img.termlens) tl.push_back((uint32_t)tl.size());
    for (auto& cb : img.cblks) {
        if (!cb.included) continue;
        if (cb.numbps > 31 || cb.numbps < 0) {
            err = "bit-plane count out of range";
            return Status::Unsupported;
        }
        int maxp = cb.numbps > 0 ? 3 * cb.numbps - 2 : 0;
        if (cb.npasses > maxp) cb.npasses = maxp;
    }
    return Status::Ok;
}

}  // namespace

int Image::max_prec() const {
    int m = 0;
    for (int p : prec) m = std::max(m, p);
    return m;
}

bool Image::all_reversible() const {
    for (const auto& t : tiles)
        for (const auto& c : t.comps)
            if (!c.reversible) return false;
    return true;
}

Status parse_codestream(const uint8_t* data, size_t len, Image& img, std::string& err, int threads) {
    return parse_impl(data, len, img, err, true, threads);
}

Status probe_codestream(const uint8_t* data, size_t len, Image& img, std::string& err) {
    return parse_impl(data, len, img, err, false, 1);
}

Status reduce_image(Image& img, int k, std::string& err) {
    if (k == 0) return Status::Ok;
    if (k < 0 || k > 32) {
        err = "reduce out of range";
        return Status::InvalidArgument;
    }
    for (const Tile& T : img.tiles)
        for (const TileComp& tc : T.comps)
            if (tc.nlev < k) {
                err = "reduce " + std::to_string(k) + " exceeds the " + std::to_string(tc.nlev) +
                      " decomposition levels of a tile-component";
                return Status::Unsupported;
            }
    // resolution coordinates k levels down: ceil(x / 2^k) (B.5)
    auto cdiv = [k](int v) { return (int)(((int64_t)v + ((int64_t)1 << k) - 1) >> k); };
    for (Tile& T : img.tiles) {
        for (TileComp& tc : T.comps) {
            tc.nlev -= k;
            tc.res.resize((size_t)tc.nlev + 1);
            const Resolution& R = tc.res[tc.nlev];
            tc.x0 = R.x0;
            tc.y0 = R.y0;
            tc.x1 = R.x1;
            tc.y1 = R.y1;
        }
        T.x0 = cdiv(T.x0);
        T.y0 = cdiv(T.y0);
        T.x1 = cdiv(T.x1);
        T.y1 = cdiv(T.y1);
    }
    for (CodeBlock& cb : img.cblks)
        if (cb.res > img.tiles[cb.tile].comps[cb.comp].nlev) cb.included = false;
    img.x0 = cdiv(img.x0);
    img.y0 = cdiv(img.y0);
    img.x1 = cdiv(img.x1);
    img.y1 = cdiv(img.y1);
    img.wx0 = img.x0;
    img.wy0 = img.y0;
    img.wx1 = img.x1;
    img.wy1 = img.y1;
    img.reduced += k;
    return Status::Ok;
}

Status window_image(Image& img, int x, int y, int w, int h, std::string& err) {
    if (img.planar()) {
        err = "region decode of sub-sampled components is not supported";
        return Status::Unsupported;
    }
    if (x < 0 || y < 0 || w <= 0 || h <= 0 || x >= img.width() || y >= img.height()) {
        err = "region outside the image";
        return Status::InvalidArgument;
    }
    img.wx0 = img.x0 + x;
    img.wy0 = img.y0 + y;
    img.wx1 = (int)std::min<int64_t>((int64_t)img.wx0 + w, img.x1);
    img.wy1 = (int)std::min<int64_t>((int64_t)img.wy0 + h, img.y1);
    // needed band ranges per tile-component and level: [a[n], b[n]) from
    // [a[n-1] - 4, b[n-1] + 4) at the next finer resolution, halved
    auto fdiv2 = [](int64_t v) { return v >= 0 ? v / 2 : -((-v + 1) / 2); };
    std::vector<std::vector<std::array<int64_t, 4>>> need(img.tiles.size());
    for (size_t t = 0; t < img.tiles.size(); t++) {
        const Tile& T = img.tiles[t];
        need[t].resize(T.comps.size() * 33);
        for (size_t c = 0; c < T.comps.size(); c++) {
            const TileComp& tc = T.comps[c];
            int64_t ax = std::max(img.wx0, tc.x0), bx = std::min(img.wx1, tc.x1);
            int64_t ay = std::max(img.wy0, tc.y0), by = std::min(img.wy1, tc.y1);
            const bool empty = ax >= bx || ay >= by;
            for (int n = 0; n <= tc.nlev; n++) {
                if (n > 0) {
                    ax = fdiv2(ax - 4);
                    bx = -fdiv2(-(bx + 4));
                    ay = fdiv2(ay - 4);
                    by = -fdiv2(-(by + 4));
                }
                need[t][c * 33 + n] = empty ? std::array<int64_t, 4>{0, 0, 0, 0}
                                            : std::array<int64_t, 4>{ax, bx, ay, by};
            }
        }
    }
    for (CodeBlock& cb : img.cblks) {
        if (!cb.included) continue;
        const TileComp& tc = img.tiles[cb.tile].comps[cb.comp];
        const int n = cb.res == 0 ? tc.nlev : tc.nlev - cb.res + 1;
        const auto& q = need[cb.tile][cb.comp * 33 + n];
        if (!(cb.x0 < q[1] && cb.x1 > q[0] && cb.y0 < q[3] && cb.y1 > q[2])) cb.included = false;
    }
    return Status::Ok;
}

const char* status_name(Status s) {
    switch (s) {
    case Status::Ok: return "ok";
    case Status::InvalidArgument: return "invalid argument";
    case Status::Truncated: return "truncated codestream";
    case Status::BadMarker: return "bad marker";
    case Status::Unsupported: return "unsupported feature";
    case Status::BadGeometry: return "bad geometry";
    case Status::OutOfMemory: return "out of memory";
    case Status::Device: return "device error";
    case Status::Capacity: return "capacity exceeded";
    }
    return "unknown";
}

}  // namespace tc
