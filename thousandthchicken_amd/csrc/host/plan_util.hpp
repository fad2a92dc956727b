// Work-list builders shared by the fused pipeline and the stage classes.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "../tc_kernels.h"
#include "codestream.hpp"

namespace tc {

// Appends the code-blocks of tile t to the tier-1 work list.  plane[c] is the
// element offset of component c's coefficient plane (pitch = tile-component
// width).  Codewords are gathered 4-byte aligned into cw; cls[0] receives
// blocks up to 32x32, cls[1] the others.
Status append_tile_cblks(const Image& im, int t, const uint64_t* plane, std::vector<uint8_t>& cw,
                         std::vector<tc_cblk_desc>& descs, std::vector<uint32_t> (&cls)[2], std::string& err);

// One codeword segment to gather: len bytes from src to offset dst of the arena.
struct CwCopy {
    const uint8_t* src;
    uint32_t len;
    uint64_t dst;
};

// Same work list as append_tile_cblks, but the codeword bytes are not moved:
// the arena grows by the block sizes (cw_size, 4-byte aligned per block) and
// the segment copies are listed for the caller to perform.
Status plan_tile_cblks(const Image& im, int t, const uint64_t* plane, uint64_t& cw_size, std::vector<CwCopy>& copies,
                       std::vector<tc_cblk_desc>& descs, std::vector<uint32_t> (&cls)[2], std::string& err);

// Orders a class of blocks for packing into waves: equal pass counts, then
// equal stripe counts (then, for the class up to 32x32, equal widths), then
// similar codeword lengths side by side, heaviest first.
void sort_for_packing(std::vector<uint32_t>& idx, const std::vector<tc_cblk_desc>& descs, bool small);

// Work-buffer sizes (elements) of the two ping-pong buffers of a tile-component.
void work_sizes(const TileComp& tc, uint64_t sz[2]);

// Job for decomposition level n (>= 1) of a tile-component.  Level n reads its
// LL from the coefficient plane (n == nlev) or from work[(nlev-n-1)%2], writes
// work[(nlev-n)%2] -- or, when raster_out is set and n == 1, the raster.
tc_dwt_job level_job(const TileComp& tc, int n, uint64_t plane, const uint64_t work[2], bool raster_out,
                     uint64_t raster_off, uint32_t raster_pitch, int tile_w, int tile_h);

}  // namespace tc
