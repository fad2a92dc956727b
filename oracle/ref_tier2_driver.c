/*
 * Test infrastructure only.  Driver around the reference's own tier-2 parser
 * (decode_codestream, TC/codestream.c:516-536, with its per-code-block
 * callback TC/codestream.c:455-470), compiled from the reference's sources by
 * oracle/Makefile into oracle/_ref/.  It prints one line per code-block:
 *   tile comp res orient tlx tly brx bry passes significant_bits length fnv1a
 * so tests can compare the reference's tier-2 tables with this repository's
 * parser on the same codestream.  The reference reads raw codestreams without
 * COM markers, one tile-part per tile, one layer and one precinct.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "codestream.h"
#include "codestream_image.h"

static uint32_t fnv1a(const unsigned char *p, unsigned n) {
    uint32_t h = 2166136261u;
    for (unsigned i = 0; i < n; i++) h = (h ^ p[i]) * 16777619u;
    return h;
}

static void on_block(type_codeblock *cb, unsigned char *bytes) {
    type_subband *sb = cb->parent_sb;
    type_res_lvl *rl = sb->parent_res_lvl;
    type_tile_comp *tc = rl->parent_tile_comp;
    printf("%u %u %u %d %u %u %u %u %u %u %u %u\n", tc->parent_tile->tile_no, tc->tile_comp_no,
           (unsigned)rl->res_lvl_no, (int)sb->orient, cb->tlx, cb->tly, cb->brx, cb->bry, cb->num_coding_passes,
           (unsigned)cb->significant_bits, cb->length, fnv1a(bytes, cb->length));
}

int main(int argc, char **argv) {
    if (argc != 2) {
        fprintf(stderr, "usage: %s file.j2k\n", argv[0]);
        return 2;
    }
    FILE *f = fopen(argv[1], "rb");
    if (!f) return 2;
    fseek(f, 0, SEEK_END);
    long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    unsigned char *data = malloc((size_t)n + 64);
    if (!data || fread(data, 1, (size_t)n, f) != (size_t)n) return 2;
    memset(data + n, 0, 64);
    fclose(f);
    type_buffer buf;
    memset(&buf, 0, sizeof buf);
    buf.data = data;
    buf.size = (unsigned long)n;
    buf.start = data;
    buf.end = data + n;
    buf.bp = data;
    type_image img;
    memset(&img, 0, sizeof img);
    codeBlockCallback = on_block;
    decode_codestream(&buf, &img);
    return 0;
}
