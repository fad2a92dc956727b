// Drop-in of the decoder-side part of TC/codestream_image_types.h:30-424: the
// image, tile and tile-component records the stage classes take.  Geometry is
// filled from the parsed codestream; img_data_d carries the tile-component's
// device plane (Mallat layout after tier-1, raster after the inverse DWT).
#pragma once
#include <stdint.h>

#include "platform.h"

struct type_image;
struct type_tile;

enum { DWT_97 = 0, DWT_53 = 1 };  // wavelet_type values (TC/codestream.c:121)
enum { UNSIGNED = 0, SIGNED = 1 };

struct type_tile_comp {
    unsigned int tile_comp_no;
    int tlx, tly, brx, bry;        // tile-component rectangle on the canvas
    unsigned short width, height;
    unsigned char num_dlvls, num_rlvls;
    unsigned char cblk_w, cblk_h;
    int reversible;                // 1: 5/3 (int32 samples), 0: 9/7 (float samples)
    int prec, sgnd;
    tDeviceMem img_data_d;         // owned by the decoder
    void* img_data_h;              // host view after mapComponentToHost
    type_tile* parent_tile;
    int stage;                     // 0 empty, 1 coefficients (Mallat), 2 samples (raster), 3 colour done
};

struct type_tile {
    unsigned int tile_no;
    int tlx, tly, brx, bry;
    unsigned short width, height;
    void* coefficients;            // unused: tier-1 writes the tile-component planes directly
    type_tile_comp* tile_comp;
    type_image* parent_img;
    int use_mct;
    const void* impl;              // parsed tile (library internal)
};

struct type_image {
    unsigned short width, height;
    unsigned short num_components;
    unsigned short depth;
    unsigned char sign;
    unsigned char num_dlvls;
    unsigned char wavelet_type;
    unsigned int num_tiles;
    unsigned char use_mct, use_part2_mct;
    unsigned short num_layers;
    type_tile* tile;
    void* impl;                    // parsed image and device state (library internal)
};
