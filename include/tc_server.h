/*
 * tc_server.h -- C ABI of the decode server: continuous batching of decode
 * requests on one MI355X.
 *
 * No counterpart in the reference (its Decoder decodes one file per call,
 * TC/Decoder.cpp:94-216).  Requests are parsed when submitted; a scheduler
 * thread admits queued requests into the next batch of a free slot while the
 * batch's device-memory footprint (an upper estimate of the coefficient
 * planes, inverse-DWT work buffers, tier-1 state, codewords, descriptors and
 * output, times the 1.25 growth of the slot's buffers) fits the slot's share
 * of the budget, and
 * launches it (one launch per stage for the whole batch) on the slot's own
 * stream, so batches of different slots overlap on the device.  Device buffers
 * of every slot stay resident between batches.  Results are copied back with
 * the batch and kept until read.  Every request is decoded independently, so
 * its result does not depend on what it was batched with.
 */
#ifndef TC_SERVER_H
#define TC_SERVER_H

#include "tc_device.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tc_server_s *tc_server;

typedef struct {
    int32_t slots;            /* concurrent batches (streams), default 2 */
    int32_t max_batch;        /* most requests per batch, default 256 */
    uint64_t memory_budget;   /* device bytes for all slots together, default 32 GiB */
    int32_t bpw_small, bpw_large; /* tier-1 packing (0: automatic) */
    int32_t dwt_shape;            /* inverse-DWT tile shape (tc_idwt_tile_shape) */
    uint64_t max_batch_bytes;     /* most codestream bytes per batch (a larger request
                                     runs alone), default 1 GiB: batches that large
                                     hold several 8K 16-bit images, whose longest
                                     block chains then bound one batch, not each */
    int32_t overlap_kernels;      /* 1: the kernels of different slots' batches may run at the
                                     same time; 0 (default): one batch's kernels after another's,
                                     with uploads, result copies and host work overlapped */
} tc_server_config;

typedef struct {
    tc_status status;
    int32_t width, height, ncomp, bytes_per_sample;
    double submit_s, start_s, done_s; /* seconds since the server started */
    uint64_t batch;                   /* batch sequence number */
    int32_t batch_size;
    uint64_t footprint;               /* device bytes accounted for the request */
} tc_request_info;

typedef struct {
    uint64_t submitted, completed, failed, batches;
    uint64_t pixels;          /* decoded pixels */
    double busy_s;            /* time with at least one batch on the device */
    uint64_t peak_bytes;      /* largest footprint of the batches holding device buffers at once
                                 (a slot plans its next batch while one runs: that one is not counted) */
    int32_t queued, in_flight;
    double plan_s;            /* summed over batches: parsing, planning and staging (host) */
    double device_s;          /* launch to results in host memory */
    double copy_s;            /* publishing results (they stay in pinned memory until read) */
    uint64_t slot_device_peak; /* most device bytes one slot's resident buffers held after a batch
                                  (a slot holding more than its share of the budget releases them) */
} tc_server_stats;

tc_status tc_server_create(tc_device dev, const tc_server_config *cfg, tc_server *out);
tc_status tc_server_destroy(tc_server srv);
/* Copies the codestream; *id identifies the request.  The main header is read
 * here (for the request's size and device footprint); the rest is parsed by
 * the worker that decodes it. */
tc_status tc_server_submit(tc_server srv, const uint8_t *cs, size_t len, uint64_t *id);
/* Same without the copy: cs must stay valid and unchanged until the request
 * has finished (tc_server_poll reports done). */
tc_status tc_server_submit_ref(tc_server srv, const uint8_t *cs, size_t len, uint64_t *id);
/* *done = 1 once the request finished (successfully or not). */
tc_status tc_server_poll(tc_server srv, uint64_t id, int *done);
/* Blocks until the request finished or timeout_s passed (TC_ERR_BUSY). */
tc_status tc_server_wait(tc_server srv, uint64_t id, double timeout_s);
/* Copies the decoded image (interleaved, as tc_decoder_decode) and releases the
 * request.  out may be NULL to read only info (the request is then kept). */
tc_status tc_server_result(tc_server srv, uint64_t id, void *out, size_t cap, tc_request_info *info);
/* Drops a finished request and its result without copying it. */
tc_status tc_server_release(tc_server srv, uint64_t id);
tc_status tc_server_get_stats(tc_server srv, tc_server_stats *stats);
/* Zeroes the counters and times of the stats (queued/in-flight stay live). */
tc_status tc_server_reset_stats(tc_server srv);
double tc_server_now(tc_server srv);

#ifdef __cplusplus
}
#endif
#endif
