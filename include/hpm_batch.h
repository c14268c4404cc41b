/*
 * hpm_batch.h — RunPatchMatch of many independent problems as one batch.
 *
 * No counterpart in the reference, which runs one problem at a time
 * (HPM::RunPatchMatch, HPM.cu:1906-1957). For many small or unevenly sized
 * problems (e.g. 32 references of 640x480 to 1920x1080 with 4-10 sources):
 * every stage of the schedule is one launch over all problems' tiles, the
 * schedule is captured once per batch shape in a hipGraph and replayed, and
 * the tile schedule (static ranges over a chosen grid, per-XCD work queues
 * with a chosen item size, or the whole schedule as one persistent
 * cooperative launch with grid barriers between stages) is picked by timing
 * the candidates on the first call of a shape (logged by hpm_batch_log). Each problem's results equal those of
 * hpm_run_patchmatch on its handle (same seeds and call index); tuning
 * trials save and restore the problems' state.
 */
#ifndef HPM_BATCH_H_
#define HPM_BATCH_H_

#include <stdint.h>

#include "hpm_capi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hpm_batch hpm_batch;

typedef struct hpm_batch_stats {
    double ms_device;        /* the replayed schedule, hipEvents */
    double ms_host;          /* the whole call, host clock (tuning and capture included on a new shape) */
    uint64_t n_eval;         /* NCC evaluations of all problems */
    int32_t problems;
    int32_t tiles_full;      /* 64x4 tiles of the full grids */
    int32_t tiles_colour;    /* 64x4 tiles of the colour grids */
    int32_t grid;            /* blocks per stage launch (autotuned) */
    int32_t graph_reused;    /* 1 when the captured schedule was replayed */
    int32_t items;           /* autotuned: 0 static tile ranges, else tiles per work-queue item */
    int32_t persistent;      /* 1: the whole schedule ran as one persistent launch */
} hpm_batch_stats;

hpm_batch* hpm_batch_create(int device);
void hpm_batch_destroy(hpm_batch* b);

/* Runs RunPatchMatch on the n handles together. Every handle must be ready
 * for hpm_run_patchmatch (device initialised, prior uploaded if enabled) and
 * on the batch's device. want_readback: fill each handle's host mirror (the
 * synchronous semantics of RunPatchMatch); 0 leaves the maps on the device.
 * Calls with the same count and sizes replay the captured schedule. */
int hpm_run_batch(hpm_batch* b, hpm_handle* const* handles, int n, int want_readback, hpm_batch_stats* stats);

/* Fixes the schedule instead of autotuning it: items 0 for static tile
 * ranges over `grid` blocks (0: 2 per CU), 1, 2 or 4 for work-queue items
 * of that many tiles; persistent (items 1 or 2): the whole schedule as one
 * cooperative launch of every resident block, grid barriers between its
 * stages; items -1 returns to autotuning. Drops the captured schedules. */
int hpm_batch_set_schedule(hpm_batch* b, int grid, int items, int persistent);
const char* hpm_batch_last_error(const hpm_batch* b);
/* one line per tuned batch shape: candidates with their times, the choice */
const char* hpm_batch_log(const hpm_batch* b);

#ifdef __cplusplus
}
#endif

#endif /* HPM_BATCH_H_ */
