/*
 * uvt_internal.h -- definitions shared by the host planner (C) and the HIP
 * side of libuvt.  Not installed; callers use include/uvt.h.
 */
#ifndef UVT_INTERNAL_H
#define UVT_INTERNAL_H

#include "uvt.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Library-internal: linked across libuvt's own sources, not exported. */
#define UVT_INTERNAL __attribute__((visibility("hidden")))

/* FP64 tile kernel configurations (rows x columns of one workgroup tile). */
enum { UVT_CFG_128x128 = 0, UVT_CFG_64x64 = 1, UVT_CFG_128x32 = 2, UVT_NUM_CFGS = 3 };
/* 128x64 tiles: only the last column block of a batch problem whose columns
 * end 33..64 past a multiple of 128 (not a tuning choice). */
enum { UVT_CFG_128x64 = 3 };
/* 256x128 tiles (16 waves, one workgroup per CU): the tile walks of large
 * problems only (plans, the server's and the BLAS trapezoid), never a batch
 * or the tuner (uvt_wide_tiles). */
enum { UVT_CFG_256x128 = 4 };
UVT_INTERNAL int32_t uvt_cfg_tile_m(int32_t cfg);
UVT_INTERNAL int32_t uvt_cfg_tile_n(int32_t cfg);

struct uvt_plan_s {
    int64_t n, m;
    uint32_t p;
    uvt_map map;
    uvt_shape shape;
    int32_t cfg;
    uvt_range *bands;
    int32_t phases;
    int64_t ntiles[2];
    uvt_tile *tiles[2];
    double flops, issued;
    /* device state, owned by the plan, created by the HIP side on first use
     * on a device: launch records of one or two launches, and, when tiles
     * write partial products, the partial tiles and the records that add
     * them to C */
    int dev_id;
    void *dev_launch[2];
    int64_t nlaunch[2];
    void *dev_sums;
    int64_t nsums;
    void *dev_parts;
    int64_t nparts;
    int64_t sk_grid; /* > 0: a stream-K launch of that many workgroups */
};

/* Planner entry used by the batch path: plan one problem with a given tile
 * configuration into caller-owned arrays (NULL out = count only). */
UVT_INTERNAL int64_t uvt_plan_trapezoid_tiles(int64_t n, int64_t m, int32_t cfg, uvt_tile *out);
/* Built-in tile-shape choice for an n x n by n x m product. */
UVT_INTERNAL int32_t uvt_default_cfg(int64_t n, int64_t m);
/* The shape a tile walk of that product runs: cfg, or 256x128 for a
 * 128x128 problem whose 256x128 tiles are all full and at least 512. */
UVT_INTERNAL int32_t uvt_wide_tiles(int32_t cfg, int64_t n, int64_t m);

/* Releases device copies of a plan (defined on the HIP side). */
UVT_INTERNAL void uvt_plan_release_device(uvt_plan plan);

#ifdef __cplusplus
}
#endif
#endif
