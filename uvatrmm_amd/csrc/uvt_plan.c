/*
 * uvt_plan.c -- band mapping and tile planning (host side of libuvt).
 *
 * The reference runs one MPI rank per row band and, inside each rank, one
 * GEMM box plus one diagonal TRMM per band in sequence
 * (UVaTRMM/src/mm_triangular.c:319-410).  Here the same bands become a tile
 * schedule for one GPU: every output tile of C is an independent work item,
 * so a band's tiles and all bands' tiles run concurrently in one launch.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "uvt_internal.h"

const char *uvt_status_string(uvt_status s)
{
    switch (s) {
    case UVT_OK: return "ok";
    case UVT_ERR_ARG: return "invalid argument";
    case UVT_ERR_SIZE: return "unsupported size";
    case UVT_ERR_MAPPING: return "mapping rejected (the reference would abort)";
    case UVT_ERR_HIP: return "HIP runtime error";
    case UVT_ERR_NOMEM: return "out of memory";
    case UVT_ERR_BUDGET: return "request exceeds the memory budget";
    case UVT_ERR_UNSUPPORTED: return "unsupported (no GPU or feature missing)";
    case UVT_ERR_STATE: return "call out of order";
    case UVT_ERR_AGAIN: return "not ready";
    }
    return "unknown status";
}

const char *uvt_version(void) { return "uvatrmm-mi355x 0.1 (gfx950)"; }

int32_t uvt_cfg_tile_m(int32_t cfg) { return cfg == UVT_CFG_64x64 ? 64 : cfg == UVT_CFG_256x128 ? 256 : 128; }
int32_t uvt_cfg_tile_n(int32_t cfg)
{
    return cfg == UVT_CFG_64x64 || cfg == UVT_CFG_128x64 ? 64 : cfg == UVT_CFG_128x32 ? 32 : 128;
}

/* Narrow B: a 128x32 tile wastes no columns; small or narrow problems: 64x64
 * tiles give enough workgroups; otherwise 128x128 (most reuse per load). */
int32_t uvt_default_cfg(int64_t n, int64_t m)
{
    if (m <= 48) return UVT_CFG_128x32;
    if (n <= 512 || m <= 96) return UVT_CFG_64x64;
    /* A problem alone on the GPU: 128x128 tiles run two workgroups per CU, so
     * with fewer than 2 x 256 of them (MI355X) half-empty CUs cost more than
     * the smaller tile's lower efficiency. */
    if (((n + 127) / 128) * ((m + 127) / 128) < 2 * 256) return UVT_CFG_64x64;
    return UVT_CFG_128x128;
}

/* A walk of 256x128 tiles (16 waves of 32x64, one workgroup per CU; results
 * bit for bit those of 128x128 tiles) moves a quarter fewer operand bytes
 * per flop.  Measured on MI355X against 128x128 tiles, same box: where every
 * tile is full (n a multiple of 256, m of 128) 0.4-0.6 % faster at N=32768,
 * M=8192, 0.3-1.7 % at 8192^2, 0.8-1.4 % at 16384 x 4096, 1.0-1.7 % at
 * 16384^2, 0.8 % at 8192 x 4096 and 3-4 % at 4096^2, 8192 x 2048 and
 * 16384 x 1024; with edge tiles 1-5 % slower in six of eight shapes (the
 * idle waves of an edge tile leave half of a one-workgroup CU empty, where
 * two 128x128 workgroups fill each other's gaps).  So: full tiles only, at
 * least two per CU (512 on MI355X). */
int32_t uvt_wide_tiles(int32_t cfg, int64_t n, int64_t m)
{
    if (cfg == UVT_CFG_128x128 && n % 256 == 0 && m % 128 == 0 && (n / 256) * (m / 128) >= 2 * 256)
        return UVT_CFG_256x128;
    return cfg;
}

/* ------------------------------------------------------------------ */
/* Mapping                                                             */
/* ------------------------------------------------------------------ */

/* Reference exit codes (error_mpi_abort exits with enum value + 2,
 * mm_triangular_common.c:2854-2863; enum include/error.h:20-39). */
#define EXIT_BLOCKS_NUM 7 /* INVALID_BLOCKS_NUM */
#define EXIT_RANGE 11     /* INVALID_RANGE */

/* Rows above band k of the balanced mapping: the largest r with
 * r(r+1)/2 <= k*ideal, evaluated exactly as get_rows_before does it
 * (mm_mapping.c:177-181): double product, sqrt, floor, halve, truncate. */
static uint64_t rows_above(double ideal, uint32_t k)
{
    double kk = (double)(int)k;
    double disc = 1.0 + 8.0 * (kk * ideal);
    double num = sqrt(disc) - 1.0;
    if (!(num > 0.0)) return 0;
    return (uint64_t)(floor(num) / 2.0);
}

static uvt_status band_regular(uint64_t n, uint32_t p, uint32_t i, uvt_range *out)
{
    /* mm_mapping.c:27-67: floor(n/p) rows each, remainder to the last band.
     * The reference's abort (:40-54) needs n < p, which callers reject. */
    uint64_t q = n / p;
    if (q == 0) q = 1;
    out->begin = (uint64_t)i * q;
    out->size = (i + 1 == p) ? q + n % p : q;
    return UVT_OK;
}

static uvt_status band_balanced(uint64_t n, uint32_t p, uint32_t i, uvt_range *out, int *ref_exit)
{
    /* mm_mapping.c:319-366.  The element count of the triangle is held in a
     * 32-bit unsigned int by the reference (:120-126); from n = 92682 on it
     * overflows, and the reference's result is undefined: reject. */
    if (n >= 92682u) return UVT_ERR_SIZE;
    uint64_t tri = n * (n + 1) / 2;
    double per_band = (double)tri / (double)p;
    if (per_band < 1.0) {
        if (ref_exit) *ref_exit = EXIT_BLOCKS_NUM;
        return UVT_ERR_MAPPING;
    }
    double ideal = (double)(uint32_t)ceil(per_band);
    uint64_t first = (uint32_t)rows_above(ideal, i);
    uint64_t last = (uint32_t)rows_above(ideal, i + 1);
    uint64_t begin = first, size = (uint32_t)(last - first);
    if (i == 0) {
        begin = 0;
        size = (p > 1) ? rows_above(ideal, 1) : n;
    }
    if (n > 1 && i + 1 == p) {
        if (begin > n) return UVT_ERR_SIZE; /* would wrap in the reference */
        size = n - begin;
    }
    if (size == 0) { /* check_range, :184-191 */
        if (ref_exit) *ref_exit = EXIT_RANGE;
        return UVT_ERR_MAPPING;
    }
    out->begin = begin;
    out->size = size;
    return UVT_OK;
}

uvt_status uvt_map_band(uint64_t n, uint32_t p, uint32_t i, uvt_map map, uvt_range *out,
                        int *ref_exit)
{
    if (ref_exit) *ref_exit = 0;
    if (!out || p == 0 || i >= p || n == 0) return UVT_ERR_ARG;
    if (n < p) { /* mm_triangular.c:79: show_usage -> exit 3 */
        if (ref_exit) *ref_exit = 3;
        return UVT_ERR_MAPPING;
    }
    switch (map) {
    case UVT_MAP_REG: return band_regular(n, p, i, out);
    case UVT_MAP_BALANCED: return band_balanced(n, p, i, out, ref_exit);
    }
    return UVT_ERR_ARG;
}

uvt_status uvt_plan_bands(uint64_t n, uint32_t p, uvt_map map, uvt_range *out, int *ref_exit)
{
    if (!out) return UVT_ERR_ARG;
    for (uint32_t i = 0; i < p; i++) {
        uvt_status s = uvt_map_band(n, p, i, map, &out[i], ref_exit);
        if (s != UVT_OK) return s;
    }
    if (p == 0) return UVT_ERR_ARG;
    return UVT_OK;
}

/* ------------------------------------------------------------------ */
/* Tiles                                                               */
/* ------------------------------------------------------------------ */

typedef struct {
    uvt_tile *v;
    int64_t len, cap;
    int oom;
} tile_vec;

static void tv_push(tile_vec *tv, int64_t r0, int64_t rows, int64_t c0, int64_t k_lo, int64_t k_hi,
                    int32_t flags, int32_t band, int32_t cfg)
{
    if (tv->oom) return;
    if (tv->len == tv->cap) {
        int64_t nc = tv->cap ? 2 * tv->cap : 1024;
        uvt_tile *nv = (uvt_tile *)realloc(tv->v, (size_t)nc * sizeof(uvt_tile));
        if (!nv) {
            tv->oom = 1;
            return;
        }
        tv->v = nv;
        tv->cap = nc;
    }
    uvt_tile *t = &tv->v[tv->len++];
    t->r0 = (int32_t)r0;
    t->rows = (int32_t)rows;
    t->c0 = (int32_t)c0;
    t->k_lo = (int32_t)k_lo;
    t->k_hi = (int32_t)k_hi;
    t->flags = flags;
    t->band = band;
    t->cfg = cfg;
}

/* Heavy tiles first (longest reduction), ties in emission order: the
 * hardware dispatcher then fills the machine longest-job-first, and tiles
 * of one row panel (same A rows) are dispatched next to each other. */
typedef struct {
    int64_t work, seq;
    uvt_tile t;
} tile_key;

static int tile_key_cmp(const void *a, const void *b)
{
    const tile_key *x = (const tile_key *)a, *y = (const tile_key *)b;
    if (x->work != y->work) return x->work > y->work ? -1 : 1;
    return (x->seq > y->seq) - (x->seq < y->seq);
}

static int sort_heavy_first(uvt_tile *v, int64_t len)
{
    if (len < 2) return 0;
    tile_key *k = (tile_key *)malloc((size_t)len * sizeof(tile_key));
    if (!k) return -1;
    for (int64_t i = 0; i < len; i++) {
        k[i].work = (int64_t)v[i].k_hi - v[i].k_lo;
        k[i].seq = i;
        k[i].t = v[i];
    }
    qsort(k, (size_t)len, sizeof(tile_key), tile_key_cmp);
    for (int64_t i = 0; i < len; i++) v[i] = k[i].t;
    free(k);
    return 0;
}

static int32_t band_of_row(const uvt_range *bands, uint32_t p, int64_t r)
{
    for (uint32_t s = 0; s < p; s++)
        if ((uint64_t)r >= bands[s].begin && (uint64_t)r < bands[s].begin + bands[s].size)
            return (int32_t)s;
    return -1;
}

int64_t uvt_plan_trapezoid_tiles(int64_t n, int64_t m, int32_t cfg, uvt_tile *out)
{
    int32_t tm = uvt_cfg_tile_m(cfg), tn = uvt_cfg_tile_n(cfg);
    int64_t nrt = (n + tm - 1) / tm, nct = (m + tn - 1) / tn, c = 0;
    for (int64_t rt = nrt - 1; rt >= 0; rt--) { /* heavy (bottom) row panels first */
        int64_t r0 = rt * tm, rows = n - r0 < tm ? n - r0 : tm;
        for (int64_t ct = 0; ct < nct; ct++, c++) {
            if (!out) continue;
            uvt_tile *t = &out[c];
            t->r0 = (int32_t)r0;
            t->rows = (int32_t)rows;
            t->c0 = (int32_t)(ct * tn);
            t->k_lo = 0;
            t->k_hi = (int32_t)(r0 + rows);
            t->flags = UVT_TILE_LOWER;
            t->band = -1;
            t->cfg = cfg;
        }
    }
    return c;
}

static double issued_flops(const uvt_tile *v, int64_t len)
{
    double s = 0;
    for (int64_t i = 0; i < len; i++) {
        int32_t tm = uvt_cfg_tile_m(v[i].cfg), tn = uvt_cfg_tile_n(v[i].cfg);
        int64_t klo = v[i].k_lo & ~(int64_t)1; /* the kernels align k down to even */
        int64_t kl = v[i].k_hi - klo;
        int64_t kpad = (kl + 15) / 16 * 16; /* 16-deep k panels */
        s += 2.0 * tm * tn * (double)kpad;
    }
    return s;
}

uvt_status uvt_plan_create(int64_t n, int64_t m, uint32_t p, uvt_map map, uvt_shape shape,
                           uvt_plan *out, int *ref_exit)
{
    if (ref_exit) *ref_exit = 0;
    if (!out || n <= 0 || m <= 0 || p == 0) return UVT_ERR_ARG;
    if (shape != UVT_SHAPE_BOXES && shape != UVT_SHAPE_TRAPEZOID) return UVT_ERR_ARG;
    if (n > 0x3FFFFFFF || m > 0x3FFFFFFF) return UVT_ERR_SIZE; /* int32 tile fields */
    struct uvt_plan_s *pl = (struct uvt_plan_s *)calloc(1, sizeof *pl);
    if (!pl) return UVT_ERR_NOMEM;
    pl->n = n;
    pl->m = m;
    pl->p = p;
    pl->map = map;
    pl->shape = shape;
    pl->cfg = uvt_wide_tiles(uvt_default_cfg(n, m), n, m);
    pl->dev_id = -1;
    pl->bands = (uvt_range *)calloc(p, sizeof(uvt_range));
    if (!pl->bands) {
        free(pl);
        return UVT_ERR_NOMEM;
    }
    uvt_status st = uvt_plan_bands((uint64_t)n, p, map, pl->bands, ref_exit);
    if (st != UVT_OK) {
        free(pl->bands);
        free(pl);
        return st;
    }
    int32_t cfg = pl->cfg, tm = uvt_cfg_tile_m(cfg), tn = uvt_cfg_tile_n(cfg);
    int64_t nct = (m + tn - 1) / tn;
    tile_vec tv[2];
    memset(tv, 0, sizeof tv);
    if (shape == UVT_SHAPE_TRAPEZOID) {
        pl->phases = 1;
        int64_t cnt = uvt_plan_trapezoid_tiles(n, m, cfg, NULL);
        tv[0].v = (uvt_tile *)malloc((size_t)cnt * sizeof(uvt_tile));
        if (!tv[0].v) tv[0].oom = 1;
        else {
            uvt_plan_trapezoid_tiles(n, m, cfg, tv[0].v);
            tv[0].len = tv[0].cap = cnt;
            for (int64_t i = 0; i < cnt; i++) tv[0].v[i].band = band_of_row(pl->bands, p, tv[0].v[i].r0);
        }
    } else {
        pl->phases = 2;
        for (uint32_t s = 0; s < p; s++) {
            int64_t b = (int64_t)pl->bands[s].begin, h = (int64_t)pl->bands[s].size;
            /* phase 0: GEMM box, rows [b, b+h), k in [0, b] (mm_triangular.c:347-360) */
            for (int64_t r0 = b; r0 < b + h; r0 += tm) {
                int64_t rows = b + h - r0 < tm ? b + h - r0 : tm;
                for (int64_t ct = 0; ct < nct; ct++)
                    tv_push(&tv[0], r0, rows, ct * tn, 0, b + 1, 0, (int32_t)s, cfg);
            }
            /* phase 1: diagonal triangle, rows and k in [b+1, b+h), lower,
             * accumulated into C (mm_triangular.c:366-393) */
            int64_t tri = h - 1 < n ? h - 1 : n;
            int64_t left = b + h - tri;
            for (int64_t r0 = left; r0 < left + tri; r0 += tm) {
                int64_t rows = left + tri - r0 < tm ? left + tri - r0 : tm;
                for (int64_t ct = 0; ct < nct; ct++)
                    tv_push(&tv[1], r0, rows, ct * tn, left, r0 + rows,
                            UVT_TILE_LOWER | UVT_TILE_ACCUM, (int32_t)s, cfg);
            }
        }
    }
    if (tv[0].oom || tv[1].oom || sort_heavy_first(tv[0].v, tv[0].len) ||
        sort_heavy_first(tv[1].v, tv[1].len)) {
        free(tv[0].v);
        free(tv[1].v);
        free(pl->bands);
        free(pl);
        return UVT_ERR_NOMEM;
    }
    for (int ph = 0; ph < 2; ph++) {
        pl->tiles[ph] = tv[ph].v;
        pl->ntiles[ph] = tv[ph].len;
    }
    pl->flops = (double)m * (double)n * (double)n;
    pl->issued = issued_flops(pl->tiles[0], pl->ntiles[0]) + issued_flops(pl->tiles[1], pl->ntiles[1]);
    *out = pl;
    return UVT_OK;
}

uvt_status uvt_plan_destroy(uvt_plan plan)
{
    if (!plan) return UVT_ERR_ARG;
    uvt_plan_release_device(plan);
    free(plan->tiles[0]);
    free(plan->tiles[1]);
    free(plan->bands);
    free(plan);
    return UVT_OK;
}

int32_t uvt_plan_phases(uvt_plan plan) { return plan ? plan->phases : 0; }

int64_t uvt_plan_num_tiles(uvt_plan plan, int32_t phase)
{
    if (!plan || phase < 0 || phase >= plan->phases) return 0;
    return plan->ntiles[phase];
}

uvt_status uvt_plan_get_tiles(uvt_plan plan, int32_t phase, uvt_tile *out, int64_t cap)
{
    if (!plan || !out || phase < 0 || phase >= plan->phases) return UVT_ERR_ARG;
    if (cap < plan->ntiles[phase]) return UVT_ERR_SIZE;
    if (plan->ntiles[phase])
        memcpy(out, plan->tiles[phase], (size_t)plan->ntiles[phase] * sizeof(uvt_tile));
    return UVT_OK;
}

uvt_status uvt_plan_get_bands(uvt_plan plan, uvt_range *out, uint32_t cap)
{
    if (!plan || !out) return UVT_ERR_ARG;
    if (cap < plan->p) return UVT_ERR_SIZE;
    memcpy(out, plan->bands, plan->p * sizeof(uvt_range));
    return UVT_OK;
}

uvt_status uvt_plan_tile_shape(uvt_plan plan, int32_t *tile_m, int32_t *tile_n)
{
    if (!plan) return UVT_ERR_ARG;
    if (tile_m) *tile_m = uvt_cfg_tile_m(plan->cfg);
    if (tile_n) *tile_n = uvt_cfg_tile_n(plan->cfg);
    return UVT_OK;
}

double uvt_plan_flops(uvt_plan plan) { return plan ? plan->flops : 0.0; }
double uvt_plan_issued_flops(uvt_plan plan) { return plan ? plan->issued : 0.0; }
