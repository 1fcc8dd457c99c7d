/*
 * patchtable_amd.h -- C ABI of the MI355X (gfx950) PatchTable library.
 *
 * This is the drop-in boundary for the reference's construct-then-lookup API:
 *
 *   reference (patchtable/patchtable.h, patchtable_params.h)        this ABI
 *   -------------------------------------------------------------   ---------------------------
 *   PatchTableParams::PatchTableParams()   patchtable.cpp:375-377     pt_params_default
 *   PatchTableParams::set_speed(int)       patchtable.cpp:138-150     pt_params_set_speed
 *   PatchTableParams::set_from_switches    patchtable.cpp:152-373     pt_params_set_switches
 *   PatchTable<>::PatchTable(p, b, allowed) patchtable.h:4397-4424    pt_table_create
 *   PatchTable<>::lookup(a, ann, prev, sv) patchtable.h:4448-4468     pt_table_lookup
 *   PatchTable<>::save(filename)           patchtable.h:4470-4490     pt_table_save
 *   PatchTableFixedN(p, filename) (-load)  patchtable.h:1593-1629     pt_table_load
 *   PatchTable<>::~PatchTable()            patchtable.h:4426-4446     pt_table_destroy
 *   ASSERT2 / exit(1) error path           solver/util.h:33           int status + pt_last_error
 *
 * Conventions
 *   - Every entry point returns PT_OK (0) or a negative PT_E* status; the message of the last
 *     failure on the calling thread is returned by pt_last_error().
 *   - Images are float32, row-major height x width x channels (HWC), exactly the layout of the
 *     reference's Array<float> (solver/array.h:251-294).
 *   - The nearest-neighbour field `ann` is (h-7) x (w-7) x 3 doubles per query image of h x w:
 *     [NNF_X, NNF_Y, NNF_DIST] = (b patch left x, b patch top y, exact RGB SSD), the layout of
 *     the reference's Array<double> ann (patchtable.h:175-177, 3500).
 *   - A table owns its device memory and one HIP stream; calls on one table are serialised by
 *     the caller; distinct tables may be used from distinct threads.
 *   - Entry points with the _dev suffix take device pointers (inputs already resident in HBM) and
 *     run on the table's stream; the others take host pointers and copy.
 */
#ifndef PATCHTABLE_AMD_H
#define PATCHTABLE_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PT_OK              0
#define PT_E_INVALID      -1   /* invalid argument or parameter value                      */
#define PT_E_UNSUPPORTED  -2   /* parameter combination outside this library's scope       */
#define PT_E_HIP          -3   /* a HIP runtime call failed                                */
#define PT_E_IO           -4   /* file could not be read or written                        */
#define PT_E_NOMEM        -5   /* allocation failed                                        */

/* Enumerations, same values as the reference (patchtable.h:198-236). */
#define PT_DT_ALGO_RASTER      0
#define PT_DT_ALGO_PROP        1
#define PT_DT_ALGO_BRUTE       2
#define PT_DT_ALGO_KDTREE      3
#define PT_DT_ALGO_DOWNSAMPLE  4
#define PT_DT_ALGO_HYBRID      5
#define PT_DT_ALGO_EUCLIDEAN   6
#define PT_COLORSPACE_YUV      0
#define PT_COLORSPACE_LAB      1
#define PT_LOOKUP_ALGO_TABLE   0
#define PT_LOOKUP_ALGO_BRUTE   1
#define PT_LOOKUP_ALGO_KDTREE  2
#define PT_LOOKUP_ALGO_PM      3
#define PT_LOOKUP_ALGO_TREECANN 4
#define PT_KCOH_ALGO_FLANN     0
#define PT_KCOH_ALGO_PM        1
#define PT_KCOH_ALGO_ANN       2
#define PT_KCOH_ALGO_TREECANN  3
#define PT_DIM_ALGO_WH         0
#define PT_DIM_ALGO_PCA        1

#define PT_MAX_DIMS 64
#define PT_MFMA_MAX_JUMPS 16   /* entries of pt_params.mfma_jumps */

/* Plain-old-data mirror of PatchTableParams (patchtable_params.h:23-161). Field names and
 * meanings follow the reference; switch names accepted by pt_params_set_switches are the
 * reference's ("-speed", "-limit", "-ndims", ...). Fields that only select code paths outside
 * this library's scope are kept so that switch strings parse the same way; pt_table_create rejects
 * values it does not implement with PT_E_UNSUPPORTED. */
typedef struct pt_params {
    int    randomize_dt;
    int    parallel_dt;
    double limit;              /* max table cells (the "-limit n" switch is in millions)  */
    int    threads;
    int    dt_threads;
    int    convert_colorspace;
    int    colorspace;
    int    regular_grid;
    int    verbose;
    int    table_knn;
    int    populate_nearest;
    int    populate_random;
    int    partition_step;
    int    check_existing;
    int    do_table_lookup;
    int    init_random;
    double treecann_eps;
    int    allowed_index;

    int    ndims;
    int    nchroma;
    int    ntables;
    int    nslices_count;      /* 0: derive slices automatically                          */
    int    nslices[PT_MAX_DIMS];
    int    populate_exptime;
    int    dt_mode;
    int    dt_algo;
    int    dt_iters;
    int    run_dt;
    int    query_step;
    int    grid_ndims;         /* -1: same as ndims                                       */

    int    prop_dist;
    int    calc_exact_dist;
    int    do_prop;
    int    prop_y_step;
    int    do_rs;
    int    spatial;
    double prob_rs;
    int    do_calc_dist;
    int    lookup_algo;

    int    is_descriptor;
    int    descriptor_padded;
    int    sanitize_input;

    char   load_filename[1024];
    char   save_filename[1024];

    int    patch_w;
    double coherence_spatial;
    double coherence_temporal;
    int    recalc_dist_temporal;
    double table_ratio;
    int    dt_knn;
    int    treecann_agrid;
    int    treecann_bgrid;
    int    flann_build_step;
    int    flann_trees;
    int    flann_checks;
    double flann_eps;
    int    flann_reset_step;
    int    query_iters;
    int    prop_iters;
    double enrich_limit;

    int    kcoherence;
    int    kcoherence_min_dist;
    int    kcoherence_algo;
    int    kcoherence_iter;
    int    kcoherence_step;
    int    kcoherence_enrich;
    int    save_kcoherence;
    int    incoherence;
    int    incoherence_min_dist;
    double triangle_factor;

    int    pm_iters;
    int    pm_knn;
    int    pm_min_dist;
    int    pm_enrich;

    int    pca_samples;
    int    pca_mean;
    int    pca_subtract_mean;
    int    dim_algo;
    int    filter_dims_count;
    int    filter_dims[PT_MAX_DIMS];
    int    auto_filter;
    int    kdtree_add_unique;
    int    ann;
    int    ann_agrid;
    int    ann_bgrid;
    double ann_eps;

    /* ---- Parameters of this library's GPU algorithms (no reference counterpart). -------------
     * kcoh_window: half-width of the sorted-key candidate window of the k-coherence kNN.
     * jf_rounds:   number of parallel propagation rounds that replace the raster-order passes.
     * mfma_rounds: 0 for the lookup above; n > 0 selects the MFMA lookup: table candidates
     *              only, then n rounds of dense int8 search on matrix cores around up to
     *              mfma_anchors shared offsets per 16 query patches, each query's best rescored
     *              with the exact SSD (image mode, query_step 1, no allowed mask).
     * mfma_anchors: most anchor offsets searched per group of 16 query patches and round.
     * mfma_jumps: the jump of each MFMA round, in patches (the distance of the neighbours whose
     *              offsets a group takes as sources); a 0 ends the list and rounds past it use 1.
     *              All 0 (the default): 16, 8, 4, 2, 1, then 1 (fewer than 5 rounds start at
     *              2^(rounds-1)). A round whose jump equals the previous round's skips the groups
     *              none of whose sources changed in that round.
     * Set by pt_params_default; accepted as switches "-kcoh_window", "-jf_rounds",
     * "-mfma_rounds", "-mfma_anchors" and "-mfma_jumps" (comma-separated, e.g. 16,4,1,1). */
    int    kcoh_window;
    int    jf_rounds;
    int    mfma_rounds;
    int    mfma_anchors;
    int    mfma_jumps[PT_MFMA_MAX_JUMPS];
} pt_params;

typedef struct pt_table pt_table;

/* ---- Parameters (patchtable.cpp:26-387) ------------------------------------------------ */
int pt_params_default(pt_params *p);                          /* PatchTableParams()          */
int pt_params_set_speed(pt_params *p, int speed);             /* set_speed(i), i in 0..9     */
int pt_params_set_switches(pt_params *p, const char *s);      /* set_from_switches(string)   */

/* ---- Table construction (PatchTableFixedN constructor, patchtable.h:2584-3373) ---------
 * b: host image h x w x c (c must be 3 unless p->is_descriptor). allowed: NULL or an
 * (h-7) x (w-7) int32 mask compared against p->allowed_index. p is read, and the fields the
 * reference writes back (grid_ndims, patchtable.h:2603) are written back. */
int pt_table_create(pt_params *p, const float *b, int h, int w, int c,
                    const int32_t *allowed, pt_table **out);
/* Same with b already in device memory (HBM). */
int pt_table_create_dev(pt_params *p, const float *b_dev, int h, int w, int c,
                        const int32_t *allowed, pt_table **out);

/* ---- Query (lookup_templated, patchtable.h:3468-4243) ----------------------------------
 * a: host image h x w x c; ann_out: host (h-7) x (w-7) x 3 doubles. ann_prev: NULL or a
 * previous field of the same shape; sv: NULL or (h-7) x (w-7) floats of per-patch temporal
 * scale. seconds (may be NULL) receives the wall time of the lookup, measured like the
 * reference (query-side colour conversion and GCK included, patchtable.h:3477, 4211). */
int pt_table_lookup(pt_table *t, const float *a, int h, int w, int c, double *ann_out,
                    const double *ann_prev, const float *sv, double *seconds);
/* Same with a, ann_out (and ann_prev, sv if given) in device memory. */
int pt_table_lookup_dev(pt_table *t, const float *a_dev, int h, int w, int c, double *ann_out_dev,
                        const double *ann_prev_dev, const float *sv_dev, double *seconds);
/* No counterpart in the reference: pt_table_lookup_dev without any device allocation, for
 * callers that cannot allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). Same arguments, results and errors as pt_table_lookup_dev; the lookup works
 * in a library-owned device array of PT_NOALLOC_MAX_BYTES per GPU (held until the process
 * ends) instead of the table's buffers. A query whose working set exceeds it fails with
 * PT_E_INVALID before any launch: about 131 bytes per query pixel at query_step 1 (a 1024 x 768
 * query needs 102 MB), plus 16 per pixel with ann_prev. Learned-key tables are not supported
 * (PT_E_UNSUPPORTED). The current device must be the table's. Calls on one GPU share the array
 * and are serialised by a lock; the call returns when the lookup is done. */
#define PT_NOALLOC_MAX_BYTES (128u << 20)
int pt_table_lookup_dev_noalloc(pt_table *t, const float *a_dev, int h, int w, int c, double *ann_out_dev,
                                const double *ann_prev_dev, const float *sv_dev, double *seconds);

/* ---- Persistence (patchtable.h:1574-1622; patchtable.cpp:499-573) ----------------------
 * The file is the reference's: "PatchTable", table, b0, wh0, nslices, PatchPartition of one
 * PatchPartition1D per grid dimension, "EndPatchTable" (strings and vectors length-prefixed
 * with int32, Arrays as their sizes vector then the data), so files move both ways.
 * pt_table_load (PatchTableFixedN(p, filename), patchtable.h:2591-2594; also reached through
 * pt_table_create when p->load_filename is set) takes the grid dimensions from the file and,
 * since the format holds no k-coherence lists, rebuilds them on the GPU when p->kcoherence > 0. */
int pt_table_save(pt_table *t, const char *filename);
int pt_table_load(pt_params *p, const char *filename, pt_table **out);
int pt_table_destroy(pt_table *t);

/* ---- Image and NNF files (solver/array.h:1637-1684, 1745-1827; solver/pfm.cpp) -----------
 * pt_image_size: height, width and channels of a .png (always 3) or .pfm file.
 * pt_image_load: load_color_image<float>: PNG as 8-bit RGB * (1/255.0), PFM as stored (rows
 *   flipped to top-first); out is h x w x c floats.
 * pt_pfm_save: save_color_image<double> to .pfm of a 1 or 3 channel h x w x c array (an NNF:
 *   c = 3), values converted to float, rows written bottom-first, little endian. */
int pt_image_size(const char *path, int *h, int *w, int *c);
int pt_image_load(const char *path, float *out, int h, int w, int c);
int pt_pfm_save(const char *path, const double *data, int h, int w, int c);

/* ---- One query against many databases ------------------------------------------------------
 * The reference's one-vs-many benchmark runs one match per (query, database) pair
 * (patchtable/bench.py:80-83); this batch keeps one table per database image in HBM.
 * pt_batch_create: n RGB databases (device pointers, h[i] x w[i] x 3), built concurrently by
 *   `threads` host threads (0: 4), each table on its own stream; p is shared and borrowed.
 * pt_batch_lookup: the query (device, h x w x 3) against every table; ann_dev[i] receives the
 *   (h-7) x (w-7) x 3 NNF for database i (device). The query descriptor is computed once; then
 *   each phase of the lookup (candidates, each propagation or MFMA round, output) is ONE launch
 *   for all pairs (a grid of tiles x pairs, every pair's arguments in a device array), on the
 *   batch's stream, replayed from one captured HIP graph after the first call with the same
 *   buffers, tile width and parameters. Every pair's NNF equals its single-table lookup's.
 *   ms_per_pair[i] (optional): device time of the whole batch (every pair ends with it);
 *   seconds (optional): wall time of the whole call.
 * pt_batch_set_tile / pt_batch_get_tile: the lattice tile of the batch's lookup kernels, 64 x 4,
 *   32 x 8 or 16 x 16 patches (tile_x 64, 32, 16; results do not depend on it).
 * pt_batch_autotune: times pt_batch_lookup of this query and these outputs with each tile width
 *   (one eager run and capture, then `reps` replays each), keeps the fastest (median) and
 *   returns it; ms_out[3] (optional): the medians for 64, 32 and 16. No counterpart in the
 *   reference (an addition of this library). */
typedef struct pt_batch pt_batch;
int pt_batch_create(pt_params *p, int n, const float *const *b_dev, const int *h, const int *w, int threads,
                    pt_batch **out);
int pt_batch_lookup(pt_batch *bt, const float *a_dev, int h, int w, double *const *ann_dev, double *ms_per_pair,
                    double *seconds);
int pt_batch_size(const pt_batch *bt);
pt_table *pt_batch_table(pt_batch *bt, int i);
int pt_batch_set_tile(pt_batch *bt, int tile_x);
int pt_batch_get_tile(const pt_batch *bt);
int pt_batch_autotune(pt_batch *bt, const float *a_dev, int h, int w, double *const *ann_dev, int reps,
                      int *tile_x_out, double *ms_out);
int pt_batch_destroy(pt_batch *bt);

/* ---- Exhaustive nearest neighbours (the reference names LOOKUP_ALGO_BRUTE, patchtable.h:230,
 * "-lookup_algo brute" patchtable.cpp:276, but implements no search for it).
 * pt_brute_lookup_dev: for each query patch of a (device RGB, ah x aw x 3), the nearest of every
 *   8x8 patch of b (device RGB, bh x bw x 3). Candidate scoring is the dense contraction
 *   |a|^2 + |b|^2 - 2 a.b over the 192-dim patch vectors (centred by -0.5) on MFMA: BF16 (round to
 *   nearest even) or FP8 e4m3 with a per-patch scale (max |v| / 448), FP32 accumulation; each
 *   query keeps its 8 best approximate scores (every lane of the search keeps the 2 best of the
 *   database columns it sees, merged per query), which are rescored with the exact FP32 SSD
 *   (patch_dist_exact's lane order, patchtable.h:330-373): the reported distance is exact.
 *   qpos_dev: nq query positions XY_TO_INT(x, y) in patch coordinates (checked), or NULL for every
 *   query patch in raster order (nq = (ah-7)*(aw-7)). Outputs (device): pos_out[nq] (XY_TO_INT of
 *   the match, patch coordinates), dist_out[nq] (exact SSD); optional cand_out[nq*8] (database
 *   patch indices y*(bw-7)+x of the survivors, best approximate score first) and
 *   score_out[nq*8] (|b|^2 - 2 a.b). stage_ms (optional, [3]): patch vectors, MFMA search,
 *   rescoring. Runs on stream 0 and returns when done.
 * pt_brute_exact_f64_dev: validation only. The exact nearest neighbour of each query patch in
 *   FP64 (sum over the 192 elements in row-major order, fused multiply-adds; ties to the smaller
 *   database index): pos_out[nq], dist_out[nq] (double). */
#define PT_BRUTE_BF16 0
#define PT_BRUTE_FP8  1
int pt_brute_lookup_dev(const float *a_dev, int ah, int aw, const float *b_dev, int bh, int bw,
                        const int32_t *qpos_dev, int nq, int precision, int32_t *pos_out_dev, float *dist_out_dev,
                        int32_t *cand_out_dev, float *score_out_dev, double *stage_ms);
int pt_brute_exact_f64_dev(const float *a_dev, int ah, int aw, const float *b_dev, int bh, int bw,
                           const int32_t *qpos_dev, int nq, int32_t *pos_out_dev, double *dist_out_dev);

/* ---- Learned patch key (no counterpart in the reference; an alternative to the GCK
 * descriptor as the table index). A 192 -> 64 -> 8 MLP with ReLU maps an 8x8 RGB patch,
 * x[dy*24 + dx*3 + c] = img(y+dy, x+dx, c) - 0.5, to an 8-dim key. Weights are BF16 for the
 * forward passes (FP32 masters for training), accumulation FP32. W1 is 64 x 192 row-major,
 * W2 8 x 64.
 * pt_key_create: He-uniform weights from `seed`, zero biases.
 * pt_key_set_weights / pt_key_get_weights: FP32 masters (host arrays); get also returns the
 *   BF16 copies (bit patterns; NULL to skip any output).
 * pt_key_forward_dev: key planes of an RGB image (device, h x w x 3) into planes_dev,
 *   8 planes of (h+7) x (w+7) floats, patch (x, y) at (y+7, x+7); other elements untouched.
 * pt_key_gradients: gradients of the triplet loss mean(max(0, |ka-kp|^2 - |ka-kn|^2 + margin))
 *   for nt triplets, pos_dev[3t + r] = XY_TO_INT(x, y) of the anchor (r=0, in a_dev) and the
 *   positive and negative (r=1, 2, in b_dev); both images h x w x 3, every position inside the
 *   patch grid (checked). Host outputs (NULL to skip).
 * pt_key_train: triplets mined from the library's own lookups of npairs (a_i, b_i) device
 *   image pairs with parameters p (tpp anchors per pair: positive = the match, negative = the
 *   match moved 2..6 px per axis or a random patch), then `steps` Adam steps of `batch`
 *   triplets; loss_per_step (optional, host, [steps]) receives each step's loss.
 * pt_table_create_keyed: a table whose 8 grid dimensions are the key of b (device image,
 *   h x w x 3), with the parameters p (its k-coherence lists are built in key space; distances
 *   stay exact RGB SSD) and, as its one further table, the GCK table of b built with p; the
 *   key network is copied, so `key` may change or go afterwards. Lookups use
 *   pt_table_lookup(_dev) with an RGB query; such a table has no file format (save fails). */
typedef struct pt_key pt_key;
int pt_key_create(uint64_t seed, pt_key **out);
int pt_key_destroy(pt_key *k);
int pt_key_set_weights(pt_key *k, const float *w1, const float *b1, const float *w2, const float *b2);
int pt_key_get_weights(const pt_key *k, float *w1, float *b1, float *w2, float *b2, uint16_t *w1_bf16,
                       uint16_t *w2_bf16);
int pt_key_forward_dev(const pt_key *k, const float *img_dev, int h, int w, float *planes_dev);
int pt_key_gradients(pt_key *k, const float *a_dev, const float *b_dev, int h, int w, const int32_t *pos_dev,
                     int nt, float margin, float *gW1, float *gb1, float *gW2, float *gb2, float *loss);
int pt_key_train(pt_key *k, pt_params *p, int npairs, const float *const *a_dev, const float *const *b_dev,
                 const int *h, const int *w, int tpp, int steps, int batch, float lr, float margin,
                 uint64_t seed, float *loss_per_step);
/* pt_key_train_exact: as pt_key_train, with each anchor's positive its exact nearest neighbour
 *   in b_i: the exhaustive search (pt_brute_lookup_dev, BF16 scoring, the 8 best rescored in
 *   exact FP32) instead of a table lookup. */
int pt_key_train_exact(pt_key *k, int npairs, const float *const *a_dev, const float *const *b_dev, const int *h,
                       const int *w, int tpp, int steps, int batch, float lr, float margin, uint64_t seed,
                       float *loss_per_step);
int pt_table_create_keyed(pt_params *p, const pt_key *key, const float *b_dev, int h, int w, pt_table **out);

/* ---- Errors ------------------------------------------------------------------------------ */
const char *pt_last_error(void);

/* ---- Introspection of a built table (for tests and the C++ header shim). ------------------
 * Shapes: descriptor planes are (h+7) x (w+7) x ndims (HWC, the reference's wh0 layout,
 * patchtable.h:1538); table is prod(nslices) int32 values XY_TO_INT(x,y) in padded coordinates
 * (patchtable.h:163-165, 3138); kcoherence is (h-7) x (w-7) x k pairs (XY_TO_INT position,
 * exact SSD) (patchtable.h:1512-1521). */
int pt_table_info(const pt_table *t, int *grid_ndims, int *nslices /* [PT_MAX_DIMS] */,
                  int64_t *ncells, int *table_add_count, int *kcoherence);
/* Database image height and width (b0) and descriptor dimensions of a built or loaded table. */
int pt_table_shape(const pt_table *t, int *h, int *w, int *ndims);
int pt_table_get_descriptor(const pt_table *t, float *wh0_out);
int pt_table_get_partition(const pt_table *t, int dim, float *min_val, float *max_val,
                           float *centers /* nslices[dim] */, int32_t *to_bin /* 16384 */);
int pt_table_get_cells(const pt_table *t, int32_t *table_out);
int pt_table_get_populated(const pt_table *t, int32_t *table_out);
int pt_table_get_kcoherence(const pt_table *t, int32_t *pos_out, float *dist_out);
/* The i-th further table of ntables > 1 (i = 0 .. ntables-2; patchtable.h:3359-3372). */
int pt_table_extra_info(const pt_table *t, int i, int *nslices /* [PT_MAX_DIMS] */, int64_t *ncells,
                        int *table_add_count);
int pt_table_get_extra_cells(const pt_table *t, int i, int32_t *table_out);
/* Query-side descriptor of the most recent lookup, (h+7) x (w+7) x ndims. */
int pt_table_get_query_descriptor(const pt_table *t, float *wh_out);
/* Per-stage device times (milliseconds) of the last create / lookup, from HIP events on the
 * table's stream: [0] gck, [1] quantiser, [2] populate, [3] dt, [4] kcoherence, [5] build
 * total, [6] query gck, [7] query candidate kernel, [8] query propagation rounds (all
 * launches), [9] query total, [10] number of propagation launches, [11] unused. */
int pt_table_stage_ms(const pt_table *t, double *ms /* [12] */);
/* Instrumentation of the MFMA lookup (mfma_rounds > 0), no counterpart in the reference: with
 * pt_table_set_mfma_stats(t, 1) every later synchronous lookup (pt_table_lookup*) also counts,
 * per round, the (group of 16 query patches, anchor) items the round scored, each 16 patches x
 * 96 candidates; pt_table_mfma_items copies up to n of the last lookup's per-round counts and
 * sets *rounds to their number (0 when nothing was counted). Counting adds one store per
 * workgroup and a copy of one int per workgroup per round back to the host. */
int pt_table_set_mfma_stats(pt_table *t, int on);
int pt_table_mfma_items(const pt_table *t, int64_t *items, int n, int *rounds);
/* Lattice tile of this table's plain lookup kernels: 64 x 4 (default), 32 x 8 or 16 x 16 patches
 * (tile_x 64, 32, 16); results do not depend on it. No counterpart in the reference. */
int pt_table_set_tile(pt_table *t, int tile_x);
/* Times lookups of this query (device, h x w x c, into ann_dev) with each tile width (one
 * warm-up and `reps` timed lookups each, device time from the table's events), keeps the
 * fastest (median) and returns it; ms_out[3] (optional): the medians for 64, 32 and 16. No
 * counterpart in the reference (an addition of this library). */
int pt_table_autotune(pt_table *t, const float *a_dev, int h, int w, int c, double *ann_dev, int reps,
                      int *tile_x_out, double *ms_out);

/* ---- Stand-alone kernels (device pointers, stream 0 of the current device) ----------------
 * Used by the parity tests and the benchmark to exercise one kernel at a time. */
/* GCK projection of an RGB (or YUV when convert_colorspace is 0) image: out is
 * (h+7) x (w+7) x (n+2*nchroma) HWC, as gck<float,float>(in,out,n,nchroma,8) (gck.h:1957). */
int pt_gck_dev(const float *img_dev, int h, int w, int n, int nchroma, int convert_yuv,
               float *out_dev);
/* Exact 8x8 RGB SSD of npairs patch pairs with the reference's lane order
 * (patch_dist_exact, patchtable.h:330-373). Patch i compares a at (ax[i],ay[i]) with b at
 * (bx[i],by[i]); images are HWC float32 with 3 channels. */
int pt_patch_ssd_dev(const float *a_dev, int aw, const float *b_dev, int bw,
                     const int32_t *ax, const int32_t *ay, const int32_t *bx, const int32_t *by,
                     int npairs, float *out_dev);
/* Descriptor-mode distance (patch_dist_approx, patchtable.h:376-437) of n descriptor pairs
 * a[i*C .. i*C+C-1], b[i*C ..] (device arrays), C >= 4, into out[n]. */
int pt_desc_dist_dev(const float *a_dev, const float *b_dev, int C, int n, float *out_dev);

/* ---- Seeded synthetic test images (the generator used by the tests and the benchmark) ----
 * a = 1/f^2 spectral noise plus 40 random coloured rectangles/ellipses; b = a warped by a
 * smooth random displacement plus Gaussian noise (sigma 0.02), clipped to [0,1]. Host memory,
 * h x w x 3 float32 each; deterministic for a given (seed, h, w). */
int pt_synth_pair(uint64_t seed, int h, int w, float *a, float *b);

#ifdef __cplusplus
}
#endif

#endif /* PATCHTABLE_AMD_H */
