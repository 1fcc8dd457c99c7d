// Host-side launchers of the HIP kernels (one .hip file each) and small device-memory helpers.
#ifndef PT_KERNELS_H
#define PT_KERNELS_H

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

#include "pt_internal.h"

// Grow-only device scratch buffers addressed by slot; freed with the owner.
struct PtScratch {
    std::vector<void *> ptr;
    std::vector<size_t> cap;
    void *get(int slot, size_t bytes) {
        if ((int)ptr.size() <= slot) {
            ptr.resize(slot + 1, nullptr);
            cap.resize(slot + 1, 0);
        }
        if (cap[slot] < bytes || !ptr[slot]) {
            if (ptr[slot]) (void)hipFree(ptr[slot]);
            ptr[slot] = nullptr;
            cap[slot] = 0;
            size_t b = bytes ? bytes : 16;
            if (hipMalloc(&ptr[slot], b) != hipSuccess) return nullptr;
            cap[slot] = b;
        }
        return ptr[slot];
    }
    void release() {
        for (void *p : ptr)
            if (p) (void)hipFree(p);
        ptr.clear();
        cap.clear();
    }
    ~PtScratch() { release(); }
};

// Owned device array, grown on demand (contents are not kept across a regrow).
template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    int ensure(size_t count) {
        if (count <= n && p) return PT_OK;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        const hipError_t e = hipMalloc((void **)&p, (count ? count : 1) * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            return pt_failf(PT_E_NOMEM, "hipMalloc of %zu bytes failed: %s", count * sizeof(T), hipGetErrorString(e));
        }
        n = count;
        return PT_OK;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
};

// Pinned host memory for small host<->device transfers that must not synchronise the stream
// (a pageable source or destination makes the copy wait). Sized once per use site: a regrow
// frees the old block, so it must not happen while a copy from or to it may be pending.
struct PinnedBuf {
    void *p = nullptr;
    size_t n = 0;
    int ensure(size_t bytes) {
        if (bytes <= n && p) return PT_OK;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        n = 0;
        const hipError_t e = hipHostMalloc(&p, bytes ? bytes : 16, hipHostMallocDefault);
        if (e != hipSuccess) {
            p = nullptr;
            return pt_failf(PT_E_NOMEM, "hipHostMalloc of %zu bytes failed: %s", bytes, hipGetErrorString(e));
        }
        n = bytes;
        return PT_OK;
    }
    template <class T>
    T *at(size_t byte_off) const {
        return (T *)((char *)p + byte_off);
    }
    PinnedBuf() = default;
    PinnedBuf(const PinnedBuf &) = delete;
    PinnedBuf &operator=(const PinnedBuf &) = delete;
    ~PinnedBuf() {
        if (p) (void)hipHostFree(p);
    }
};

// Quantiser state on the device: per grid dimension min, range, 16384-entry bin map, stride.
struct DevQuant {
    int G = 0;
    const float *min_val = nullptr;
    const float *range = nullptr;
    const int32_t *to_bin = nullptr;
    const int64_t *stride = nullptr;
};

// gck.hip
int pt_gck_supported(int n, int nchroma);
int pt_gck_launch(const float *img, int h, int w, int n, int nchroma, int convert_yuv, float *yuv_planes,
                  float *temp, float *desc, hipStream_t s);

// quant.hip
// allowed: NULL or the (h-7) x (w-7) device mask; only patches equal to aidx take part.
int pt_minmax_launch(const float *desc, size_t plane, int Wp, int h, int w, int G, const int32_t *allowed, int aidx,
                     float *partials, int nblocks, hipStream_t s);
// The same order statistics for every dimension d with nslices[d] > 0 at once (an exact radix
// select), into values_dev[d * ld + j]; needs nslices[d] + 1 <= PT_RS_SLOTS. ambig_dev[d] = 1
// where a value is a zero whose sign the select cannot tell (the dimension holds -0.0 and +0.0):
// such a dimension takes pt_quantile_values instead.
// The 16384-entry value->bin maps of G dimensions from their slice centres (centers[d * ldc + k],
// k < nslices[d]) and min/range (device arrays), into to_bin[d * PT_NMAP + i].
int pt_to_bin_launch(const float *centers, int ldc, const int *nslices, const float *min_val, const float *range, int G,
                     int32_t *to_bin, hipStream_t s);
constexpr int PT_RS_SLOTS = 48;
constexpr int PT_RS_MAX_WIDTH = 65535;   // partition grid columns the select handles
int pt_quantile_select(const float *desc, size_t plane, int Wp, int h, int w, int step, int G, const int *nslices,
                       const int32_t *allowed, int aidx, int64_t npart, int ld, float *values_dev, int *ambig_dev,
                       PtScratch &sc, hipStream_t s);
// The values at ranks (npart-1)*j/nslices, j = 0..nslices, of dimension `dim` over the partition
// grid (ascending order, masked patches last), written to values_dev[0..nslices].
int pt_quantile_values(const float *desc, size_t plane, int Wp, int h, int w, int step, int dim, const int32_t *allowed,
                       int aidx, int64_t npart, int nslices, float *values_dev, PtScratch &sc, hipStream_t s);

// populate.hip
int pt_list_launch(int bew, int n, int32_t *list, hipStream_t s);
int pt_shuffle_ranks(const uint32_t *jump_dev, const uint32_t *state_dev, int n, int populate_random, PtScratch &sc,
                     int32_t *rank, hipStream_t s);
// rand.cpp: glibc generator state before its first output and jump matrices A^(L*2^i).
void pt_glibc_initial_state(uint32_t seed, uint32_t state[31]);
const uint32_t *pt_glibc_jump_matrices(int L, int levels);
#define PT_RAND_CHUNK 256
#define PT_RAND_LEVELS 24
int pt_populate_launch(const float *desc, size_t plane, int Wp, const int32_t *list_xy, const int32_t *rank, int n,
                       const DevQuant &q, int32_t *table, int64_t ncells, int64_t *cell_scratch,
                       uint8_t *win_scratch, int *count_dev, hipStream_t s);
int pt_fill_launch(int32_t *p, int64_t n, int32_t v, hipStream_t s);

// dt.hip
// key: ncells + 1 eight-byte words of scratch (the keys, then a flag). flag_host (pinned, or
// null): receives the flag on the stream; when it reads nonzero after the stream has
// synchronised, pt_dt_prop_fallback_launch completes the fill (key must still hold the keys).
int pt_dt_prop_launch(int32_t *table, int G, const int *nslices, int64_t ncells, int dt_iters, int32_t fill,
                      uint64_t *key, hipStream_t s, int *flag_host);
int pt_dt_prop_fallback_launch(int32_t *table, int G, const int *nslices, int64_t ncells, int dt_iters, int32_t fill,
                               uint64_t *key, hipStream_t s);
int pt_fill_unused_launch(int32_t *table, int64_t ncells, int32_t fill, hipStream_t s);
int pt_dt_downsample_launch(int32_t *table, int G, const int *nslices, PtScratch &sc, hipStream_t s);
int pt_dt_euclidean_launch(int32_t *table, int G, const int *nslices, int64_t ncells, float *dist_scratch,
                           hipStream_t s);

// kcoh.hip
// b_descI: NULL in image mode (lists re-scored by exact RGB SSD on b0), else the interleaved
// descriptor (re-scored by patch_dist_approx). allowed: NULL or the (bh-7) x (bw-7) device mask.
int pt_kcoherence_launch(const float *desc, size_t plane, int Wp, int ndims, const float *b0, const float *b_descI,
                         int bw, int bh, const int32_t *allowed, int aidx, const DevQuant &q, const int *nslices_host,
                         int K, int window, int min_dist2, int32_t *kpos, float *kdist, PtScratch &sc, hipStream_t s);

// lookup.hip
#define PT_LOOKUP_CMAX 40    // descriptor dims supported in descriptor mode
#define PT_MAX_XTABLES 15    // further tables (ntables - 1)
#define PT_LOOKUP_HIST 16    // candidates remembered per lattice patch during a lookup
struct LookupArgs {
    const float *a = nullptr;       // query RGB, HWC (image mode)
    int ah = 0, aw = 0;             // query "image" size (descriptor size + 7 in descriptor mode)
    int ann_w = 0, ann_h = 0;       // query patches
    const float *a_desc = nullptr;  // query descriptor planes, (ah+7) x (aw+7) each
    const float *b = nullptr;       // database RGB, HWC (image mode)
    int bh = 0, bw = 0;
    const int32_t *table = nullptr;
    DevQuant q;
    int nxt = 0;                    // further tables
    const int32_t *xtable[PT_MAX_XTABLES] = {};
    DevQuant xq[PT_MAX_XTABLES];
    const float *xa_desc = nullptr; // query descriptor planes indexing the further tables (a_desc,
                                    // or the GCK planes of a learned-key table's fallback)
    const int32_t *kpos = nullptr;
    const float *kdist = nullptr;
    int K = 0;
    int kcoh_iter0 = 1;             // k-coherence walk in the candidate pass
    int kcoh_rounds = 0;            // ... and again in the propagation rounds
    double triangle_factor = 2.0;
    int spatial = 1;
    int spatial_cand = 0;           // spatial window in the candidate pass (no rounds)
    int spatial_from = 2;           // first round (1-based) with the spatial window
    int rs = 0;                     // random search (do_rs without spatial)
    int do_table_lookup = 1;
    int rounds = 4;
    int qs = 1;                     // query_step: lattice spacing of the processed patches
    int gqs = 1;                    // query_step seen by the gather pass
    int vw = 0, vh = 0;             // lattice size
    int C = 0;                      // descriptor mode: descriptor dims (0: image mode)
    const float *a_descI = nullptr; // descriptor mode: query descriptor, interleaved
    const float *b_descI = nullptr; // descriptor mode: database descriptor, interleaved
    const int32_t *allowed = nullptr;   // (bh-7) x (bw-7) mask of allowed database patches
    int allowed_index = 1;
    const int32_t *init_xy = nullptr;   // initial field from a previous one (per patch)
    const float *init_d = nullptr;
    int32_t *hist = nullptr;        // per lattice patch: candidates evaluated so far (16 x nv, entry-major)
    uint8_t *hn = nullptr;          // ... and how many were recorded
    int tile_x = 64;                // lattice tile of a workgroup: 64 x 4, 32 x 8 or 16 x 16 patches
                                    // (image mode; results do not depend on it)
    // MFMA form: the candidate pass scores by int8 SSD (image mode, query_step 1), from the int8
    // RGB0 query (ah x aw), the padded int8 copy of b and its norm keys; the pass writes |q|^2 of
    // every query patch to qnorm (read again by the rounds); a previous field's initial score is
    // its int8 SSD x (init_sv ? init_sv[i] : init_ts)
    int int8 = 0;
    const uint32_t *aq8 = nullptr, *bq8 = nullptr, *bnk = nullptr;
    const int32_t *qnorm = nullptr;
    const float *init_sv = nullptr;
    double init_ts = 1.0;
};
// One nearest-neighbour field over the lattice: XY_TO_INT(b patch x, y), distance, and flags
// (1: the spatial window around the match was searched without improvement; 2: the match changed
// in the round that produced the field; 4: the match's k-coherence list was walked; 8, MFMA form:
// the match is not the initial one and the distance is its int8 score, not yet the exact SSD).
struct NnfField {
    int32_t *xy = nullptr;
    float *d = nullptr;
    uint8_t *fl = nullptr;
};
// Jump of propagation round r (0-based) of `rounds`: 8, 4, 2, 1, then unit jumps (fewer rounds
// start lower: 3 rounds 4, 2, 1; 2 rounds 2, 1; 1 round 1).
static inline int pt_jump_at(int rounds, int r) {
    const int big = rounds >= 4 ? 8 : (rounds == 3 ? 4 : (rounds == 2 ? 2 : 1));
    const int j = big >> (r < 30 ? r : 30);
    return j > 1 ? j : 1;
}
// mfma.hip: the MFMA lookup rounds (dense int8 scoring of 16 queries x 32 x (2R+1) candidates
// per anchor offset, exact rescoring of each query's best).
#define PT_MFMA_BAND 1          // R: candidate rows above and below an anchor
#define PT_MFMA_MAX_ANCHORS 16
// Margin of zeros around the int8 copy of b and its norm table: every band a round reads lies
// inside it (anchor offsets come from matches inside b and sources within 16 + 15 columns and 16
// rows of the group), so the band loads need no bounds checks.
#define PT_MFMA_PAD 48
// Norm keys of the MFMA scoring: (|p|^2 + PT_MFMA_OFF) << 7 for a database patch p inside b (int8
// RGB0, |p|^2 <= PT_MFMA_OFF = 192 * 128^2), PT_MFMA_NK_PAD in the margins.
#define PT_MFMA_OFF (192u * 128u * 128u)
#define PT_MFMA_NK_PAD 0xC4000000u
#define PT_FIELD_MOVED 8
struct MfmaArgs {
    const float *a = nullptr, *b = nullptr;         // RGB float images (exact rescoring)
    const uint32_t *aq = nullptr;                   // int8 RGB0 query, ah x aw
    const uint32_t *bq = nullptr;                   // int8 RGB0 database, (bh + 2 PAD) x (bw + 2 PAD)
    const uint32_t *bnorm = nullptr;                // norm key of every database patch, (beh + 2 PAD) x (bew + 2 PAD)
    int ah = 0, aw = 0, bh = 0, bw = 0;             // image sizes
    int ann_w = 0, ann_h = 0;                       // query patches (query_step 1)
    int nanchors = 0;
    const int32_t *qnorm = nullptr;                 // |q|^2 of every query patch (int8 RGB0)
    // optional instrumentation: per round, the items (group, anchor) of every workgroup, one entry
    // per workgroup (round r at items_out + r * workgroups); null: not written
    int32_t *items_out = nullptr;
    // every 16-patch group's anchors in the last round that searched it: pt_mfma_anchor_words
    // packed offsets (x in the low 16 bits), PT_MFMA_NO_ANCHOR past the group's count. Each round
    // writes it; a round with `seen` set skips an anchor found there (see mfma_round_body).
    int32_t *anch = nullptr;
};
constexpr int32_t PT_MFMA_NO_ANCHOR = (int32_t)0x80008000u;   // no offset packs to this (|x| < 32768)
static inline size_t pt_mfma_anchor_words(int ann_w, int ann_h, int nanchors) {
    return (size_t)((ann_w + 15) / 16) * (size_t)ann_h * (size_t)nanchors;
}
// Workgroups of one MFMA round launch.
long pt_mfma_round_workgroups(int ann_w, int ann_h);
// Jump of MFMA round r of `rounds`: jumps[r] when the list (pt_params.mfma_jumps, 0-terminated)
// is given, 1 past its end; with an empty list 16, 8, 4, 2, 1, then unit jumps (fewer rounds
// start lower).
static inline int pt_mfma_jump_at(const int *jumps, int rounds, int r) {
    if (jumps && jumps[0] > 0) {
        for (int k = 0; k < PT_MFMA_MAX_JUMPS && jumps[k] > 0; k++)
            if (k == r) return jumps[k];
        return 1;
    }
    const int big = 1 << (rounds >= 5 ? 4 : (rounds >= 1 ? rounds - 1 : 0));
    const int j = big >> (r < 30 ? r : 30);
    return j > 1 ? j : 1;
}
int pt_quant_rgb0_launch(const float *img, long npx, uint32_t *out, hipStream_t s);

// The padded int8 copy of b and its patch norm keys (margins written by the caller: zeros in the
// copy, PT_MFMA_NK_PAD in the norm keys).
int pt_mfma_prepare_launch(const float *b, int h, int w, uint32_t *bq, uint32_t *bnorm, hipStream_t s);
// Runs `rounds` rounds from f0; *which = 1 if the result is in f1. seen_ok: every match's score
// is at most the int8 SSD of the match (no previous field, or one weighted by a temporal scale of
// at most 1 and no per-patch scales), so rounds after the first may skip the anchors a group
// searched in its last round (M.anch must then be set).
int pt_mfma_rounds_launch(const MfmaArgs &M, int rounds, const int *jumps, bool seen_ok, NnfField f0, NnfField f1,
                          hipStream_t s, int *which);
// One pair of a batch (one query against many tables), read by the batched kernels from a device
// array indexed by the pair: the lookup's arguments (L; G: the same for the query_step > 1 gather,
// whose lanes are single patches), its two fields, its gathered field and its NNF output.
struct PairLookup {
    LookupArgs L, G;
    MfmaArgs M;
    NnfField f[2];
    int32_t *gxy = nullptr;
    float *gd = nullptr;
    double *ann = nullptr;
};
// Candidates, propagation rounds, gather and output of every pair of a batch, one launch per phase
// (P_dev: device copy of the pairs; P0: host copy of pair 0, whose shape and parameters every pair
// shares): lookup.hip.
int pt_lookup_batch_launch(const PairLookup *P_dev, const PairLookup &P0, int npairs, hipStream_t s,
                           hipEvent_t ev_after_rounds, int *launches);
// The output phase alone, from field `src` of every pair.
int pt_lookup_batch_output(const PairLookup *P_dev, const PairLookup &P0, int npairs, int src, hipStream_t s,
                           hipEvent_t ev_before_write, int *launches);
// The table candidates of every pair (MFMA batches): field 0 of each pair.
int pt_lookup_batch_candidates(const PairLookup *P_dev, const PairLookup &P0, int npairs, hipStream_t s);
// MFMA rounds of every pair of a batch, one launch per round; *which = the field with the result.
int pt_mfma_rounds_batch_launch(const PairLookup *P_dev, const PairLookup &P0, int npairs, int rounds, const int *jumps,
                                hipStream_t s,
                                int *which);
// MFMA form's output (lookup.hip): NNF_X, NNF_Y and NNF_DIST as doubles, NNF_DIST the exact FP32
// SSD (the reference's lane order) of every match other than the initial one (field flag
// F_MOVED), the initial distance for a patch that kept its initial match, and the temporal
// rescaling of a previous field; the batch form writes every pair's output (no previous field).
int pt_lookup_exact_out(const LookupArgs &L, NnfField f, double *ann_out, const double *prev, double temporal_scale,
                        int temporal_fix, hipStream_t s);
int pt_lookup_batch_exact_out(const PairLookup *P_dev, const PairLookup &P0, int npairs, int src, hipStream_t s);
// Tile widths the lookup kernels are built for.
int pt_lookup_tile_ok(int tile_x);
// Output stage of a lookup (gather for query_step > 1, then the double field): lookup.hip.
int pt_lookup_output(const LookupArgs &L, NnfField res, int32_t *gxy, float *gd, double *ann_out, const double *prev,
                     const float *init_d, double temporal_scale, int temporal_fix, hipStream_t s, int *launches,
                     hipEvent_t ev_before_write = nullptr);
// Candidate pass only (tables, k-coherence walk per L) into f0: lookup.hip.
int pt_lookup_candidates(const LookupArgs &L, NnfField f0, hipStream_t s);

// Device buffers one lookup works in: the table's own (grown on demand), or carved from a
// caller's workspace (pt_lookup_workspace_layout).
struct LookupBufs {
    float *adesc = nullptr, *adescI = nullptr, *yuv = nullptr, *temp = nullptr;
    int32_t *nxy = nullptr, *nxy2 = nullptr, *gxy = nullptr, *ixy = nullptr, *hist = nullptr;
    float *nd = nullptr, *nd2 = nullptr, *gd = nullptr, *id = nullptr;
    uint8_t *nfl = nullptr, *nfl2 = nullptr, *hn = nullptr;
    int *bad = nullptr;
    uint32_t *aq = nullptr;   // int8 RGB0 query (MFMA lookup)
    int32_t *qnorm = nullptr; // ... and |q|^2 of its patches
    int32_t *anch = nullptr;  // ... and each group's anchors of its last searched round
};
struct pt_table;
// Enqueues one lookup on the table's stream (table.cpp). ws: the buffers to work in, or NULL for
// the table's own.
int pt_lookup_enqueue(pt_table *t, const float *a_dev, int h, int w, int c, double *ann_dev, const double *prev_dev,
                      const float *sv_dev, const float *adesc_pre, bool timing, int *launches_out,
                      const LookupBufs *ws = nullptr);
// Device the table was built on (its stream's device).
int pt_table_device(const pt_table *t);
// Bytes of the file-scope device array behind pt_table_lookup_dev_noalloc (noalloc.hip).
#define PT_NOALLOC_SCRATCH_BYTES PT_NOALLOC_MAX_BYTES
// Bytes of workspace a lookup of an h x w x c query needs (with a previous field or not), and,
// when base is not NULL, the buffers carved from base (256-byte aligned pieces).
int pt_lookup_workspace_layout(const pt_table *t, int h, int w, int c, bool prev, unsigned char *base,
                               size_t *bytes, LookupBufs *bufs);
// One lookup in a caller's workspace (no device allocation), synchronous like pt_table_lookup_dev.
int pt_lookup_in_workspace(pt_table *t, const float *a_dev, int h, int w, int c, double *ann_dev,
                           const double *prev_dev, const float *sv_dev, unsigned char *ws, size_t ws_bytes,
                           double *seconds);
int pt_lookup_init_prev(const LookupArgs &L, const double *prev, const float *sv, int recalc, double temporal_scale,
                        int sanitize, int32_t *ixy, float *id, int *bad_dev, hipStream_t s);
// Candidates, propagation rounds, gather (query_step > 1, into gxy/gd) and the double output.
int pt_lookup_launch(const LookupArgs &L, NnfField f0, NnfField f1, int32_t *gxy, float *gd, double *ann_out,
                     const double *prev, const float *init_d, double temporal_scale, int temporal_fix, hipStream_t s,
                     hipEvent_t ev_after_candidates, hipEvent_t ev_after_rounds, int *launches);
int pt_desc_pad_launch(const float *in, int h_in, int w_in, int C, int padded, int Hp, int Wp, float *planes,
                       float *inter, hipStream_t s);
int pt_ssd_pairs_launch(const float *a, int aw, const float *b, int bw, const int32_t *ax, const int32_t *ay,
                        const int32_t *bx, const int32_t *by, int n, float *out, hipStream_t s);
int pt_desc_dist_pairs_launch(const float *a, const float *b, int C, int n, float *out, hipStream_t s);

// key.hip: the learned patch key (192 -> 64 -> 8 MLP, BF16 weights, FP32 accumulate).
#define PT_KEY_IN 192
#define PT_KEY_HIDDEN 64
#define PT_KEY_OUT 8
#define PT_KEY_NPARAM (PT_KEY_HIDDEN * PT_KEY_IN + PT_KEY_HIDDEN + PT_KEY_OUT * PT_KEY_HIDDEN + PT_KEY_OUT)
struct KeyParamsDev {          // FP32 masters, BF16 copies, Adam moments (m, v: PT_KEY_NPARAM each)
    float *w1, *b1, *w2, *b2;
    unsigned short *w1bf, *w2bf;
    float *m, *v;
};
struct KeyWorkDev {            // per-batch training buffers (n = 3 x triplets samples)
    float *Z, *K, *dK, *dH;
    float *gW1, *gb1, *gW2, *gb2;
    float *loss;
    float *part;               // per-chunk partial gradients, pt_key_grad_part_floats(n) floats
};
int pt_key_grad_chunks(int nsamples);
size_t pt_key_grad_part_floats(int nsamples);
struct KeyNetDev {             // BF16 weights and FP32 biases of one key network
    const unsigned short *w1 = nullptr;
    const float *b1 = nullptr;
    const unsigned short *w2 = nullptr;
    const float *b2 = nullptr;
};
// brute.hip: exhaustive search (MFMA candidate scoring + exact rescoring; FP64 validation).
int pt_brute_vectors_launch(const float *img, int w, const int32_t *pos, long n, int ew, int fp8, unsigned char *vec,
                            float *norm, float *scale, hipStream_t s);
int pt_brute_mfma_launch(const unsigned char *qvec, const float *qscale, int nq, const unsigned char *dvec,
                         const float *dnorm, const float *dscale, int nd, int fp8, int32_t *cand, float *cscore,
                         hipStream_t s);
int pt_brute_rescore_launch(const float *a, int aw, const int32_t *qpos, int nq, int qew, const float *b, int bw,
                            int dew, const int32_t *cand, int32_t *out_pos, float *out_dist, hipStream_t s);
int pt_brute_f64_blocks(long nd, int *per_block);
int pt_brute_f64_launch(const float *a, int aw, const int32_t *qpos, int nq, int qew, const float *b, int bw, int dew,
                        long nd, double *pdist, int32_t *pidx, int32_t *out_pos, double *out_dist, hipStream_t s);

struct pt_key;
int pt_key_snapshot(const pt_key *k, unsigned short *w1, float *b1, unsigned short *w2, float *b2, hipStream_t s);
int pt_key_forward_launch(const float *img, int h, int w, const unsigned short *w1, const float *b1,
                          const unsigned short *w2, const float *b2, float *planes, hipStream_t s);
int pt_key_extract_launch(const float *a, const float *b, int w, const int32_t *pos, const int64_t *slot, int nt,
                          unsigned short *X, hipStream_t s);
// One training step on nt triplets (lr 0: gradients only).
int pt_key_train_step(const unsigned short *X, int nt, float margin, const KeyParamsDev &P, const KeyWorkDev &W,
                      float lr, float beta1, float beta2, float eps, int step, float wd, hipStream_t s);
int pt_key_to_bf16_launch(const float *in, int n, unsigned short *out, hipStream_t s);

#endif
