// The PatchTable object behind the C ABI: construction (PatchTableFixedN constructor,
// patchtable.h:2584-3373) and lookup (lookup_templated, patchtable.h:3468-4243) orchestrated on
// one HIP stream per table; all per-patch and per-cell work runs in the HIP kernels.
#include "pt_internal.h"
#include "pt_kernels.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>


#define TRY(expr)                          \
    do {                                   \
        int rc_ = (expr);                  \
        if (rc_ != PT_OK) return rc_;      \
    } while (0)

#define HIPTRY(expr)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            return pt_failf(PT_E_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));     \
    } while (0)

namespace {

struct HostPartition {   // PatchPartition1D fields (patchtable.h:1133-1141)
    std::vector<int32_t> to_bin;
    std::vector<float> centers, lsizes, rsizes;
    float min_val = 0, max_val = 0, range = 0;
    int nslices = 0;
};

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

struct pt_table {
    pt_params *p = nullptr;   // the caller's parameters, read at lookup time (patchtable.h:2602)
    pt_params build;          // parameters the table was built with
    hipStream_t stream = nullptr;
    int device = -1;          // the device current when the table was made (its stream's)
    hipEvent_t ev[12] = {};
    int h = 0, w = 0;         // "image" size: b, or the descriptor size + 7 in descriptor mode
    int ndims = 0, nchroma = 0, G = 0;
    bool desc_mode = false;   // is_descriptor: b is a descriptor image, distances are descriptor L2
    std::vector<int> nslices;
    std::vector<int64_t> stride;
    int64_t ncells = 0;
    std::vector<HostPartition> part;
    int table_add_count = 0;
    PinnedBuf pin;                   // staging of the build's small transfers
    bool pin_count_valid = false;    // the populate count was copied into pin
    bool dt_check = false;           // ... and the 32-bit DT's flag beside it
    int dt_iters = -1;
    hipStream_t aux = nullptr;       // build: populate's shuffle, beside the descriptor work
    bool shuffle_queued = false;     // ... enqueued there; ev[11] marks its end
    long rand_used = 0;       // rand() outputs consumed by this table's shuffle and the earlier ones
    int K = 0;
    double stage_ms[12] = {};
    std::vector<pt_table *> extra;   // the ntables-1 further tables (patchtable.h:3359-3372)

    DevBuf<float> b0, desc, descI;   // descI: interleaved descriptor (descriptor mode)
    DevBuf<int32_t> allowed;         // (h-7) x (w-7) mask, empty without one
    int allowed_index = 1;
    DevBuf<float> qmin, qrange;
    DevBuf<int32_t> qtobin;
    DevBuf<int32_t> qcent;   // slice centres and counts the device bin maps are made from
    DevBuf<int64_t> qstride;
    DevBuf<int32_t> table;
    DevBuf<int32_t> list_xy, rank;
    DevBuf<uint32_t> rnd;   // generator state + jump matrices
    DevBuf<int64_t> cells;
    DevBuf<uint8_t> win;
    DevBuf<int> counter;
    DevBuf<uint64_t> dtkey;
    DevBuf<int32_t> kpos;
    DevBuf<float> kdist;
    DevBuf<uint32_t> bq;      // MFMA lookup: int8 RGB0 copy of b0
    DevBuf<uint32_t> bnorm;   // ... and the norm key of every database patch
    bool mfma_stats = false;              // count the MFMA rounds' items (instrumentation)
    DevBuf<int32_t> mfma_items;           // per round, per workgroup
    std::vector<int64_t> mfma_items_sum;  // per round, of the last synchronous lookup
    int n_pop = 0;          // patches listed for populate (allowed ones)
    // query side
    DevBuf<float> a, adesc, adescI, yuv, temp;
    DevBuf<uint32_t> aq;
    DevBuf<int32_t> qnorm;   // MFMA lookup: |q|^2 of every query patch (int8 RGB0)
    DevBuf<int32_t> anch;    // ... and each 16-patch group's anchors of its last searched round
    DevBuf<int32_t> nxy, nxy2, gxy, ixy;
    DevBuf<float> nd, nd2, gd, id;
    DevBuf<uint8_t> nfl, nfl2;
    DevBuf<double> ann, prevbuf;
    DevBuf<float> svbuf;
    DevBuf<int> bad;
    DevBuf<int32_t> hist;   // lookup: candidates evaluated per lattice patch (PT_LOOKUP_HIST each)
    DevBuf<uint8_t> hn;
    int qh = 0, qw = 0;
    int tile_x = 64;        // lookup kernels' lattice tile: 64 x 4, 32 x 8 or 16 x 16 patches
    PtScratch sc;
    // learned-key table: the key network (copied at build) indexes the table; extra[0] is the GCK
    // fallback table, indexed by the query's GCK planes (xadesc)
    bool keyed = false;
    KeyNetDev knet;
    DevBuf<unsigned short> kw1, kw2;
    DevBuf<float> kb1, kb2, xadesc;

    DevQuant quant() const {
        DevQuant q;
        q.G = G;
        q.min_val = qmin.p;
        q.range = qrange.p;
        q.to_bin = qtobin.p;
        q.stride = qstride.p;
        return q;
    }
    ~pt_table() {
        for (pt_table *e : extra) delete e;
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
        if (aux) (void)hipStreamDestroy(aux);
    }
};

namespace {

int init_stream(pt_table *t) {
    if (hipGetDevice(&t->device) != hipSuccess) return pt_fail(PT_E_HIP, "hipGetDevice failed");
    if (hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking) != hipSuccess)
        return pt_fail(PT_E_HIP, "hipStreamCreate failed");
    for (int i = 0; i < 12; i++)
        if (hipEventCreate(&t->ev[i]) != hipSuccess) return pt_fail(PT_E_HIP, "hipEventCreate failed");
    return PT_OK;
}

// Descriptor sizes the reference instantiates (DECLARE_CONSTRUCTOR, patchtable.h:4405-4423),
// from 4 on (patch_dist_approx reads four values at once).
bool desc_dims_supported(int C) { return C >= 4 && C <= PT_LOOKUP_CMAX; }

int check_build_params(const pt_params *p, int c) {
    if (p->lookup_algo != PT_LOOKUP_ALGO_TABLE) return pt_fail(PT_E_UNSUPPORTED, "only lookup_algo table is supported");
    if (p->patch_w != 8) return pt_fail(PT_E_UNSUPPORTED, "gck: patch_w unsupported (only 8)");
    if (p->filter_dims_count) return pt_fail(PT_E_UNSUPPORTED, "filter_dims is not supported");
    if (p->ntables < 1 || p->ntables > PT_MAX_XTABLES + 1)
        return pt_failf(PT_E_UNSUPPORTED, "ntables must be in 1..%d", PT_MAX_XTABLES + 1);
    if (p->ntables > 1 && !(p->table_ratio > 0)) return pt_fail(PT_E_INVALID, "table_ratio must be > 0");
    if (p->kcoherence && p->kcoherence_algo != PT_KCOH_ALGO_FLANN)
        return pt_fail(PT_E_UNSUPPORTED, "only the default kcoherence_algo is supported");
    if (p->kcoherence_step != 1) return pt_fail(PT_E_UNSUPPORTED, "kcoherence_step > 1 is not supported");
    if (p->incoherence || p->kcoherence_enrich) return pt_fail(PT_E_UNSUPPORTED, "incoherence/kcoherence_enrich unsupported");
    if (p->randomize_dt) return pt_fail(PT_E_UNSUPPORTED, "randomize_dt is not supported");
    if (p->run_dt && !(p->dt_algo == PT_DT_ALGO_PROP || p->dt_algo == PT_DT_ALGO_DOWNSAMPLE ||
                       p->dt_algo == PT_DT_ALGO_EUCLIDEAN))
        return pt_fail(PT_E_UNSUPPORTED, "dt_algo must be prop, downsample or euclidean");
    if (p->regular_grid) return pt_fail(PT_E_UNSUPPORTED, "regular_grid is not supported");
    if (p->kcoherence < 0 || p->kcoherence > 32) return pt_fail(PT_E_UNSUPPORTED, "kcoherence must be in 0..32");
    if (p->partition_step < 1) return pt_fail(PT_E_INVALID, "partition_step must be >= 1");
    if (p->is_descriptor) {
        if (!desc_dims_supported(c))
            return pt_failf(PT_E_UNSUPPORTED, "descriptor mode: %d channels unsupported (4..%d)", c, PT_LOOKUP_CMAX);
        int gd = p->grid_ndims < 0 ? c : p->grid_ndims;
        if (gd < 1 || gd > c) return pt_fail(PT_E_INVALID, "invalid grid_ndims");
        return PT_OK;
    }
    if (p->dim_algo != PT_DIM_ALGO_WH) return pt_fail(PT_E_UNSUPPORTED, "only dim_algo wh is supported");
    if (p->convert_colorspace && p->colorspace != PT_COLORSPACE_YUV)
        return pt_fail(PT_E_UNSUPPORTED, "only the yuv and rgb colorspaces are supported");
    if (c != 3) return pt_fail(PT_E_INVALID, "expected 3 channel image as input to PatchTable");
    int gd = p->grid_ndims < 0 ? p->ndims : p->grid_ndims;
    if (p->ndims < 1 || p->ndims > PT_MAX_DIMS || gd < 1 || gd > p->ndims)
        return pt_fail(PT_E_INVALID, "invalid ndims/grid_ndims");
    if (!pt_gck_supported(p->ndims - 2 * p->nchroma, p->nchroma))
        return pt_failf(PT_E_UNSUPPORTED, "gck: n, nchroma unsupported: %d, %d", p->ndims - 2 * p->nchroma, p->nchroma);
    return PT_OK;
}

// scale_slices_to_limit (patchtable.cpp:611-642) with product()/multiply() (util.cpp:186-206).
std::vector<int> scale_slices(std::vector<int> ns, double limit) {
    auto product = [](const std::vector<int> &L) {
        double a = 1.0;
        for (int v : L) a *= v;
        return a;
    };
    if (limit > 0) {
        std::vector<int> ns0(ns);
        double factor = 1.0;
        for (auto &v : ns) v = std::max(v, 1);
        while (product(ns) > limit) {
            factor *= 0.995;
            for (size_t i = 0; i < ns.size(); i++) ns[i] = int(ns0[i] * factor + 0.5);
            for (auto &v : ns) v = std::max(v, 1);
        }
        for (;;) {
            ns0 = ns;
            bool ok = false;
            for (size_t i = 0; i < ns.size(); i++) {
                ns[i]++;
                if (product(ns) > limit) ns = ns0;
                else {
                    ok = true;
                    ns0 = ns;
                }
            }
            if (!ok) break;
        }
    }
    return ns;
}

// Centres and half sizes of one dimension (PatchPartition1D constructor, patchtable.h:1193-1231)
// from the order statistics q of the sorted values: q[i] is the value at rank (n-1)*i/nslices,
// i = 0..nslices. The value->bin map is made on the device (pt_to_bin_launch) or by host_to_bin.
void make_partition(HostPartition &P, const std::vector<float> &q) {
    const int ns = P.nslices;
    P.centers.resize(ns);
    P.lsizes.resize(ns);
    P.rsizes.resize(ns);
    for (int i = 0; i < ns; i++) {
        float cell_min = q[i], cell_max = q[i + 1];
        P.centers[i] = (cell_min + cell_max) / 2;
    }
    for (int i = 0; i < ns; i++) {
        float prev_center = i - 1 >= 0 ? P.centers[i - 1] : P.min_val;
        float next_center = i + 1 < ns ? P.centers[i + 1] : P.max_val;
        P.lsizes[i] = P.centers[i] - prev_center;
        if (i - 1 >= 0) P.lsizes[i] /= 2;
        P.rsizes[i] = next_center - P.centers[i];
        if (i + 1 < ns) P.rsizes[i] /= 2;
    }
    P.to_bin.clear();
}

// The 16384-entry value->bin map on the host (patchtable.h:1233-1251), for dimensions with more
// centres than the device path stages.
void host_to_bin(HostPartition &P) {
    const int ns = P.nslices;
    P.to_bin.resize(PT_NMAP);
    int icenter = 0;
    for (int i = 0; i < PT_NMAP; i++) {
        double coord = P.min_val + (i + 0.5) * P.range / (PT_NMAP);
        while (coord > P.centers[icenter]) {
            if (icenter + 1 < ns) icenter++;
            else break;
        }
        int prev_center = icenter - 1;
        if (prev_center < 0) prev_center = 0;
        int next_center = icenter;
        if (next_center == 0) next_center++;
        if (ns == 1) next_center = prev_center = 0;
        double mid = (P.centers[prev_center] + P.centers[next_center]) / 2.0;
        P.to_bin[i] = (coord < mid) ? prev_center : next_center;
    }
}

// Value->bin map of dimension d into out (PT_NMAP entries): the host copy where there is one
// (a loaded table), else the device's.
int fetch_to_bin(const pt_table *t, int d, int32_t *out) {
    const HostPartition &P = t->part[d];
    if ((int)P.to_bin.size() == PT_NMAP) {
        std::copy(P.to_bin.begin(), P.to_bin.end(), out);
        return PT_OK;
    }
    HIPTRY(hipStreamSynchronize(t->stream));
    HIPTRY(hipMemcpy(out, t->qtobin.p + (size_t)d * PT_NMAP, PT_NMAP * 4, hipMemcpyDeviceToHost));
    return PT_OK;
}

// Byte offsets of the quantiser upload in the table's pinned staging buffer: min, range and
// stride per dimension, then the bin maps; the populate counter lands after them.
size_t pin_upload_bytes(int G) { return (size_t)G * 16 + (size_t)G * PT_NMAP * 4; }

// Device copy of the quantiser (min, range, strides and bin maps), staged through pinned memory
// so that the copies do not wait for the stream. The bin maps are made on the device from the
// centres; a loaded table brings its maps along, and a dimension with PT_NMAP or more centres
// gets its map on the host. Called once per table: the staging buffer is not rewritten while
// these copies may still be pending.
int upload_quant(pt_table *t, hipStream_t s) {
    const int G = t->G;
    int ldc = 1;
    bool host_maps = true;
    for (int i = 0; i < G; i++) {
        ldc = std::max(ldc, t->part[i].nslices);
        host_maps &= (int)t->part[i].to_bin.size() == PT_NMAP;
    }
    if (!host_maps && ldc + 1 > PT_NMAP) {
        for (int i = 0; i < G; i++) host_to_bin(t->part[i]);
        host_maps = true;
    }
    TRY(t->pin.ensure(pin_upload_bytes(G) + 16));
    float *mn = t->pin.at<float>(0), *rg = t->pin.at<float>((size_t)G * 4);
    int64_t *st = t->pin.at<int64_t>((size_t)G * 8);
    int32_t *tb = t->pin.at<int32_t>((size_t)G * 16);   // G * PT_NMAP words
    for (int i = 0; i < G; i++) {
        const HostPartition &P = t->part[i];
        mn[i] = P.min_val;
        rg[i] = P.range;
        st[i] = t->stride[i];
        if (host_maps) {
            std::copy(P.to_bin.begin(), P.to_bin.end(), tb + (size_t)i * PT_NMAP);
        } else {   // centres [G][ldc], then the slice counts
            std::copy(P.centers.begin(), P.centers.end(), (float *)tb + (size_t)i * ldc);
            tb[(size_t)G * ldc + i] = P.nslices;
        }
    }
    TRY(t->qmin.ensure(G));
    TRY(t->qrange.ensure(G));
    TRY(t->qtobin.ensure((size_t)G * PT_NMAP));
    TRY(t->qstride.ensure(G));
    HIPTRY(hipMemcpyAsync(t->qmin.p, mn, G * 4, hipMemcpyHostToDevice, s));
    HIPTRY(hipMemcpyAsync(t->qrange.p, rg, G * 4, hipMemcpyHostToDevice, s));
    HIPTRY(hipMemcpyAsync(t->qstride.p, st, G * 8, hipMemcpyHostToDevice, s));
    if (host_maps) {
        HIPTRY(hipMemcpyAsync(t->qtobin.p, tb, (size_t)G * PT_NMAP * 4, hipMemcpyHostToDevice, s));
    } else {
        const size_t words = (size_t)G * ldc + G;
        TRY(t->qcent.ensure(words));
        HIPTRY(hipMemcpyAsync(t->qcent.p, tb, words * 4, hipMemcpyHostToDevice, s));
        TRY(pt_to_bin_launch((const float *)t->qcent.p, ldc, t->qcent.p + (size_t)G * ldc, t->qmin.p, t->qrange.p, G,
                             t->qtobin.p, s));
    }
    return PT_OK;
}

// After the build's final synchronisation: the populate count arrives in pinned memory; a DT whose
// 32-bit keys left cells empty is completed by the 64-bit passes; the DT work array (8 bytes per
// cell) is released (hipFree waits for the device, so not earlier).
int finish_grid(pt_table *t, hipStream_t s) {
    if (t->pin.p && t->pin_count_valid) t->table_add_count = *t->pin.at<int>(pin_upload_bytes(t->G));
    if (t->dt_check && t->pin.p && t->pin.at<int>(pin_upload_bytes(t->G))[1] != 0) {
        TRY(pt_dt_prop_fallback_launch(t->table.p, t->G, t->nslices.data(), t->ncells, t->dt_iters,
                                       PT_XY_TO_INT(7, 7), t->dtkey.p, s));
        HIPTRY(hipStreamSynchronize(s));
    }
    t->dt_check = false;
    t->dtkey.release();
    return PT_OK;
}

// The glibc generator state after `skip` outputs from the state of a fresh process.
void advanced_rand_state(long skip, uint32_t st[31]) {
    pt_glibc_initial_state(1, st);
    const uint32_t *jm = pt_glibc_jump_matrices(PT_RAND_CHUNK, PT_RAND_LEVELS);
    long q = skip / PT_RAND_CHUNK, r = skip % PT_RAND_CHUNK;
    for (int lv = 0; q; lv++, q >>= 1) {
        if (!(q & 1)) continue;
        const uint32_t *P = jm + (size_t)lv * 961;
        uint32_t nx[31];
        for (int i = 0; i < 31; i++) {
            uint32_t acc = 0;
            for (int k = 0; k < 31; k++) acc += P[i * 31 + k] * st[k];
            nx[i] = acc;
        }
        std::memcpy(st, nx, sizeof(nx));
    }
    for (long k = 0; k < r; k++) {
        uint32_t v = st[0] + st[28];
        std::memmove(st, st + 1, 30 * 4);
        st[30] = v;
    }
}

// Populate's random order (patchtable.h:3121-3159): the allowed patches in raster order and
// their shuffle ranks, on stream q. It depends only on the patch count and the rand() position,
// not on the descriptors or the quantiser.
int enqueue_shuffle(pt_table *t, const std::vector<int32_t> &mask_host, long rand_skip, hipStream_t q) {
    const pt_params &p = t->build;
    const int bew = t->w - 7, beh = t->h - 7, aidx = t->allowed_index;
    int n = bew * beh;
    TRY(t->list_xy.ensure(n));
    if (mask_host.empty()) {
        TRY(pt_list_launch(bew, n, t->list_xy.p, q));
    } else {
        std::vector<int32_t> lst;
        lst.reserve(n);
        for (int y = 0; y < beh; y++)
            for (int x = 0; x < bew; x++)
                if (mask_host[(size_t)y * bew + x] == aidx) lst.push_back(PT_XY_TO_INT(x + 7, y + 7));
        n = (int)lst.size();
        HIPTRY(hipMemcpyAsync(t->list_xy.p, lst.data(), lst.size() * 4, hipMemcpyHostToDevice, q));
        HIPTRY(hipStreamSynchronize(q));
    }
    t->n_pop = n;
    t->rand_used = rand_skip;
    if (p.populate_random) {
        if ((int64_t)n > ((int64_t)PT_RAND_CHUNK << PT_RAND_LEVELS))
            return pt_fail(PT_E_UNSUPPORTED, "too many database patches for the rand() jump tables");
        uint32_t st[31];
        advanced_rand_state(rand_skip, st);
        const uint32_t *jm = pt_glibc_jump_matrices(PT_RAND_CHUNK, PT_RAND_LEVELS);
        TRY(t->rnd.ensure(31 + (size_t)PT_RAND_LEVELS * 961));
        HIPTRY(hipMemcpyAsync(t->rnd.p, st, 31 * 4, hipMemcpyHostToDevice, q));
        HIPTRY(hipMemcpyAsync(t->rnd.p + 31, jm, (size_t)PT_RAND_LEVELS * 961 * 4, hipMemcpyHostToDevice, q));
        if (n > 1) t->rand_used += n - 1;
    }
    TRY(t->rank.ensure(n));
    return pt_shuffle_ranks(t->rnd.p + 31, t->rnd.p, n, p.populate_random, t->sc, t->rank.p, q);
}

// The main table's shuffle goes on its second stream before the descriptor work; the GPU runs
// it beside the GCK's sequential recurrences and the quantiser's host round trips.
int start_shuffle(pt_table *t, const std::vector<int32_t> &mask_host) {
    if (!t->aux && hipStreamCreateWithFlags(&t->aux, hipStreamNonBlocking) != hipSuccess)
        return pt_fail(PT_E_HIP, "hipStreamCreate failed");
    TRY(enqueue_shuffle(t, mask_host, 0, t->aux));
    HIPTRY(hipEventRecord(t->ev[11], t->aux));
    t->shuffle_queued = true;
    return PT_OK;
}

// Quantiser, populate and DT of table t (PatchTableFixedN constructor, patchtable.h:2567-3295)
// from the descriptor planes `desc` (owned by the main table). `mask_host` is the allowed mask
// (or empty); `rand_skip` the rand() outputs used by earlier tables' shuffles. Records events
// ev[1..4] of `evt` when given.
int build_grid(pt_table *t, const float *desc, const std::vector<int32_t> &mask_host, const int32_t *mask_dev,
               long rand_skip, hipStream_t s, const hipEvent_t *evt) {
    const pt_params &p = t->build;
    const int h = t->h, w = t->w, Wp = w + 7;
    const size_t plane = (size_t)(h + 7) * Wp;
    const int aidx = t->allowed_index;
    if (!t->shuffle_queued) TRY(enqueue_shuffle(t, mask_host, rand_skip, s));
    // quantiser: min/max (patchtable.h:1758-1778), slices, order statistics, to_bin
    const int G = t->G, NB = 512;
    float *partials = (float *)t->sc.get(40, (size_t)NB * G * 2 * 4);
    if (!partials) return pt_fail(PT_E_NOMEM, "min/max scratch allocation failed");
    TRY(pt_minmax_launch(desc, plane, Wp, h, w, G, mask_dev, aidx, partials, NB, s));
    // one pinned block for the build's small transfers (the quantiser upload is the largest)
    TRY(t->pin.ensure(std::max((size_t)NB * G * 2 * 4, pin_upload_bytes(G) + 16)));
    const float *hp = t->pin.at<float>(0);
    HIPTRY(hipMemcpyAsync(t->pin.p, partials, (size_t)NB * G * 2 * 4, hipMemcpyDeviceToHost, s));
    HIPTRY(hipStreamSynchronize(s));
    std::vector<float> mn(G, (float)1e20), mx(G, (float)-1e20), range(G);
    for (int bk = 0; bk < NB; bk++)
        for (int i = 0; i < G; i++) {
            float a = hp[((size_t)bk * G + i) * 2], b = hp[((size_t)bk * G + i) * 2 + 1];
            mn[i] = mn[i] < a ? mn[i] : a;
            mx[i] = mx[i] > b ? mx[i] : b;
        }
    for (int i = 0; i < G; i++) range[i] = mx[i] - mn[i];
    std::vector<int> ns(G);
    if (p.nslices_count == 0) {   // set_default_nslices (patchtable.h:1780-1810)
        float max_range = range[0];
        for (int i = 1; i < G; i++) max_range = std::max(max_range, range[i]);
        double scale = 200.0 / max_range;
        for (int i = 0; i < G; i++) ns[i] = max_range > 0 ? int(range[i] * scale) : 0;
    } else {
        if (p.nslices_count != G) return pt_fail(PT_E_INVALID, "expected nslices size to match GRID_NCHANNELS");
        for (int i = 0; i < G; i++) ns[i] = p.nslices[i];
    }
    t->nslices = scale_slices(ns, p.limit);
    t->stride.assign(G, 0);
    int64_t acc = 1;
    for (int i = G - 1; i >= 0; i--) {
        if (t->nslices[i] < 1) return pt_fail(PT_E_INVALID, "expected nslices >= 1");
        t->stride[i] = acc;
        acc *= t->nslices[i];
    }
    t->ncells = acc;
    if (acc > ((int64_t)1 << 31) - 1) return pt_fail(PT_E_UNSUPPORTED, "table larger than 2^31 cells");
    const int step = p.partition_step;
    const int bew = w - 7, beh = h - 7;
    int64_t npart = (int64_t)((bew + step - 1) / step) * ((beh + step - 1) / step);
    if (!mask_host.empty()) {   // allowed patches of the partition grid (patchtable.h:1357-1378)
        npart = 0;
        for (int y = 0; y < beh; y += step)
            for (int x = 0; x < bew; x += step) npart += mask_host[(size_t)y * bew + x] == aidx;
        if (npart == 0) return pt_fail(PT_E_INVALID, "no allowed patch on the partition grid");
    }
    t->part.assign(G, HostPartition());
    {
        // Every dimension's order statistics, one copy back: the radix select for dimensions of
        // up to PT_RS_SLOTS - 1 slices, a sort for the others and for a select that met zeros of
        // both signs at a rank.
        const size_t ld = (size_t)*std::max_element(t->nslices.begin(), t->nslices.end()) + 1;
        float *vals = (float *)t->sc.get(42, (size_t)G * ld * 4 + (size_t)G * 4);
        if (!vals) return pt_fail(PT_E_NOMEM, "quantile scratch allocation failed");
        int *amb = (int *)(vals + (size_t)G * ld);
        std::vector<int> sel(G, 0), sorted;
        const bool narrow = (bew + step - 1) / step <= PT_RS_MAX_WIDTH;
        for (int i = 0; i < G; i++) {
            if (narrow && t->nslices[i] + 1 <= PT_RS_SLOTS) sel[i] = t->nslices[i];
            else sorted.push_back(i);
        }
        TRY(pt_quantile_select(desc, plane, Wp, h, w, step, G, sel.data(), mask_dev, aidx, npart, (int)ld, vals, amb,
                               t->sc, s));
        for (int i : sorted)
            TRY(pt_quantile_values(desc, plane, Wp, h, w, step, i, mask_dev, aidx, npart, t->nslices[i], vals + i * ld,
                                   t->sc, s));
        // buffers of the populate and DT stages, allocated while the GPU runs the select
        TRY(t->table.ensure((size_t)t->ncells));
        TRY(t->cells.ensure(t->n_pop));
        TRY(t->win.ensure(t->n_pop));
        TRY(t->counter.ensure(1));
        if (p.run_dt && p.dt_algo == PT_DT_ALGO_PROP) TRY(t->dtkey.ensure((size_t)t->ncells + 1));
        const size_t back = (size_t)G * ld * 4 + (size_t)G * 4;
        TRY(t->pin.ensure(std::max(back, pin_upload_bytes(G) + 16)));
        HIPTRY(hipMemcpyAsync(t->pin.p, vals, back, hipMemcpyDeviceToHost, s));
        HIPTRY(hipStreamSynchronize(s));
        const int *hamb = t->pin.at<int>((size_t)G * ld * 4);
        std::vector<int> redo;
        for (int i = 0; i < G; i++)
            if (sel[i] > 0 && hamb[i]) redo.push_back(i);
        if (!redo.empty()) {
            for (int i : redo)
                TRY(pt_quantile_values(desc, plane, Wp, h, w, step, i, mask_dev, aidx, npart, t->nslices[i],
                                       vals + i * ld, t->sc, s));
            HIPTRY(hipMemcpyAsync(t->pin.p, vals, (size_t)G * ld * 4, hipMemcpyDeviceToHost, s));
            HIPTRY(hipStreamSynchronize(s));
        }
        for (int i = 0; i < G; i++) {
            HostPartition &P = t->part[i];
            P.nslices = t->nslices[i];
            P.min_val = mn[i];
            P.max_val = mx[i];
            P.range = mx[i] - mn[i];
            const float *q = t->pin.at<float>(i * ld * 4);
            make_partition(P, std::vector<float>(q, q + P.nslices + 1));
        }
    }
    TRY(upload_quant(t, s));
    if (evt) HIPTRY(hipEventRecord(evt[2], s));
    // populate (patchtable.h:3121-3159): the shuffled allowed patches claim their cells
    if (t->shuffle_queued) HIPTRY(hipStreamWaitEvent(s, t->ev[11], 0));
    const int n = t->n_pop;
    TRY(t->table.ensure((size_t)t->ncells));
    TRY(t->cells.ensure(n));
    TRY(t->win.ensure(n));
    TRY(t->counter.ensure(1));
    TRY(pt_populate_launch(desc, plane, Wp, t->list_xy.p, t->rank.p, n, t->quant(), t->table.p, t->ncells, t->cells.p,
                           t->win.p, t->counter.p, s));
    // read by finish_grid() once the build has synchronised
    HIPTRY(hipMemcpyAsync(t->pin.at<int>(pin_upload_bytes(G)), t->counter.p, sizeof(int), hipMemcpyDeviceToHost, s));
    t->pin_count_valid = true;
    if (evt) HIPTRY(hipEventRecord(evt[3], s));
    // distance transform (patchtable.h:3254-3295) and fill_zeros (:2063-2079)
    const int32_t vfill = PT_XY_TO_INT(7, 7);
    if (p.run_dt && p.dt_algo == PT_DT_ALGO_PROP) {
        TRY(t->dtkey.ensure((size_t)t->ncells + 1));
        int *flag_host = t->pin.at<int>(pin_upload_bytes(G)) + 1;   // beside the populate count
        TRY(pt_dt_prop_launch(t->table.p, G, t->nslices.data(), t->ncells, p.dt_iters, vfill, t->dtkey.p, s,
                              flag_host));
        t->dt_check = true;
        t->dt_iters = p.dt_iters;
    } else if (p.run_dt && p.dt_algo == PT_DT_ALGO_DOWNSAMPLE) {
        TRY(pt_dt_downsample_launch(t->table.p, G, t->nslices.data(), t->sc, s));
    } else if (p.run_dt && p.dt_algo == PT_DT_ALGO_EUCLIDEAN) {
        DevBuf<float> dist;
        TRY(dist.ensure((size_t)t->ncells));
        TRY(pt_dt_euclidean_launch(t->table.p, G, t->nslices.data(), t->ncells, dist.p, s));
        HIPTRY(hipStreamSynchronize(s));
    } else {
        // run_dt off: fill_zeros replaces every empty cell (patchtable.h:3289-3293)
        TRY(pt_fill_unused_launch(t->table.p, t->ncells, vfill, s));
    }
    if (evt) HIPTRY(hipEventRecord(evt[4], s));
    return PT_OK;
}

// PatchTableFixedN constructor (patchtable.h:2584-3373) from a device image (RGB, or a
// descriptor image in descriptor mode) of h_in x w_in x c.
// The MFMA lookup's view of b: an int8 RGB0 copy of b0 and the norm key of every patch, both
// inside margins of PT_MFMA_PAD (zeros in the copy, PT_MFMA_NK_PAD in the norm keys).
int prepare_mfma(pt_table *t, hipStream_t s) {
    const size_t nq = (size_t)(t->h + 2 * PT_MFMA_PAD) * (t->w + 2 * PT_MFMA_PAD);
    const size_t nn = (size_t)(t->h - 7 + 2 * PT_MFMA_PAD) * (t->w - 7 + 2 * PT_MFMA_PAD);
    TRY(t->bq.ensure(nq));
    TRY(t->bnorm.ensure(nn));
    HIPTRY(hipMemsetAsync(t->bq.p, 0, nq * 4, s));      // the margins stay zero
    TRY(pt_fill_launch((int32_t *)t->bnorm.p, (int64_t)nn, (int32_t)PT_MFMA_NK_PAD, s));
    TRY(pt_mfma_prepare_launch(t->b0.p, t->h, t->w, t->bq.p, t->bnorm.p, s));
    return PT_OK;
}

int build_table(pt_table *t, const float *b_dev, int h_in, int w_in, int c, const std::vector<int32_t> &mask_host) {
    const pt_params &p = t->build;
    hipStream_t s = t->stream;
    const int h = t->h, w = t->w, Hp = h + 7, Wp = w + 7;
    const size_t plane = (size_t)Hp * Wp;
    HIPTRY(hipEventRecord(t->ev[0], s));
    // 1. descriptor (get_padded_descriptor, patchtable.h:2492-2546) and the copy of b (:2628)
    TRY(t->desc.ensure(plane * t->ndims));
    if (t->desc_mode) {
        TRY(t->descI.ensure(plane * t->ndims));
        TRY(pt_desc_pad_launch(b_dev, h_in, w_in, c, p.descriptor_padded, Hp, Wp, t->desc.p, t->descI.p, s));
    } else {
        TRY(t->b0.ensure((size_t)h * w * 3));
        HIPTRY(hipMemcpyAsync(t->b0.p, b_dev, (size_t)h * w * 3 * 4, hipMemcpyDeviceToDevice, s));
        TRY(t->yuv.ensure((size_t)h * w * 3));
        TRY(t->temp.ensure(plane * 3));
        TRY(pt_gck_launch(t->b0.p, h, w, t->ndims - 2 * t->nchroma, t->nchroma, p.convert_colorspace, t->yuv.p,
                          t->temp.p, t->desc.p, s));
    }
    HIPTRY(hipEventRecord(t->ev[1], s));
    TRY(start_shuffle(t, mask_host));   // its allocations overlap the descriptor kernels
    if (!mask_host.empty()) {
        TRY(t->allowed.ensure(mask_host.size()));
        HIPTRY(hipMemcpyAsync(t->allowed.p, mask_host.data(), mask_host.size() * 4, hipMemcpyHostToDevice, s));
    }
    const int32_t *mask_dev = mask_host.empty() ? nullptr : t->allowed.p;
    // 2-4. quantiser, populate, DT
    TRY(build_grid(t, t->desc.p, mask_host, mask_dev, 0, s, t->ev));
    // 5. k-coherence (patchtable.h:2653-2987); the MFMA lookup walks no lists and takes the int8
    // copy of b instead
    t->K = p.mfma_rounds > 0 && !t->desc_mode ? 0 : p.kcoherence;
    if (p.mfma_rounds > 0 && !t->desc_mode) TRY(prepare_mfma(t, s));
    if (t->K > 0) {
        const size_t n = (size_t)(w - 7) * (h - 7);
        TRY(t->kpos.ensure(n * t->K));
        TRY(t->kdist.ensure(n * t->K));
        TRY(pt_kcoherence_launch(t->desc.p, plane, Wp, t->ndims, t->b0.p, t->desc_mode ? t->descI.p : nullptr, w, h,
                                 mask_dev, t->allowed_index, t->quant(), t->nslices.data(), t->K, p.kcoh_window,
                                 p.kcoherence_min_dist * p.kcoherence_min_dist, t->kpos.p, t->kdist.p, t->sc, s));
    }
    // 6. further tables (patchtable.h:3359-3372): limit / table_ratio^i, no k-coherence; their
    // shuffles continue the rand() stream after the earlier tables'.
    long used = t->rand_used;
    for (int i = 1; i < p.ntables; i++) {
        pt_table *E = new pt_table();
        t->extra.push_back(E);
        E->build = p;
        E->build.limit = (double)(int)(p.limit * (1.0 / pow(p.table_ratio, i)));
        E->build.kcoherence = 0;
        E->build.ntables = 1;
        E->h = h;
        E->w = w;
        E->ndims = t->ndims;
        E->nchroma = t->nchroma;
        E->G = t->G;
        E->desc_mode = t->desc_mode;
        E->allowed_index = t->allowed_index;
        TRY(build_grid(E, t->desc.p, mask_host, mask_dev, used, s, nullptr));
        used = E->rand_used;
        HIPTRY(hipStreamSynchronize(s));
        TRY(finish_grid(E, s));
        E->list_xy.release();
        E->rank.release();
        E->cells.release();
        E->win.release();
        E->rnd.release();
        E->sc.release();
    }
    HIPTRY(hipEventRecord(t->ev[5], s));
    HIPTRY(hipStreamSynchronize(s));
    TRY(finish_grid(t, s));
    float ms = 0;
    for (int k = 0; k < 5; k++) {
        HIPTRY(hipEventElapsedTime(&ms, t->ev[k], t->ev[k + 1]));
        t->stage_ms[k] = ms;
    }
    HIPTRY(hipEventElapsedTime(&ms, t->ev[0], t->ev[5]));
    t->stage_ms[5] = ms;
    t->rnd.release();
    return PT_OK;
}

int create_common(pt_params *p, const float *b, bool b_on_device, int h, int w, int c, const int32_t *allowed,
                  pt_table **out) {
    if (!p || !out) return pt_fail(PT_E_INVALID, "pt_table_create: null argument");
    *out = nullptr;
    if (p->load_filename[0]) return pt_table_load(p, p->load_filename, out);
    if (!b) return pt_fail(PT_E_INVALID, "pt_table_create: null image");
    TRY(check_build_params(p, c));
    int vh = h, vw = w;   // "image" size
    if (p->is_descriptor) {
        vh = p->descriptor_padded ? h - 7 : h + 7;
        vw = p->descriptor_padded ? w - 7 : w + 7;
    }
    if (vh < 8 || vw < 8) return pt_fail(PT_E_INVALID, "database image must be at least 8x8 (descriptor: 1x1)");
    if (!PT_SIZE_OK(vh, vw))
        return pt_failf(PT_E_INVALID, "image too large for XY_TO_INT packing (height <= %d, width <= %d)", PT_MAX_H, PT_MAX_W);
    if (p->is_descriptor) {   // patchtable.h:4398-4401, written back as the reference does
        p->ndims = c;
        p->nchroma = 0;
    }
    if (p->grid_ndims < 0) p->grid_ndims = p->ndims;   // patchtable.h:2603
    std::vector<int32_t> mask;
    if (allowed) {
        mask.assign(allowed, allowed + (size_t)(vh - 7) * (vw - 7));
        bool any = false;
        for (int32_t v : mask) any |= v == p->allowed_index;
        if (!any) return pt_fail(PT_E_INVALID, "allowed_patches allows no patch");
    }
    pt_table *t = new pt_table();
    t->p = p;
    t->build = *p;
    t->h = vh;
    t->w = vw;
    t->ndims = p->ndims;
    t->nchroma = p->nchroma;
    t->G = p->grid_ndims;
    t->desc_mode = p->is_descriptor != 0;
    t->allowed_index = p->allowed_index;
    int rc = init_stream(t);
    DevBuf<float> staging;
    const float *bd = b;
    if (rc == PT_OK && !b_on_device) {
        rc = staging.ensure((size_t)h * w * c);
        if (rc == PT_OK && hipMemcpy(staging.p, b, (size_t)h * w * c * 4, hipMemcpyHostToDevice) != hipSuccess)
            rc = pt_fail(PT_E_HIP, "hipMemcpy of the database image failed");
        bd = staging.p;
    }
    if (rc == PT_OK) rc = build_table(t, bd, h, w, c, mask);
    if (rc == PT_OK && p->save_filename[0]) rc = pt_table_save(t, p->save_filename);
    if (rc != PT_OK) {
        delete t;
        return rc;
    }
    *out = t;
    return PT_OK;
}

// Query dims of a lookup input of h x w x c: "image" size and patch counts.
int query_shape(const pt_table *t, int h, int w, int c, int &ah, int &aw) {
    if (t->desc_mode) {
        if (c != t->ndims) return pt_failf(PT_E_INVALID, "expected a %d channel query descriptor", t->ndims);
        ah = t->build.descriptor_padded ? h - 7 : h + 7;
        aw = t->build.descriptor_padded ? w - 7 : w + 7;
    } else {
        if (c != 3) return pt_fail(PT_E_INVALID, "expected 3 channel query image");
        ah = h;
        aw = w;
    }
    if (ah < 8 || aw < 8) return pt_fail(PT_E_INVALID, "query image must be at least 8x8 (descriptor: 1x1)");
    if (!PT_SIZE_OK(ah, aw))
        return pt_failf(PT_E_INVALID, "query image too large (height <= %d, width <= %d)", PT_MAX_H, PT_MAX_W);
    return PT_OK;
}

}  // namespace

// Enqueues one lookup on the table's stream. adesc_pre: the query descriptor planes when the
// caller already computed them (same parameters), else NULL. timing: record the stage events
// (off while a graph is being captured). ws: buffers to work in (no allocation), else the
// table's own, grown as needed.
namespace {

// Checks a lookup of an h x w x c query on table t with parameters p (prev: a previous field is
// given; ws: in a caller's workspace) and sets the query "image" size; makes the table's int8
// copy of b when an MFMA lookup needs it first.
int check_lookup(pt_table *t, const pt_params &p, int h, int w, int c, bool prev, const LookupBufs *ws, int &ah,
                 int &aw) {
    TRY(query_shape(t, h, w, c, ah, aw));
    if (!p.do_table_lookup && !prev)
        return pt_fail(PT_E_UNSUPPORTED, "do_table_lookup 0 requires a previous field (ann_prev)");
    if (p.init_random) return pt_fail(PT_E_UNSUPPORTED, "init_random is not supported");
    if (p.query_step < 1) return pt_fail(PT_E_INVALID, "query_step must be >= 1");
    if (p.prop_dist != 0) return pt_fail(PT_E_UNSUPPORTED, "prop_dist != 0 is not supported");
    if (p.prop_iters < 1) return pt_fail(PT_E_UNSUPPORTED, "prop_iters must be >= 1");
    if (p.coherence_spatial != 0.0) return pt_fail(PT_E_UNSUPPORTED, "coherence_spatial != 0 is not supported");
    if (p.kcoherence > t->K && p.mfma_rounds <= 0)
        return pt_fail(PT_E_INVALID, "kcoherence exceeds the table's k-coherence lists");
    if (p.spatial < 0) return pt_fail(PT_E_INVALID, "spatial must be >= 0");
    if (ws && t->keyed) return pt_fail(PT_E_UNSUPPORTED, "learned-key tables look up in their own buffers");
    const bool mfma = p.mfma_rounds > 0;
    if (mfma) {
        if (t->desc_mode) return pt_fail(PT_E_UNSUPPORTED, "the MFMA lookup scores RGB patches (is_descriptor 0)");
        if (p.query_step != 1) return pt_fail(PT_E_UNSUPPORTED, "the MFMA lookup needs query_step 1");
        if (t->allowed.p) return pt_fail(PT_E_UNSUPPORTED, "the MFMA lookup takes no allowed-patch mask");
        if (t->w > 32768 || w > 32768)
            return pt_fail(PT_E_UNSUPPORTED, "the MFMA lookup packs offsets in 16 bits: widths up to 32768");
        if (p.mfma_anchors < 1 || p.mfma_anchors > PT_MFMA_MAX_ANCHORS)
            return pt_failf(PT_E_INVALID, "mfma_anchors must be in 1..%d", PT_MFMA_MAX_ANCHORS);
        if (!t->bq.p) {
            if (ws) return pt_fail(PT_E_INVALID, "the table has no int8 copy of b: build it with mfma_rounds > 0");
            TRY(prepare_mfma(t, t->stream));
        }
    }
    return PT_OK;
}

// The lookup arguments every form shares: the table(s), the parameters of the candidate pass and
// the rounds, the lattice, the mask. The caller sets the query images (a / descriptors).
void fill_lookup_args(const pt_table *t, const pt_params &p, int ah, int aw, const float *adesc, const float *xadesc,
                      const LookupBufs &B, LookupArgs &L) {
    const int qs = p.query_step;
    const int ann_w = aw - 7, ann_h = ah - 7;
    const bool mfma = p.mfma_rounds > 0;
    L.ah = ah;
    L.aw = aw;
    L.ann_w = ann_w;
    L.ann_h = ann_h;
    L.a_desc = adesc;
    L.xa_desc = xadesc;
    L.bh = t->h;
    L.bw = t->w;
    L.table = t->table.p;
    L.q = t->quant();
    L.nxt = (int)t->extra.size();
    for (int i = 0; i < L.nxt; i++) {
        L.xtable[i] = t->extra[i]->table.p;
        L.xq[i] = t->extra[i]->quant();
    }
    L.kpos = t->kpos.p;
    L.kdist = t->kdist.p;
    L.hist = B.hist;
    L.hn = B.hn;
    L.K = p.kcoherence;
    L.kcoh_iter0 = p.kcoherence > 0 && (p.kcoherence_iter < 0 || p.kcoherence_iter == 0);
    // k-coherence walks in the rounds when kcoherence_iter asks for them after the first pass, and
    // without table candidates (the walk then has to follow the propagated matches).
    L.kcoh_rounds = p.kcoherence > 0 && (p.kcoherence_iter < 0 || (p.kcoherence_iter >= 1 && p.kcoherence_iter < p.prop_iters) ||
                                         !p.do_table_lookup);
    L.triangle_factor = p.triangle_factor;
    L.spatial = p.spatial;
    L.rs = !p.spatial && p.do_rs;
    L.do_table_lookup = p.do_table_lookup;
    L.rounds = p.do_prop ? p.jf_rounds : 0;
    L.spatial_cand = L.rounds == 0;
    L.spatial_from = L.rounds >= 2 ? 2 : 1;
    L.qs = qs;
    L.vw = (ann_w + qs - 1) / qs;
    L.vh = (ann_h + qs - 1) / qs;
    L.tile_x = t->tile_x;
    if (t->allowed.p) {
        L.allowed = t->allowed.p;
        L.allowed_index = p.allowed_index;
    }
    if (mfma) {   // candidates: the cells of every table only; then the MFMA rounds
        L.K = 0;
        L.kcoh_iter0 = false;
        L.kcoh_rounds = false;
        L.spatial_cand = false;
        L.rs = false;
        L.rounds = 0;
        L.hist = nullptr;
        L.hn = nullptr;
    }
}

// The MFMA form's candidate pass scores by int8 SSD: the int8 query, the table's int8 copy of b
// and norm keys, the query norms; a previous field's initial score is its int8 SSD times the
// temporal weight (sv per patch, else ts).
void set_int8_args(const pt_table *t, const uint32_t *aq, const int32_t *qnorm, const float *sv, double ts,
                   LookupArgs &L) {
    L.int8 = 1;
    L.aq8 = aq;
    L.bq8 = t->bq.p;
    L.bnk = t->bnorm.p;
    L.qnorm = qnorm;
    L.init_sv = sv;
    L.init_ts = ts;
}

void fill_mfma_args(const pt_table *t, const pt_params &p, const float *a_dev, const uint32_t *aq, int ah, int aw,
                    MfmaArgs &M) {
    M.a = a_dev;
    M.b = t->b0.p;
    M.aq = aq;
    M.bq = t->bq.p;
    M.bnorm = t->bnorm.p;
    M.ah = ah;
    M.aw = aw;
    M.bh = t->h;
    M.bw = t->w;
    M.ann_w = aw - 7;
    M.ann_h = ah - 7;
    M.nanchors = p.mfma_anchors;
}

}  // namespace

int pt_lookup_enqueue(pt_table *t, const float *a_dev, int h, int w, int c, double *ann_dev, const double *prev_dev,
                      const float *sv_dev, const float *adesc_pre, bool timing, int *launches_out,
                      const LookupBufs *ws) {
    const pt_params &p = *t->p;
    int ah = 0, aw = 0;
    TRY(check_lookup(t, p, h, w, c, prev_dev != nullptr, ws, ah, aw));
    const bool mfma = p.mfma_rounds > 0;
    hipStream_t s = t->stream;
    const int Hp = ah + 7, Wp = aw + 7;
    const size_t plane = (size_t)Hp * Wp;
    const int ann_w = aw - 7, ann_h = ah - 7;
    const long n = (long)ann_w * ann_h;
    const int qs = p.query_step;
    const int vw = (ann_w + qs - 1) / qs, vh = (ann_h + qs - 1) / qs;
    const long nv = (long)vw * vh;
    LookupBufs B;
    if (ws) {
        B = *ws;
    } else {
        TRY(t->adesc.ensure(plane * t->ndims));
        TRY(t->nxy.ensure(nv));
        TRY(t->nd.ensure(nv));
        TRY(t->nfl.ensure(nv));
        TRY(t->nxy2.ensure(nv));
        TRY(t->nd2.ensure(nv));
        TRY(t->nfl2.ensure(nv));
        TRY(t->hist.ensure((size_t)nv * PT_LOOKUP_HIST));
        TRY(t->hn.ensure(nv));
        if (qs > 1) {
            TRY(t->gxy.ensure(n));
            TRY(t->gd.ensure(n));
        }
        if (t->desc_mode) TRY(t->adescI.ensure(plane * t->ndims));
        if (!t->desc_mode) {
            TRY(t->yuv.ensure((size_t)ah * aw * 3));
            TRY(t->temp.ensure(plane * 3));
        }
        if (prev_dev) {
            TRY(t->ixy.ensure(n));
            TRY(t->id.ensure(n));
            TRY(t->bad.ensure(1));
        }
        if (mfma) {
            TRY(t->aq.ensure((size_t)ah * aw));
            TRY(t->qnorm.ensure(n));
            TRY(t->anch.ensure(pt_mfma_anchor_words(ann_w, ann_h, p.mfma_anchors)));
        }
        B.aq = t->aq.p;
        B.qnorm = t->qnorm.p;
        B.anch = t->anch.p;
        B.adesc = t->adesc.p;
        B.adescI = t->adescI.p;
        B.yuv = t->yuv.p;
        B.temp = t->temp.p;
        B.nxy = t->nxy.p;
        B.nxy2 = t->nxy2.p;
        B.gxy = t->gxy.p;
        B.ixy = t->ixy.p;
        B.hist = t->hist.p;
        B.nd = t->nd.p;
        B.nd2 = t->nd2.p;
        B.gd = t->gd.p;
        B.id = t->id.p;
        B.nfl = t->nfl.p;
        B.nfl2 = t->nfl2.p;
        B.hn = t->hn.p;
        B.bad = t->bad.p;
    }
    if (!ws) {   // what pt_table_get_query_descriptor reads back
        t->qh = ah;
        t->qw = aw;
    }
    if (timing) HIPTRY(hipEventRecord(t->ev[6], s));
    LookupArgs L;
    if (t->keyed) {
        // the key planes index the main table, the GCK planes the fallback table
        const pt_table *E = t->extra[0];
        HIPTRY(hipMemsetAsync(B.adesc, 0, plane * t->ndims * sizeof(float), s));
        TRY(pt_key_forward_launch(a_dev, ah, aw, t->knet.w1, t->knet.b1, t->knet.w2, t->knet.b2, B.adesc, s));
        TRY(t->xadesc.ensure(plane * E->ndims));
        TRY(pt_gck_launch(a_dev, ah, aw, E->ndims - 2 * E->nchroma, E->nchroma, E->build.convert_colorspace, B.yuv,
                          B.temp, t->xadesc.p, s));
        L.a = a_dev;
        L.b = t->b0.p;
    } else if (adesc_pre && !t->desc_mode) {
        L.a = a_dev;
        L.b = t->b0.p;
    } else if (t->desc_mode) {
        TRY(pt_desc_pad_launch(a_dev, h, w, c, t->build.descriptor_padded, Hp, Wp, B.adesc, B.adescI, s));
        L.C = t->ndims;
        L.a_descI = B.adescI;
        L.b_descI = t->descI.p;
    } else {
        TRY(pt_gck_launch(a_dev, ah, aw, t->ndims - 2 * t->nchroma, t->nchroma, t->build.convert_colorspace, B.yuv,
                          B.temp, B.adesc, s));
        L.a = a_dev;
        L.b = t->b0.p;
    }
    {
        const float *adesc = adesc_pre && !t->desc_mode && !t->keyed ? adesc_pre : B.adesc;
        fill_lookup_args(t, p, ah, aw, adesc, t->keyed ? t->xadesc.p : adesc, B, L);
    }
    const double temporal_scale = 1.0 / (1 + p.coherence_temporal);
    if (prev_dev) {
        HIPTRY(hipMemsetAsync(B.bad, 0, sizeof(int), s));
        TRY(pt_lookup_init_prev(L, prev_dev, sv_dev, p.recalc_dist_temporal, temporal_scale, p.sanitize_input, B.ixy,
                                B.id, B.bad, s));
        L.init_xy = B.ixy;
        L.init_d = B.id;
    }
    if (timing) HIPTRY(hipEventRecord(t->ev[7], s));
    NnfField f0, f1;
    f0.xy = B.nxy;
    f0.d = B.nd;
    f0.fl = B.nfl;
    f1.xy = B.nxy2;
    f1.d = B.nd2;
    f1.fl = B.nfl2;
    int launches = 0;
    const int tfix = prev_dev && p.coherence_temporal != 0.0;
    if (mfma) {
        TRY(pt_quant_rgb0_launch(a_dev, (long)ah * aw, B.aq, s));
        set_int8_args(t, B.aq, B.qnorm, sv_dev, temporal_scale, L);
        TRY(pt_lookup_candidates(L, f0, s));
        if (timing) HIPTRY(hipEventRecord(t->ev[8], s));
        MfmaArgs M;
        fill_mfma_args(t, p, a_dev, B.aq, ah, aw, M);
        M.qnorm = B.qnorm;
        M.anch = B.anch;
        // a previous field's scores are its int8 SSDs times the temporal weight: at most the SSD
        // when that weight is 1 / (1 + coherence_temporal) with coherence_temporal >= 0
        const bool seen_ok = !prev_dev || (!sv_dev && p.coherence_temporal >= 0.0);
        if (t->mfma_stats && !ws) {
            const size_t n = (size_t)p.mfma_rounds * pt_mfma_round_workgroups(ann_w, ann_h);
            TRY(t->mfma_items.ensure(n));
            M.items_out = t->mfma_items.p;
        }
        int which = 0;
        TRY(pt_mfma_rounds_launch(M, p.mfma_rounds, p.mfma_jumps, seen_ok, f0, f1, s, &which));
        launches = p.mfma_rounds;
        if (timing) HIPTRY(hipEventRecord(t->ev[10], s));
        // the exact distances and the double output in one pass (k_init_prev's initial distances,
        // L.init_d, serve the patches that kept a previous field's match)
        TRY(pt_lookup_exact_out(L, which ? f1 : f0, ann_dev, prev_dev, temporal_scale, tfix, s));
    } else {
        TRY(pt_lookup_launch(L, f0, f1, B.gxy, B.gd, ann_dev, prev_dev, B.id, temporal_scale, tfix, s,
                             timing ? t->ev[8] : nullptr, timing ? t->ev[10] : nullptr, &launches));
    }
    if (timing) HIPTRY(hipEventRecord(t->ev[9], s));
    if (launches_out) *launches_out = launches;
    return PT_OK;
}

namespace {

int lookup_common(pt_table *t, const float *a_dev, int h, int w, int c, double *ann_dev, const double *prev_dev,
                  const float *sv_dev, double *seconds) {
    const pt_params &p = *t->p;
    hipStream_t s = t->stream;
    double T0 = now_s();
    int launches = 0;
    TRY(pt_lookup_enqueue(t, a_dev, h, w, c, ann_dev, prev_dev, sv_dev, nullptr, true, &launches));
    HIPTRY(hipStreamSynchronize(s));
    double T1 = now_s();
    if (prev_dev && !p.sanitize_input) {
        int bad = 0;
        HIPTRY(hipMemcpy(&bad, t->bad.p, sizeof(int), hipMemcpyDeviceToHost));
        if (bad) return pt_fail(PT_E_INVALID, "expected xsrc, ysrc in bounds in prev_nnf mode (sanitize_input 0)");
    }
    if (seconds) *seconds = T1 - T0;
    t->mfma_items_sum.clear();
    if (t->mfma_stats && p.mfma_rounds > 0) {
        const long nwg = pt_mfma_round_workgroups(t->qw - 7, t->qh - 7);
        std::vector<int32_t> h((size_t)p.mfma_rounds * nwg);
        HIPTRY(hipMemcpy(h.data(), t->mfma_items.p, h.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
        for (int r = 0; r < p.mfma_rounds; r++) {
            int64_t sum = 0;
            for (long i = 0; i < nwg; i++) sum += h[(size_t)r * nwg + i];
            t->mfma_items_sum.push_back(sum);
        }
    }
    float ms = 0;
    HIPTRY(hipEventElapsedTime(&ms, t->ev[6], t->ev[7]));
    t->stage_ms[6] = ms;
    HIPTRY(hipEventElapsedTime(&ms, t->ev[7], t->ev[8]));
    t->stage_ms[7] = ms;
    HIPTRY(hipEventElapsedTime(&ms, t->ev[8], t->ev[10]));
    t->stage_ms[8] = ms;
    t->stage_ms[10] = launches;
    HIPTRY(hipEventElapsedTime(&ms, t->ev[6], t->ev[9]));
    t->stage_ms[9] = ms;
    return PT_OK;
}

}  // namespace

int pt_table_device(const pt_table *t) { return t->device; }

// The buffers pt_lookup_enqueue works in, laid out one after another from base in 256-byte
// aligned pieces: query descriptor planes (and the interleaved copy in descriptor mode), the
// GCK's colour planes and running sums (image mode), two lattice fields, the candidate history,
// the gather output (query_step > 1) and the initial field (with a previous field).
int pt_lookup_workspace_layout(const pt_table *t, int h, int w, int c, bool prev, unsigned char *base, size_t *bytes,
                               LookupBufs *bufs) {
    int ah = 0, aw = 0;
    TRY(query_shape(t, h, w, c, ah, aw));
    const pt_params &p = *t->p;
    if (p.query_step < 1) return pt_fail(PT_E_INVALID, "query_step must be >= 1");
    const size_t plane = (size_t)(ah + 7) * (aw + 7);
    const int ann_w = aw - 7, ann_h = ah - 7, qs = p.query_step;
    const size_t n = (size_t)ann_w * ann_h;
    const size_t nv = (size_t)((ann_w + qs - 1) / qs) * ((ann_h + qs - 1) / qs);
    size_t off = 0;
    LookupBufs B;
    auto take = [&](size_t nbytes) -> unsigned char * {
        unsigned char *q = base ? base + off : nullptr;
        off += (nbytes + 255) & ~(size_t)255;
        return q;
    };
    B.adesc = (float *)take(plane * t->ndims * 4);
    if (t->desc_mode) {
        B.adescI = (float *)take(plane * t->ndims * 4);
    } else {
        B.yuv = (float *)take((size_t)ah * aw * 3 * 4);
        B.temp = (float *)take(plane * 3 * 4);
    }
    B.nxy = (int32_t *)take(nv * 4);
    B.nxy2 = (int32_t *)take(nv * 4);
    B.nd = (float *)take(nv * 4);
    B.nd2 = (float *)take(nv * 4);
    B.nfl = take(nv);
    B.nfl2 = take(nv);
    B.hist = (int32_t *)take(nv * PT_LOOKUP_HIST * 4);
    B.hn = take(nv);
    if (qs > 1) {
        B.gxy = (int32_t *)take(n * 4);
        B.gd = (float *)take(n * 4);
    }
    if (prev) {
        B.ixy = (int32_t *)take(n * 4);
        B.id = (float *)take(n * 4);
        B.bad = (int *)take(sizeof(int));
    }
    if (p.mfma_rounds > 0) {
        B.aq = (uint32_t *)take((size_t)ah * aw * 4);
        B.qnorm = (int32_t *)take(n * 4);
        B.anch = (int32_t *)take(pt_mfma_anchor_words(ann_w, ann_h, p.mfma_anchors) * 4);
    }
    if (bytes) *bytes = off;
    if (bufs) *bufs = B;
    return PT_OK;
}

// One lookup in the caller's workspace ws (ws_bytes long, 256-byte aligned): no device memory is
// allocated or freed. Synchronous, like pt_table_lookup_dev.
int pt_lookup_in_workspace(pt_table *t, const float *a_dev, int h, int w, int c, double *ann_dev,
                           const double *prev_dev, const float *sv_dev, unsigned char *ws, size_t ws_bytes,
                           double *seconds) {
    if (t->keyed) return pt_fail(PT_E_UNSUPPORTED, "learned-key tables look up in their own buffers");
    size_t need = 0;
    TRY(pt_lookup_workspace_layout(t, h, w, c, prev_dev != nullptr, nullptr, &need, nullptr));
    if (need > ws_bytes)
        return pt_failf(PT_E_INVALID, "query needs %zu bytes of lookup workspace, more than the %zu available", need,
                        ws_bytes);
    LookupBufs B;
    TRY(pt_lookup_workspace_layout(t, h, w, c, prev_dev != nullptr, ws, &need, &B));
    const pt_params &p = *t->p;
    const double T0 = now_s();
    TRY(pt_lookup_enqueue(t, a_dev, h, w, c, ann_dev, prev_dev, sv_dev, nullptr, false, nullptr, &B));
    HIPTRY(hipStreamSynchronize(t->stream));
    if (seconds) *seconds = now_s() - T0;
    if (prev_dev && !p.sanitize_input) {
        int bad = 0;
        HIPTRY(hipMemcpy(&bad, B.bad, sizeof(int), hipMemcpyDeviceToHost));
        if (bad) return pt_fail(PT_E_INVALID, "expected xsrc, ysrc in bounds in prev_nnf mode (sanitize_input 0)");
    }
    return PT_OK;
}

namespace {

// Shape of the NNF of a query of h x w x c (reference: (h-7) x (w-7), or h x w in descriptor mode).
int ann_size(const pt_table *t, int h, int w, int c, long &n) {
    int ah = 0, aw = 0;
    TRY(query_shape(t, h, w, c, ah, aw));
    n = (long)(ah - 7) * (aw - 7);
    return PT_OK;
}

}  // namespace

extern "C" int pt_table_create(pt_params *p, const float *b, int h, int w, int c, const int32_t *allowed,
                               pt_table **out) {
    return create_common(p, b, false, h, w, c, allowed, out);
}

extern "C" int pt_table_create_dev(pt_params *p, const float *b_dev, int h, int w, int c, const int32_t *allowed,
                                   pt_table **out) {
    return create_common(p, b_dev, true, h, w, c, allowed, out);
}

namespace {

// Learned-key table: main table over the 8 key planes of b (k-coherence in key space), then the
// GCK table of b with the caller's parameters as the further table; its shuffle continues the
// rand() stream after the key table's.
int build_keyed(pt_table *t, const pt_key *key, const float *b_dev, int h, int w) {
    const pt_params &p = t->build;
    hipStream_t s = t->stream;
    const int Hp = h + 7, Wp = w + 7;
    const size_t plane = (size_t)Hp * Wp;
    TRY(t->kw1.ensure(PT_KEY_HIDDEN * PT_KEY_IN));
    TRY(t->kb1.ensure(PT_KEY_HIDDEN));
    TRY(t->kw2.ensure(PT_KEY_OUT * PT_KEY_HIDDEN));
    TRY(t->kb2.ensure(PT_KEY_OUT));
    TRY(pt_key_snapshot(key, t->kw1.p, t->kb1.p, t->kw2.p, t->kb2.p, s));
    t->knet.w1 = t->kw1.p;
    t->knet.b1 = t->kb1.p;
    t->knet.w2 = t->kw2.p;
    t->knet.b2 = t->kb2.p;
    HIPTRY(hipEventRecord(t->ev[0], s));
    TRY(t->b0.ensure((size_t)h * w * 3));
    HIPTRY(hipMemcpyAsync(t->b0.p, b_dev, (size_t)h * w * 3 * 4, hipMemcpyDeviceToDevice, s));
    TRY(t->desc.ensure(plane * t->ndims));
    HIPTRY(hipMemsetAsync(t->desc.p, 0, plane * t->ndims * sizeof(float), s));
    TRY(pt_key_forward_launch(t->b0.p, h, w, t->knet.w1, t->knet.b1, t->knet.w2, t->knet.b2, t->desc.p, s));
    HIPTRY(hipEventRecord(t->ev[1], s));
    TRY(start_shuffle(t, std::vector<int32_t>()));
    const std::vector<int32_t> no_mask;
    TRY(build_grid(t, t->desc.p, no_mask, nullptr, 0, s, t->ev));
    t->K = p.mfma_rounds > 0 ? 0 : p.kcoherence;   // the MFMA lookup walks no lists
    if (p.mfma_rounds > 0) TRY(prepare_mfma(t, s));
    if (t->K > 0) {
        const size_t n = (size_t)(w - 7) * (h - 7);
        TRY(t->kpos.ensure(n * t->K));
        TRY(t->kdist.ensure(n * t->K));
        TRY(pt_kcoherence_launch(t->desc.p, plane, Wp, t->ndims, t->b0.p, nullptr, w, h, nullptr, t->allowed_index,
                                 t->quant(), t->nslices.data(), t->K, p.kcoh_window,
                                 p.kcoherence_min_dist * p.kcoherence_min_dist, t->kpos.p, t->kdist.p, t->sc, s));
    }
    pt_table *E = new pt_table();
    t->extra.push_back(E);
    E->build = *t->p;
    E->build.kcoherence = 0;
    E->build.ntables = 1;
    E->h = h;
    E->w = w;
    E->ndims = E->build.ndims;
    E->nchroma = E->build.nchroma;
    E->G = E->build.grid_ndims;
    E->allowed_index = t->allowed_index;
    TRY(E->desc.ensure(plane * E->ndims));
    TRY(t->yuv.ensure((size_t)h * w * 3));
    TRY(t->temp.ensure(plane * 3));
    TRY(pt_gck_launch(t->b0.p, h, w, E->ndims - 2 * E->nchroma, E->nchroma, E->build.convert_colorspace, t->yuv.p,
                      t->temp.p, E->desc.p, s));
    TRY(build_grid(E, E->desc.p, no_mask, nullptr, t->rand_used, s, nullptr));
    HIPTRY(hipEventRecord(t->ev[5], s));
    HIPTRY(hipStreamSynchronize(s));
    TRY(finish_grid(t, s));
    TRY(finish_grid(E, s));
    for (DevBuf<float> *b : {&E->desc}) b->release();
    E->list_xy.release();
    E->rank.release();
    E->cells.release();
    E->win.release();
    E->rnd.release();
    E->sc.release();
    float ms = 0;
    for (int k = 0; k < 5; k++) {
        HIPTRY(hipEventElapsedTime(&ms, t->ev[k], t->ev[k + 1]));
        t->stage_ms[k] = ms;
    }
    HIPTRY(hipEventElapsedTime(&ms, t->ev[0], t->ev[5]));
    t->stage_ms[5] = ms;
    t->rnd.release();
    return PT_OK;
}

}  // namespace

extern "C" int pt_table_create_keyed(pt_params *p, const pt_key *key, const float *b_dev, int h, int w,
                                     pt_table **out) {
    if (!p || !key || !b_dev || !out) return pt_fail(PT_E_INVALID, "pt_table_create_keyed: null argument");
    *out = nullptr;
    if (p->is_descriptor) return pt_fail(PT_E_UNSUPPORTED, "a learned-key table indexes RGB images (is_descriptor 0)");
    if (p->ntables != 1) return pt_fail(PT_E_UNSUPPORTED, "a learned-key table has its own further table (ntables 1)");
    if (p->load_filename[0] || p->save_filename[0])
        return pt_fail(PT_E_UNSUPPORTED, "a learned-key table has no file format (load/save)");
    TRY(check_build_params(p, 3));
    if (h < 8 || w < 8) return pt_fail(PT_E_INVALID, "database image must be at least 8x8");
    if (!PT_SIZE_OK(h, w))
        return pt_failf(PT_E_INVALID, "image too large for XY_TO_INT packing (height <= %d, width <= %d)", PT_MAX_H, PT_MAX_W);
    if (p->grid_ndims < 0) p->grid_ndims = p->ndims;   // patchtable.h:2603
    pt_table *t = new pt_table();
    t->p = p;
    t->build = *p;
    t->build.ndims = PT_KEY_OUT;   // the key table: 8 grid dimensions, default slices
    t->build.grid_ndims = PT_KEY_OUT;
    t->build.nchroma = 0;
    t->build.nslices_count = 0;
    t->h = h;
    t->w = w;
    t->ndims = PT_KEY_OUT;
    t->nchroma = 0;
    t->G = PT_KEY_OUT;
    t->keyed = true;
    t->allowed_index = p->allowed_index;
    int rc = init_stream(t);
    if (rc == PT_OK) rc = build_keyed(t, key, b_dev, h, w);
    if (rc != PT_OK) {
        delete t;
        return rc;
    }
    *out = t;
    return PT_OK;
}

extern "C" int pt_table_lookup(pt_table *t, const float *a, int h, int w, int c, double *ann_out,
                               const double *ann_prev, const float *sv, double *seconds) {
    if (!t || !a || !ann_out) return pt_fail(PT_E_INVALID, "pt_table_lookup: null argument");
    long n = 0;
    TRY(ann_size(t, h, w, c, n));
    TRY(t->a.ensure((size_t)h * w * c));
    TRY(t->ann.ensure((size_t)n * 3));
    HIPTRY(hipMemcpyAsync(t->a.p, a, (size_t)h * w * c * 4, hipMemcpyHostToDevice, t->stream));
    if (ann_prev) {
        TRY(t->prevbuf.ensure((size_t)n * 3));
        HIPTRY(hipMemcpyAsync(t->prevbuf.p, ann_prev, (size_t)n * 3 * 8, hipMemcpyHostToDevice, t->stream));
    }
    if (sv) {
        TRY(t->svbuf.ensure((size_t)n));
        HIPTRY(hipMemcpyAsync(t->svbuf.p, sv, (size_t)n * 4, hipMemcpyHostToDevice, t->stream));
    }
    TRY(lookup_common(t, t->a.p, h, w, c, t->ann.p, ann_prev ? t->prevbuf.p : nullptr, sv ? t->svbuf.p : nullptr,
                      seconds));
    HIPTRY(hipMemcpyAsync(ann_out, t->ann.p, (size_t)n * 3 * 8, hipMemcpyDeviceToHost, t->stream));
    HIPTRY(hipStreamSynchronize(t->stream));
    return PT_OK;
}

extern "C" int pt_table_lookup_dev(pt_table *t, const float *a_dev, int h, int w, int c, double *ann_out_dev,
                                   const double *ann_prev_dev, const float *sv_dev, double *seconds) {
    if (!t || !a_dev || !ann_out_dev) return pt_fail(PT_E_INVALID, "pt_table_lookup_dev: null argument");
    return lookup_common(t, a_dev, h, w, c, ann_out_dev, ann_prev_dev, sv_dev, seconds);
}

extern "C" int pt_table_destroy(pt_table *t) {
    delete t;
    return PT_OK;
}

extern "C" int pt_table_shape(const pt_table *t, int *h, int *w, int *ndims) {
    if (!t) return pt_fail(PT_E_INVALID, "null table");
    if (h) *h = t->h;
    if (w) *w = t->w;
    if (ndims) *ndims = t->ndims;
    return PT_OK;
}

static int planes_to_hwc(const float *planes, int Hp, int Wp, int C, float *out) {
    std::vector<float> host((size_t)Hp * Wp * C);
    HIPTRY(hipMemcpy(host.data(), planes, host.size() * 4, hipMemcpyDeviceToHost));
    const size_t plane = (size_t)Hp * Wp;
    for (size_t i = 0; i < plane; i++)
        for (int c = 0; c < C; c++) out[i * C + c] = host[c * plane + i];
    return PT_OK;
}

// ---------------------------------------------------------------------------------------------
// Persistence: the reference's binary format (save/load of basic types, Arrays and strings,
// patchtable.cpp:499-573; PatchTableFixedN::save/load, patchtable.h:1574-1593; PatchPartition
// and PatchPartition1D, patchtable.h:1145-1170, 1334-1353).
// ---------------------------------------------------------------------------------------------
namespace {

struct Writer {
    FILE *f;
    bool ok = true;
    void raw(const void *p, size_t bytes) {
        if (ok && bytes && fwrite(p, 1, bytes, f) != bytes) ok = false;
    }
    void i32(int32_t v) { raw(&v, 4); }
    void f32(float v) { raw(&v, 4); }
    void str(const char *s) {
        i32((int32_t)strlen(s));
        raw(s, strlen(s));
    }
    template <class T>
    void vec(const std::vector<T> &v) {
        i32((int32_t)v.size());
        raw(v.data(), v.size() * sizeof(T));
    }
    template <class T>
    void array(const std::vector<int> &sizes, const T *data, size_t n) {
        vec(sizes);
        raw(data, n * sizeof(T));
    }
};

struct Reader {
    FILE *f;
    const char *path;
    int raw(void *p, size_t bytes) {
        if (bytes && fread(p, 1, bytes, f) != bytes) return pt_failf(PT_E_IO, "%s: truncated PatchTable file", path);
        return PT_OK;
    }
    int i32(int32_t &v) { return raw(&v, 4); }
    int f32(float &v) { return raw(&v, 4); }
    int id(const char *expect) {   // load_id (patchtable.cpp:560-567)
        int32_t n = 0;
        TRY(i32(n));
        if (n < 0 || n > 256) return pt_failf(PT_E_IO, "%s: expected |%s|", path, expect);
        std::string s((size_t)n, '\0');
        TRY(raw(&s[0], (size_t)n));
        if (s != expect) return pt_failf(PT_E_IO, "%s: load_id expected |%s|, received |%s|", path, expect, s.c_str());
        return PT_OK;
    }
    template <class T>
    int vec(std::vector<T> &v, int32_t max_n) {
        int32_t n = 0;
        TRY(i32(n));
        if (n < 0 || n > max_n) return pt_failf(PT_E_IO, "%s: bad vector length %d", path, n);
        v.resize((size_t)n);
        return raw(v.data(), (size_t)n * sizeof(T));
    }
    template <class T>
    int array(std::vector<int> &sizes, std::vector<T> &data) {
        TRY(vec(sizes, 64));
        size_t n = 1;
        for (int d : sizes) {
            if (d < 0) return pt_failf(PT_E_IO, "%s: bad array size", path);
            n *= (size_t)d;
        }
        if (n > ((size_t)1 << 34)) return pt_failf(PT_E_IO, "%s: array too large", path);
        data.resize(n);
        return raw(data.data(), n * sizeof(T));
    }
};

int hwc_to_planes(const float *hwc, size_t plane, int C, std::vector<float> &planes) {
    planes.resize(plane * C);
    for (size_t i = 0; i < plane; i++)
        for (int c = 0; c < C; c++) planes[c * plane + i] = hwc[i * C + c];
    return PT_OK;
}

}  // namespace

extern "C" int pt_table_save(pt_table *t, const char *filename) {
    if (t && t->keyed) return pt_fail(PT_E_UNSUPPORTED, "a learned-key table has no file format");
    if (!t || !filename) return pt_fail(PT_E_INVALID, "pt_table_save: null argument");
    const int Hp = t->h + 7, Wp = t->w + 7;
    std::vector<int32_t> cells((size_t)t->ncells);
    std::vector<float> b0((size_t)t->h * t->w * 3), wh0((size_t)Hp * Wp * t->ndims);
    HIPTRY(hipStreamSynchronize(t->stream));
    HIPTRY(hipMemcpy(cells.data(), t->table.p, cells.size() * 4, hipMemcpyDeviceToHost));
    if (!t->desc_mode)   // descriptor mode keeps no RGB copy; the reference's b0 is unset there
        HIPTRY(hipMemcpy(b0.data(), t->b0.p, b0.size() * 4, hipMemcpyDeviceToHost));
    TRY(planes_to_hwc(t->desc.p, Hp, Wp, t->ndims, wh0.data()));
    std::vector<std::vector<int32_t>> maps(t->G, std::vector<int32_t>(PT_NMAP));
    for (int d = 0; d < t->G; d++) TRY(fetch_to_bin(t, d, maps[d].data()));
    FILE *f = fopen(filename, "wb");
    if (!f) return pt_failf(PT_E_IO, "PatchTable: could not save to %s", filename);
    Writer W{f};
    W.str("PatchTable");
    W.array(t->nslices, cells.data(), cells.size());
    W.array(std::vector<int>{t->h, t->w, 3}, b0.data(), b0.size());
    W.array(std::vector<int>{Hp, Wp, t->ndims}, wh0.data(), wh0.size());
    W.vec(t->nslices);
    W.str("PatchPartition");
    W.i32(t->G);
    for (const HostPartition &P : t->part) {
        W.str("PatchPartition1D");
        W.vec(maps[&P - t->part.data()]);
        W.vec(P.centers);
        W.vec(P.lsizes);
        W.vec(P.rsizes);
        W.f32(P.min_val);
        W.f32(P.max_val);
        W.f32(P.range);
        W.i32(P.nslices);
        W.str("EndPatchPartition1D");
    }
    W.str("EndPatchPartition");
    W.str("EndPatchTable");
    const bool ok = W.ok && fclose(f) == 0;
    if (!ok) return pt_failf(PT_E_IO, "PatchTable: write to %s failed", filename);
    return PT_OK;
}

namespace {

int load_file(pt_table *t, const char *filename) {
    FILE *f = fopen(filename, "rb");
    if (!f) return pt_failf(PT_E_IO, "PatchTable: could not load from %s", filename);
    struct Closer {
        FILE *f;
        ~Closer() { fclose(f); }
    } closer{f};
    Reader R{f, filename};
    std::vector<int> tsz, bsz, wsz, ns;
    std::vector<int32_t> cells;
    std::vector<float> b0, wh0;
    TRY(R.id("PatchTable"));
    TRY(R.array(tsz, cells));
    TRY(R.array(bsz, b0));
    TRY(R.array(wsz, wh0));
    TRY(R.vec(ns, PT_MAX_DIMS));
    if (bsz.size() != 3 || bsz[2] != 3 || wsz.size() != 3 || wsz[0] != bsz[0] + 7 || wsz[1] != bsz[1] + 7)
        return pt_failf(PT_E_IO, "%s: inconsistent image and descriptor sizes", filename);
    if (tsz != ns || ns.empty()) return pt_failf(PT_E_IO, "%s: table sizes do not match nslices", filename);
    t->h = bsz[0];
    t->w = bsz[1];
    t->ndims = wsz[2];
    t->G = (int)ns.size();
    if (t->h < 8 || t->w < 8 || !PT_SIZE_OK(t->h, t->w) || t->ndims < 1 || t->ndims > PT_MAX_DIMS || t->G > t->ndims)
        return pt_failf(PT_E_IO, "%s: bad table dimensions", filename);
    t->nslices = ns;
    t->stride.assign(t->G, 0);
    int64_t acc = 1;
    for (int i = t->G - 1; i >= 0; i--) {
        if (ns[i] < 1) return pt_failf(PT_E_IO, "%s: bad nslices", filename);
        t->stride[i] = acc;
        acc *= ns[i];
    }
    t->ncells = acc;
    TRY(R.id("PatchPartition"));
    int32_t np = 0;
    TRY(R.i32(np));
    if (np != t->G) return pt_failf(PT_E_IO, "%s: partition count %d != grid dims %d", filename, np, t->G);
    t->part.assign(t->G, HostPartition());
    for (int i = 0; i < t->G; i++) {
        HostPartition &P = t->part[i];
        TRY(R.id("PatchPartition1D"));
        TRY(R.vec(P.to_bin, PT_NMAP));
        TRY(R.vec(P.centers, 1 << 24));
        TRY(R.vec(P.lsizes, 1 << 24));
        TRY(R.vec(P.rsizes, 1 << 24));
        TRY(R.f32(P.min_val));
        TRY(R.f32(P.max_val));
        TRY(R.f32(P.range));
        TRY(R.i32(P.nslices));
        TRY(R.id("EndPatchPartition1D"));
        if ((int)P.to_bin.size() != PT_NMAP || P.nslices != ns[i])
            return pt_failf(PT_E_IO, "%s: bad partition of dimension %d", filename, i);
        for (int32_t v : P.to_bin)
            if (v < 0 || v >= P.nslices) return pt_failf(PT_E_IO, "%s: bin index out of range", filename);
    }
    TRY(R.id("EndPatchPartition"));
    TRY(R.id("EndPatchTable"));
    const int bew = t->w - 7, beh = t->h - 7;
    for (int32_t v : cells) {
        const int x = PT_INT_TO_X(v) - 7, y = PT_INT_TO_Y(v) - 7;
        if (x < 0 || y < 0 || x >= bew || y >= beh) return pt_failf(PT_E_IO, "%s: table entry out of range", filename);
    }
    // upload
    hipStream_t s = t->stream;
    const size_t plane = (size_t)(t->h + 7) * (t->w + 7);
    std::vector<float> planes;
    hwc_to_planes(wh0.data(), plane, t->ndims, planes);
    TRY(t->desc.ensure(planes.size()));
    TRY(t->table.ensure(cells.size()));
    if (t->desc_mode) {   // descriptor mode: distances use the interleaved descriptor (= wh0)
        TRY(t->descI.ensure(wh0.size()));
        HIPTRY(hipMemcpyAsync(t->descI.p, wh0.data(), wh0.size() * 4, hipMemcpyHostToDevice, s));
    } else {
        TRY(t->b0.ensure(b0.size()));
        HIPTRY(hipMemcpyAsync(t->b0.p, b0.data(), b0.size() * 4, hipMemcpyHostToDevice, s));
    }
    HIPTRY(hipMemcpyAsync(t->desc.p, planes.data(), planes.size() * 4, hipMemcpyHostToDevice, s));
    HIPTRY(hipMemcpyAsync(t->table.p, cells.data(), cells.size() * 4, hipMemcpyHostToDevice, s));
    TRY(upload_quant(t, s));
    return PT_OK;
}

}  // namespace

extern "C" int pt_table_load(pt_params *p, const char *filename, pt_table **out) {
    if (!p || !filename || !out) return pt_fail(PT_E_INVALID, "pt_table_load: null argument");
    *out = nullptr;
    pt_table *t = new pt_table();
    t->p = p;
    t->build = *p;
    t->desc_mode = p->is_descriptor != 0;
    int rc = init_stream(t);
    if (rc == PT_OK) rc = load_file(t, filename);
    if (rc == PT_OK && t->desc_mode && !desc_dims_supported(t->ndims))
        rc = pt_failf(PT_E_UNSUPPORTED, "descriptor mode: %d channels unsupported", t->ndims);
    if (rc == PT_OK && !t->desc_mode && t->ndims != p->ndims)
        rc = pt_failf(PT_E_INVALID, "%s holds %d-dimensional descriptors, parameters say ndims %d", filename, t->ndims,
                      p->ndims);
    if (rc == PT_OK) {
        t->nchroma = p->nchroma;
        t->build.grid_ndims = t->G;
        // The format holds no k-coherence lists; rebuild them from the loaded b0 and wh0.
        t->K = p->kcoherence;
        if (t->K < 0 || t->K > 32) rc = pt_fail(PT_E_UNSUPPORTED, "kcoherence must be in 0..32");
    }
    if (rc == PT_OK && t->K > 0) {
        const size_t n = (size_t)(t->w - 7) * (t->h - 7);
        const size_t plane = (size_t)(t->h + 7) * (t->w + 7);
        rc = t->kpos.ensure(n * t->K);
        if (rc == PT_OK) rc = t->kdist.ensure(n * t->K);
        if (rc == PT_OK)
            rc = pt_kcoherence_launch(t->desc.p, plane, t->w + 7, t->ndims, t->b0.p,
                                      t->desc_mode ? t->descI.p : nullptr, t->w, t->h, nullptr, 1, t->quant(),
                                      t->nslices.data(), t->K, p->kcoh_window,
                                      p->kcoherence_min_dist * p->kcoherence_min_dist, t->kpos.p, t->kdist.p, t->sc,
                                      t->stream);
    }
    if (rc == PT_OK && hipStreamSynchronize(t->stream) != hipSuccess) rc = pt_fail(PT_E_HIP, "load: stream sync failed");
    if (rc != PT_OK) {
        delete t;
        return rc;
    }
    *out = t;
    return PT_OK;
}

// ---------------------------------------------------------------------------------------------
// Introspection
// ---------------------------------------------------------------------------------------------
extern "C" int pt_table_info(const pt_table *t, int *grid_ndims, int *nslices, int64_t *ncells, int *table_add_count,
                             int *kcoherence) {
    if (!t) return pt_fail(PT_E_INVALID, "null table");
    if (grid_ndims) *grid_ndims = t->G;
    if (nslices)
        for (int i = 0; i < t->G; i++) nslices[i] = t->nslices[i];
    if (ncells) *ncells = t->ncells;
    if (table_add_count) *table_add_count = t->table_add_count;
    if (kcoherence) *kcoherence = t->K;
    return PT_OK;
}

extern "C" int pt_table_get_descriptor(const pt_table *t, float *out) {
    if (!t || !out) return pt_fail(PT_E_INVALID, "null argument");
    return planes_to_hwc(t->desc.p, t->h + 7, t->w + 7, t->ndims, out);
}

extern "C" int pt_table_get_query_descriptor(const pt_table *t, float *out) {
    if (!t || !out || !t->qh) return pt_fail(PT_E_INVALID, "no query descriptor");
    return planes_to_hwc(t->adesc.p, t->qh + 7, t->qw + 7, t->ndims, out);
}

extern "C" int pt_table_get_partition(const pt_table *t, int dim, float *min_val, float *max_val, float *centers,
                                      int32_t *to_bin) {
    if (!t || dim < 0 || dim >= t->G) return pt_fail(PT_E_INVALID, "bad dim");
    const HostPartition &P = t->part[dim];
    if (min_val) *min_val = P.min_val;
    if (max_val) *max_val = P.max_val;
    if (centers) std::copy(P.centers.begin(), P.centers.end(), centers);
    if (to_bin) TRY(fetch_to_bin(t, dim, to_bin));
    return PT_OK;
}

extern "C" int pt_table_get_cells(const pt_table *t, int32_t *out) {
    if (!t || !out) return pt_fail(PT_E_INVALID, "null argument");
    HIPTRY(hipMemcpy(out, t->table.p, (size_t)t->ncells * 4, hipMemcpyDeviceToHost));
    return PT_OK;
}

extern "C" int pt_table_get_populated(const pt_table *t, int32_t *out) {
    if (!t || !out) return pt_fail(PT_E_INVALID, "null argument");
    if (!t->cells.p) return pt_fail(PT_E_INVALID, "no populate record (a loaded table)");
    const long n = t->n_pop;
    std::vector<int64_t> cells(n);
    std::vector<uint8_t> win(n);
    std::vector<int32_t> xy(n);
    HIPTRY(hipMemcpy(cells.data(), t->cells.p, n * 8, hipMemcpyDeviceToHost));
    HIPTRY(hipMemcpy(win.data(), t->win.p, n, hipMemcpyDeviceToHost));
    HIPTRY(hipMemcpy(xy.data(), t->list_xy.p, n * 4, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < t->ncells; i++) out[i] = PT_TABLE_UNUSED;
    for (long k = 0; k < n; k++)
        if (win[k]) out[cells[k]] = xy[k];
    return PT_OK;
}

extern "C" int pt_table_extra_info(const pt_table *t, int i, int *nslices, int64_t *ncells, int *table_add_count) {
    if (!t || i < 0 || i >= (int)t->extra.size()) return pt_fail(PT_E_INVALID, "no such further table");
    const pt_table *E = t->extra[i];
    if (nslices)
        for (int d = 0; d < E->G; d++) nslices[d] = E->nslices[d];
    if (ncells) *ncells = E->ncells;
    if (table_add_count) *table_add_count = E->table_add_count;
    return PT_OK;
}

extern "C" int pt_table_get_extra_cells(const pt_table *t, int i, int32_t *out) {
    if (!t || !out || i < 0 || i >= (int)t->extra.size()) return pt_fail(PT_E_INVALID, "no such further table");
    HIPTRY(hipMemcpy(out, t->extra[i]->table.p, (size_t)t->extra[i]->ncells * 4, hipMemcpyDeviceToHost));
    return PT_OK;
}

extern "C" int pt_table_get_kcoherence(const pt_table *t, int32_t *pos_out, float *dist_out) {
    if (!t) return pt_fail(PT_E_INVALID, "null table");
    const size_t n = (size_t)(t->w - 7) * (t->h - 7) * t->K;
    if (n == 0) return PT_OK;
    if (pos_out) HIPTRY(hipMemcpy(pos_out, t->kpos.p, n * 4, hipMemcpyDeviceToHost));
    if (dist_out) HIPTRY(hipMemcpy(dist_out, t->kdist.p, n * 4, hipMemcpyDeviceToHost));
    return PT_OK;
}

extern "C" int pt_table_set_mfma_stats(pt_table *t, int on) {
    if (!t) return pt_fail(PT_E_INVALID, "null argument");
    t->mfma_stats = on != 0;
    return PT_OK;
}

extern "C" int pt_table_set_tile(pt_table *t, int tile_x) {
    if (!t) return pt_fail(PT_E_INVALID, "null argument");
    if (!pt_lookup_tile_ok(tile_x)) return pt_failf(PT_E_INVALID, "tile width %d is not one of 64, 32, 16", tile_x);
    t->tile_x = tile_x;
    return PT_OK;
}

extern "C" int pt_table_autotune(pt_table *t, const float *a_dev, int h, int w, int c, double *ann_dev, int reps,
                                 int *tile_x_out, double *ms_out) {
    if (!t || !a_dev || !ann_dev || reps < 1) return pt_fail(PT_E_INVALID, "pt_table_autotune: bad arguments");
    static const int shapes[3] = {64, 32, 16};
    const int keep = t->tile_x;
    double best = 0;
    int best_tx = keep;
    for (int k = 0; k < 3; k++) {
        t->tile_x = shapes[k];
        int rc = lookup_common(t, a_dev, h, w, c, ann_dev, nullptr, nullptr, nullptr);   // warm-up
        std::vector<double> ms;
        for (int r = 0; rc == PT_OK && r < reps; r++) {
            rc = lookup_common(t, a_dev, h, w, c, ann_dev, nullptr, nullptr, nullptr);
            ms.push_back(t->stage_ms[9]);   // device time of the whole lookup
        }
        if (rc != PT_OK) {
            t->tile_x = keep;
            return rc;
        }
        std::sort(ms.begin(), ms.end());
        const double med = ms[ms.size() / 2];
        if (ms_out) ms_out[k] = med;
        if (k == 0 || med < best) {
            best = med;
            best_tx = shapes[k];
        }
    }
    t->tile_x = best_tx;
    if (tile_x_out) *tile_x_out = best_tx;
    return PT_OK;
}

extern "C" int pt_table_mfma_items(const pt_table *t, int64_t *items, int n, int *rounds) {
    if (!t || !rounds) return pt_fail(PT_E_INVALID, "null argument");
    *rounds = (int)t->mfma_items_sum.size();
    for (int r = 0; r < n && r < *rounds; r++) items[r] = t->mfma_items_sum[r];
    return PT_OK;
}

extern "C" int pt_table_stage_ms(const pt_table *t, double *ms) {
    if (!t || !ms) return pt_fail(PT_E_INVALID, "null argument");
    for (int i = 0; i < 12; i++) ms[i] = t->stage_ms[i];
    return PT_OK;
}

// ---------------------------------------------------------------------------------------------
// Stand-alone kernels
// ---------------------------------------------------------------------------------------------
extern "C" int pt_gck_dev(const float *img_dev, int h, int w, int n, int nchroma, int convert_yuv, float *out_dev) {
    if (!img_dev || !out_dev || h < 1 || w < 1) return pt_fail(PT_E_INVALID, "pt_gck_dev: bad arguments");
    DevBuf<float> yuv, temp;
    TRY(yuv.ensure((size_t)h * w * 3));
    TRY(temp.ensure((size_t)(h + 7) * (w + 7) * 3));
    TRY(pt_gck_launch(img_dev, h, w, n, nchroma, convert_yuv, yuv.p, temp.p, out_dev, 0));
    HIPTRY(hipStreamSynchronize(0));
    return PT_OK;
}

extern "C" int pt_patch_ssd_dev(const float *a_dev, int aw, const float *b_dev, int bw, const int32_t *ax,
                                const int32_t *ay, const int32_t *bx, const int32_t *by, int npairs, float *out_dev) {
    TRY(pt_ssd_pairs_launch(a_dev, aw, b_dev, bw, ax, ay, bx, by, npairs, out_dev, 0));
    HIPTRY(hipStreamSynchronize(0));
    return PT_OK;
}

extern "C" int pt_desc_dist_dev(const float *a_dev, const float *b_dev, int C, int n, float *out_dev) {
    if (!a_dev || !b_dev || !out_dev || C < 4 || n < 0) return pt_fail(PT_E_INVALID, "pt_desc_dist_dev: bad arguments");
    TRY(pt_desc_dist_pairs_launch(a_dev, b_dev, C, n, out_dev, 0));
    HIPTRY(hipStreamSynchronize(0));
    return PT_OK;
}

// ---------------------------------------------------------------------------------------------
// One query against many databases (the reference's one-vs-many benchmark runs its front end
// once per pair, bench.py:80-83). The tables are built concurrently from host threads, each on
// its own stream; a lookup computes the query descriptor once, then every table's lookup runs on
// its own stream, replayed from a captured HIP graph after the first call with the same buffers.
// ---------------------------------------------------------------------------------------------
struct pt_batch {
    pt_params *p = nullptr;
    std::vector<pt_table *> tabs;
    hipStream_t qs = nullptr;     // the batch's stream: the query's descriptor and every phase
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipGraphExec_t gexec = nullptr;
    std::vector<const void *> gkey;   // buffers and tile width the graph was captured with
    pt_params gpar;                   // ... and the parameters (byte copy): the caller may change
                                      // p between lookups
    int tile_x = 64;                  // lattice tile of the lookup kernels (pt_batch_autotune)
    DevBuf<float> adesc, yuv, temp;
    DevBuf<uint32_t> aq;              // int8 RGB0 query (MFMA lookups)
    DevBuf<int32_t> qnorm;            // ... and |q|^2 of its patches
    DevBuf<PairLookup> pairs;         // every pair's arguments, read by the batched kernels
    std::vector<PairLookup> hpairs;   // ... and their host copy
    ~pt_batch() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (qs) (void)hipStreamDestroy(qs);
        for (pt_table *t : tabs) delete t;
    }
};

extern "C" int pt_batch_create(pt_params *p, int n, const float *const *b_dev, const int *h, const int *w,
                               int threads, pt_batch **out) {
    if (!p || !out || n < 1 || !b_dev || !h || !w) return pt_fail(PT_E_INVALID, "pt_batch_create: bad arguments");
    if (p->is_descriptor) return pt_fail(PT_E_UNSUPPORTED, "pt_batch: descriptor mode is not supported");
    if (n > 65535) return pt_fail(PT_E_INVALID, "pt_batch_create: at most 65535 databases");
    *out = nullptr;
    pt_batch *bt = new pt_batch();
    bt->p = p;
    bt->tabs.assign(n, nullptr);
    // Parameters are shared; the first create writes grid_ndims back, so do it once up front.
    if (p->grid_ndims < 0) p->grid_ndims = p->ndims;
    std::vector<int> rc(n, PT_OK);
    std::vector<std::string> err(n);
    const int nt = std::max(1, std::min(threads > 0 ? threads : 4, n));
    std::vector<std::thread> pool;
    for (int k = 0; k < nt; k++)
        pool.emplace_back([&, k]() {
            for (int i = k; i < n; i += nt) {
                rc[i] = create_common(p, b_dev[i], true, h[i], w[i], 3, nullptr, &bt->tabs[i]);
                if (rc[i] != PT_OK) err[i] = pt_last_error();
            }
        });
    for (auto &th : pool) th.join();
    for (int i = 0; i < n; i++)
        if (rc[i] != PT_OK) {
            delete bt;
            return pt_failf(rc[i], "database %d: %s", i, err[i].c_str());
        }
    if (hipStreamCreateWithFlags(&bt->qs, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&bt->e0) != hipSuccess ||
        hipEventCreate(&bt->e1) != hipSuccess) {
        delete bt;
        return pt_fail(PT_E_HIP, "pt_batch_create: stream/event creation failed");
    }
    *out = bt;
    return PT_OK;
}

extern "C" int pt_batch_size(const pt_batch *bt) { return bt ? (int)bt->tabs.size() : 0; }

extern "C" pt_table *pt_batch_table(pt_batch *bt, int i) {
    return bt && i >= 0 && i < (int)bt->tabs.size() ? bt->tabs[i] : nullptr;
}

namespace {

// Every pair's arguments for a lookup of the h x w query a_dev (buffers grown on each table), and
// their device copy; the tables' builds and earlier work on their streams are waited for.
int batch_prepare(pt_batch *bt, const float *a_dev, int h, int w, double *const *ann_dev) {
    const pt_params &p = *bt->p;
    const int n = (int)bt->tabs.size();
    pt_table *t0 = bt->tabs[0];
    const size_t plane = (size_t)(h + 7) * (w + 7);
    TRY(bt->adesc.ensure(plane * t0->ndims));
    TRY(bt->yuv.ensure((size_t)h * w * 3));
    TRY(bt->temp.ensure(plane * 3));
    const bool mfma = p.mfma_rounds > 0;
    if (mfma) {
        TRY(bt->aq.ensure((size_t)h * w));
        TRY(bt->qnorm.ensure((size_t)(h - 7) * (w - 7)));
    }
    bt->hpairs.assign(n, PairLookup());
    for (int i = 0; i < n; i++) {
        pt_table *t = bt->tabs[i];
        if (t->ndims != t0->ndims || t->nchroma != t0->nchroma || t->extra.size() != t0->extra.size())
            return pt_failf(PT_E_INVALID, "database %d: tables of a batch share their descriptor", i);
        int ah = 0, aw = 0;
        TRY(check_lookup(t, p, h, w, 3, false, nullptr, ah, aw));
        const int ann_w = aw - 7, ann_h = ah - 7, qs = p.query_step;
        const long nv = (long)((ann_w + qs - 1) / qs) * ((ann_h + qs - 1) / qs), np = (long)ann_w * ann_h;
        TRY(t->nxy.ensure(nv));
        TRY(t->nd.ensure(nv));
        TRY(t->nfl.ensure(nv));
        TRY(t->nxy2.ensure(nv));
        TRY(t->nd2.ensure(nv));
        TRY(t->nfl2.ensure(nv));
        if (!mfma) {
            TRY(t->hist.ensure((size_t)nv * PT_LOOKUP_HIST));
            TRY(t->hn.ensure(nv));
        } else {
            TRY(t->anch.ensure(pt_mfma_anchor_words(ann_w, ann_h, p.mfma_anchors)));
        }
        if (qs > 1) {
            TRY(t->gxy.ensure(np));
            TRY(t->gd.ensure(np));
        }
        LookupBufs B;
        B.hist = t->hist.p;
        B.hn = t->hn.p;
        PairLookup &P = bt->hpairs[i];
        fill_lookup_args(t, p, ah, aw, bt->adesc.p, bt->adesc.p, B, P.L);
        P.L.a = a_dev;
        P.L.b = t->b0.p;
        P.L.tile_x = bt->tile_x;
        P.G = P.L;
        P.G.gqs = qs;
        P.G.qs = 1;   // gather lanes are single patches
        if (mfma) {
            set_int8_args(t, bt->aq.p, bt->qnorm.p, nullptr, 1.0, P.L);
            fill_mfma_args(t, p, a_dev, bt->aq.p, ah, aw, P.M);
            P.M.qnorm = bt->qnorm.p;
            P.M.anch = t->anch.p;
        }
        P.f[0].xy = t->nxy.p;
        P.f[0].d = t->nd.p;
        P.f[0].fl = t->nfl.p;
        P.f[1].xy = t->nxy2.p;
        P.f[1].d = t->nd2.p;
        P.f[1].fl = t->nfl2.p;
        P.gxy = t->gxy.p;
        P.gd = t->gd.p;
        P.ann = ann_dev[i];
        HIPTRY(hipStreamSynchronize(t->stream));   // its build (and any single lookup) is done
    }
    TRY(bt->pairs.ensure(n));
    HIPTRY(hipMemcpy(bt->pairs.p, bt->hpairs.data(), n * sizeof(PairLookup), hipMemcpyHostToDevice));
    return PT_OK;
}

// The launches of one batched lookup, on the batch's stream (captured into a graph after the first
// eager run): the query's descriptor once, then each phase once for all pairs.
int batch_enqueue(pt_batch *bt, const float *a_dev, int h, int w, int *launches) {
    const pt_params &p = *bt->p;
    const int n = (int)bt->tabs.size();
    pt_table *t0 = bt->tabs[0];
    const PairLookup &P0 = bt->hpairs[0];
    hipStream_t s = bt->qs;
    TRY(pt_gck_launch(a_dev, h, w, t0->ndims - 2 * t0->nchroma, t0->nchroma, t0->build.convert_colorspace, bt->yuv.p,
                      bt->temp.p, bt->adesc.p, s));
    int nl = 1;
    if (p.mfma_rounds > 0) {
        TRY(pt_quant_rgb0_launch(a_dev, (long)h * w, bt->aq.p, s));
        TRY(pt_lookup_batch_candidates(bt->pairs.p, P0, n, s));
        int which = 0;
        TRY(pt_mfma_rounds_batch_launch(bt->pairs.p, P0, n, p.mfma_rounds, p.mfma_jumps, s, &which));
        TRY(pt_lookup_batch_exact_out(bt->pairs.p, P0, n, which, s));
        nl += 4 + p.mfma_rounds;
    } else {
        int l = 0;
        TRY(pt_lookup_batch_launch(bt->pairs.p, P0, n, s, nullptr, &l));
        nl += l + 1;
    }
    if (launches) *launches = nl;
    return PT_OK;
}

}  // namespace

extern "C" int pt_batch_lookup(pt_batch *bt, const float *a_dev, int h, int w, double *const *ann_dev,
                               double *ms_per_pair, double *seconds) {
    if (!bt || !a_dev || !ann_dev) return pt_fail(PT_E_INVALID, "pt_batch_lookup: null argument");
    const int n = (int)bt->tabs.size();
    for (int i = 0; i < n; i++)
        if (!ann_dev[i]) return pt_fail(PT_E_INVALID, "pt_batch_lookup: null output");
    if (h < 8 || w < 8) return pt_fail(PT_E_INVALID, "query image must be at least 8x8");
    if (!PT_SIZE_OK(h, w)) return pt_fail(PT_E_INVALID, "query image too large for XY_TO_INT packing");
    const double T0 = now_s();
    std::vector<const void *> key = {a_dev, (const void *)(intptr_t)h, (const void *)(intptr_t)w,
                                     (const void *)(intptr_t)bt->tile_x};
    for (int i = 0; i < n; i++) key.push_back(ann_dev[i]);
    hipStream_t s = bt->qs;
    if (bt->gexec && bt->gkey == key && std::memcmp(&bt->gpar, bt->p, sizeof(pt_params)) == 0) {
        HIPTRY(hipEventRecord(bt->e0, s));
        HIPTRY(hipGraphLaunch(bt->gexec, s));
        HIPTRY(hipEventRecord(bt->e1, s));
    } else {
        // eager run (sizes the buffers, uploads the pairs), then capture the same launches
        if (bt->gexec) {
            HIPTRY(hipGraphExecDestroy(bt->gexec));
            bt->gexec = nullptr;
        }
        TRY(batch_prepare(bt, a_dev, h, w, ann_dev));
        HIPTRY(hipEventRecord(bt->e0, s));
        TRY(batch_enqueue(bt, a_dev, h, w, nullptr));
        HIPTRY(hipEventRecord(bt->e1, s));
        hipGraph_t g = nullptr;
        HIPTRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        const int rc = batch_enqueue(bt, a_dev, h, w, nullptr);
        const hipError_t ec = hipStreamEndCapture(s, &g);
        if (rc != PT_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (ec != hipSuccess) return pt_failf(PT_E_HIP, "hipStreamEndCapture failed: %s", hipGetErrorString(ec));
        const hipError_t ei = hipGraphInstantiate(&bt->gexec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) {
            bt->gexec = nullptr;
            return pt_failf(PT_E_HIP, "hipGraphInstantiate failed: %s", hipGetErrorString(ei));
        }
        bt->gkey = key;
        std::memcpy(&bt->gpar, bt->p, sizeof(pt_params));
    }
    HIPTRY(hipStreamSynchronize(s));
    const double T1 = now_s();
    if (ms_per_pair) {
        float ms = 0;
        HIPTRY(hipEventElapsedTime(&ms, bt->e0, bt->e1));
        for (int i = 0; i < n; i++) ms_per_pair[i] = ms;   // every pair ends with the batch
    }
    if (seconds) *seconds = T1 - T0;
    return PT_OK;
}

extern "C" int pt_batch_set_tile(pt_batch *bt, int tile_x) {
    if (!bt) return pt_fail(PT_E_INVALID, "pt_batch_set_tile: null batch");
    if (!pt_lookup_tile_ok(tile_x)) return pt_failf(PT_E_INVALID, "tile width %d is not one of 64, 32, 16", tile_x);
    bt->tile_x = tile_x;
    return PT_OK;
}

extern "C" int pt_batch_get_tile(const pt_batch *bt) { return bt ? bt->tile_x : 0; }

extern "C" int pt_batch_autotune(pt_batch *bt, const float *a_dev, int h, int w, double *const *ann_dev, int reps,
                                 int *tile_x_out, double *ms_out) {
    if (!bt || reps < 1) return pt_fail(PT_E_INVALID, "pt_batch_autotune: bad arguments");
    static const int shapes[3] = {64, 32, 16};
    double best = 0;
    int best_tx = 64;
    for (int k = 0; k < 3; k++) {
        bt->tile_x = shapes[k];
        TRY(pt_batch_lookup(bt, a_dev, h, w, ann_dev, nullptr, nullptr));   // eager run + capture
        std::vector<double> ms;
        for (int r = 0; r < reps; r++) {
            TRY(pt_batch_lookup(bt, a_dev, h, w, ann_dev, nullptr, nullptr));
            float e = 0;
            HIPTRY(hipEventElapsedTime(&e, bt->e0, bt->e1));
            ms.push_back(e);
        }
        std::sort(ms.begin(), ms.end());
        const double med = ms[ms.size() / 2];
        if (ms_out) ms_out[k] = med;
        if (k == 0 || med < best) {
            best = med;
            best_tx = shapes[k];
        }
    }
    bt->tile_x = best_tx;
    if (tile_x_out) *tile_x_out = best_tx;
    return PT_OK;
}

extern "C" int pt_batch_destroy(pt_batch *bt) {
    delete bt;
    return PT_OK;
}
