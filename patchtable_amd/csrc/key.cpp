// Host side of the learned patch key: parameters, training (triplet mining from the library's
// own lookups, then Adam steps on the GPU) and the forward pass over an image.
#include "pt_internal.h"
#include "pt_kernels.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
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

constexpr int KI = PT_KEY_IN, KH = PT_KEY_HIDDEN, KO = PT_KEY_OUT;

// PCG32 (O'Neill), the generator of the weight initialisation and of the triplet sampling.
struct Pcg {
    uint64_t s, inc;
    explicit Pcg(uint64_t seed) : s(0), inc((seed << 1) | 1u) {
        next();
        s += 0x853c49e6748fea9bULL ^ seed;
        next();
    }
    uint32_t next() {
        uint64_t old = s;
        s = old * 6364136223846793005ULL + inc;
        uint32_t x = (uint32_t)(((old >> 18) ^ old) >> 27), r = (uint32_t)(old >> 59);
        return (x >> r) | (x << ((32 - r) & 31));
    }
    float uniform() { return (next() >> 8) * (1.0f / 16777216.0f); }
    int below(int n) { return (int)((uint64_t)next() * (uint64_t)n >> 32); }
};

}  // namespace

struct pt_key {
    DevBuf<float> w1, b1, w2, b2, m, v;
    DevBuf<unsigned short> w1bf, w2bf;
    hipStream_t s = nullptr;
    KeyParamsDev dev() {
        KeyParamsDev P;
        P.w1 = w1.p;
        P.b1 = b1.p;
        P.w2 = w2.p;
        P.b2 = b2.p;
        P.w1bf = w1bf.p;
        P.w2bf = w2bf.p;
        P.m = m.p;
        P.v = v.p;
        return P;
    }
    ~pt_key() {
        if (s) (void)hipStreamDestroy(s);
    }
};

namespace {

int upload_weights(pt_key *k, const float *w1, const float *b1, const float *w2, const float *b2) {
    HIPTRY(hipMemcpyAsync(k->w1.p, w1, KH * KI * 4, hipMemcpyHostToDevice, k->s));
    HIPTRY(hipMemcpyAsync(k->b1.p, b1, KH * 4, hipMemcpyHostToDevice, k->s));
    HIPTRY(hipMemcpyAsync(k->w2.p, w2, KO * KH * 4, hipMemcpyHostToDevice, k->s));
    HIPTRY(hipMemcpyAsync(k->b2.p, b2, KO * 4, hipMemcpyHostToDevice, k->s));
    TRY(pt_key_to_bf16_launch(k->w1.p, KH * KI, k->w1bf.p, k->s));
    TRY(pt_key_to_bf16_launch(k->w2.p, KO * KH, k->w2bf.p, k->s));
    HIPTRY(hipMemsetAsync(k->m.p, 0, PT_KEY_NPARAM * 4, k->s));
    HIPTRY(hipMemsetAsync(k->v.p, 0, PT_KEY_NPARAM * 4, k->s));
    HIPTRY(hipStreamSynchronize(k->s));
    return PT_OK;
}

struct Work {
    DevBuf<float> Z, K, dK, dH, gW1, gb1, gW2, gb2, loss, part;
    int ensure(int nt) {
        const size_t n = (size_t)3 * nt;
        TRY(part.ensure(pt_key_grad_part_floats(3 * nt)));
        TRY(Z.ensure(n * KH));
        TRY(K.ensure(n * KO));
        TRY(dK.ensure(n * KO));
        TRY(dH.ensure(n * KH));
        TRY(gW1.ensure(KH * KI));
        TRY(gb1.ensure(KH));
        TRY(gW2.ensure(KO * KH));
        TRY(gb2.ensure(KO));
        TRY(loss.ensure(1));
        return PT_OK;
    }
    KeyWorkDev dev() {
        KeyWorkDev W;
        W.Z = Z.p;
        W.K = K.p;
        W.dK = dK.p;
        W.dH = dH.p;
        W.gW1 = gW1.p;
        W.gb1 = gb1.p;
        W.gW2 = gW2.p;
        W.gb2 = gb2.p;
        W.loss = loss.p;
        W.part = part.p;
        return W;
    }
};

}  // namespace

extern "C" int pt_key_create(uint64_t seed, pt_key **out) {
    if (!out) return pt_fail(PT_E_INVALID, "pt_key_create: null argument");
    *out = nullptr;
    pt_key *k = new pt_key();
    int rc = PT_OK;
    if (hipStreamCreateWithFlags(&k->s, hipStreamNonBlocking) != hipSuccess) rc = pt_fail(PT_E_HIP, "stream");
    if (rc == PT_OK) rc = k->w1.ensure(KH * KI);
    if (rc == PT_OK) rc = k->b1.ensure(KH);
    if (rc == PT_OK) rc = k->w2.ensure(KO * KH);
    if (rc == PT_OK) rc = k->b2.ensure(KO);
    if (rc == PT_OK) rc = k->m.ensure(PT_KEY_NPARAM);
    if (rc == PT_OK) rc = k->v.ensure(PT_KEY_NPARAM);
    if (rc == PT_OK) rc = k->w1bf.ensure(KH * KI);
    if (rc == PT_OK) rc = k->w2bf.ensure(KO * KH);
    if (rc == PT_OK) {
        // He-uniform initialisation
        Pcg g(seed);
        std::vector<float> w1(KH * KI), b1(KH, 0.f), w2(KO * KH), b2(KO, 0.f);
        const float a1 = std::sqrt(6.f / KI), a2 = std::sqrt(6.f / KH);
        for (auto &x : w1) x = (2.f * g.uniform() - 1.f) * a1;
        for (auto &x : w2) x = (2.f * g.uniform() - 1.f) * a2;
        rc = upload_weights(k, w1.data(), b1.data(), w2.data(), b2.data());
    }
    if (rc != PT_OK) {
        delete k;
        return rc;
    }
    *out = k;
    return PT_OK;
}

extern "C" int pt_key_destroy(pt_key *k) {
    delete k;
    return PT_OK;
}

extern "C" int pt_key_set_weights(pt_key *k, const float *w1, const float *b1, const float *w2, const float *b2) {
    if (!k || !w1 || !b1 || !w2 || !b2) return pt_fail(PT_E_INVALID, "pt_key_set_weights: null argument");
    return upload_weights(k, w1, b1, w2, b2);
}

extern "C" int pt_key_get_weights(const pt_key *k, float *w1, float *b1, float *w2, float *b2, uint16_t *w1bf,
                                  uint16_t *w2bf) {
    if (!k) return pt_fail(PT_E_INVALID, "pt_key_get_weights: null key");
    HIPTRY(hipStreamSynchronize(k->s));
    if (w1) HIPTRY(hipMemcpy(w1, k->w1.p, KH * KI * 4, hipMemcpyDeviceToHost));
    if (b1) HIPTRY(hipMemcpy(b1, k->b1.p, KH * 4, hipMemcpyDeviceToHost));
    if (w2) HIPTRY(hipMemcpy(w2, k->w2.p, KO * KH * 4, hipMemcpyDeviceToHost));
    if (b2) HIPTRY(hipMemcpy(b2, k->b2.p, KO * 4, hipMemcpyDeviceToHost));
    if (w1bf) HIPTRY(hipMemcpy(w1bf, k->w1bf.p, KH * KI * 2, hipMemcpyDeviceToHost));
    if (w2bf) HIPTRY(hipMemcpy(w2bf, k->w2bf.p, KO * KH * 2, hipMemcpyDeviceToHost));
    return PT_OK;
}

// Device copy of the network (BF16 weights, FP32 biases) into caller buffers, ordered on s.
int pt_key_snapshot(const pt_key *k, unsigned short *w1, float *b1, unsigned short *w2, float *b2, hipStream_t s) {
    HIPTRY(hipStreamSynchronize(k->s));
    HIPTRY(hipMemcpyAsync(w1, k->w1bf.p, KH * KI * 2, hipMemcpyDeviceToDevice, s));
    HIPTRY(hipMemcpyAsync(b1, k->b1.p, KH * 4, hipMemcpyDeviceToDevice, s));
    HIPTRY(hipMemcpyAsync(w2, k->w2bf.p, KO * KH * 2, hipMemcpyDeviceToDevice, s));
    HIPTRY(hipMemcpyAsync(b2, k->b2.p, KO * 4, hipMemcpyDeviceToDevice, s));
    return PT_OK;
}

// Key planes of an image (device), 8 planes of (h+7) x (w+7): the descriptor layout of a table.
int pt_key_forward(const pt_key *k, const float *img_dev, int h, int w, float *planes_dev, hipStream_t s) {
    return pt_key_forward_launch(img_dev, h, w, k->w1bf.p, k->b1.p, k->w2bf.p, k->b2.p, planes_dev, s);
}

extern "C" int pt_key_forward_dev(const pt_key *k, const float *img_dev, int h, int w, float *planes_dev) {
    if (!k || !img_dev || !planes_dev || h < 8 || w < 8 || !PT_SIZE_OK(h, w)) return pt_fail(PT_E_INVALID, "pt_key_forward_dev: bad arguments");
    TRY(pt_key_forward(k, img_dev, h, w, planes_dev, k->s));
    HIPTRY(hipStreamSynchronize(k->s));
    return PT_OK;
}

// Gradients of the triplet loss on one batch (no update), for tests: triplet patch positions
// pos[3*t + role] as XY_TO_INT (anchor in a, positive and negative in b; both h x w). The
// positions are checked on the host before the gather.
extern "C" int pt_key_gradients(pt_key *k, const float *a_dev, const float *b_dev, int h, int w, const int32_t *pos_dev,
                                int nt, float margin, float *gW1, float *gb1, float *gW2, float *gb2, float *loss) {
    if (!k || !a_dev || !b_dev || !pos_dev || nt < 1 || h < 8 || w < 8 || !PT_SIZE_OK(h, w))
        return pt_fail(PT_E_INVALID, "pt_key_gradients: bad arguments");
    std::vector<int32_t> hp((size_t)nt * 3);
    HIPTRY(hipMemcpy(hp.data(), pos_dev, hp.size() * 4, hipMemcpyDeviceToHost));
    for (int32_t v : hp)
        if (v < 0 || PT_INT_TO_X(v) > w - 8 || PT_INT_TO_Y(v) > h - 8)
            return pt_fail(PT_E_INVALID, "pt_key_gradients: triplet position outside the patch grid");
    DevBuf<unsigned short> X;
    TRY(X.ensure((size_t)nt * 3 * KI));
    Work W;
    TRY(W.ensure(nt));
    TRY(pt_key_extract_launch(a_dev, b_dev, w, pos_dev, nullptr, nt, X.p, k->s));
    HIPTRY(hipMemsetAsync(W.loss.p, 0, 4, k->s));
    TRY(pt_key_train_step(X.p, nt, margin, k->dev(), W.dev(), 0.f, 0.9f, 0.999f, 1e-8f, 1, 0.f, k->s));
    HIPTRY(hipStreamSynchronize(k->s));
    if (gW1) HIPTRY(hipMemcpy(gW1, W.gW1.p, KH * KI * 4, hipMemcpyDeviceToHost));
    if (gb1) HIPTRY(hipMemcpy(gb1, W.gb1.p, KH * 4, hipMemcpyDeviceToHost));
    if (gW2) HIPTRY(hipMemcpy(gW2, W.gW2.p, KO * KH * 4, hipMemcpyDeviceToHost));
    if (gb2) HIPTRY(hipMemcpy(gb2, W.gb2.p, KO * 4, hipMemcpyDeviceToHost));
    if (loss) HIPTRY(hipMemcpy(loss, W.loss.p, 4, hipMemcpyDeviceToHost));
    return PT_OK;
}

// Training: for every pair (a_i, b_i), `tpp` anchors drawn uniformly from the query patches;
// the positive is the anchor's match in b_i -- the library's own lookup with parameters p, or
// (exact) the exhaustive search's nearest neighbour (BF16 scoring on the matrix cores, the 8 best
// rescored in exact FP32) -- and the negative the positive moved by 2..6 pixels in each axis
// (half the triplets) or a uniformly random database patch. The triplets of all pairs are
// shuffled, then `steps` Adam steps of `batch` triplets each (epochs wrap around) minimise the
// margin loss on squared key distances.
namespace {

int key_train(pt_key *k, pt_params *p, bool exact, int npairs, const float *const *a_dev, const float *const *b_dev,
              const int *h, const int *w, int tpp, int steps, int batch, float lr, float margin, uint64_t seed,
              float *loss_per_step) {
    if (!k || (!exact && !p) || npairs < 1 || !a_dev || !b_dev || !h || !w || tpp < 1 || steps < 0 || batch < 1)
        return pt_fail(PT_E_INVALID, "pt_key_train: bad arguments");
    for (int pi = 0; pi < npairs; pi++)
        if (h[pi] < 8 || w[pi] < 8 || !PT_SIZE_OK(h[pi], w[pi]))
            return pt_fail(PT_E_INVALID, "pt_key_train: bad image size");
    Pcg g(seed);
    const long T = (long)npairs * tpp;
    DevBuf<unsigned short> X;
    TRY(X.ensure((size_t)T * 3 * KI));
    // triplets of all pairs go to shuffled slots of X
    std::vector<int64_t> slot(T);
    for (long i = 0; i < T; i++) slot[i] = i;
    for (long i = T - 1; i > 0; i--) std::swap(slot[i], slot[(long)(g.next() % (uint64_t)(i + 1))]);
    DevBuf<double> ann;
    DevBuf<int32_t> pos, qpos, npos;
    DevBuf<float> ndist;
    DevBuf<int64_t> dslot;
    TRY(dslot.ensure(T));
    HIPTRY(hipMemcpy(dslot.p, slot.data(), T * 8, hipMemcpyHostToDevice));
    std::vector<double> hann;
    std::vector<int32_t> hq, hn;
    for (int pi = 0; pi < npairs; pi++) {
        const int aw = w[pi] - 7, ah = h[pi] - 7;
        std::vector<int32_t> hp((size_t)tpp * 3);
        auto negative = [&](int bx, int by, int &nx, int &ny) {
            if (g.next() & 1u) {
                const int dx = (2 + g.below(5)) * (g.next() & 1u ? 1 : -1), dy = (2 + g.below(5)) * (g.next() & 1u ? 1 : -1);
                nx = std::min(std::max(bx + dx, 0), aw - 1);
                ny = std::min(std::max(by + dy, 0), ah - 1);
                if (nx == bx && ny == by) nx = (bx + aw / 2) % aw;
            } else {
                nx = g.below(aw);
                ny = g.below(ah);
            }
        };
        if (!exact) {
            pt_table *t = nullptr;
            TRY(pt_table_create_dev(p, b_dev[pi], h[pi], w[pi], 3, nullptr, &t));
            TRY(ann.ensure((size_t)aw * ah * 3));
            int rc = pt_table_lookup_dev(t, a_dev[pi], h[pi], w[pi], 3, ann.p, nullptr, nullptr, nullptr);
            pt_table_destroy(t);
            if (rc != PT_OK) return rc;
            hann.resize((size_t)aw * ah * 3);
            HIPTRY(hipMemcpy(hann.data(), ann.p, hann.size() * 8, hipMemcpyDeviceToHost));
            for (int t2 = 0; t2 < tpp; t2++) {
                const int x = g.below(aw), y = g.below(ah);
                const double *q = &hann[((size_t)y * aw + x) * 3];
                const int bx = (int)q[0], by = (int)q[1];
                int nx, ny;
                negative(bx, by, nx, ny);
                hp[(size_t)t2 * 3 + 0] = PT_XY_TO_INT(x, y);
                hp[(size_t)t2 * 3 + 1] = PT_XY_TO_INT(bx, by);
                hp[(size_t)t2 * 3 + 2] = PT_XY_TO_INT(nx, ny);
            }
        } else {
            hq.resize(tpp);
            for (int t2 = 0; t2 < tpp; t2++) {
                const int x = g.below(aw), y = g.below(ah);
                hq[t2] = PT_XY_TO_INT(x, y);
            }
            TRY(qpos.ensure(tpp));
            TRY(npos.ensure(tpp));
            TRY(ndist.ensure(tpp));
            HIPTRY(hipMemcpy(qpos.p, hq.data(), (size_t)tpp * 4, hipMemcpyHostToDevice));
            TRY(pt_brute_lookup_dev(a_dev[pi], h[pi], w[pi], b_dev[pi], h[pi], w[pi], qpos.p, tpp, PT_BRUTE_BF16,
                                    npos.p, ndist.p, nullptr, nullptr, nullptr));
            hn.resize(tpp);
            HIPTRY(hipMemcpy(hn.data(), npos.p, (size_t)tpp * 4, hipMemcpyDeviceToHost));
            for (int t2 = 0; t2 < tpp; t2++) {
                const int bx = PT_INT_TO_X(hn[t2]), by = PT_INT_TO_Y(hn[t2]);
                int nx, ny;
                negative(bx, by, nx, ny);
                hp[(size_t)t2 * 3 + 0] = hq[t2];
                hp[(size_t)t2 * 3 + 1] = hn[t2];
                hp[(size_t)t2 * 3 + 2] = PT_XY_TO_INT(nx, ny);
            }
        }
        TRY(pos.ensure(hp.size()));
        HIPTRY(hipMemcpy(pos.p, hp.data(), hp.size() * 4, hipMemcpyHostToDevice));
        TRY(pt_key_extract_launch(a_dev[pi], b_dev[pi], w[pi], pos.p, dslot.p + (size_t)pi * tpp, tpp, X.p, k->s));
        HIPTRY(hipStreamSynchronize(k->s));
    }
    Work W;
    const int nb = (int)std::min<long>(batch, T);
    TRY(W.ensure(nb));
    long cursor = 0;
    for (int st = 0; st < steps; st++) {
        if (cursor + nb > T) cursor = 0;
        HIPTRY(hipMemsetAsync(W.loss.p, 0, 4, k->s));
        TRY(pt_key_train_step(X.p + (size_t)cursor * 3 * KI, nb, margin, k->dev(), W.dev(), lr, 0.9f, 0.999f, 1e-8f,
                              st + 1, 1e-4f, k->s));
        cursor += nb;
        if (loss_per_step) {
            float l = 0.f;
            HIPTRY(hipMemcpyAsync(&l, W.loss.p, 4, hipMemcpyDeviceToHost, k->s));
            HIPTRY(hipStreamSynchronize(k->s));
            loss_per_step[st] = l;
        }
    }
    HIPTRY(hipStreamSynchronize(k->s));
    return PT_OK;
}

}  // namespace

extern "C" int pt_key_train(pt_key *k, pt_params *p, int npairs, const float *const *a_dev, const float *const *b_dev,
                            const int *h, const int *w, int tpp, int steps, int batch, float lr, float margin,
                            uint64_t seed, float *loss_per_step) {
    return key_train(k, p, false, npairs, a_dev, b_dev, h, w, tpp, steps, batch, lr, margin, seed, loss_per_step);
}

extern "C" int pt_key_train_exact(pt_key *k, int npairs, const float *const *a_dev, const float *const *b_dev,
                                  const int *h, const int *w, int tpp, int steps, int batch, float lr, float margin,
                                  uint64_t seed, float *loss_per_step) {
    return key_train(k, nullptr, true, npairs, a_dev, b_dev, h, w, tpp, steps, batch, lr, margin, seed,
                     loss_per_step);
}
