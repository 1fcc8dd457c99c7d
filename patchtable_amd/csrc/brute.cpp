// Host side of the exhaustive search: argument checks, buffers, launches and stage timing.
#include "pt_internal.h"
#include "pt_kernels.h"

#include <hip/hip_runtime.h>

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

// Query list: nq positions (device, XY_TO_INT patch coordinates) checked against the ah x aw
// query, or every query patch in raster order when qpos_dev is NULL (nq must then match).
int check_queries(const int32_t *qpos_dev, int nq, int ah, int aw) {
    const long all = (long)(ah - 7) * (aw - 7);
    if (ah < 8 || aw < 8) return pt_fail(PT_E_INVALID, "query image must be at least 8x8");
    if (nq < 1) return pt_fail(PT_E_INVALID, "no query patches");
    if (!qpos_dev) {
        if (nq != all) return pt_fail(PT_E_INVALID, "without positions, nq must be (h-7)*(w-7)");
        return PT_OK;
    }
    std::vector<int32_t> h(nq);
    HIPTRY(hipMemcpy(h.data(), qpos_dev, (size_t)nq * 4, hipMemcpyDeviceToHost));
    for (int32_t v : h)
        if (v < 0 || PT_INT_TO_X(v) > aw - 8 || PT_INT_TO_Y(v) > ah - 8)
            return pt_fail(PT_E_INVALID, "query position outside the patch grid");
    return PT_OK;
}

}  // namespace

extern "C" int pt_brute_lookup_dev(const float *a_dev, int ah, int aw, const float *b_dev, int bh, int bw,
                                   const int32_t *qpos_dev, int nq, int precision, int32_t *pos_out_dev,
                                   float *dist_out_dev, int32_t *cand_out_dev, float *score_out_dev, double *stage_ms) {
    if (!a_dev || !b_dev || !pos_out_dev || !dist_out_dev) return pt_fail(PT_E_INVALID, "pt_brute_lookup_dev: null argument");
    if (precision != PT_BRUTE_BF16 && precision != PT_BRUTE_FP8)
        return pt_fail(PT_E_INVALID, "precision must be PT_BRUTE_BF16 or PT_BRUTE_FP8");
    if (bh < 8 || bw < 8) return pt_fail(PT_E_INVALID, "database image must be at least 8x8");
    if (!PT_SIZE_OK(bh, bw) || !PT_SIZE_OK(ah, aw))
        return pt_fail(PT_E_INVALID, "image too large for XY_TO_INT packing");
    TRY(check_queries(qpos_dev, nq, ah, aw));
    const long nd = (long)(bh - 7) * (bw - 7);
    if (nd > 0x7fffffffL / 8) return pt_fail(PT_E_INVALID, "database too large");
    const int fp8 = precision == PT_BRUTE_FP8;
    const size_t eb = fp8 ? 1 : 2;
    hipStream_t s = nullptr;
    DevBuf<unsigned char> qv, dv;
    DevBuf<float> qn, qsc, dn, dsc, cscore;
    DevBuf<int32_t> cand;
    TRY(qv.ensure((size_t)nq * 192 * eb));
    TRY(dv.ensure((size_t)nd * 192 * eb));
    TRY(qn.ensure(nq));
    TRY(dn.ensure(nd));
    TRY(qsc.ensure(nq));
    TRY(dsc.ensure(nd));
    int32_t *cd = cand_out_dev;
    float *cs = score_out_dev;
    if (!cd) {
        TRY(cand.ensure((size_t)nq * 8));
        cd = cand.p;
    }
    if (!cs) {
        TRY(cscore.ensure((size_t)nq * 8));
        cs = cscore.p;
    }
    hipEvent_t ev[4];
    for (auto &e : ev) HIPTRY(hipEventCreate(&e));
    int rc = PT_OK;
    do {
        if (hipEventRecord(ev[0], s) != hipSuccess) { rc = pt_fail(PT_E_HIP, "event"); break; }
        rc = pt_brute_vectors_launch(a_dev, aw, qpos_dev, nq, aw - 7, fp8, qv.p, qn.p, qsc.p, s);
        // FP8 database vectors are stored negated with a doubled scale (the search then yields -2 a.b)
        if (rc == PT_OK) rc = pt_brute_vectors_launch(b_dev, bw, nullptr, nd, bw - 7, fp8 ? 2 : 0, dv.p, dn.p, dsc.p, s);
        if (rc != PT_OK) break;
        if (hipEventRecord(ev[1], s) != hipSuccess) { rc = pt_fail(PT_E_HIP, "event"); break; }
        rc = pt_brute_mfma_launch(qv.p, qsc.p, nq, dv.p, dn.p, dsc.p, (int)nd, fp8, cd, cs, s);
        if (rc != PT_OK) break;
        if (hipEventRecord(ev[2], s) != hipSuccess) { rc = pt_fail(PT_E_HIP, "event"); break; }
        rc = pt_brute_rescore_launch(a_dev, aw, qpos_dev, nq, aw - 7, b_dev, bw, bw - 7, cd, pos_out_dev, dist_out_dev, s);
        if (rc != PT_OK) break;
        if (hipEventRecord(ev[3], s) != hipSuccess) { rc = pt_fail(PT_E_HIP, "event"); break; }
        if (hipEventSynchronize(ev[3]) != hipSuccess) { rc = pt_fail(PT_E_HIP, "brute-force search failed"); break; }
        if (stage_ms)
            for (int k = 0; k < 3; k++) {
                float ms = 0.f;
                (void)hipEventElapsedTime(&ms, ev[k], ev[k + 1]);
                stage_ms[k] = ms;
            }
    } while (0);
    for (auto &e : ev) (void)hipEventDestroy(e);
    return rc;
}

extern "C" int pt_brute_exact_f64_dev(const float *a_dev, int ah, int aw, const float *b_dev, int bh, int bw,
                                      const int32_t *qpos_dev, int nq, int32_t *pos_out_dev, double *dist_out_dev) {
    if (!a_dev || !b_dev || !pos_out_dev || !dist_out_dev)
        return pt_fail(PT_E_INVALID, "pt_brute_exact_f64_dev: null argument");
    if (bh < 8 || bw < 8) return pt_fail(PT_E_INVALID, "database image must be at least 8x8");
    if (!PT_SIZE_OK(bh, bw) || !PT_SIZE_OK(ah, aw))
        return pt_fail(PT_E_INVALID, "image too large for XY_TO_INT packing");
    TRY(check_queries(qpos_dev, nq, ah, aw));
    const long nd = (long)(bh - 7) * (bw - 7);
    int per = 0;
    const int nblk = pt_brute_f64_blocks(nd, &per);
    DevBuf<double> pd;
    DevBuf<int32_t> pi;
    TRY(pd.ensure((size_t)nblk * nq));
    TRY(pi.ensure((size_t)nblk * nq));
    TRY(pt_brute_f64_launch(a_dev, aw, qpos_dev, nq, aw - 7, b_dev, bw, bw - 7, nd, pd.p, pi.p, pos_out_dev,
                            dist_out_dev, nullptr));
    HIPTRY(hipStreamSynchronize(nullptr));
    return PT_OK;
}
