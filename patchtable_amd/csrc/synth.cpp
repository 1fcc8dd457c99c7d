// Seeded synthetic image pairs for tests and benchmarks.
//
// a: per channel a 1/f^2-power (1/|f| amplitude) Gaussian random field, min-max normalised to
//    [0,1], then 40 random opaque rectangles/ellipses (colour U[0,1]^3, centre uniform, each
//    extent U[4%,30%] of min(h,w)).
// b: a warped by d(x,y) = (sx + 8 sin(2 pi y / l1 + p1), sy + 8 sin(2 pi x / l2 + p2)) with
//    sx, sy ~ U[-20,20] px and l ~ U[64,256] px, bilinear sampling with edge clamp, plus
//    N(0, 0.02^2) noise, clipped to [0,1].
// The random field is drawn directly in the frequency domain (independent complex Gaussian
// coefficients scaled by 1/|f|, DC zeroed) and brought to the image domain with one radix-2
// inverse FFT on the enclosing power-of-two square; the real part is kept. This has the same
// distribution as filtering white noise. All arithmetic is double precision and the file is
// compiled without floating-point contraction, so a given (seed, h, w) yields identical images
// on every x86-64 host.
#include "pt_internal.h"

#include <cmath>
#include <complex>
#include <vector>

namespace {

struct Pcg32 {
    uint64_t state, inc;
    explicit Pcg32(uint64_t seed, uint64_t stream = 54u) {
        state = 0u;
        inc = (stream << 1u) | 1u;
        next();
        state += seed;
        next();
    }
    uint32_t next() {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xorshifted >> rot) | (xorshifted << ((32u - rot) & 31u));
    }
    double uniform() { return (next() + 0.5) * (1.0 / 4294967296.0); }   // (0,1)
    double uniform(double lo, double hi) { return lo + (hi - lo) * uniform(); }
    double gauss() {
        double u1 = uniform(), u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
};

typedef std::complex<double> cd;

void fft1d(cd *v, int n, int stride, bool inverse) {
    // Iterative radix-2 Cooley-Tukey on v[0], v[stride], ..., v[(n-1)*stride].
    for (int i = 1, j = 0; i < n; i++) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(v[i * stride], v[j * stride]);
    }
    for (int len = 2; len <= n; len <<= 1) {
        double ang = 2.0 * M_PI / len * (inverse ? 1.0 : -1.0);
        cd wl(std::cos(ang), std::sin(ang));
        for (int i = 0; i < n; i += len) {
            cd w(1.0, 0.0);
            for (int k = 0; k < len / 2; k++) {
                cd u = v[(i + k) * stride];
                cd t = v[(i + k + len / 2) * stride] * w;
                v[(i + k) * stride] = u + t;
                v[(i + k + len / 2) * stride] = u - t;
                w *= wl;
            }
        }
    }
}

void spectral_noise(Pcg32 &rng, int h, int w, std::vector<double> &out) {
    int n = 1;
    while (n < h || n < w) n <<= 1;
    std::vector<cd> g((size_t)n * n);
    for (int fy = 0; fy < n; fy++) {
        int ky = fy <= n / 2 ? fy : fy - n;
        for (int fx = 0; fx < n; fx++) {
            int kx = fx <= n / 2 ? fx : fx - n;
            double re = rng.gauss(), im = rng.gauss();
            double f = std::sqrt((double)kx * kx + (double)ky * ky);
            double s = f > 0 ? 1.0 / f : 0.0;
            g[(size_t)fy * n + fx] = cd(re * s, im * s);
        }
    }
    for (int y = 0; y < n; y++) fft1d(&g[(size_t)y * n], n, 1, true);
    for (int x = 0; x < n; x++) fft1d(&g[x], n, n, true);
    out.resize((size_t)h * w);
    double lo = 1e300, hi = -1e300;
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            double v = g[(size_t)y * n + x].real();
            out[(size_t)y * w + x] = v;
            lo = v < lo ? v : lo;
            hi = v > hi ? v : hi;
        }
    double r = hi > lo ? hi - lo : 1.0;
    for (auto &v : out) v = (v - lo) / r;
}

}  // namespace

extern "C" int pt_synth_pair(uint64_t seed, int h, int w, float *a, float *b) {
    if (h < 1 || w < 1 || !a || !b) return pt_fail(PT_E_INVALID, "pt_synth_pair: bad arguments");
    Pcg32 rng(seed);
    size_t npx = (size_t)h * w;
    std::vector<double> img(npx * 3);
    std::vector<double> plane;
    for (int c = 0; c < 3; c++) {
        spectral_noise(rng, h, w, plane);
        for (size_t i = 0; i < npx; i++) img[i * 3 + c] = plane[i];
    }
    double m = (double)(h < w ? h : w);
    for (int s = 0; s < 40; s++) {
        bool ellipse = rng.uniform() < 0.5;
        double col[3] = {rng.uniform(), rng.uniform(), rng.uniform()};
        double cx = rng.uniform(0.0, (double)w), cy = rng.uniform(0.0, (double)h);
        double rx = 0.5 * rng.uniform(0.04, 0.30) * m, ry = 0.5 * rng.uniform(0.04, 0.30) * m;
        int x0 = (int)std::floor(cx - rx), x1 = (int)std::ceil(cx + rx);
        int y0 = (int)std::floor(cy - ry), y1 = (int)std::ceil(cy + ry);
        if (x0 < 0) x0 = 0;
        if (y0 < 0) y0 = 0;
        if (x1 > w - 1) x1 = w - 1;
        if (y1 > h - 1) y1 = h - 1;
        for (int y = y0; y <= y1; y++)
            for (int x = x0; x <= x1; x++) {
                double dx = (x + 0.5 - cx) / rx, dy = (y + 0.5 - cy) / ry;
                bool inside = ellipse ? (dx * dx + dy * dy <= 1.0) : (std::fabs(dx) <= 1.0 && std::fabs(dy) <= 1.0);
                if (inside)
                    for (int c = 0; c < 3; c++) img[((size_t)y * w + x) * 3 + c] = col[c];
            }
    }
    for (size_t i = 0; i < npx * 3; i++) a[i] = (float)img[i];

    double sx = rng.uniform(-20.0, 20.0), sy = rng.uniform(-20.0, 20.0);
    double l1 = rng.uniform(64.0, 256.0), l2 = rng.uniform(64.0, 256.0);
    double p1 = rng.uniform(0.0, 2.0 * M_PI), p2 = rng.uniform(0.0, 2.0 * M_PI);
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            double fx = x + sx + 8.0 * std::sin(2.0 * M_PI * y / l1 + p1);
            double fy = y + sy + 8.0 * std::sin(2.0 * M_PI * x / l2 + p2);
            if (fx < 0) fx = 0;
            if (fy < 0) fy = 0;
            if (fx > w - 1) fx = w - 1;
            if (fy > h - 1) fy = h - 1;
            int ix = (int)std::floor(fx), iy = (int)std::floor(fy);
            int ix1 = ix + 1 < w ? ix + 1 : ix, iy1 = iy + 1 < h ? iy + 1 : iy;
            double tx = fx - ix, ty = fy - iy;
            for (int c = 0; c < 3; c++) {
                double v00 = img[((size_t)iy * w + ix) * 3 + c], v01 = img[((size_t)iy * w + ix1) * 3 + c];
                double v10 = img[((size_t)iy1 * w + ix) * 3 + c], v11 = img[((size_t)iy1 * w + ix1) * 3 + c];
                double v = (v00 * (1 - tx) + v01 * tx) * (1 - ty) + (v10 * (1 - tx) + v11 * tx) * ty;
                v += 0.02 * rng.gauss();
                if (v < 0) v = 0;
                if (v > 1) v = 1;
                b[((size_t)y * w + x) * 3 + c] = (float)v;
            }
        }
    return PT_OK;
}
