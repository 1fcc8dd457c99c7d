/*
 * patchtable/patchtable.h -- C++ drop-in for the reference's construct-then-lookup API, over the
 * C ABI of the MI355X library (include/patchtable_amd.h).
 *
 * Mirrors, with the same names and argument meaning:
 *   Array<T>          solver/array.h:251-294 (row-major, sizes/stride/nelems/data; HWC images)
 *   PatchTableParams  patchtable_params.h:23-161, patchtable.cpp:26-387
 *   PatchTable<>      patchtable.h:4359-4491 (constructor, lookup, save)
 *   NNF_X/NNF_Y/NNF_DIST, DIST_INFINITY    patchtable.h:173-177
 * Errors follow the reference: a failing call prints the message to stderr and exits with
 * status 1 (ASSERT2, solver/util.h:33).
 *
 * Build:  g++ app.cpp -I<repo>/include -L<repo>/patchtable_amd -lpatchtable_hip
 */
#ifndef PATCHTABLE_AMD_PATCHTABLE_H
#define PATCHTABLE_AMD_PATCHTABLE_H

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../patchtable_amd.h"

#define NNF_X 0
#define NNF_Y 1
#define NNF_DIST 2
#define DIST_INFINITY 100000

namespace patchtable_amd_detail {
inline void check(int rc) {
    if (rc != PT_OK) {
        fprintf(stderr, "%s\n", pt_last_error());
        exit(1);
    }
}
}  // namespace patchtable_amd_detail

/* Minimal N-d array with the reference's layout: row-major, last dimension fastest. */
template <class real>
class Array {
  public:
    real *data = nullptr;
    std::vector<int> sizes;
    std::vector<int> stride;
    int nelems = 0;

    Array() {}
    Array(int h, int w) { resize(h, w); }
    Array(int h, int w, int c) { resize(h, w, c); }
    Array(const Array &o) { assign(o); }
    Array &operator=(const Array &o) {
        if (this != &o) assign(o);
        return *this;
    }
    ~Array() { delete[] data; }

    /* Changes array sizes; contents are lost (as the reference's resize, array.h:261-294). */
    void resize(const std::vector<int> &s) {
        if (s == sizes && data) return;
        delete[] data;
        sizes = s;
        stride.assign(s.size(), 0);
        nelems = 1;
        for (int i = (int)s.size() - 1; i >= 0; i--) {
            stride[i] = nelems;
            nelems *= s[i];
        }
        data = new real[nelems];
    }
    void resize(int h, int w) { resize(std::vector<int>{h, w}); }
    void resize(int h, int w, int c) { resize(std::vector<int>{h, w, c}); }
    void assign(const Array &o) {
        resize(o.sizes);
        if (nelems) memcpy(data, o.data, sizeof(real) * nelems);
    }
    int height() const { return sizes.size() > 0 ? sizes[0] : 1; }
    int width() const { return sizes.size() > 1 ? sizes[1] : 1; }
    int channels() const { return sizes.size() > 2 ? sizes[2] : 1; }
    int dimensions() const { return (int)sizes.size(); }
    real &operator()(int y, int x) { return data[y * stride[0] + x * stride[1]]; }
    real &operator()(int y, int x, int c) { return data[y * stride[0] + x * stride[1] + c]; }
    const real &operator()(int y, int x) const { return data[y * stride[0] + x * stride[1]]; }
    const real &operator()(int y, int x, int c) const { return data[y * stride[0] + x * stride[1] + c]; }
    void clear(real v = real(0)) {
        for (int i = 0; i < nelems; i++) data[i] = v;
    }
};

/* load_color_image<real> (solver/array.h:1637-1684): .png as 8-bit RGB * (1/255.0), .pfm as
 * stored; h x w x 3 (h x w x 1 for a one-channel .pfm). */
template <class real>
Array<real> load_color_image(const char *filename) {
    int h = 0, w = 0, c = 0;
    patchtable_amd_detail::check(pt_image_size(filename, &h, &w, &c));
    std::vector<float> f((size_t)h * w * c);
    patchtable_amd_detail::check(pt_image_load(filename, f.data(), h, w, c));
    Array<real> ans(h, w, c);
    for (size_t i = 0; i < f.size(); i++) ans.data[i] = (real)f[i];
    return ans;
}

/* save_color_image (solver/array.h:1745-1827) for .pfm files of 1 or 3 channels (an NNF). */
template <class real>
void save_color_image(const Array<real> &a, const char *filename) {
    std::vector<double> d((size_t)a.nelems);
    for (size_t i = 0; i < d.size(); i++) d[i] = (double)a.data[i];
    patchtable_amd_detail::check(pt_pfm_save(filename, d.data(), a.height(), a.width(), a.channels()));
}

/* Parameters: same field names as the reference; defaults, speed presets and switch parsing are
 * the library's restatement of patchtable.cpp:26-387. */
class PatchTableParams : public pt_params {
  public:
    PatchTableParams() { patchtable_amd_detail::check(pt_params_default(this)); }
    PatchTableParams(int argc, const char **argv) {
        patchtable_amd_detail::check(pt_params_default(this));
        set_from_switches(argc, argv);
    }
    PatchTableParams(int argc, char **argv) : PatchTableParams(argc, (const char **)argv) {}
    void set_defaults() { patchtable_amd_detail::check(pt_params_default(this)); }
    void set_speed(int i) { patchtable_amd_detail::check(pt_params_set_speed(this, i)); }
    void set_from_switches(std::string s) { patchtable_amd_detail::check(pt_params_set_switches(this, s.c_str())); }
    void set_from_switches(int argc, const char **argv) {
        /* argv[0] is the program name and is skipped, as parse_switches does. */
        std::string s;
        for (int i = 0; i < argc; i++) {
            if (i) s += " ";
            s += argv[i];
        }
        set_from_switches(s);
    }
    void set_from_switches(int argc, char **argv) { set_from_switches(argc, (const char **)argv); }
};

/* PatchTable<real=float, in_type=float, itype=int32_t> (patchtable.h:4359-4491). p is borrowed
 * and must outlive the table; parameters changed after construction apply to later lookups. */
template <class real = float, class in_type = float, class itype = int32_t>
class PatchTable {
  public:
    pt_table *t = nullptr;
    PatchTableParams *p = nullptr;

    PatchTable(PatchTableParams *p_, const Array<in_type> &b, Array<itype> *allowed_patches = NULL) : p(p_) {
        /* allowed_patches may be larger than the patch grid (the reference asserts >=,
         * patchtable.h:2643-2646); its top-left patch-grid region is what is read. */
        std::vector<int32_t> mask;
        if (allowed_patches) {
            const int pad = p->is_descriptor ? (p->descriptor_padded ? 14 : 0) : 7;
            const int bh = b.height() - pad, bw = b.width() - pad;
            if (allowed_patches->height() < bh || allowed_patches->width() < bw) {
                fprintf(stderr, "expected allowed_patches to be at least the patch grid size %dx%d\n", bw, bh);
                exit(1);
            }
            mask.resize((size_t)bh * bw);
            for (int y = 0; y < bh; y++)
                for (int x = 0; x < bw; x++) mask[(size_t)y * bw + x] = (int32_t)(*allowed_patches)(y, x);
        }
        patchtable_amd_detail::check(pt_table_create(p, (const float *)b.data, b.height(), b.width(), b.channels(),
                                                     allowed_patches ? mask.data() : nullptr, &t));
    }
    ~PatchTable() {
        if (t) pt_table_destroy(t);
    }
    PatchTable(const PatchTable &) = delete;
    PatchTable &operator=(const PatchTable &) = delete;

    /* Fills ann ((h-7) x (w-7) x 3: NNF_X, NNF_Y, NNF_DIST; h x w in descriptor mode); returns the
     * lookup time in seconds. */
    double lookup(const Array<in_type> &a, Array<double> &ann, Array<double> *ann_prev = NULL,
                  Array<float> *coherence_temporal_sv = NULL) {
        const int pad = p->is_descriptor ? (p->descriptor_padded ? 14 : 0) : 7;
        ann.resize(a.height() - pad, a.width() - pad, 3);
        double secs = 0;
        patchtable_amd_detail::check(pt_table_lookup(t, (const float *)a.data, a.height(), a.width(), a.channels(),
                                                     ann.data, ann_prev ? ann_prev->data : nullptr,
                                                     coherence_temporal_sv ? coherence_temporal_sv->data : nullptr,
                                                     &secs));
        return secs;
    }
    void save(std::string filename) { patchtable_amd_detail::check(pt_table_save(t, filename.c_str())); }
};

#endif
