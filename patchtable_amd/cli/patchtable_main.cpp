// patchtable -- command-line front end of the PatchTable drop-in (the match and check_dist modes of
// patchtable_main.cpp:257-390), built on the C++ header shim over the MI355X library.
//
//   patchtable match a.png b.png out.pfm [switches] [-prev_nnf prev.pfm] [-allowed_patches mask.pfm]
//   patchtable check_dist a.png b.png nnf.pfm [nnf_dist_recalc.pfm] [-check 1]
//   patchtable test_descriptor
//
// Output lines and exit codes follow the reference: match prints
// "total_time: T, query_time: Q, mean_dist: M, mean_dist_recomputed: R" (or the verbose report),
// check_dist prints the recomputed mean distance; errors print a message and exit with status 1.
#include <patchtable/patchtable.h>

#include <chrono>
#include <cmath>
#include <map>
#include <random>
#include <string>

namespace {

double wall_time() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// parse_switches (patchtable.cpp:262-296): "-name value" pairs after the positional arguments.
std::map<std::string, std::string> parse_switches(int argc, char **argv) {
    std::map<std::string, std::string> sw;
    for (int i = 0; i < argc; i++)
        if (argv[i][0] == '-' && i + 1 < argc) sw[argv[i]] = argv[i + 1];
    return sw;
}

// get_mean_dist (patchtable_main.cpp:90-101): mean of sqrt(max(NNF_DIST, 0)).
double get_mean_dist(const Array<double> &ann) {
    double s = 0;
    for (int y = 0; y < ann.height(); y++)
        for (int x = 0; x < ann.width(); x++) {
            double d = ann(y, x, NNF_DIST);
            if (d < 0) d = 0;
            s += std::sqrt(d);
        }
    return s / double(ann.height() * ann.width());
}

// recompute_mean_dist (patchtable_main.cpp:103-172): FP64 distance of every match, clamped into
// the database; optional |recomputed - stored| <= 1e-4 check and overwrite of NNF_DIST.
double recompute_mean_dist(bool is_descriptor, const Array<float> &a, const Array<float> &b, Array<double> &ann,
                           int patch_w, bool check, bool overwrite) {
    int ann_w = a.width() - patch_w + 1, ann_h = a.height() - patch_w + 1;
    int bew = b.width() - patch_w + 1, beh = b.height() - patch_w + 1;
    if (is_descriptor) {
        ann_w = a.width();
        ann_h = a.height();
        bew = b.width();
        beh = b.height();
    }
    if (ann.height() != ann_h || ann.width() != ann_w) {
        fprintf(stderr, "expected ann size to match a size (after removing patch_w-1 from a size)\n");
        exit(1);
    }
    double sum = 0;
    for (int ay = 0; ay < ann_h; ay++)
        for (int ax = 0; ax < ann_w; ax++) {
            int bx = (int)ann(ay, ax, NNF_X), by = (int)ann(ay, ax, NNF_Y);
            if (bx < 0) bx = 0;
            else if (bx >= bew) bx = bew - 1;
            if (by < 0) by = 0;
            else if (by >= beh) by = beh - 1;
            double dist = 0;
            if (!is_descriptor) {
                for (int dy = 0; dy < patch_w; dy++)
                    for (int dx = 0; dx < patch_w; dx++)
                        for (int c = 0; c < a.channels(); c++) {
                            double d = (double)a(ay + dy, ax + dx, c) - (double)b(by + dy, bx + dx, c);
                            dist += d * d;
                        }
            } else {
                for (int c = 0; c < a.channels(); c++) {
                    double d = (double)a(ay, ax, c) - (double)b(by, bx, c);
                    dist += d * d;
                }
            }
            if (overwrite) ann(ay, ax, NNF_DIST) = dist;
            if (check && std::fabs(dist - ann(ay, ax, NNF_DIST)) > 1e-4) {
                fprintf(stderr, "recomputed distance %f differs from existing distance %f at %d, %d\n", dist,
                        ann(ay, ax, NNF_DIST), ax, ay);
                exit(1);
            }
            sum += std::sqrt(dist);
        }
    return sum / double(ann_w * ann_h);
}

void usage() {
    printf("patchtable match a.png b.png out.pfm [-prev_nnf prev.pfm] [-allowed_patches mask.pfm] [-is_descriptor 0|1]\n");
    printf("patchtable check_dist a.png b.png nnf.pfm [nnf_dist_recalc.pfm] -- Prints average distance for a -> b nnf\n");
    printf("patchtable test_descriptor\n");
    printf("\n");
    printf("Switches are those of PatchTableParams (-speed 0..9, -limit, -kcoherence, ...).\n");
    printf("\n");
    printf("  Match with accuracy roughly similar to PatchMatch:\n");
    printf("  $ ./patchtable match vidpair0/a.png vidpair0/b.png out.pfm\n");
    exit(1);
}

void test_match(const char *a_filename, const char *b_filename, const char *out_filename, int argc, char **argv) {
    PatchTableParams *p = new PatchTableParams(argc, argv);
    p->calc_exact_dist = true;
    std::map<std::string, std::string> sw = parse_switches(argc, argv);
    bool is_descriptor = false;
    if (sw.count("-is_descriptor")) is_descriptor = atoi(sw["-is_descriptor"].c_str()) != 0;
    p->is_descriptor = is_descriptor;
    Array<double> *prev_nnf = nullptr;
    if (sw.count("-prev_nnf")) prev_nnf = new Array<double>(load_color_image<double>(sw["-prev_nnf"].c_str()));
    Array<int32_t> *allowed = nullptr;
    if (sw.count("-allowed_patches")) {
        Array<double> af(load_color_image<double>(sw["-allowed_patches"].c_str()));
        allowed = new Array<int32_t>(af.height(), af.width());
        for (int y = 0; y < af.height(); y++)
            for (int x = 0; x < af.width(); x++) (*allowed)(y, x) = (int32_t)af(y, x, 0);
    }
    Array<float> a(load_color_image<float>(a_filename));
    Array<float> b(load_color_image<float>(b_filename));
    if (p->verbose) {
        printf("a size: width %d x height %d x channels %d, %s\n", a.width(), a.height(), a.channels(), a_filename);
        printf("b size: width %d x height %d x channels %d, %s\n", b.width(), b.height(), b.channels(), b_filename);
    }
    if (is_descriptor && a.channels() == 3) {
        fprintf(stderr, "-is_descriptor 1 with RGB images (descriptors from gck_no_pad) is not supported\n");
        exit(1);
    }
    double T0 = wall_time();
    PatchTable<> table(p, b, allowed);
    double T1 = wall_time();
    Array<double> ann;
    double latest_time = 1e100;
    for (int iter = 0; iter < p->query_iters; iter++) latest_time = table.lookup(a, ann, prev_nnf);
    save_color_image(ann, out_filename);
    double total_time = latest_time + (T1 - T0);
    double mean_dist = get_mean_dist(ann);
    double mean_dist_recomputed = recompute_mean_dist(is_descriptor, a, b, ann, p->patch_w, p->check_existing, false);
    if (p->is_descriptor) mean_dist = mean_dist_recomputed;
    if (p->verbose) {
        printf("Combined precompute and lookup time: %f\n", total_time);
        printf("mean_dist: %f\n", mean_dist);
        printf("mean_dist_recomputed: %f\n", mean_dist_recomputed);
    } else {
        printf("total_time: %f, query_time: %f, mean_dist: %f, mean_dist_recomputed: %f\n", total_time, latest_time,
               mean_dist, mean_dist_recomputed);
    }
    delete prev_nnf;
    delete allowed;
    // p must outlive the table (the table reads it at lookup); both end with the process here.
}

void check_dist(const char *a_filename, const char *b_filename, const char *nnf_filename, int argc, char **argv) {
    Array<float> a(load_color_image<float>(a_filename));
    Array<float> b(load_color_image<float>(b_filename));
    Array<double> ann(load_color_image<double>(nnf_filename));
    if (ann.channels() != 3) {
        fprintf(stderr, "expected ann to have 3 channels\n");
        exit(1);
    }
    PatchTableParams p(argc, argv);
    std::map<std::string, std::string> sw = parse_switches(argc, argv);
    bool check = sw.count("-check") && atoi(sw["-check"].c_str()) != 0;
    bool overwrite = argc > 4;
    printf("%f\n", recompute_mean_dist(false, a, b, ann, p.patch_w, check, overwrite));
    if (overwrite) save_color_image(ann, argv[4]);
}

// test_descriptor (patchtable_main.cpp:174-240): descriptor mode on uniform random 8-channel
// descriptors with default parameters -- all 8 dims as the grid, then 6 of 8, then 6 of 6 -- must
// give mean distances (over the full 8 dims) that grow in that order, the first below 0.77.
void test_descriptor(int w = 400, int h = 300, int k = 8) {
    Array<float> a0(h, w, k), b0(h, w, k);
    std::mt19937 rng(1);
    std::uniform_real_distribution<float> U(0.f, 1.f);
    for (int i = 0; i < a0.nelems; i++) a0.data[i] = U(rng);
    for (int i = 0; i < b0.nelems; i++) b0.data[i] = U(rng);
    Array<float> a(a0), b(b0);
    PatchTableParams *p = new PatchTableParams();
    p->is_descriptor = true;
    double dL[3];
    for (int reduce = 0; reduce < 3; reduce++) {
        if (reduce == 1) p->grid_ndims = k - 2;
        else if (reduce == 2) {
            p->grid_ndims = p->ndims = k - 2;
            Array<float> ap(h, w, k - 2), bp(h, w, k - 2);
            for (int y = 0; y < h; y++)
                for (int x = 0; x < w; x++)
                    for (int c = 0; c < k - 2; c++) {
                        ap(y, x, c) = a0(y, x, c);
                        bp(y, x, c) = b0(y, x, c);
                    }
            a.assign(ap);
            b.assign(bp);
        }
        printf("test_descriptor: reduce=%d, ndims=%d, b channels: %d\n", reduce, p->grid_ndims, b.channels());
        PatchTable<> btable(p, b);
        Array<double> ann;
        btable.lookup(a, ann);
        if (ann.width() != a.width() || ann.height() != a.height()) {
            fprintf(stderr, "expected ann size %dx%d to match a size %dx%d\n", ann.width(), ann.height(), a.width(),
                    a.height());
            exit(1);
        }
        for (int y = 0; y < ann.height(); y++)
            for (int x = 0; x < ann.width(); x++) {
                int bx = (int)ann(y, x, NNF_X), by = (int)ann(y, x, NNF_Y);
                if ((unsigned)bx >= (unsigned)b.width() || (unsigned)by >= (unsigned)b.height()) {
                    fprintf(stderr, "expected ax, ay=%d, %d => bx, by=%d, %d to be in bounds %dx%d\n", x, y, bx, by,
                            b.width(), b.height());
                    exit(1);
                }
            }
        dL[reduce] = recompute_mean_dist(true, a0, b0, ann, p->patch_w, false, false);
        printf("mean_dist: %f (reduce=%d)\n", dL[reduce], reduce);
    }
    if (dL[0] < dL[1] && dL[1] < dL[2] && dL[0] < 0.77) {
        printf("test_descriptor: OK\n");
    } else {
        fprintf(stderr, "test_descriptor: Failed, distances not as expected\n");
        exit(1);
    }
}

}  // namespace

int main(int argc0, char **argv0) {
    int argc = argc0 - 1;
    char **argv = argv0 + 1;
    if (argc < 1) usage();
    if (!strcmp(argv[0], "match")) {
        if (argc < 4) {
            fprintf(stderr, "expected 4 args in match mode\n");
            exit(1);
        }
        test_match(argv[1], argv[2], argv[3], argc, argv);
    } else if (!strcmp(argv[0], "test_descriptor")) {
        test_descriptor();
    } else if (!strcmp(argv[0], "check_dist")) {
        if (argc <= 3) {
            fprintf(stderr, "expected 3 args in check_dist mode\n");
            exit(1);
        }
        check_dist(argv[1], argv[2], argv[3], argc, argv);
    } else {
        usage();
    }
    return 0;
}
