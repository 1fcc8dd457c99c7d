// Parameters of the PatchTable drop-in: defaults, speed presets and switch parsing.
//
// Mirrors PatchTableParams (patchtable/patchtable_params.h:23-161) and its methods in
// patchtable/patchtable.cpp:10-387 with the same switch names, defaults and presets. The
// reference's quirks are kept because they decide the effective parameters of every preset:
//   - set_from_switches(string) builds an argv from the tokens and parse_switches() starts at
//     index 1 (patchtable.cpp:12, 152-161), so the first token of the string is dropped. Every
//     speed preset therefore leaves ndims unchanged ("-ndims" is its first switch).
//   - "-limit n" stores int(1e6 * n) (patchtable.cpp:171).
// Errors that the reference turns into fprintf + exit(1) return PT_E_INVALID here.
#include "pt_internal.h"

#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

using std::map;
using std::string;
using std::vector;

namespace {

vector<string> split_on(const string &s, char c) {
    // Same splitting as str_split (solver/util.cpp:42-59): empty fields are kept.
    vector<string> out;
    string::size_type pos = 0, last = 0;
    for (;;) {
        pos = s.find(c, last);
        if (pos == string::npos) {
            out.push_back(s.substr(last));
            break;
        }
        out.push_back(s.substr(last, pos - last));
        last = pos + 1;
    }
    return out;
}

map<string, string> switch_map(const vector<string> &argv) {
    // parse_switches (patchtable.cpp:10-20): argv[0] is skipped; "-name value" pairs.
    map<string, string> ans;
    for (size_t i = 1; i < argv.size(); i++) {
        if (!argv[i].empty() && argv[i][0] == '-' && i + 1 < argv.size()) {
            ans[argv[i]] = argv[i + 1];
            i += 1;
        }
    }
    return ans;
}

int apply_switches(pt_params *p, const vector<string> &argv);

const char *speed_preset(int i) {
    // patchtable.cpp:139-148, verbatim switch strings of the ten presets.
    static const char *presets[10] = {
        "-ndims 8 -nchroma 1 -limit 10.000000 -do_rs 1 -run_dt 1 -partition_step 1 -dt_iters -1 -do_prop 1 -query_step 1 -kcoherence 20 -kcoherence_iter -1 -prop_iters 2 -spatial 1",
        "-ndims 6 -nchroma 1 -limit 100.000000 -do_rs 1 -run_dt 1 -partition_step 1 -dt_iters -1 -do_prop 1 -query_step 1 -kcoherence 10 -kcoherence_iter 0 -prop_iters 2 -spatial 1",
        "-ndims 6 -nchroma 1 -limit 100.000000 -do_rs 1 -run_dt 1 -partition_step 1 -dt_iters -1 -do_prop 1 -query_step 1 -kcoherence 10 -kcoherence_iter 0 -prop_iters 2 -spatial 0",
        "-ndims 6 -nchroma 1 -limit 100.000000 -do_rs 0 -run_dt 1 -partition_step 1 -dt_iters -1 -do_prop 1 -query_step 1 -kcoherence 2 -kcoherence_iter -1 -prop_iters 2 -spatial 0",
        "-ndims 7 -nchroma 1 -limit 10.000000 -do_rs 1 -run_dt 1 -partition_step 1 -dt_iters -1 -do_prop 1 -query_step 2 -kcoherence 20 -kcoherence_iter 0 -prop_iters 1 -spatial 1",
        "-ndims 8 -nchroma 1 -limit 100.000000 -do_rs 1 -run_dt 1 -partition_step 1 -dt_iters -1 -do_prop 1 -query_step 2 -kcoherence 4 -kcoherence_iter 0 -prop_iters 1 -spatial 1",
        "-ndims 6 -nchroma 1 -limit 100.000000 -do_rs 0 -run_dt 1 -partition_step 1 -dt_iters -1 -do_prop 1 -query_step 2 -kcoherence 2 -kcoherence_iter -1 -prop_iters 1 -spatial 0",
        "-ndims 7 -nchroma 1 -limit 1.000000 -do_rs 1 -run_dt 1 -partition_step 1 -dt_iters -1 -do_prop 1 -query_step 4 -kcoherence 3 -kcoherence_iter 0 -prop_iters 1 -spatial 1",
        "-ndims 6 -nchroma 1 -limit 100.000000 -do_rs 1 -run_dt 1 -partition_step 1 -dt_iters -1 -do_prop 1 -query_step 6 -kcoherence 5 -kcoherence_iter 0 -prop_iters 1 -spatial 0",
        "-ndims 6 -nchroma 1 -limit 100.000000 -do_rs 0 -run_dt 1 -partition_step 1 -dt_iters -1 -do_prop 1 -query_step 10 -kcoherence 5 -kcoherence_iter 0 -prop_iters 1 -spatial 0",
    };
    return (i >= 0 && i <= 9) ? presets[i] : nullptr;
}

int set_from_string(pt_params *p, const char *s) {
    vector<string> tokens = split_on(string(s ? s : ""), ' ');
    vector<string> argv;
    for (auto &t : tokens) {
        if (!t.empty() && !(t.size() == 1 && t[0] == ' ')) argv.push_back(t);
    }
    return apply_switches(p, argv);
}

int to_int(const string &s) { return atoi(s.c_str()); }
double to_double(const string &s) { return atof(s.c_str()); }

int apply_switches(pt_params *p, const vector<string> &argv) {
    map<string, string> sw = switch_map(argv);
    auto has = [&](const char *k) { return sw.count(k) > 0; };
    if (has("-speed")) {
        int rc = pt_params_set_speed(p, to_int(sw["-speed"]));
        if (rc != PT_OK) return rc;
    }
    if (has("-limit")) p->limit = int(1000 * 1000 * to_double(sw["-limit"]));
    if (has("-enrich_limit")) p->enrich_limit = int(1000 * 1000 * to_double(sw["-enrich_limit"]));
    if (has("-regular_grid")) p->regular_grid = to_int(sw["-regular_grid"]);
    if (has("-table_knn")) p->table_knn = to_int(sw["-table_knn"]);
    if (has("-randomize_dt")) p->randomize_dt = to_int(sw["-randomize_dt"]) != 0;
    if (has("-parallel_dt")) p->parallel_dt = to_int(sw["-parallel_dt"]) != 0;
    if (has("-dt_threads")) p->dt_threads = to_int(sw["-dt_threads"]);
    if (has("-threads")) p->threads = to_int(sw["-threads"]);
    if (has("-verbose")) p->verbose = to_int(sw["-verbose"]);
    if (has("-populate_nearest")) p->populate_nearest = to_int(sw["-populate_nearest"]) != 0;
    if (has("-populate_random")) p->populate_random = to_int(sw["-populate_random"]) != 0;
    if (has("-partition_step")) p->partition_step = to_int(sw["-partition_step"]);
    if (has("-ndims")) p->ndims = to_int(sw["-ndims"]);
    if (has("-nchroma")) p->nchroma = to_int(sw["-nchroma"]);
    if (has("-ntables")) p->ntables = to_int(sw["-ntables"]);
    if (has("-nslices")) {
        vector<string> nl = split_on(sw["-nslices"], ',');
        if ((int)nl.size() > PT_MAX_DIMS) return pt_fail(PT_E_INVALID, "-nslices: too many dimensions");
        p->ndims = (int)nl.size();
        p->nslices_count = (int)nl.size();
        for (size_t i = 0; i < nl.size(); i++) p->nslices[i] = to_int(nl[i]);
    }
    if (has("-run_dt")) p->run_dt = to_int(sw["-run_dt"]) != 0;
    if (has("-do_prop")) p->do_prop = to_int(sw["-do_prop"]);
    if (has("-prop_y_step")) p->prop_y_step = (int)to_double(sw["-prop_y_step"]);
    if (has("-do_rs")) p->do_rs = to_int(sw["-do_rs"]);
    if (has("-spatial")) p->spatial = to_int(sw["-spatial"]);
    if (has("-prob_rs")) p->prob_rs = to_double(sw["-prob_rs"]);
    if (has("-load")) {
        if (sw["-load"].size() >= sizeof(p->load_filename)) return pt_fail(PT_E_INVALID, "-load: path too long");
        strcpy(p->load_filename, sw["-load"].c_str());
    }
    if (has("-save")) {
        if (sw["-save"].size() >= sizeof(p->save_filename)) return pt_fail(PT_E_INVALID, "-save: path too long");
        strcpy(p->save_filename, sw["-save"].c_str());
    }
    if (has("-check_existing")) p->check_existing = to_int(sw["-check_existing"]) != 0;
    if (has("-patch_w")) p->patch_w = to_int(sw["-patch_w"]);
    if (has("-pca_samples")) p->pca_samples = to_int(sw["-pca_samples"]);
    if (has("-pca_mean")) p->pca_mean = to_int(sw["-pca_mean"]) != 0;
    if (has("-pca_subtract_mean")) p->pca_subtract_mean = to_int(sw["-pca_subtract_mean"]) != 0;
    if (has("-ann")) p->ann = to_int(sw["-ann"]) != 0;
    if (has("-ann_agrid")) p->ann_agrid = to_int(sw["-ann_agrid"]);
    if (has("-ann_bgrid")) p->ann_bgrid = to_int(sw["-ann_bgrid"]);
    if (has("-ann_eps")) p->ann_eps = to_double(sw["-ann_eps"]);
    if (has("-dt_knn")) p->dt_knn = to_int(sw["-dt_knn"]);
    if (has("-prop_iters")) p->prop_iters = to_int(sw["-prop_iters"]);
    if (has("-dt_iters")) p->dt_iters = to_int(sw["-dt_iters"]);
    if (has("-query_step")) p->query_step = to_int(sw["-query_step"]);
    if (has("-coherence_spatial")) p->coherence_spatial = to_double(sw["-coherence_spatial"]);
    if (has("-coherence_temporal")) p->coherence_temporal = to_double(sw["-coherence_temporal"]);
    if (has("-recalc_dist_temporal")) p->recalc_dist_temporal = to_int(sw["-recalc_dist_temporal"]) != 0;
    if (has("-dt_algo")) {
        const string &m = sw["-dt_algo"];
        if (m == "raster") p->dt_algo = PT_DT_ALGO_RASTER;
        else if (m == "prop") p->dt_algo = PT_DT_ALGO_PROP;
        else if (m == "brute") p->dt_algo = PT_DT_ALGO_BRUTE;
        else if (m == "kdtree") p->dt_algo = PT_DT_ALGO_KDTREE;
        else if (m == "downsample") p->dt_algo = PT_DT_ALGO_DOWNSAMPLE;
        else if (m == "hybrid") p->dt_algo = PT_DT_ALGO_HYBRID;
        else if (m == "euclidean") p->dt_algo = PT_DT_ALGO_EUCLIDEAN;
        else return pt_fail(PT_E_INVALID, "invalid dt_algo");
    }
    if (has("-colorspace")) {
        const string &m = sw["-colorspace"];
        if (m == "yuv") p->colorspace = PT_COLORSPACE_YUV;
        else if (m == "lab") p->colorspace = PT_COLORSPACE_LAB;
        else if (m == "rgb") p->convert_colorspace = 0;
        else return pt_fail(PT_E_INVALID, "unsupported colorspace");
    }
    if (has("-lookup_algo")) {
        const string &m = sw["-lookup_algo"];
        if (m == "table") p->lookup_algo = PT_LOOKUP_ALGO_TABLE;
        else if (m == "brute") p->lookup_algo = PT_LOOKUP_ALGO_BRUTE;
        else if (m == "kdtree") p->lookup_algo = PT_LOOKUP_ALGO_KDTREE;
        else if (m == "pm") p->lookup_algo = PT_LOOKUP_ALGO_PM;
        else if (m == "treecann") p->lookup_algo = PT_LOOKUP_ALGO_TREECANN;
        else return pt_fail(PT_E_INVALID, "unsupported lookup_algo");
    }
    if (has("-kcoherence_algo")) {
        const string &m = sw["-kcoherence_algo"];
        if (m == "flann") p->kcoherence_algo = PT_KCOH_ALGO_FLANN;
        else if (m == "pm") p->kcoherence_algo = PT_KCOH_ALGO_PM;
        else if (m == "ann") p->kcoherence_algo = PT_KCOH_ALGO_ANN;
        else if (m == "treecann") p->kcoherence_algo = PT_KCOH_ALGO_TREECANN;
        else return pt_fail(PT_E_INVALID, "unsupported lookup_algo");
    }
    if (has("-dim_algo")) {
        const string &m = sw["-dim_algo"];
        if (m == "pca") p->dim_algo = PT_DIM_ALGO_PCA;
        else if (m == "wh") p->dim_algo = PT_DIM_ALGO_WH;
        else return pt_fail(PT_E_INVALID, "unsupported dim_algo");
    }
    if (has("-flann_trees")) p->flann_trees = to_int(sw["-flann_trees"]);
    if (has("-flann_checks")) p->flann_checks = to_int(sw["-flann_checks"]);
    if (has("-flann_build_step")) p->flann_build_step = to_int(sw["-flann_build_step"]);
    if (has("-flann_eps")) p->flann_eps = to_double(sw["-flann_eps"]);
    if (has("-treecann_eps")) p->treecann_eps = to_double(sw["-treecann_eps"]);
    if (has("-flann_reset_step")) p->flann_reset_step = to_int(sw["-flann_reset_step"]);
    if (has("-query_iters")) p->query_iters = to_int(sw["-query_iters"]);
    if (has("-treecann_agrid")) p->treecann_agrid = to_int(sw["-treecann_agrid"]);
    if (has("-treecann_bgrid")) p->treecann_bgrid = to_int(sw["-treecann_bgrid"]);
    if (has("-kcoherence")) p->kcoherence = to_int(sw["-kcoherence"]);
    if (has("-kcoherence_iter")) p->kcoherence_iter = to_int(sw["-kcoherence_iter"]);
    if (has("-kcoherence_min_dist")) p->kcoherence_min_dist = to_int(sw["-kcoherence_min_dist"]);
    if (has("-kcoherence_step")) p->kcoherence_step = to_int(sw["-kcoherence_step"]);
    if (has("-kcoherence_enrich")) p->kcoherence_enrich = to_int(sw["-kcoherence_enrich"]);
    if (has("-incoherence")) p->incoherence = to_int(sw["-incoherence"]) != 0;
    if (has("-incoherence_min_dist")) p->incoherence_min_dist = to_int(sw["-incoherence_min_dist"]);
    if (has("-save_kcoherence")) p->save_kcoherence = to_int(sw["-save_kcoherence"]) != 0;
    if (has("-triangle_factor")) p->triangle_factor = to_double(sw["-triangle_factor"]);
    if (has("-init_random")) p->init_random = to_int(sw["-init_random"]) != 0;
    if (has("-pm_iters")) p->pm_iters = to_int(sw["-pm_iters"]);
    if (has("-pm_enrich")) p->pm_enrich = to_int(sw["-pm_enrich"]) != 0;
    if (has("-sanitize_input")) p->sanitize_input = to_int(sw["-sanitize_input"]) != 0;
    if (has("-do_table_lookup")) p->do_table_lookup = to_int(sw["-do_table_lookup"]) != 0;
    if (has("-grid_ndims")) p->grid_ndims = to_int(sw["-grid_ndims"]);
    if (has("-table_ratio")) p->table_ratio = to_double(sw["-table_ratio"]);
    if (has("-prop_dist")) p->prop_dist = to_int(sw["-prop_dist"]);
    if (has("-kdtree_add_unique")) p->kdtree_add_unique = to_int(sw["-kdtree_add_unique"]) != 0;
    if (has("-filter_dims")) {
        vector<string> nl = split_on(sw["-filter_dims"], ',');
        if ((int)nl.size() > PT_MAX_DIMS) return pt_fail(PT_E_INVALID, "-filter_dims: too many dimensions");
        p->ndims = (int)nl.size();
        p->filter_dims_count = (int)nl.size();
        for (size_t i = 0; i < nl.size(); i++) p->filter_dims[i] = to_int(nl[i]);
    }
    if (has("-auto_filter")) p->auto_filter = to_int(sw["-auto_filter"]) != 0;
    if (p->auto_filter) {
        // patchtable.cpp:365-372
        static const int f6[] = {2, 1, 5, 0, 3, 10};
        static const int f7[] = {5, 1, 3, 10, 0, 6, 2};
        static const int f8[] = {2, 1, 6, 3, 0, 5, 19, 10};
        static const int f10[] = {2, 6, 19, 10, 3, 0, 1, 15, 5, 13};
        static const int f11[] = {13, 19, 6, 1, 3, 30, 10, 0, 5, 28, 2};
        const int *f = nullptr;
        if (p->ndims == 6) f = f6;
        else if (p->ndims == 7) f = f7;
        else if (p->ndims == 8) f = f8;
        else if (p->ndims == 10) f = f10;
        else if (p->ndims == 11) f = f11;
        else return pt_fail(PT_E_INVALID, "auto_filter -ndims not supported");
        p->filter_dims_count = p->ndims;
        for (int i = 0; i < p->ndims; i++) p->filter_dims[i] = f[i];
    }
    // Switches of this library's GPU algorithms (no counterpart in the reference).
    if (has("-kcoh_window")) p->kcoh_window = to_int(sw["-kcoh_window"]);
    if (has("-jf_rounds")) p->jf_rounds = to_int(sw["-jf_rounds"]);
    if (has("-mfma_rounds")) p->mfma_rounds = to_int(sw["-mfma_rounds"]);
    if (has("-mfma_anchors")) p->mfma_anchors = to_int(sw["-mfma_anchors"]);
    if (has("-mfma_jumps")) {
        // comma-separated jumps, at most PT_MFMA_MAX_JUMPS, each in 1..1024
        int v[PT_MFMA_MAX_JUMPS] = {0}, n = 0;
        const string &js = sw["-mfma_jumps"];
        size_t i = 0;
        while (i <= js.size()) {
            size_t j = js.find(',', i);
            if (j == string::npos) j = js.size();
            const string tok = js.substr(i, j - i);
            char *end = nullptr;
            const long x = strtol(tok.c_str(), &end, 10);
            if (tok.empty() || *end != '\0' || x < 1 || x > 1024 || n == PT_MFMA_MAX_JUMPS)
                return pt_failf(PT_E_INVALID, "-mfma_jumps: expected up to %d comma-separated jumps in 1..1024",
                                PT_MFMA_MAX_JUMPS);
            v[n++] = (int)x;
            i = j + 1;
        }
        for (int k = 0; k < PT_MFMA_MAX_JUMPS; k++) p->mfma_jumps[k] = v[k];
    }
    return PT_OK;
}

}  // namespace

extern "C" int pt_params_set_speed(pt_params *p, int i) {
    if (!p) return pt_fail(PT_E_INVALID, "pt_params_set_speed: null params");
    const char *s = speed_preset(i);
    if (!s) return pt_fail(PT_E_INVALID, "Error: set_speed(i), expected integer in 0...9");
    return set_from_string(p, s);
}

extern "C" int pt_params_set_switches(pt_params *p, const char *s) {
    if (!p) return pt_fail(PT_E_INVALID, "pt_params_set_switches: null params");
    return set_from_string(p, s);
}

extern "C" int pt_params_default(pt_params *p) {
    // set_defaults (patchtable.cpp:26-136); ends with set_speed(3).
    if (!p) return pt_fail(PT_E_INVALID, "pt_params_default: null params");
    memset(p, 0, sizeof(*p));
    p->randomize_dt = 0;
    p->parallel_dt = 0;
    p->limit = 0.1e6;
    p->threads = 1;
    p->dt_threads = 1;
    p->convert_colorspace = 1;
    p->colorspace = PT_COLORSPACE_YUV;
    p->regular_grid = 0;
    p->verbose = 1;
    p->table_knn = 1;
    p->populate_nearest = 1;
    p->partition_step = 1;
    p->ndims = 6;
    p->nchroma = 1;
    p->ntables = 1;
    p->populate_exptime = 0;
    p->init_random = 0;
    p->dt_mode = 0;
    p->dt_algo = PT_DT_ALGO_PROP;
    p->dt_iters = -1;
    p->query_step = 1;
    p->run_dt = 1;
    p->calc_exact_dist = 0;
    p->do_prop = 1;
    p->prop_y_step = 1;
    p->do_rs = 1;
    p->allowed_index = 1;
    p->spatial = 0;
    p->prob_rs = 1.0;
    p->do_calc_dist = 1;
    p->patch_w = 8;
    p->kcoherence_step = 1;
    p->coherence_spatial = 0.0;
    p->coherence_temporal = 0.0;
    p->table_ratio = 2.0;
    p->treecann_agrid = 1;
    p->treecann_bgrid = 1;
    p->recalc_dist_temporal = 1;
    p->is_descriptor = 0;
    p->check_existing = 0;
    p->descriptor_padded = 0;
    p->lookup_algo = PT_LOOKUP_ALGO_TABLE;
    p->flann_build_step = 1;
    p->flann_trees = 4;
    p->flann_checks = 1;
    p->flann_eps = 500000;
    p->treecann_eps = 0;
    p->flann_reset_step = 10;
    p->query_iters = 2;
    p->dt_knn = 1;
    p->prop_iters = 0;
    p->enrich_limit = 0.1e6;
    p->kcoherence = 0;
    p->kcoherence_min_dist = 2;
    p->pm_iters = 5;
    p->pm_knn = 0;
    p->pm_min_dist = 2;
    p->pm_enrich = 0;
    p->kcoherence_algo = PT_KCOH_ALGO_FLANN;
    p->pca_samples = 1000;
    p->dim_algo = PT_DIM_ALGO_WH;
    p->pca_mean = 0;
    p->pca_subtract_mean = 1;
    p->auto_filter = 0;
    p->kdtree_add_unique = 0;
    p->populate_random = 1;
    p->kcoherence_iter = -1;
    p->kcoherence_enrich = 0;
    p->grid_ndims = -1;
    p->prop_dist = 0;
    p->incoherence = 0;
    p->incoherence_min_dist = 2;
    p->save_kcoherence = 0;
    p->triangle_factor = 2.0;
    p->sanitize_input = 1;
    p->do_table_lookup = 1;
    p->ann = 0;
    p->ann_agrid = 4;
    p->ann_bgrid = 4;
    p->ann_eps = 3.0;
    p->kcoh_window = 8;
    p->jf_rounds = 6;
    p->mfma_rounds = 0;
    p->mfma_anchors = 6;
    for (int k = 0; k < PT_MFMA_MAX_JUMPS; k++) p->mfma_jumps[k] = 0;
    return pt_params_set_speed(p, 3);
}
