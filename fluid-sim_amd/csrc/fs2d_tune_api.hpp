// fs2d_tune_api.hpp -- launch shapes that change no result, and the
// autotuner that picks them (fs2d_*_tuning, fs2d_autotune,
// fs2d_batch_autotune of include/fs2d.h).  Included at the end of
// fs2d_capi.hip (it uses the handles and the stage code).
//
// Knobs:
//   pcg_grid        physical workgroups of the PCG SpMV and axpy launches.
//                   Their loops run a fixed list of virtual workgroups (each
//                   writes its own partial sums), so every choice gives the
//                   same bits; a multiple of 8 keeps the XCD mapping.
//   sweep_prefetch  (ensembles) whether the batched solves' loaders issue
//                   their first loads before the skip test: sooner data for
//                   members still iterating, wasted loads for converged ones.
//   fused_pcg       (single handles) a PCG iteration as two launches
//                   (fwd_axpy_kernel, bwd_spmv_kernel) or as four.  The
//                   SpMV's virtual workgroups, each entry's update and the
//                   maximum of |r| are the same either way, so the bits are.
//
// The tuner times each candidate on the handle's own state: `iters` PCG
// iterations of the projection (tolerance 0, so every member runs them all),
// best of three after a warm-up, and keeps the fastest.  The choice is
// recorded in a text file under a key of the device and the problem (grid
// sizes, mode), and read from there by later calls with the same key.

namespace {

// Workgroups of `kernel` (block threads, dynamic LDS) that the device holds at
// once, a multiple of 8.
template <class K>
int resident_groups(K kernel, int block, size_t lds) {
    int dev = 0, per_cu = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, lds) != hipSuccess) return 0;
    return std::max(8, per_cu * prop.multiProcessorCount / 8 * 8);
}

std::string device_tag() {
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return "unknown";
    char buf[160];
    snprintf(buf, sizeof buf, "%s/%dcu", prop.gcnArchName, prop.multiProcessorCount);
    for (char* c = buf; *c; c++)
        if (*c == ' ' || *c == ':') *c = '_';
    return buf;
}

// The cache file: one line per key, "key pcg_grid sweep_prefetch fused_pcg"
// (a line without the last value keeps t's fused_pcg, the handle's default);
// the last line of a key wins.
bool cache_lookup(const char* path, const std::string& key, fs2d_tuning* t) {
    if (!path) return false;
    FILE* f = fopen(path, "r");
    if (!f) return false;
    char line[768], k[512];
    bool found = false;
    while (fgets(line, sizeof line, f)) {
        int g = 0, p = 0, fu = t->fused_pcg;
        if (sscanf(line, "%511s %d %d %d", k, &g, &p, &fu) >= 3 && key == k) {
            t->pcg_grid = g;
            t->sweep_prefetch = p;
            t->fused_pcg = fu;
            found = true;
        }
    }
    fclose(f);
    return found;
}
void cache_store(const char* path, const std::string& key, const fs2d_tuning& t) {
    if (!path) return;
    FILE* f = fopen(path, "a");
    if (!f) return;
    fprintf(f, "%s %d %d %d\n", key.c_str(), t.pcg_grid, t.sweep_prefetch, t.fused_pcg);
    fclose(f);
}

double now_s() {
    timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

// Candidate workgroup counts of a single handle's SpMV / axpy below `full`
// (0: one per virtual workgroup).
std::vector<int> grid_candidates(int full, int resident) {
    std::vector<int> c = {0};
    for (int g : {resident, resident / 2, 2 * resident, 1024, 512, 256}) {
        g = g / 8 * 8;
        if (g >= 8 && g < full && std::find(c.begin(), c.end(), g) == c.end()) c.push_back(g);
    }
    return c;
}
// Candidates of a batch, whose loops run up to `full` virtual workgroups: one
// workgroup per virtual workgroup lets the hardware balance members of
// different weights; fewer make iterations after most members have converged
// cheaper.
std::vector<int> batch_grid_candidates(int full, int resident) {
    std::vector<int> c;
    for (int g : {full, full / 2, full / 4, 4 * resident, 2 * resident, resident}) {
        g = (g + 7) / 8 * 8;
        if (g >= 8 && g <= (full + 7) / 8 * 8 && std::find(c.begin(), c.end(), g) == c.end()) c.push_back(g);
    }
    return c;
}

}  // namespace

extern "C" {

fs2d_status fs2d_get_tuning(fs2d_sim* s, fs2d_tuning* t) {
    if (!s || !t) return FS2D_ERR_ARG;
    t->pcg_grid = s->pcg_grid;
    t->sweep_prefetch = 0;
    t->fused_pcg = s->fused_pcg ? 1 : 0;
    return FS2D_OK;
}

fs2d_status fs2d_set_tuning(fs2d_sim* s, const fs2d_tuning* t) {
    if (!s || !t || t->pcg_grid < 0 || (t->pcg_grid & 7) || (t->fused_pcg != 0 && t->fused_pcg != 1))
        return FS2D_ERR_ARG;
    HIPCHK(s, hipStreamSynchronize(s->stream));
    if ((t->fused_pcg != 0) != s->fused_pcg || t->pcg_grid != s->pcg_grid) {
        // captured steps hold the old launches
        for (auto& g : s->graphs) graph_free(g);
        s->graphs.clear();
    }
    s->pcg_grid = t->pcg_grid;
    s->fused_pcg = t->fused_pcg != 0;
    return FS2D_OK;
}

fs2d_status fs2d_batch_get_tuning(fs2d_batch* B, fs2d_tuning* t) {
    if (!B || !t) return FS2D_ERR_ARG;
    t->pcg_grid = B->pcg_grid;
    t->sweep_prefetch = B->sweep_prefetch ? 1 : 0;
    t->fused_pcg = B->fused_pcg ? 1 : 0;
    return FS2D_OK;
}

fs2d_status fs2d_batch_set_tuning(fs2d_batch* B, const fs2d_tuning* t) {
    if (!B || !t || t->pcg_grid < 8 || (t->pcg_grid & 7) || (t->fused_pcg != 0 && t->fused_pcg != 1))
        return FS2D_ERR_ARG;
    BHIPCHK(B, hipStreamSynchronize(B->stream));
    B->pcg_grid = t->pcg_grid;
    B->sweep_prefetch = t->sweep_prefetch != 0;
    B->fused_pcg = t->fused_pcg != 0;
    B->tuned = true;
    return FS2D_OK;
}

fs2d_status fs2d_autotune(fs2d_sim* s, int iters, const char* cache_path, fs2d_tuning* out, int* from_cache) {
    if (!s || iters < 1) return FS2D_ERR_ARG;
    char kb[256];
    snprintf(kb, sizeof kb, "%s|single|%dx%d|mode%d|flags%d", device_tag().c_str(), s->nx, s->ny, s->mode,
             s->flags);
    const std::string key = kb;
    fs2d_tuning best{s->pcg_grid, 0, s->fused_pcg ? 1 : 0};
    if (from_cache) *from_cache = 0;
    if (cache_lookup(cache_path, key, &best)) {
        if (from_cache) *from_cache = 1;
        TRY(fs2d_set_tuning(s, &best));
        if (out) *out = best;
        return FS2D_OK;
    }
    const double tol = s->tol;
    const int max_iters = s->max_iters, hint = s->last_iters_hint;
    s->tol = 0.0;
    s->max_iters = iters;
    const int resident = resident_groups(pcg_spmv_sk_kernel, kBlock, 0);
    double best_t = 1e30;
    fs2d_status st = FS2D_OK;
    for (int fu = 1; fu >= 0 && st >= 0; fu--)
        for (int g : fu ? std::vector<int>{0} : grid_candidates(pcg_blocks(s), resident)) {
            s->pcg_grid = g;
            s->fused_pcg = fu != 0;
            double t_min = 1e30;
            for (int rep = 0; rep < 4 && st >= 0; rep++) {
                HIPCHK(s, hipStreamSynchronize(s->stream));
                const double t0 = now_s();
                st = fs2d_run_stage(s, FS2D_STAGE_APPLY_PROJECTION);
                const double t = now_s() - t0;
                if (rep > 0) t_min = std::min(t_min, t);
            }
            if (st < 0) break;
            if (t_min < best_t) {
                best_t = t_min;
                best.pcg_grid = g;
                best.fused_pcg = fu;
            }
        }
    s->tol = tol;
    s->max_iters = max_iters;
    s->last_iters_hint = hint;
    s->pcg_grid = best.pcg_grid;
    s->fused_pcg = best.fused_pcg != 0;
    if (st < 0) return st;
    cache_store(cache_path, key, best);
    if (out) *out = best;
    return FS2D_OK;
}

fs2d_status fs2d_batch_autotune(fs2d_batch* B, int iters, const char* cache_path, fs2d_tuning* out,
                                int* from_cache) {
    if (!B || iters < 1) return FS2D_ERR_ARG;
    // the key: device, mode and the members' sizes (FNV-1a of the list)
    unsigned long long h = 1469598103934665603ull;
    for (auto* s : B->sims)
        for (int v : {s->nx, s->ny}) {
            h ^= (unsigned)v;
            h *= 1099511628211ull;
        }
    char kb[256];
    snprintf(kb, sizeof kb, "%s|batch|n%d|sizes%016llx|mode%d|flags%d", device_tag().c_str(), B->n, h, B->mode,
             B->flags);
    const std::string key = kb;
    fs2d_tuning best{B->pcg_grid, B->sweep_prefetch ? 1 : 0, B->fused_pcg ? 1 : 0};
    if (from_cache) *from_cache = 0;
    if (cache_lookup(cache_path, key, &best)) {
        if (from_cache) *from_cache = 1;
        TRY(fs2d_batch_set_tuning(B, &best));
        if (out) *out = best;
        return FS2D_OK;
    }
    TRY(refresh_inst(B));
    const double tol = B->tol;
    const int max_iters = B->max_iters, hint = B->last_iters_hint;
    // tolerance 0 through the members' scalars (b_pcg_init_kernel sets it)
    B->tol = 0.0;
    B->max_iters = iters;
    const int resident = resident_groups(b_spmv_kernel, kBlock, 0);
    int full = 0;
    for (auto* s : B->sims) full += (pcg_blocks(s) + 7) / 8 * 8;
    double best_t = 1e30;
    fs2d_status st = FS2D_OK;
    for (int fu = 0; fu < 2 && st >= 0; fu++)
    for (int pf = 0; pf < 2 && st >= 0; pf++)
        for (int g : batch_grid_candidates(full, resident)) {
            B->pcg_grid = g;
            B->sweep_prefetch = pf != 0;
            B->fused_pcg = fu != 0;
            double t_min = 1e30;
            for (int rep = 0; rep < 4 && st >= 0; rep++) {
                BHIPCHK(B, hipStreamSynchronize(B->stream));
                const double t0 = now_s();
                st = fs2d_batch_run_stage(B, FS2D_STAGE_APPLY_PROJECTION);
                const double t = now_s() - t0;
                if (rep > 0) t_min = std::min(t_min, t);
            }
            if (st < 0) break;
            if (t_min < best_t) {
                best_t = t_min;
                best.pcg_grid = g;
                best.sweep_prefetch = pf;
                best.fused_pcg = fu;
            }
        }
    B->tol = tol;
    B->max_iters = max_iters;
    B->last_iters_hint = hint;
    B->pcg_grid = best.pcg_grid;
    B->sweep_prefetch = best.sweep_prefetch != 0;
    B->fused_pcg = best.fused_pcg != 0;
    B->tuned = true;
    if (st < 0) return st;
    cache_store(cache_path, key, best);
    if (out) *out = best;
    return FS2D_OK;
}

}  // extern "C"
