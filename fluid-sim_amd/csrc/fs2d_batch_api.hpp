// fs2d_batch_api.hpp -- the fs2d_batch_* entry points of include/fs2d.h: an
// ensemble of independent simulations of different sizes stepped together,
// each kernel of the step launched once for all of them (fs2d_batch.hpp).
// Included at the end of fs2d_capi.hip (it uses that file's handles, stage
// helpers and allocation code).
//
// Each member is an ordinary handle (fs2d_create): its state, fields and
// statistics are reached through the fs2d_* calls on fs2d_batch_sim(b, i).
// The batch owns the hand-off rows and tickets of its wavefront passes, so
// batched and single launches never share them.  The PCG loop runs until
// every member has converged; converged members' workgroups return at once.

struct fs2d_batch {
    int n = 0;
    std::vector<fs2d_sim*> sims;
    hipStream_t stream = nullptr;
    int mode = 0, flags = 0, max_iters = 0;
    double tol = 0;
    bool any_particles = false;
    std::vector<InstBufs> h_ins;
    InstBufs* d_ins = nullptr;
    int* d_maps = nullptr;
    BlockMap m_cells{}, m_faces{}, m_sl{}, m_part{}, m_diag{}, m_sk{}, m_tiles{}, m_one{}, m_xrows{}, m_xcols{}, m_pcg{};
    int g_cells = 0, g_faces = 0, g_sl = 0, g_part = 0, g_diag = 0, g_sk = 0, g_tiles = 0, g_one = 0, g_xrows = 0,
        g_xcols = 0;
    StripMap strips{};
    int* werr = nullptr;
    unsigned long long* sk_ticket = nullptr;
    int* d_status = nullptr;
    int* h_status = nullptr;  // pinned: per member done, iterations, NaN flag, CFL warning
    // The PCG loop's SpMV and axpy are persistent launches of pcg_grid
    // workgroups over the members still iterating (a plan formed once per
    // iteration from the members' done flags, b_pcg_plan_kernel, with
    // plan_lds bytes of LDS); the sweeps keep the full strip lists, since
    // their epochs count launches of a fixed strip total.  all_done_since:
    // the first iteration at which no member iterated (INT_MAX while some do).
    int pcg_grid = 2048;
    bool tuned = false;  // pcg_grid set by fs2d_batch_set_tuning / autotune
    size_t plan_lds = 0;
    PcgPlanBufs plan{};
    bool sweep_prefetch = true;
    // the p/r update with the forward solves in one launch (b_fwd_axpy_kernel,
    // its own ticket; the stopping test in the backward launch).  Off by
    // default: on the 256-member ensemble the batched forward solves already
    // fill the device (1233 strips), leaving no idle CUs for the update to
    // hide in (107.4 against 103.6 ms per step); the autotuner tries both.
    bool fused_pcg = false;
    unsigned long long* fa_ticket = nullptr;
    int* all_done_since = nullptr;
    const int* pcg_order = nullptr;  // members by SpMV/axpy virtual-workgroup weight, heaviest first (in d_maps)
    int* h_err = nullptr;
    bool sorted_valid = false;
    int last_iters_hint = 8;
    // Captured steps (fs2d_batch_set_graph), as for a single handle.  Kernel
    // arguments are the descriptors' and maps' device addresses, so a capture
    // stays valid until the maps are rebuilt (a member's particle count
    // changed) or a launch parameter changes.
    bool use_graph = false;
    int graph_budget = 0;
    struct Captured {
        const int* maps;
        int budget, pcg_grid, prefetch, fused;
        hipGraphExec_t pre = nullptr, post = nullptr;
    };
    std::vector<Captured> graphs;
    std::vector<void*> allocs;
    std::string err;

    unsigned long long timing = 0;  // kernel classes timed (bit k: class k)
    std::vector<TimedLaunch> pending;
    std::vector<hipEvent_t> event_pool;
    double ktime[K_COUNT] = {};
    long long kcount[K_COUNT] = {};
};

namespace {

fs2d_status bfail(fs2d_batch* B, fs2d_status st, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    B->err = buf;
    return st;
}

#define BHIPCHK(B, x)                                                                                 \
    do {                                                                                              \
        hipError_t e_ = (x);                                                                          \
        if (e_ != hipSuccess)                                                                         \
            return bfail(B, e_ == hipErrorOutOfMemory ? FS2D_ERR_OOM : FS2D_ERR_DEVICE, "%s: %s", #x, \
                         hipGetErrorString(e_));                                                      \
    } while (0)

template <class T>
fs2d_status balloc(fs2d_batch* B, T** p, size_t n) {
    void* ptr = nullptr;
    const size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
    hipError_t e = hipMalloc(&ptr, bytes);
    if (e != hipSuccess) return bfail(B, FS2D_ERR_OOM, "hipMalloc(%zu bytes): %s", bytes, hipGetErrorString(e));
    B->allocs.push_back(ptr);
    *p = (T*)ptr;
    BHIPCHK(B, hipMemsetAsync(ptr, 0, bytes, B->stream));
    return FS2D_OK;
}

// Descriptor of member s (the hand-off rows and run flags are the batch's).
InstBufs inst_bufs(const fs2d_sim* s, unsigned long long* sk_handoff, unsigned* fa_flags,
                   unsigned long long* fa_handoff) {
    InstBufs I;
    memset(&I, 0, sizeof I);  // padding too: descriptors are compared bytewise
    I.nx = s->nx; I.ny = s->ny; I.mode = s->mode; I.flags = s->flags; I.np = s->np;
    I.dt = s->dt; I.dx = s->dx; I.dr = s->dr; I.rho = s->rho; I.gx = s->gx; I.gy = s->gy; I.alpha = s->alpha;
    I.tau = s->tau; I.sigma = s->sigma; I.max_iters = s->max_iters;
    I.pmass = s->k > 0 ? s->rho * s->dx * s->dx / (s->k * s->k) : 0.0;
    I.rhs_scale = 1.0 / (s->dx * solver_scale(s));
    I.rhs_zero = 1e-12 / solver_scale(s);
    I.geom = s->geom;
    I.u = s->u; I.v = s->v; I.u2 = s->u2; I.v2 = s->v2; I.nu = s->nu; I.nv = s->nv; I.dens = s->dens; I.dens2 = s->dens2;
    I.cell = s->cell; I.cell2 = s->cell2; I.phi = s->phi; I.phi2 = s->phi2; I.t0 = s->t0; I.t1 = s->t1; I.ls_chg = s->ls_chg;
    I.p = s->p; I.r = s->r; I.z = s->z; I.q = s->q; I.w = s->w; I.s0 = s->s[0]; I.s1 = s->s[1]; I.pk = s->pk;
    I.Adiag = s->Adiag; I.P = s->P; I.fx = s->fx; I.fy = s->fy; I.gx_ = s->gx_; I.gy_ = s->gy_; I.cpl = s->cpl;
    I.sc = s->sc; I.qq = s->qq; I.ws_part = s->ws_part; I.rmax_part = s->rmax_part;
    I.mu = s->mu; I.mv = s->mv; I.fu0 = s->fu0; I.fu1 = s->fu1; I.fv0 = s->fv0; I.fv1 = s->fv1; I.fcount = s->fcount;
    I.fpos = s->fpos; I.fsrc = s->fsrc; I.fkv = s->fkv;
    I.px = s->px; I.py = s->py; I.pvx = s->pvx; I.pvy = s->pvy;
    I.pkey = s->pkey; I.counts = s->counts; I.starts = s->starts; I.cursor = s->cursor; I.sorted = s->sorted;
    I.tile_sums = s->tile_sums; I.dg = s->dg; I.partials = s->partials;
    I.sk_handoff = sk_handoff;
    I.fa_flags = fa_flags;
    I.fa_handoff = fa_handoff;
    return I;
}

// Appends a block map for per-member workgroup counts `cnt` to H; member
// ranges start at multiples of 8.  Returns the launch's workgroup count.
struct MapOff {
    size_t inst, first, count;
};
int add_map(std::vector<int>& H, const std::vector<int>& cnt, MapOff& off) {
    const int n = (int)cnt.size();
    std::vector<int> first(n);
    int total = 0;
    for (int b = 0; b < n; b++) {
        total = (total + 7) & ~7;
        first[b] = total;
        total += cnt[b];
    }
    off.inst = H.size();
    H.resize(H.size() + total, -1);
    for (int b = 0; b < n; b++)
        for (int k = 0; k < cnt[b]; k++) H[off.inst + first[b] + k] = b;
    off.first = H.size();
    H.insert(H.end(), first.begin(), first.end());
    off.count = H.size();
    H.insert(H.end(), cnt.begin(), cnt.end());
    return total;
}

// Launch maps: every kernel's per-member workgroup count is the count of that
// member's single launch (blocks_for / sk_blocks of the same sizes).
fs2d_status build_maps(fs2d_batch* B) {
    const int n = B->n;
    std::vector<int> c_cells(n), c_faces(n), c_sl(n), c_part(n), c_diag(n), c_sk(n), c_tiles(n), c_one(n, 1);
    std::vector<int> c_xrows(n), c_xcols(n), c_pcg(n);
    std::vector<int> nstrips(n);
    B->any_particles = false;
    for (int b = 0; b < n; b++) {
        fs2d_sim* s = B->sims[b];
        const long nc = n_c(s), nf = n_u(s) + n_v(s);
        c_cells[b] = blocks_for(nc);
        c_faces[b] = blocks_for(nf);
        c_sl[b] = sl_tiles(s->nx, s->ny);
        c_part[b] = s->np > 0 ? blocks_for(s->np) : 0;
        c_diag[b] = blocks_for(nc + s->np);
        c_sk[b] = sk_blocks(s);
        c_pcg[b] = pcg_blocks(s);
        c_tiles[b] = (int)((nc + 1023) / 1024);
        c_xrows[b] = extrap_row_blocks(s->ny);
        c_xcols[b] = extrap_col_blocks(s->nx);
        nstrips[b] = (s->ny + 63) / 64;
        B->any_particles = B->any_particles || s->np > 0;
    }
    std::vector<int> H;
    MapOff o_cells, o_faces, o_sl, o_part, o_diag, o_sk, o_tiles, o_one, o_xrows, o_xcols, o_pcg;
    B->g_cells = add_map(H, c_cells, o_cells);
    B->g_faces = add_map(H, c_faces, o_faces);
    B->g_sl = add_map(H, c_sl, o_sl);
    B->g_part = add_map(H, c_part, o_part);
    B->g_diag = add_map(H, c_diag, o_diag);
    B->g_sk = add_map(H, c_sk, o_sk);
    B->g_tiles = add_map(H, c_tiles, o_tiles);
    B->g_one = add_map(H, c_one, o_one);
    B->g_xrows = add_map(H, c_xrows, o_xrows);
    B->g_xcols = add_map(H, c_xcols, o_xcols);
    (void)add_map(H, c_pcg, o_pcg);
    // Strip list of the wavefront passes: each member's strips contiguous and
    // in order (a strip's upstream neighbour is taken earlier), members with
    // the longest strip chains first, so that their pipelines start in the
    // first round of residency and the short members fill in behind them.
    // Which workgroup runs a strip does not change any result.
    const size_t strip_inst = H.size();
    int total = 0;
    std::vector<int> sfirst(n), order(n);
    for (int b = 0; b < n; b++) order[b] = b;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        const fs2d_sim *sa = B->sims[a], *sb = B->sims[b];
        return nstrips[a] != nstrips[b] ? nstrips[a] > nstrips[b] : sa->nx > sb->nx;
    });
    for (int b : order) {
        sfirst[b] = total;
        for (int k = 0; k < nstrips[b]; k++) H.push_back(b);
        total += nstrips[b];
    }
    const size_t strip_first = H.size();
    H.insert(H.end(), sfirst.begin(), sfirst.end());
    // members for the PCG plan: entries per SpMV/axpy virtual workgroup,
    // descending (ties by index)
    std::vector<int> porder(n);
    for (int b = 0; b < n; b++) porder[b] = b;
    std::stable_sort(porder.begin(), porder.end(), [&](int a, int b) {
        return B->sims[a]->geom.size() * c_pcg[b] > B->sims[b]->geom.size() * c_pcg[a];
    });
    const size_t order_at = H.size();
    H.insert(H.end(), porder.begin(), porder.end());
    int* maps = nullptr;
    TRY(balloc(B, &maps, H.size()));
    BHIPCHK(B, hipMemcpyAsync(maps, H.data(), H.size() * sizeof(int), hipMemcpyHostToDevice, B->stream));
    BHIPCHK(B, hipStreamSynchronize(B->stream));
    B->d_maps = maps;
    auto mk = [&](const MapOff& o) { return BlockMap{maps + o.inst, maps + o.first, maps + o.count}; };
    B->m_cells = mk(o_cells); B->m_faces = mk(o_faces); B->m_sl = mk(o_sl); B->m_part = mk(o_part);
    B->m_diag = mk(o_diag); B->m_sk = mk(o_sk); B->m_tiles = mk(o_tiles); B->m_one = mk(o_one);
    B->m_xrows = mk(o_xrows); B->m_xcols = mk(o_xcols); B->m_pcg = mk(o_pcg);
    B->strips = StripMap{maps + strip_inst, maps + strip_first, total};
    B->pcg_order = maps + order_at;
    int* plan = nullptr;
    int plan_len = 0;
    for (int b = 0; b < n; b++) plan_len += (c_pcg[b] + 7) & ~7;
    TRY(balloc(B, &plan, (size_t)plan_len + 2 * (size_t)n + 1));
    B->plan = PcgPlanBufs{plan, plan + plan_len, plan + plan_len + n, plan + plan_len + n + 1};
    // untuned: one workgroup per virtual workgroup of a full iteration (the
    // hardware then balances members of different weights), at most 32768
    if (!B->tuned) B->pcg_grid = std::max(8, std::min(32768, plan_len));
    return FS2D_OK;
}

// Re-reads members' buffer pointers (a single-simulation SL stage swaps them)
// and uploads the descriptors if any changed.
fs2d_status refresh_inst(fs2d_batch* B) {
    bool changed = false, counts = false;
    for (int b = 0; b < B->n; b++) {
        InstBufs I = inst_bufs(B->sims[b], B->h_ins[b].sk_handoff, B->h_ins[b].fa_flags, B->h_ins[b].fa_handoff);
        if (memcmp(&I, &B->h_ins[b], sizeof I) != 0) {
            counts = counts || I.np != B->h_ins[b].np;
            B->h_ins[b] = I;
            changed = true;
        }
    }
    if (counts) TRY(build_maps(B));
    if (changed) {
        BHIPCHK(B, hipMemcpyAsync(B->d_ins, B->h_ins.data(), B->n * sizeof(InstBufs), hipMemcpyHostToDevice,
                                  B->stream));
        BHIPCHK(B, hipStreamSynchronize(B->stream));
    }
    return FS2D_OK;
}

template <class K, class... A>
void blaunch(fs2d_batch* B, K kernel, int grid, int block, A... args) {
    if (grid > 0) hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0, B->stream, args...);
}

// ---- stages, in the order of fs2d_step ------------------------------------
void bstage_sort(fs2d_batch* B) {
    if (B->sorted_valid) return;
    Timed t(B, K_PARTICLE_SORT);
    blaunch(B, b_fill_kernel, B->g_cells, kBlock, B->d_ins, B->m_cells, (int)BF_COUNTS);
    if (B->any_particles) blaunch(B, b_particle_keys_kernel, B->g_part, kBlock, B->d_ins, B->m_part);
    blaunch(B, b_scan_tiles_kernel, B->g_tiles, kBlock, B->d_ins, B->m_tiles);
    blaunch(B, b_scan_sums_kernel, B->n, 64, B->d_ins, B->n);
    blaunch(B, b_scan_add_kernel, B->g_cells, kBlock, B->d_ins, B->m_cells);
    if (B->any_particles) {
        blaunch(B, b_copy_kernel, B->g_cells, kBlock, B->d_ins, B->m_cells, (int)BC_CURSOR_FROM_STARTS);
        blaunch(B, b_particle_scatter_kernel, B->g_part, kBlock, B->d_ins, B->m_part);
        blaunch(B, b_segment_sort_kernel, B->g_cells, kBlock, B->d_ins, B->m_cells);
    }
    B->sorted_valid = true;
}

void bstage_diagnostics(fs2d_batch* B) {
    Timed t(B, K_DIAGNOSTICS);
    blaunch(B, b_diagnostics_kernel, B->g_diag, kBlock, B->d_ins, B->m_diag);
    blaunch(B, b_diag_finalize_kernel, B->n, kBlock, B->d_ins, B->n, B->m_diag);
}

void bstage_level_set(fs2d_batch* B) {
    if (B->flags & FS2D_FLAG_PARTICLE_CELLS) {
        Timed t(B, K_LS_OTHER);
        blaunch(B, b_pc_clear_kernel, B->g_cells, kBlock, B->d_ins, B->m_cells);
        if (B->any_particles) blaunch(B, b_pc_mark_kernel, B->g_part, kBlock, B->d_ins, B->m_part);
        blaunch(B, b_pc_fill_kernel, B->g_cells, kBlock, B->d_ins, B->m_cells);
        blaunch(B, b_pc_copy_kernel, B->g_cells, kBlock, B->d_ins, B->m_cells);
    } else if (!(B->flags & FS2D_FLAG_STATIC_CELLS)) {
        bstage_sort(B);
        {
            Timed t(B, K_LS_OTHER);
            blaunch(B, b_ls_seed_kernel, B->g_cells, kBlock, B->d_ins, B->m_cells);
        }
        const int S = B->strips.total;
        // closest-particle passes on the sweep engine, as the single path:
        // candidates in SK4 order of the pass's direction from the previous
        // pass's output, the pass, and the last output back to t0; a member's
        // passes after its first one that changes nothing return at once
        {
            Timed t(B, K_LS_OTHER);
            blaunch(B, b_ls_chg_clear_kernel, (16 * B->n + kBlock - 1) / kBlock, kBlock, B->d_ins, B->n);
        }
        for (int k = 0; k < 16; k++) {
            const int d = k & 3;
            {
                // four cells per thread (the tile map): a member whose passes
                // have stopped returns at once, and after the first few passes
                // all do, so the launch is mostly its workgroups' start-up
                Timed t(B, K_LS_OTHER);
                blaunch(B, b_ls_cand_sk_kernel, B->g_tiles, kBlock, B->d_ins, B->m_tiles, d, k == 0 ? 1 : 0,
                        (k + 3) & 3, k);
            }
            {
                Timed t(B, K_LS_SWEEP);
                switch (d) {
                    case 0: blaunch(B, b_sweep_kernel<ClosestSweepOp<0, 0>, false, ClosestSweepMaker<0, 0>>, S,
                                    kSweepThreads, B->d_ins, B->strips, B->sk_ticket, B->werr,
                                    ClosestSweepMaker<0, 0>{k}, (const int*)nullptr, 0, (const int*)nullptr); break;
                    case 1: blaunch(B, b_sweep_kernel<ClosestSweepOp<1, 0>, false, ClosestSweepMaker<1, 0>>, S,
                                    kSweepThreads, B->d_ins, B->strips, B->sk_ticket, B->werr,
                                    ClosestSweepMaker<1, 0>{k}, (const int*)nullptr, 0, (const int*)nullptr); break;
                    case 2: blaunch(B, b_sweep_kernel<ClosestSweepOp<0, 1>, false, ClosestSweepMaker<0, 1>>, S,
                                    kSweepThreads, B->d_ins, B->strips, B->sk_ticket, B->werr,
                                    ClosestSweepMaker<0, 1>{k}, (const int*)nullptr, 0, (const int*)nullptr); break;
                    default: blaunch(B, b_sweep_kernel<ClosestSweepOp<1, 1>, false, ClosestSweepMaker<1, 1>>, S,
                                     kSweepThreads, B->d_ins, B->strips, B->sk_ticket, B->werr,
                                     ClosestSweepMaker<1, 1>{k}, (const int*)nullptr, 0, (const int*)nullptr); break;
                }
            }
        }
        {
            Timed t(B, K_LS_OTHER);
            blaunch(B, b_closest_t_kernel, B->g_cells, kBlock, B->d_ins, B->m_cells);
        }
        {
            Timed t(B, K_LS_OTHER);
            blaunch(B, b_ls_phi_kernel, B->g_cells, kBlock, B->d_ins, B->m_cells);
            blaunch(B, b_ls_surface_kernel, B->g_cells, kBlock, B->d_ins, B->m_cells);
        }
        // eikonal passes on the sweep engine, as in the single path: phi (parity
        // 0) into SK4 order of direction 0, 16 passes (direction d = k mod 4:
        // fx = d & 1, fy = d >> 1) with each output moved to the next pass's
        // order, and the last one back to phi
        {
            Timed t(B, K_LS_OTHER);
            blaunch(B, b_eik_to_sk_kernel, B->g_sk, kBlock, B->d_ins, B->m_sk, 0, 0, 0);
        }
        for (int k = 0; k < 16; k++) {
            const int d = k & 3, d2 = (k + 1) & 3;
            {
                Timed t(B, K_LS_SWEEP);
                blaunch(B, b_sweep_kernel<EikSweepOp, false, EikSweepMaker>, S, kSweepThreads, B->d_ins, B->strips,
                        B->sk_ticket, B->werr, EikSweepMaker{}, (const int*)nullptr, 0, (const int*)nullptr);
            }
            Timed t(B, K_LS_OTHER);
            if (k < 15)
                blaunch(B, b_eik_sk_to_sk_kernel, B->g_sk, kBlock, B->d_ins, B->m_sk, d & 1, d >> 1, d2 & 1, d2 >> 1);
            else
                blaunch(B, b_eik_from_sk_kernel, B->g_cells, kBlock, B->d_ins, B->m_cells, 1, d & 1, d >> 1);
        }
        {
            Timed t(B, K_LS_OTHER);
            blaunch(B, b_ls_smooth_kernel, B->g_cells, kBlock, B->d_ins, B->m_cells, 0);
            blaunch(B, b_ls_smooth_kernel, B->g_cells, kBlock, B->d_ins, B->m_cells, 1);
            blaunch(B, b_classify_kernel, B->g_cells, kBlock, B->d_ins, B->m_cells);
        }
    }
    bstage_diagnostics(B);
}

void bstage_extrapolate(fs2d_batch* B, int which) {
    Timed t(B, K_EXTRAPOLATE);
    blaunch(B, b_extrap_kernel<BX_ROWS>, B->g_xrows, kBlock, B->d_ins, B->m_xrows, which);
    blaunch(B, b_extrap_cols_kernel, B->g_xcols, kExtraColThreads, B->d_ins, B->m_xcols, which);
    blaunch(B, b_extrap_kernel<BX_HIST>, B->g_faces, kBlock, B->d_ins, B->m_faces, which);
    blaunch(B, b_extrap_scan_kernel, 2 * B->n, kBlock, B->d_ins, B->n, which);
    blaunch(B, b_extrap_kernel<BX_SCATTER>, B->g_faces, kBlock, B->d_ins, B->m_faces, which);
    blaunch(B, b_extrap_kernel<BX_LINK>, B->g_faces, kBlock, B->d_ins, B->m_faces, which);
    blaunch(B, b_extrap_values_kernel, 2 * B->n, 1024, B->d_ins, B->n, which);
}

void bstage_p2g(fs2d_batch* B) {
    bstage_sort(B);
    {
        Timed t(B, K_P2G);
        blaunch(B, b_p2g_kernel, B->g_faces, kBlock, B->d_ins, B->m_faces);
    }
    bstage_extrapolate(B, 0);
}

void bstage_sl(fs2d_batch* B, bool with_gravity) {
    Timed t(B, K_SL_ADVECT);
    blaunch(B, b_sl_advect_kernel, B->g_sl, kBlock, B->d_ins, B->m_sl, with_gravity ? 1 : 0);
    blaunch(B, b_copy_kernel, B->g_faces, kBlock, B->d_ins, B->m_faces, (int)BC_UV_FROM_U2);
    if (B->flags & FS2D_FLAG_DENSITY)
        blaunch(B, b_copy_kernel, B->g_cells, kBlock, B->d_ins, B->m_cells, (int)BC_DENS_FROM_DENS2);
}

void bstage_gravity(fs2d_batch* B) {
    Timed t(B, K_GRAVITY);
    blaunch(B, b_gravity_kernel, B->g_faces, kBlock, B->d_ins, B->m_faces);
}

fs2d_status gather_status(fs2d_batch* B) {
    blaunch(B, b_gather_status_kernel, (B->n + 255) / 256, 256, B->d_ins, B->n, B->d_status);
    BHIPCHK(B, hipMemcpyAsync(B->h_status, B->d_status, 4 * B->n * sizeof(int), hipMemcpyDeviceToHost, B->stream));
    BHIPCHK(B, hipStreamSynchronize(B->stream));
    return FS2D_OK;
}

// Iterations h0 .. h1-1 of the batched PCG loop (FluidSim2D.cpp:431-463):
// every launch has a fixed shape, and members that have converged (or, after
// all have, whole launches) return at once.
template <class Fwd, class Bwd>
void bpcg_iterations(fs2d_batch* B, int h0, int h1, Fwd fwd, Bwd bwd) {
    const int S = B->strips.total;
    for (int h = h0; h < h1; h++) {
        {
            Timed t(B, K_PCG_SPMV);
            hipLaunchKernelGGL(b_pcg_plan_kernel, dim3(B->n), dim3(kBlock), B->plan_lds, B->stream, B->d_ins, B->n,
                               B->m_pcg.count, B->pcg_order, B->plan, h, B->all_done_since);
            blaunch(B, b_spmv_kernel, B->pcg_grid, kBlock, B->d_ins, B->m_pcg.count, B->plan, h);
        }
        if (B->fused_pcg) {
            {
                Timed t(B, K_PCG_AXPY);
                blaunch(B, b_ws_reduce_kernel, B->n, 64, B->d_ins, B->n, B->m_pcg.count);
            }
            {
                // the p/r update beside the forward solves; one update workgroup per CU
                Timed t(B, K_PCG_FWD);
                const int M = B->sims[0]->ncu > 0 ? B->sims[0]->ncu : 256;
                if (B->sweep_prefetch)
                    blaunch(B, b_fwd_axpy_kernel<true>, M + S, kSweepThreads, B->d_ins, B->m_pcg.count, B->plan,
                            B->strips, B->fa_ticket, B->werr, M, (const int*)B->all_done_since, h);
                else
                    blaunch(B, b_fwd_axpy_kernel<false>, M + S, kSweepThreads, B->d_ins, B->m_pcg.count, B->plan,
                            B->strips, B->fa_ticket, B->werr, M, (const int*)B->all_done_since, h);
            }
            Timed t(B, K_PCG_BWD);
            blaunch(B, bwd, S, kSweepThreads, B->d_ins, B->strips, B->sk_ticket, B->werr,
                    BwdMaker{h, B->m_pcg.count, 1}, (const int*)B->all_done_since, h, (const int*)nullptr);
            continue;
        }
        {
            Timed t(B, K_PCG_AXPY);
            blaunch(B, b_ws_reduce_kernel, B->n, 64, B->d_ins, B->n, B->m_pcg.count);
            blaunch(B, b_axpy_kernel, B->pcg_grid, kBlock, B->d_ins, B->m_pcg.count, B->plan, h);
        }
        {
            Timed t(B, K_PCG_FWD);
            blaunch(B, fwd, S, kSweepThreads, B->d_ins, B->strips, B->sk_ticket, B->werr,
                    FwdMaker{B->m_pcg.count, 1}, (const int*)B->all_done_since, h, (const int*)B->plan.active);
        }
        {
            Timed t(B, K_PCG_BWD);
            blaunch(B, bwd, S, kSweepThreads, B->d_ins, B->strips, B->sk_ticket, B->werr,
                    BwdMaker{h, nullptr, 0}, (const int*)B->all_done_since, h, (const int*)B->plan.active);
        }
    }
}
void bpcg_iterations(fs2d_batch* B, int h0, int h1) {
    if (B->sweep_prefetch)
        bpcg_iterations(B, h0, h1, b_sweep_kernel<FwdOp, false, FwdMaker, true>,
                        b_sweep_kernel<BwdOp, true, BwdMaker, true>);
    else
        bpcg_iterations(B, h0, h1, b_sweep_kernel<FwdOp, false, FwdMaker, false>,
                        b_sweep_kernel<BwdOp, true, BwdMaker, false>);
}

// Assembly, MIC(0) factor and the first preconditioner application of every
// member (FluidSim2D.cpp:252-430).
void bpcg_begin(fs2d_batch* B) {
    blaunch(B, b_pcg_init_kernel, B->n, 64, B->d_ins, B->n, B->tol, B->all_done_since);
    {
        Timed t(B, K_ASSEMBLE);
        blaunch(B, b_assemble_kernel, B->g_sk, kBlock, B->d_ins, B->m_sk);
    }
    const int S = B->strips.total;
    {
        Timed t(B, K_FACTOR);
        blaunch(B, b_sweep_kernel<FactorOp, false, FactorMaker>, S, kSweepThreads, B->d_ins, B->strips,
                B->sk_ticket, B->werr, FactorMaker{B->m_sk.count}, (const int*)nullptr, 0, (const int*)nullptr);
    }
    {
        Timed t(B, K_PCG_FWD);
        blaunch(B, b_sweep_kernel<FwdOp, false, FwdMaker>, S, kSweepThreads, B->d_ins, B->strips, B->sk_ticket,
                B->werr, FwdMaker{B->m_sk.count, 0}, (const int*)nullptr, 0, (const int*)nullptr);
    }
    {
        Timed t(B, K_PCG_BWD);
        blaunch(B, b_sweep_kernel<BwdOp, true, BwdMaker>, S, kSweepThreads, B->d_ins, B->strips, B->sk_ticket,
                B->werr, BwdMaker{-1, nullptr, 0}, (const int*)nullptr, 0, (const int*)nullptr);
    }
}

bool all_members_done(const fs2d_batch* B) {
    for (int b = 0; b < B->n; b++)
        if (B->h_status[4 * b] == 0) return false;
    return true;
}

// Queues chunks of iterations from h until every member has converged or
// max_iters is reached (one host round trip per chunk).
fs2d_status bpcg_continue(fs2d_batch* B, int h, int chunk) {
    while (h < B->max_iters) {
        const int nchunk = std::min(chunk, B->max_iters - h);
        bpcg_iterations(B, h, h + nchunk);
        h += nchunk;
        TRY(gather_status(B));
        if (all_members_done(B)) break;
        chunk = 8;
    }
    return FS2D_OK;
}

void bpcg_end(fs2d_batch* B) {
    Timed t(B, K_MISC);
    blaunch(B, b_sk_to_rowmajor_kernel, B->g_cells, kBlock, B->d_ins, B->m_cells);
}

// Iteration counts from the members' scalars (after gather_status).
fs2d_status bpcg_status(fs2d_batch* B) {
    int iters = 0;
    bool hit = false;
    for (int b = 0; b < B->n; b++) {
        iters = std::max(iters, B->h_status[4 * b + 1]);
        hit = hit || B->h_status[4 * b + 1] >= B->max_iters;
        B->sims[b]->last_iters_hint = std::max(2, B->h_status[4 * b + 1] + 1);
    }
    B->last_iters_hint = std::max(2, iters + 1);
    return hit ? FS2D_WARN_MAXITER : FS2D_OK;
}

fs2d_status bstage_projection(fs2d_batch* B) {
    bpcg_begin(B);
    TRY(bpcg_continue(B, 0, std::max(2, std::min(B->last_iters_hint, B->max_iters))));
    bpcg_end(B);
    TRY(gather_status(B));
    return bpcg_status(B);
}

void bstage_update_velocity(fs2d_batch* B) {
    {
        Timed t(B, K_UPDATE_VEL);
        blaunch(B, b_update_velocity_kernel, B->g_faces, kBlock, B->d_ins, B->m_faces);
    }
    bstage_extrapolate(B, 1);
    if (B->mode == FS2D_SEMILAGRANGIAN) {
        Timed t(B, K_MISC);
        blaunch(B, b_copy_kernel, B->g_faces, kBlock, B->d_ins, B->m_faces, (int)BC_UV_FROM_NEW);
    }
}

void bstage_g2p(fs2d_batch* B) {
    if (B->any_particles) {
        Timed t(B, K_G2P);
        blaunch(B, b_g2p_kernel, B->g_part, kBlock, B->d_ins, B->m_part);
    }
    Timed t(B, K_MISC);
    blaunch(B, b_copy_kernel, B->g_faces, kBlock, B->d_ins, B->m_faces, (int)BC_UV_FROM_NEW);
}

void bstage_particle_advect(fs2d_batch* B) {
    if (!B->any_particles) return;
    Timed t(B, K_PARTICLE_ADVECT);
    blaunch(B, b_particle_advect_kernel, B->g_part, kBlock, B->d_ins, B->m_part);
    blaunch(B, b_cfl_finalize_kernel, B->n, kBlock, B->d_ins, B->n, B->m_part);
    B->sorted_valid = false;
}

fs2d_status bstage(fs2d_batch* B, int stage, bool fuse_gravity, fs2d_status* worst) {
    fs2d_status st = FS2D_OK;
    switch (stage) {
        case FS2D_STAGE_INIT: break;
        case FS2D_STAGE_CREATE_WATER_LEVEL_SET: bstage_level_set(B); break;
        case FS2D_STAGE_TRANSFER_VELOCITY_TO_GRID: bstage_p2g(B); break;
        case FS2D_STAGE_SEMI_LAGRANGIAN_ADVECTION: bstage_sl(B, fuse_gravity); break;
        case FS2D_STAGE_APPLY_GRAVITY: bstage_gravity(B); break;
        case FS2D_STAGE_CREATE_SOLID_LEVEL_SET: break;
        case FS2D_STAGE_APPLY_PROJECTION: st = bstage_projection(B); break;
        case FS2D_STAGE_UPDATE_VELOCITY: bstage_update_velocity(B); break;
        case FS2D_STAGE_UPDATE_PARTICLE_VELOCITIES: bstage_g2p(B); break;
        case FS2D_STAGE_APPLY_ADVECTION: bstage_particle_advect(B); break;
        default: return bfail(B, FS2D_ERR_ARG, "unknown stage %d", stage);
    }
    if (st < 0) return st;
    if (st > *worst) *worst = st;
    return FS2D_OK;
}

// One host round trip: NaN, hang and CFL status of every member.
fs2d_status bfinish(fs2d_batch* B, fs2d_status st) {
    TRY(gather_status(B));
    BHIPCHK(B, hipMemcpyAsync(B->h_err, B->werr, sizeof(int), hipMemcpyDeviceToHost, B->stream));
    BHIPCHK(B, hipStreamSynchronize(B->stream));
    BHIPCHK(B, hipGetLastError());
    if (B->timing) resolve_timing(B);
    for (auto* s : B->sims) s->sorted_valid = false;
    if (*B->h_err) return bfail(B, FS2D_ERR_HANG, "a wavefront hand-off wait expired");
    for (int b = 0; b < B->n; b++)
        if (B->h_status[4 * b + 2]) return bfail(B, FS2D_ERR_NAN, "member %d: advected particle position is NaN", b);
    if (st == FS2D_WARN_MAXITER) {
        bfail(B, st, "PCG: maximum iteration limit (%d) exceeded by a member", B->max_iters);
        return st;
    }
    for (int b = 0; b < B->n; b++)
        if (B->h_status[4 * b + 3]) {
            bfail(B, FS2D_WARN_CFL, "member %d: particle velocity violates the CFL condition", b);
            return FS2D_WARN_CFL;
        }
    return st;
}

void bgraph_free(fs2d_batch::Captured& G) {
    if (G.pre) (void)hipGraphExecDestroy(G.pre);
    if (G.post) (void)hipGraphExecDestroy(G.post);
    G.pre = G.post = nullptr;
}

fs2d_status bend_capture(fs2d_batch* B, hipGraphExec_t* exec) {
    hipGraph_t g = nullptr;
    BHIPCHK(B, hipStreamEndCapture(B->stream, &g));
    hipError_t e = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return bfail(B, FS2D_ERR_DEVICE, "hipGraphInstantiate: %s", hipGetErrorString(e));
    return FS2D_OK;
}

void bstep_stages(const fs2d_batch* B, std::vector<int>& before, std::vector<int>& after) {
    if (B->mode == FS2D_SEMILAGRANGIAN) {
        before = {FS2D_STAGE_CREATE_WATER_LEVEL_SET, FS2D_STAGE_SEMI_LAGRANGIAN_ADVECTION,
                  FS2D_STAGE_CREATE_SOLID_LEVEL_SET};
        after = {FS2D_STAGE_UPDATE_VELOCITY, FS2D_STAGE_APPLY_ADVECTION};
    } else {
        before = {FS2D_STAGE_CREATE_WATER_LEVEL_SET, FS2D_STAGE_TRANSFER_VELOCITY_TO_GRID, FS2D_STAGE_APPLY_GRAVITY,
                  FS2D_STAGE_CREATE_SOLID_LEVEL_SET};
        after = {FS2D_STAGE_UPDATE_VELOCITY, FS2D_STAGE_UPDATE_PARTICLE_VELOCITIES, FS2D_STAGE_APPLY_ADVECTION};
    }
}

// `pre`: the stages before the projection, the assembly, factor and
// graph_budget PCG iterations of every member; `post`: the pressures' copy
// and the stages after the projection.  Captured with the batch's host flags
// as fs2d_batch_step sets them (sort invalid at the start of a step).
fs2d_status bcapture_step(fs2d_batch* B, fs2d_batch::Captured& G) {
    G.maps = B->d_maps;
    G.budget = B->graph_budget;
    G.pcg_grid = B->pcg_grid;
    G.prefetch = B->sweep_prefetch ? 1 : 0;
    G.fused = B->fused_pcg ? 1 : 0;
    std::vector<int> st_before, st_after;
    bstep_stages(B, st_before, st_after);
    const bool fuse = B->mode == FS2D_SEMILAGRANGIAN;
    fs2d_status worst = FS2D_OK, st = FS2D_OK;
    B->sorted_valid = false;
    BHIPCHK(B, hipStreamBeginCapture(B->stream, hipStreamCaptureModeThreadLocal));
    for (int stg : st_before)
        if (st == FS2D_OK) st = bstage(B, stg, fuse, &worst);
    if (st == FS2D_OK) {
        bpcg_begin(B);
        bpcg_iterations(B, 0, G.budget);
    }
    fs2d_status ce = bend_capture(B, &G.pre);
    if (st == FS2D_OK && ce == FS2D_OK) {
        BHIPCHK(B, hipStreamBeginCapture(B->stream, hipStreamCaptureModeThreadLocal));
        bpcg_end(B);
        for (int stg : st_after)
            if (st == FS2D_OK) st = bstage(B, stg, fuse, &worst);
        ce = bend_capture(B, &G.post);
    }
    B->sorted_valid = false;
    if (st != FS2D_OK || ce != FS2D_OK) {
        bgraph_free(G);
        return st != FS2D_OK ? st : ce;
    }
    return FS2D_OK;
}

void bupdate_graph_budget(fs2d_batch* B, int needed) {
    const int want = std::min(B->max_iters, std::max(16, (needed + needed / 16 + 4 + 31) / 32 * 32));
    if (B->graph_budget <= 0 || needed > B->graph_budget || 2 * needed < B->graph_budget) B->graph_budget = want;
    B->graph_budget = std::min(B->graph_budget, B->max_iters);
}

// fs2d_batch_step by replaying a captured step: one host check of every
// member's convergence after the budget, chunks if some member needs more.
fs2d_status bstep_graph(fs2d_batch* B) {
    if (B->graph_budget <= 0) bupdate_graph_budget(B, B->last_iters_hint);
    fs2d_batch::Captured* G = nullptr;
    for (auto& g : B->graphs)
        if (g.maps == B->d_maps && g.budget == B->graph_budget && g.pcg_grid == B->pcg_grid &&
            g.prefetch == (B->sweep_prefetch ? 1 : 0) && g.fused == (B->fused_pcg ? 1 : 0))
            G = &g;
    if (!G) {
        if (B->graphs.size() >= 2) {
            bgraph_free(B->graphs.front());
            B->graphs.erase(B->graphs.begin());
        }
        fs2d_batch::Captured ng;
        TRY(bcapture_step(B, ng));
        B->graphs.push_back(ng);
        G = &B->graphs.back();
    }
    BHIPCHK(B, hipGraphLaunch(G->pre, B->stream));
    TRY(gather_status(B));
    if (!all_members_done(B) && G->budget < B->max_iters) TRY(bpcg_continue(B, G->budget, 8));
    BHIPCHK(B, hipGraphLaunch(G->post, B->stream));
    TRY(gather_status(B));
    const fs2d_status pst = bpcg_status(B);
    bupdate_graph_budget(B, B->last_iters_hint - 1);
    for (auto* s : B->sims) {
        s->current_time += s->dt;
        s->steps++;
    }
    return bfinish(B, pst);
}

}  // namespace

extern "C" {

fs2d_status fs2d_batch_set_graph(fs2d_batch* B, int enable) {
    if (!B) return FS2D_ERR_ARG;
    BHIPCHK(B, hipStreamSynchronize(B->stream));
    B->use_graph = enable != 0;
    if (!B->use_graph) {
        for (auto& g : B->graphs) bgraph_free(g);
        B->graphs.clear();
    }
    return FS2D_OK;
}

fs2d_status fs2d_batch_create(const fs2d_config* cfgs, int n, fs2d_batch** out) {
    if (!cfgs || !out || n < 1) return FS2D_ERR_ARG;
    *out = nullptr;
    for (int b = 1; b < n; b++) {
        const fs2d_config &a = cfgs[0], &c = cfgs[b];
        if (c.mode != a.mode || c.flags != a.flags || c.tau != a.tau || c.sigma != a.sigma ||
            c.pcg_tol != a.pcg_tol || c.pcg_max_iters != a.pcg_max_iters) {
            g_create_error = "batch members must share mode, flags and solver parameters (member " +
                             std::to_string(b) + " differs from member 0)";
            return FS2D_ERR_CONFIG;
        }
    }
    fs2d_batch* B = new fs2d_batch();
    B->n = n;
    B->mode = cfgs[0].mode;
    B->flags = cfgs[0].flags;
    B->tol = cfgs[0].pcg_tol;
    B->max_iters = cfgs[0].pcg_max_iters;
    auto bail = [&](fs2d_status st) {
        g_create_error = B->err;
        fs2d_batch_destroy(B);
        return st;
    };
    if (hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking) != hipSuccess) {
        B->err = "HIP device unavailable";
        return bail(FS2D_ERR_DEVICE);
    }
    for (int b = 0; b < n; b++) {
        fs2d_sim* s = nullptr;
        fs2d_status st = fs2d_create(&cfgs[b], &s);
        if (st != FS2D_OK) {
            B->err = "member " + std::to_string(b) + ": " + g_create_error;
            return bail(st);
        }
        B->sims.push_back(s);
    }
    // hand-off rows and descriptors, then the launch maps
    B->h_ins.resize(n);
    fs2d_status st = FS2D_OK;
    for (int b = 0; b < n; b++) {
        fs2d_sim* s = B->sims[b];
        const int ns = (s->ny + 63) / 64;
        unsigned long long *skh = nullptr, *fah = nullptr;
        unsigned* faf = nullptr;
        if ((st = balloc(B, &skh, (size_t)ns * sk_handoff_stride(s->nx, 3))) != FS2D_OK ||
            (st = balloc(B, &fah, (size_t)ns * sk_handoff_stride(s->nx, FwdFusedOp::NW))) != FS2D_OK ||
            (st = balloc(B, &faf, (size_t)kMaxBlocks)) != FS2D_OK)
            return bail(st);
        B->h_ins[b] = inst_bufs(s, skh, faf, fah);
    }
    if ((st = balloc(B, &B->d_ins, (size_t)n)) != FS2D_OK) return bail(st);
    if ((st = build_maps(B)) != FS2D_OK) return bail(st);
    if ((st = balloc(B, &B->werr, 1)) != FS2D_OK || (st = balloc(B, &B->all_done_since, 1)) != FS2D_OK ||
        (st = balloc(B, &B->sk_ticket, 1)) != FS2D_OK || (st = balloc(B, &B->fa_ticket, 1)) != FS2D_OK ||
        (st = balloc(B, &B->d_status, 4 * (size_t)n)) != FS2D_OK)
        return bail(st);
    B->plan_lds = sizeof(int) * (size_t)n;
    if (hipHostMalloc((void**)&B->h_status, 4 * n * sizeof(int), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&B->h_err, sizeof(int), hipHostMallocDefault) != hipSuccess) {
        B->err = "hipHostMalloc failed";
        return bail(FS2D_ERR_OOM);
    }
    if (hipMemcpyAsync(B->d_ins, B->h_ins.data(), n * sizeof(InstBufs), hipMemcpyHostToDevice, B->stream) !=
            hipSuccess ||
        hipStreamSynchronize(B->stream) != hipSuccess) {
        B->err = "batch upload failed";
        return bail(FS2D_ERR_DEVICE);
    }
    *out = B;
    return FS2D_OK;
}

void fs2d_batch_destroy(fs2d_batch* B) {
    if (!B) return;
    if (B->stream) (void)hipStreamSynchronize(B->stream);
    for (auto& g : B->graphs) bgraph_free(g);
    for (auto* s : B->sims) fs2d_destroy(s);
    for (void* p : B->allocs) (void)hipFree(p);
    for (auto& t : B->pending) {
        (void)hipEventDestroy(t.a);
        (void)hipEventDestroy(t.b);
    }
    for (auto e : B->event_pool) (void)hipEventDestroy(e);
    if (B->h_status) (void)hipHostFree(B->h_status);
    if (B->h_err) (void)hipHostFree(B->h_err);
    if (B->stream) (void)hipStreamDestroy(B->stream);
    delete B;
}

int fs2d_batch_size(const fs2d_batch* B) { return B ? B->n : 0; }

fs2d_sim* fs2d_batch_sim(fs2d_batch* B, int i) { return (B && i >= 0 && i < B->n) ? B->sims[i] : nullptr; }

fs2d_status fs2d_batch_step(fs2d_batch* B) {
    if (!B) return FS2D_ERR_ARG;
    TRY(refresh_inst(B));
    if (B->use_graph && !B->timing) return bstep_graph(B);
    B->sorted_valid = false;
    fs2d_status worst = FS2D_OK, st;
    if (B->mode == FS2D_SEMILAGRANGIAN) {
        const int stages[] = {FS2D_STAGE_CREATE_WATER_LEVEL_SET, FS2D_STAGE_SEMI_LAGRANGIAN_ADVECTION,
                              FS2D_STAGE_CREATE_SOLID_LEVEL_SET, FS2D_STAGE_APPLY_PROJECTION,
                              FS2D_STAGE_UPDATE_VELOCITY, FS2D_STAGE_APPLY_ADVECTION};
        for (int stg : stages)
            if ((st = bstage(B, stg, true, &worst)) != FS2D_OK) return st;
    } else {
        const int stages[] = {FS2D_STAGE_CREATE_WATER_LEVEL_SET, FS2D_STAGE_TRANSFER_VELOCITY_TO_GRID,
                              FS2D_STAGE_APPLY_GRAVITY, FS2D_STAGE_CREATE_SOLID_LEVEL_SET,
                              FS2D_STAGE_APPLY_PROJECTION, FS2D_STAGE_UPDATE_VELOCITY,
                              FS2D_STAGE_UPDATE_PARTICLE_VELOCITIES, FS2D_STAGE_APPLY_ADVECTION};
        for (int stg : stages)
            if ((st = bstage(B, stg, false, &worst)) != FS2D_OK) return st;
    }
    for (auto* s : B->sims) {
        s->current_time += s->dt;
        s->steps++;
    }
    return bfinish(B, worst);
}

fs2d_status fs2d_batch_run_stage(fs2d_batch* B, int stage) {
    if (!B) return FS2D_ERR_ARG;
    TRY(refresh_inst(B));
    if (stage == FS2D_STAGE_CREATE_WATER_LEVEL_SET || stage == FS2D_STAGE_TRANSFER_VELOCITY_TO_GRID) {
        // a member's particles may have been changed through its handle
        bool valid = true;
        for (auto* s : B->sims) valid = valid && s->sorted_valid;
        if (!valid) B->sorted_valid = false;
    }
    fs2d_status worst = FS2D_OK;
    fs2d_status st = bstage(B, stage, false, &worst);
    if (st != FS2D_OK) return st;
    const bool keep = B->sorted_valid;
    st = bfinish(B, worst);
    if (keep)
        for (auto* s : B->sims) s->sorted_valid = true;
    return st;
}

fs2d_status fs2d_batch_set_timing(fs2d_batch* B, int enable) {
    if (!B) return FS2D_ERR_ARG;
    BHIPCHK(B, hipStreamSynchronize(B->stream));
    resolve_timing(B);
    B->timing = enable ? ((1ull << K_COUNT) - 1ull) : 0ull;
    return FS2D_OK;
}

fs2d_status fs2d_batch_get_kernel_time(fs2d_batch* B, int k, double* total_ms, long long* launches) {
    if (!B || k < 0 || k >= K_COUNT || !total_ms || !launches) return FS2D_ERR_ARG;
    BHIPCHK(B, hipStreamSynchronize(B->stream));
    resolve_timing(B);
    *total_ms = B->ktime[k];
    *launches = B->kcount[k];
    return FS2D_OK;
}

fs2d_status fs2d_batch_reset_kernel_times(fs2d_batch* B) {
    if (!B) return FS2D_ERR_ARG;
    BHIPCHK(B, hipStreamSynchronize(B->stream));
    resolve_timing(B);
    for (int k = 0; k < K_COUNT; k++) {
        B->ktime[k] = 0;
        B->kcount[k] = 0;
    }
    return FS2D_OK;
}

const char* fs2d_batch_last_error(const fs2d_batch* B) { return B ? B->err.c_str() : g_create_error.c_str(); }

}  // extern "C"
