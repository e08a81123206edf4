// fs2d_adjoint_api.hpp -- the fs2d_adjoint_* entry points of include/fs2d.h:
// checkpointed reverse-mode differentiation of a rollout of the all-fluid
// semi-Lagrangian step.  Included at the end of fs2d_capi.hip (it uses that
// file's stage functions and allocation helpers).
//
// forward(n) runs n steps from the handle's state, keeping the state before
// steps 0, k, 2k, ... (k = checkpoint_every).  backward() walks the segments
// from last to first: it restores the segment's checkpoint, recomputes the
// segment's steps keeping each step's input state, then applies the adjoint
// step (fs2d_adjoint.hpp) in reverse order.  Recomputed steps are bit-identical
// to the recorded ones: every kernel of the step is deterministic.

struct ExtraAdjHost {
    std::vector<int> tgt, off, zero;
    std::vector<int4> src;
    std::vector<float4> wt;
};

struct fs2d_adjoint {
    fs2d_sim* sim = nullptr;
    int every = 1, max_steps = 0, recorded = -1;
    bool dens = false;
    std::vector<float*> ck_u, ck_v, ck_d;  // state before steps 0, k, 2k, ...
    std::vector<float*> tr_u, tr_v, tr_d;  // input states of the segment being reversed
    float *fin_u = nullptr, *fin_v = nullptr, *fin_d = nullptr;
    double fin_time = 0;
    long long fin_steps = 0;
    float *gu = nullptr, *gv = nullptr, *gd = nullptr;  // adjoint of the state
    float* lam = nullptr;
    unsigned long long *au = nullptr, *av = nullptr, *ad = nullptr;
    unsigned* gmax = nullptr;
    // extrapolation transpose lists (u then v), rebuilt by forward()
    int* ex_buf = nullptr;
    size_t ex_cap = 0;
    ExtraAdj ex[2]{};
    std::vector<void*> allocs;
};

namespace {

fs2d_status adj_alloc(fs2d_adjoint* a, void** p, size_t bytes) {
    fs2d_sim* s = a->sim;
    *p = nullptr;
    hipError_t e = hipMalloc(p, std::max<size_t>(bytes, 4));
    if (e != hipSuccess) return fail(s, FS2D_ERR_OOM, "hipMalloc(%zu bytes): %s", bytes, hipGetErrorString(e));
    a->allocs.push_back(*p);
    HIPCHK(s, hipMemsetAsync(*p, 0, std::max<size_t>(bytes, 4), s->stream));
    return FS2D_OK;
}
template <class T>
fs2d_status adj_alloc_n(fs2d_adjoint* a, T** p, size_t n) {
    return adj_alloc(a, (void**)p, n * sizeof(T));
}

// Scope of the adjoint: SL mode, static flags without EMPTY cells, no particles.
fs2d_status adj_check_scope(fs2d_sim* s, std::vector<uint8_t>* cells_out) {
    if (s->mode != FS2D_SEMILAGRANGIAN) return fail(s, FS2D_ERR_CONFIG, "adjoint: needs FS2D_SEMILAGRANGIAN");
    if (!(s->flags & FS2D_FLAG_STATIC_CELLS)) return fail(s, FS2D_ERR_CONFIG, "adjoint: needs FS2D_FLAG_STATIC_CELLS");
    if (s->np != 0) return fail(s, FS2D_ERR_CONFIG, "adjoint: the handle has %ld particles", s->np);
    std::vector<uint8_t> c((size_t)n_c(s));
    HIPCHK(s, hipStreamSynchronize(s->stream));
    TRY(d2h(s, c.data(), s->cell, c.size()));
    for (size_t k = 0; k < c.size(); k++)
        if (c[k] == FS2D_EMPTY)
            return fail(s, FS2D_ERR_CONFIG, "adjoint: EMPTY cell at (%d,%d)", (int)(k % s->nx), (int)(k / s->nx));
    if (cells_out) cells_out->swap(c);
    return FS2D_OK;
}

// Layers of the faces exactly as update_velocity_kernel marks them and
// extrap_layer1/extrap_layers_kernel number them, then the per-stage gather
// lists of adj_extrap_kernel (stage order L = Lmax .. 1).
void build_extrap_lists(const std::vector<uint8_t>& cell, int nx, int ny, bool is_u, ExtraAdjHost& H) {
    const int w = is_u ? nx + 1 : nx, h = is_u ? ny : ny + 1;
    const uint32_t UNK = 0xFFFFFFFFu;
    std::vector<uint32_t> m((size_t)w * h, 0);
    auto C = [&](int i, int j) { return cell[(size_t)j * nx + i]; };
    for (int j = 0; j < h; j++)
        for (int i = 0; i < w; i++) {
            bool adj;
            if (is_u) {
                if (i >= nx) continue;
                adj = (i > 0 && C(i - 1, j) == FS2D_FLUID) || C(i, j) == FS2D_FLUID;
            } else {
                if (j >= ny) continue;
                adj = (j > 0 && C(i, j - 1) == FS2D_FLUID) || C(i, j) == FS2D_FLUID;
            }
            if (!adj) m[(size_t)j * w + i] = UNK;
        }
    auto nbrs = [&](int f, int out[4]) {
        const int i = f % w, j = f / w;
        out[0] = i > 0 ? f - 1 : -1;
        out[1] = i < w - 1 ? f + 1 : -1;
        out[2] = j > 0 ? f - w : -1;
        out[3] = j < h - 1 ? f + w : -1;
    };
    std::vector<std::vector<int>> layers(1);
    for (int f = 0; f < w * h; f++) {
        if (m[f] == 0) continue;
        int nb[4];
        nbrs(f, nb);
        bool k = false;
        for (int q = 0; q < 4; q++) k = k || (nb[q] >= 0 && m[nb[q]] == 0);
        if (k) layers[0].push_back(f);
    }
    for (int f : layers[0]) m[f] = 1;
    while (!layers.back().empty()) {
        const uint32_t L = (uint32_t)layers.size();
        std::vector<int> next;
        for (int f : layers.back()) {
            int nb[4];
            nbrs(f, nb);
            for (int q = 0; q < 4; q++)
                if (nb[q] >= 0 && m[nb[q]] == UNK) {
                    m[nb[q]] = L + 1;
                    next.push_back(nb[q]);
                }
        }
        layers.push_back(std::move(next));
    }
    layers.pop_back();  // the empty last frontier
    H = ExtraAdjHost{};
    H.off.push_back(0);
    for (int L = (int)layers.size(); L >= 1; L--) {
        std::vector<int> tg;
        for (int f : layers[L - 1]) {
            int nb[4];
            nbrs(f, nb);
            for (int q = 0; q < 4; q++)
                if (nb[q] >= 0 && m[nb[q]] == (uint32_t)(L - 1)) tg.push_back(nb[q]);
        }
        std::sort(tg.begin(), tg.end());
        tg.erase(std::unique(tg.begin(), tg.end()), tg.end());
        for (int t : tg) {
            int nb[4], sv[4] = {-1, -1, -1, -1};
            float wv[4] = {0, 0, 0, 0};
            nbrs(t, nb);
            for (int q = 0; q < 4; q++) {
                const int f = nb[q];
                if (f < 0 || m[f] != (uint32_t)L) continue;
                int fb[4], cnt = 0;
                nbrs(f, fb);
                for (int r = 0; r < 4; r++) cnt += (fb[r] >= 0 && m[fb[r]] < (uint32_t)L);
                sv[q] = f;
                wv[q] = 1.0f / (float)cnt;  // the forward divides the sum by cnt
            }
            H.tgt.push_back(t);
            H.src.push_back(make_int4(sv[0], sv[1], sv[2], sv[3]));
            H.wt.push_back(make_float4(wv[0], wv[1], wv[2], wv[3]));
        }
        H.off.push_back((int)H.tgt.size());
    }
    for (auto& lay : layers)
        for (int f : lay) H.zero.push_back(f);
}

fs2d_status adj_upload_lists(fs2d_adjoint* a, const std::vector<uint8_t>& cells) {
    fs2d_sim* s = a->sim;
    ExtraAdjHost H[2];
    build_extrap_lists(cells, s->nx, s->ny, true, H[0]);
    build_extrap_lists(cells, s->nx, s->ny, false, H[1]);
    // one buffer: per array tgt | off | zero (ints), src (int4), wt (float4)
    size_t words = 0;
    for (auto& h : H) words += h.tgt.size() * 9 + h.off.size() + h.zero.size() + 8;
    if (words > a->ex_cap) {
        TRY(adj_alloc_n(a, &a->ex_buf, words));
        a->ex_cap = words;
    }
    std::vector<int> buf(words, 0);
    size_t pos = 0;
    float* gs[2] = {a->gu, a->gv};
    for (int k = 0; k < 2; k++) {
        auto& h = H[k];
        ExtraAdj& X = a->ex[k];
        X.g = gs[k];
        pos = (pos + 3) & ~(size_t)3;  // int4 alignment
        X.src = (const int4*)(a->ex_buf + pos);
        memcpy(&buf[pos], h.src.data(), h.src.size() * sizeof(int4));
        pos += h.src.size() * 4;
        X.wt = (const float4*)(a->ex_buf + pos);
        memcpy(&buf[pos], h.wt.data(), h.wt.size() * sizeof(float4));
        pos += h.wt.size() * 4;
        X.tgt = a->ex_buf + pos;
        std::copy(h.tgt.begin(), h.tgt.end(), buf.begin() + pos);
        pos += h.tgt.size();
        X.stage_off = a->ex_buf + pos;
        std::copy(h.off.begin(), h.off.end(), buf.begin() + pos);
        pos += h.off.size();
        X.nstages = (int)h.off.size() - 1;
        X.zero = a->ex_buf + pos;
        std::copy(h.zero.begin(), h.zero.end(), buf.begin() + pos);
        pos += h.zero.size();
        X.nzero = (int)h.zero.size();
    }
    return h2d(s, a->ex_buf, buf.data(), pos * sizeof(int));
}

void adj_save(fs2d_adjoint* a, float* u, float* v, float* d) {
    fs2d_sim* s = a->sim;
    dcopy(s, u, s->u, (size_t)n_u(s));
    dcopy(s, v, s->v, (size_t)n_v(s));
    if (a->dens) dcopy(s, d, s->dens, (size_t)n_c(s));
}
void adj_restore(fs2d_adjoint* a, const float* u, const float* v, const float* d) {
    fs2d_sim* s = a->sim;
    dcopy(s, s->u, u, (size_t)n_u(s));
    dcopy(s, s->v, v, (size_t)n_v(s));
    if (a->dens) dcopy(s, s->dens, d, (size_t)n_c(s));
}

// One adjoint step at the handle's current state: (gu, gv, gd) hold the
// adjoints of the step's outputs and receive those of its inputs.
fs2d_status adjoint_step_impl(fs2d_adjoint* a) {
    fs2d_sim* s = a->sim;
    // updateVelocity's extrapolation, transposed
    {
        Timed t(s, K_ADJ_OTHER);
        hipLaunchKernelGGL(adj_extrap_kernel, dim3(2), dim3(1024), 0, s->stream, a->ex[0], a->ex[1]);
    }
    // pressure: p-bar, then lam = A^-1 p-bar with the forward solver (A = A^T);
    // the solver holds A / scaleA, so its rhs is p-bar / scaleA:
    // (dt / (rho dx)) / (dt / (rho dx^2)) = dx
    assemble(s);
    {
        Timed t(s, K_ADJ_OTHER);
        hipLaunchKernelGGL(adj_pressure_rhs_kernel, dim3(sk_blocks(s)), dim3(kBlock), 0, s->stream, s->geom, s->cell,
                           a->gu, a->gv, s->r, s->pk, s->s[0], s->dx, s->rmax_part);
    }
    fs2d_status st = pcg_solve(s, a->lam);
    if (st < 0) return st;
    {
        Timed t(s, K_ADJ_OTHER);
        hipLaunchKernelGGL(adj_update_velocity_kernel, dim3(blocks_for(n_u(s) + n_v(s))), dim3(kBlock), 0, s->stream,
                           s->cell, a->lam, a->gu, a->gv, s->nx, s->ny, 1.0 / s->dx);
    }
    // applyGravity: identity.  Semi-Lagrangian advection, transposed.
    {
        Timed t(s, K_ADJ_SL);
        const long npts = n_u(s) + n_v(s) + (a->dens ? n_c(s) : 0);
        const double terms = (double)npts * 7.0 * 16.0;
        AdjSl p{mac_view(s), a->dens ? s->dens : nullptr, a->gu, a->gv, a->gd, a->au, a->av, a->ad, a->gmax,
                s->dt, clamp_box(s->nx, s->ny, s->dx), (int)std::ceil(std::log2(terms))};
        fill(s, a->gmax, 0u, 1);
        hipLaunchKernelGGL(adj_sl_kernel<false>, dim3(blocks_for(npts)), dim3(kBlock), 0, s->stream, p);
        hipLaunchKernelGGL(adj_sl_kernel<true>, dim3(blocks_for(npts)), dim3(kBlock), 0, s->stream, p);
        hipLaunchKernelGGL(adj_fixed_to_float_kernel, dim3(blocks_for(npts)), dim3(kBlock), 0, s->stream, p, a->gu,
                           a->gv, a->gd);
    }
    return st;
}

float* adj_field(fs2d_adjoint* a, int field) {
    switch (field) {
        case FS2D_FIELD_U: return a->gu;
        case FS2D_FIELD_V: return a->gv;
        case FS2D_FIELD_DENSITY: return a->dens ? a->gd : nullptr;
        default: return nullptr;
    }
}

}  // namespace

extern "C" {

fs2d_status fs2d_adjoint_create(fs2d_sim* s, int checkpoint_every, int max_steps, fs2d_adjoint** out) {
    if (!s || !out) return FS2D_ERR_ARG;
    *out = nullptr;
    if (checkpoint_every < 1 || max_steps < 1)
        return fail(s, FS2D_ERR_ARG, "adjoint: checkpoint_every %d and max_steps %d must be >= 1", checkpoint_every,
                    max_steps);
    TRY(adj_check_scope(s, nullptr));
    fs2d_adjoint* a = new fs2d_adjoint();
    a->sim = s;
    a->every = std::min(checkpoint_every, max_steps);
    a->max_steps = max_steps;
    a->dens = (s->flags & FS2D_FLAG_DENSITY) != 0;
    const size_t nU = n_u(s), nV = n_v(s), nC = n_c(s), nD = a->dens ? nC : 1;
    const int nck = (max_steps + a->every - 1) / a->every;
    fs2d_status st = FS2D_OK;
    auto grab = [&](float** p, size_t n) {
        if (st == FS2D_OK) st = adj_alloc_n(a, p, n);
    };
    a->ck_u.assign(nck, nullptr); a->ck_v.assign(nck, nullptr); a->ck_d.assign(nck, nullptr);
    a->tr_u.assign(a->every, nullptr); a->tr_v.assign(a->every, nullptr); a->tr_d.assign(a->every, nullptr);
    for (int k = 0; k < nck; k++) { grab(&a->ck_u[k], nU); grab(&a->ck_v[k], nV); grab(&a->ck_d[k], nD); }
    for (int k = 0; k < a->every; k++) { grab(&a->tr_u[k], nU); grab(&a->tr_v[k], nV); grab(&a->tr_d[k], nD); }
    grab(&a->fin_u, nU); grab(&a->fin_v, nV); grab(&a->fin_d, nD);
    grab(&a->gu, nU); grab(&a->gv, nV); grab(&a->gd, nD); grab(&a->lam, nC);
    if (st == FS2D_OK) st = adj_alloc_n(a, &a->au, nU);
    if (st == FS2D_OK) st = adj_alloc_n(a, &a->av, nV);
    if (st == FS2D_OK) st = adj_alloc_n(a, &a->ad, nD);
    if (st == FS2D_OK) st = adj_alloc_n(a, &a->gmax, 1);
    if (st == FS2D_OK && hipStreamSynchronize(s->stream) != hipSuccess)
        st = fail(s, FS2D_ERR_DEVICE, "adjoint: device synchronisation failed");
    if (st != FS2D_OK) {
        fs2d_adjoint_destroy(a);
        return st;
    }
    *out = a;
    return FS2D_OK;
}

void fs2d_adjoint_destroy(fs2d_adjoint* a) {
    if (!a) return;
    if (a->sim && a->sim->stream) (void)hipStreamSynchronize(a->sim->stream);
    for (void* p : a->allocs) (void)hipFree(p);
    delete a;
}

fs2d_status fs2d_adjoint_forward(fs2d_adjoint* a, int nsteps) {
    if (!a) return FS2D_ERR_ARG;
    fs2d_sim* s = a->sim;
    if (nsteps < 1 || nsteps > a->max_steps)
        return fail(s, FS2D_ERR_ARG, "adjoint: nsteps %d outside [1, %d]", nsteps, a->max_steps);
    std::vector<uint8_t> cells;
    TRY(adj_check_scope(s, &cells));
    TRY(adj_upload_lists(a, cells));
    a->recorded = -1;
    fs2d_status worst = FS2D_OK;
    for (int t = 0; t < nsteps; t++) {
        if (t % a->every == 0) adj_save(a, a->ck_u[t / a->every], a->ck_v[t / a->every], a->ck_d[t / a->every]);
        fs2d_status st = fs2d_step(s);
        if (st < 0) return st;
        worst = std::max(worst, st);
    }
    adj_save(a, a->fin_u, a->fin_v, a->fin_d);
    a->fin_time = s->current_time;
    a->fin_steps = s->steps;
    HIPCHK(s, hipStreamSynchronize(s->stream));
    a->recorded = nsteps;
    return worst;
}

fs2d_status fs2d_adjoint_backward(fs2d_adjoint* a) {
    if (!a) return FS2D_ERR_ARG;
    fs2d_sim* s = a->sim;
    if (a->recorded < 1) return fail(s, FS2D_ERR_ARG, "adjoint: backward() needs a forward() first");
    const int n = a->recorded, k = a->every;
    fs2d_status worst = FS2D_OK;
    for (int seg = (n + k - 1) / k - 1; seg >= 0; seg--) {
        const int t0 = seg * k, t1 = std::min(n, t0 + k);
        adj_restore(a, a->ck_u[seg], a->ck_v[seg], a->ck_d[seg]);
        for (int t = t0; t < t1; t++) {
            adj_save(a, a->tr_u[t - t0], a->tr_v[t - t0], a->tr_d[t - t0]);
            if (t + 1 < t1) {
                fs2d_status st = fs2d_step(s);
                if (st < 0) return st;
            }
        }
        for (int t = t1 - 1; t >= t0; t--) {
            adj_restore(a, a->tr_u[t - t0], a->tr_v[t - t0], a->tr_d[t - t0]);
            fs2d_status st = finish(s, adjoint_step_impl(a));
            if (st < 0) return st;
            worst = std::max(worst, st);
        }
    }
    adj_restore(a, a->fin_u, a->fin_v, a->fin_d);
    s->current_time = a->fin_time;
    s->steps = a->fin_steps;
    HIPCHK(s, hipStreamSynchronize(s->stream));
    return worst;
}

fs2d_status fs2d_adjoint_step(fs2d_adjoint* a) {
    if (!a) return FS2D_ERR_ARG;
    fs2d_sim* s = a->sim;
    std::vector<uint8_t> cells;
    TRY(adj_check_scope(s, &cells));
    TRY(adj_upload_lists(a, cells));
    return finish(s, adjoint_step_impl(a));
}

fs2d_status fs2d_adjoint_get_field(fs2d_adjoint* a, int field, int dtype, void* host, size_t count) {
    if (!a || !host) return FS2D_ERR_ARG;
    fs2d_sim* s = a->sim;
    float* f = adj_field(a, field);
    if (!f) return fail(s, FS2D_ERR_ARG, "adjoint: no field %d", field);
    const size_t n = fs2d_field_count(s, field);
    if (count != n) return fail(s, FS2D_ERR_ARG, "field %d has %zu elements, buffer has %zu", field, n, count);
    if (dtype != FS2D_F32 && dtype != FS2D_F64) return fail(s, FS2D_ERR_ARG, "dtype %d", dtype);
    HIPCHK(s, hipStreamSynchronize(s->stream));
    std::vector<float> tmp(n);
    TRY(d2h(s, tmp.data(), f, n * sizeof(float)));
    if (dtype == FS2D_F32) memcpy(host, tmp.data(), n * 4);
    else for (size_t i = 0; i < n; i++) ((double*)host)[i] = tmp[i];
    return FS2D_OK;
}

fs2d_status fs2d_adjoint_set_field(fs2d_adjoint* a, int field, int dtype, const void* host, size_t count) {
    if (!a || !host) return FS2D_ERR_ARG;
    fs2d_sim* s = a->sim;
    float* f = adj_field(a, field);
    if (!f) return fail(s, FS2D_ERR_ARG, "adjoint: no field %d", field);
    const size_t n = fs2d_field_count(s, field);
    if (count != n) return fail(s, FS2D_ERR_ARG, "field %d has %zu elements, buffer has %zu", field, n, count);
    if (dtype != FS2D_F32 && dtype != FS2D_F64) return fail(s, FS2D_ERR_ARG, "dtype %d", dtype);
    std::vector<float> tmp(n);
    for (size_t i = 0; i < n; i++)
        tmp[i] = dtype == FS2D_F32 ? ((const float*)host)[i] : (float)((const double*)host)[i];
    HIPCHK(s, hipStreamSynchronize(s->stream));
    return h2d(s, f, tmp.data(), n * sizeof(float));
}

fs2d_status fs2d_adjoint_device_field(fs2d_adjoint* a, int field, void** dptr) {
    if (!a || !dptr) return FS2D_ERR_ARG;
    float* f = adj_field(a, field);
    if (!f) return fail(a->sim, FS2D_ERR_ARG, "adjoint: no field %d", field);
    *dptr = f;
    return FS2D_OK;
}

}  // extern "C"
