// fs3d_api.hpp -- C ABI of include/fs3d.h: the 3D MAC grid (include/MACGrid3D.h,
// include/Array3D.h) on gfx950.  Included at the end of fs2d_capi.hip (one
// translation unit, one code object).
#include "../../include/fs3d.h"
#include "fs3d_kernels.hpp"

struct fs3d_grid {
    int nx = 0, ny = 0, nz = 0;
    float* f[3] = {nullptr, nullptr, nullptr};  // u, v, w
    hipStream_t stream = nullptr;
    unsigned* range_err = nullptr;  // samples whose taps left the array
    double* partials = nullptr;     // one per reduction workgroup
    double* red_out = nullptr;
    void* stage = nullptr;          // device staging of host-memory calls
    size_t stage_bytes = 0;
    std::string err;
};

namespace {

constexpr int kRedBlocks = 1024;  // a multiple of 8: xcd_loop's contiguous chunks
static thread_local std::string g3_create_error;

fs2d_status fail3(fs3d_grid* g, fs2d_status st, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (g) g->err = buf;
    else g3_create_error = buf;
    return st;
}

#define HIPCHK3(g, x)                                                                                  \
    do {                                                                                               \
        hipError_t e_ = (x);                                                                           \
        if (e_ != hipSuccess)                                                                          \
            return fail3(g, e_ == hipErrorOutOfMemory ? FS2D_ERR_OOM : FS2D_ERR_DEVICE, "%s: %s", #x, \
                         hipGetErrorString(e_));                                                       \
    } while (0)

size_t count3(const fs3d_grid* g, int f) {
    return (size_t)(g->nx + (f == 0)) * (g->ny + (f == 1)) * (g->nz + (f == 2));
}
size_t cells3(const fs3d_grid* g) { return (size_t)g->nx * g->ny * g->nz; }

fs3d::Grid3 grid3(const fs3d_grid* g) { return fs3d::Grid3{g->f[0], g->f[1], g->f[2], g->nx, g->ny, g->nz}; }

int stream_blocks(size_t n) {  // grid-stride launches: a multiple of 8, at most 8192 workgroups
    const size_t b = (n + kBlock - 1) / kBlock;
    const size_t r = (std::min<size_t>(b, 8192) + 7) / 8 * 8;
    return (int)std::max<size_t>(r, 8);
}

fs2d_status ensure_stage(fs3d_grid* g, size_t bytes) {
    if (bytes <= g->stage_bytes) return FS2D_OK;
    if (g->stage) {
        HIPCHK3(g, hipStreamSynchronize(g->stream));
        HIPCHK3(g, hipFree(g->stage));
        g->stage = nullptr;
        g->stage_bytes = 0;
    }
    HIPCHK3(g, hipMalloc(&g->stage, bytes));
    g->stage_bytes = bytes;
    return FS2D_OK;
}

// FP32 device results (count values at d) to host memory of type t.
fs2d_status download3(fs3d_grid* g, const float* d, size_t count, fs2d_dtype t, void* host) {
    if (t == FS2D_F32) {
        HIPCHK3(g, hipMemcpyAsync(host, d, count * sizeof(float), hipMemcpyDeviceToHost, g->stream));
        HIPCHK3(g, hipStreamSynchronize(g->stream));
        return FS2D_OK;
    }
    std::vector<float> tmp(count);
    HIPCHK3(g, hipMemcpyAsync(tmp.data(), d, count * sizeof(float), hipMemcpyDeviceToHost, g->stream));
    HIPCHK3(g, hipStreamSynchronize(g->stream));
    double* o = (double*)host;
    for (size_t e = 0; e < count; e++) o[e] = tmp[e];
    return FS2D_OK;
}

fs2d_status check_out(fs3d_grid* g, fs2d_dtype t, const void* host, size_t count, size_t want) {
    if (!g) return FS2D_ERR_ARG;
    if (!host && want) return fail3(g, FS2D_ERR_ARG, "null host buffer");
    if (t != FS2D_F32 && t != FS2D_F64) return fail3(g, FS2D_ERR_ARG, "dtype %d: FS2D_F32 or FS2D_F64", (int)t);
    if (count != want) return fail3(g, FS2D_ERR_ARG, "count %zu, expected %zu", count, want);
    return FS2D_OK;
}

}  // namespace

extern "C" {

fs2d_status fs3d_create(int nx, int ny, int nz, fs3d_grid** out) {
    if (!out) return FS2D_ERR_ARG;
    *out = nullptr;
    if (nx < 1 || ny < 1 || nz < 1 || (double)(nx + 1) * (ny + 1) * (nz + 1) > 4e9)
        return fail3(nullptr, FS2D_ERR_CONFIG, "grid %d x %d x %d: each side >= 1, (nx+1)(ny+1)(nz+1) <= 4e9", nx,
                     ny, nz);
    fs3d_grid* g = new fs3d_grid();
    g->nx = nx;
    g->ny = ny;
    g->nz = nz;
    auto bail = [&](fs2d_status st) {
        g3_create_error = g->err;
        fs3d_destroy(g);
        return st;
    };
    hipError_t e = hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking);
    for (int f = 0; f < 3 && e == hipSuccess; f++) e = hipMalloc((void**)&g->f[f], count3(g, f) * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&g->range_err, sizeof(unsigned));
    if (e == hipSuccess) e = hipMalloc((void**)&g->partials, kRedBlocks * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void**)&g->red_out, sizeof(double));
    for (int f = 0; f < 3 && e == hipSuccess; f++)
        e = hipMemsetAsync(g->f[f], 0, count3(g, f) * sizeof(float), g->stream);
    if (e == hipSuccess) e = hipMemsetAsync(g->range_err, 0, sizeof(unsigned), g->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
    if (e != hipSuccess) {
        g->err = std::string("fs3d_create: ") + hipGetErrorString(e);
        return bail(e == hipErrorOutOfMemory ? FS2D_ERR_OOM : FS2D_ERR_DEVICE);
    }
    *out = g;
    return FS2D_OK;
}

void fs3d_destroy(fs3d_grid* g) {
    if (!g) return;
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    for (int f = 0; f < 3; f++)
        if (g->f[f]) (void)hipFree(g->f[f]);
    if (g->range_err) (void)hipFree(g->range_err);
    if (g->partials) (void)hipFree(g->partials);
    if (g->red_out) (void)hipFree(g->red_out);
    if (g->stage) (void)hipFree(g->stage);
    if (g->stream) (void)hipStreamDestroy(g->stream);
    delete g;
}

const char* fs3d_last_error(const fs3d_grid* g) { return g ? g->err.c_str() : g3_create_error.c_str(); }

void* fs3d_stream(fs3d_grid* g) { return g ? (void*)g->stream : nullptr; }

size_t fs3d_field_count(const fs3d_grid* g, fs3d_field f) {
    if (!g || (int)f < 0 || (int)f > 2) return 0;
    return count3(g, (int)f);
}

float* fs3d_device_field(fs3d_grid* g, fs3d_field f) {
    if (!g || (int)f < 0 || (int)f > 2) return nullptr;
    return g->f[(int)f];
}

fs2d_status fs3d_set_field(fs3d_grid* g, fs3d_field f, fs2d_dtype t, const void* host, size_t count) {
    if (!g) return FS2D_ERR_ARG;
    if ((int)f < 0 || (int)f > 2) return fail3(g, FS2D_ERR_ARG, "field %d", (int)f);
    fs2d_status st = check_out(g, t, host, count, count3(g, (int)f));
    if (st != FS2D_OK) return st;
    if (t == FS2D_F32) {
        HIPCHK3(g, hipMemcpyAsync(g->f[(int)f], host, count * sizeof(float), hipMemcpyHostToDevice, g->stream));
        HIPCHK3(g, hipStreamSynchronize(g->stream));
        return FS2D_OK;
    }
    std::vector<float> tmp(count);
    const double* h = (const double*)host;
    for (size_t e = 0; e < count; e++) tmp[e] = (float)h[e];
    HIPCHK3(g, hipMemcpyAsync(g->f[(int)f], tmp.data(), count * sizeof(float), hipMemcpyHostToDevice, g->stream));
    HIPCHK3(g, hipStreamSynchronize(g->stream));
    return FS2D_OK;
}

fs2d_status fs3d_get_field(fs3d_grid* g, fs3d_field f, fs2d_dtype t, void* host, size_t count) {
    if (!g) return FS2D_ERR_ARG;
    if ((int)f < 0 || (int)f > 2) return fail3(g, FS2D_ERR_ARG, "field %d", (int)f);
    fs2d_status st = check_out(g, t, host, count, count3(g, (int)f));
    if (st != FS2D_OK) return st;
    return download3(g, g->f[(int)f], count, t, host);
}

fs2d_status fs3d_vel_interp_device(fs3d_grid* g, const float* pos, size_t n, float* out, fs3d_sample mode) {
    if (!g) return FS2D_ERR_ARG;
    if (mode != FS3D_SAMPLE_REFERENCE && mode != FS3D_SAMPLE_VELOCITY)
        return fail3(g, FS2D_ERR_ARG, "sample mode %d", (int)mode);
    if (n == 0) return FS2D_OK;
    if (!pos || !out) return fail3(g, FS2D_ERR_ARG, "null device buffer");
    const int blocks = stream_blocks(n);
    if (mode == FS3D_SAMPLE_REFERENCE)
        hipLaunchKernelGGL((fs3d::vel_interp_kernel<float, true>), dim3(blocks), dim3(kBlock), 0, g->stream,
                           grid3(g), pos, (long)n, out, g->range_err);
    else
        hipLaunchKernelGGL((fs3d::vel_interp_kernel<float, false>), dim3(blocks), dim3(kBlock), 0, g->stream,
                           grid3(g), pos, (long)n, out, g->range_err);
    HIPCHK3(g, hipGetLastError());
    return FS2D_OK;
}

long fs3d_range_errors(fs3d_grid* g) {
    if (!g) return -1;
    unsigned h = 0;
    if (hipMemcpyAsync(&h, g->range_err, sizeof h, hipMemcpyDeviceToHost, g->stream) != hipSuccess ||
        hipMemsetAsync(g->range_err, 0, sizeof(unsigned), g->stream) != hipSuccess ||
        hipStreamSynchronize(g->stream) != hipSuccess)
        return -1;
    return (long)h;
}

fs2d_status fs3d_vel_interp(fs3d_grid* g, const double* pos, size_t n, double* out, fs3d_sample mode) {
    if (!g) return FS2D_ERR_ARG;
    if (mode != FS3D_SAMPLE_REFERENCE && mode != FS3D_SAMPLE_VELOCITY)
        return fail3(g, FS2D_ERR_ARG, "sample mode %d", (int)mode);
    if (n == 0) return FS2D_OK;
    if (!pos || !out) return fail3(g, FS2D_ERR_ARG, "null host buffer");
    if (fs3d_range_errors(g) < 0) return fail3(g, FS2D_ERR_DEVICE, "range-error counter unreadable");
    const size_t bytes = 3 * n * sizeof(double);
    fs2d_status st = ensure_stage(g, 2 * bytes);
    if (st != FS2D_OK) return st;
    double* dpos = (double*)g->stage;
    double* dout = dpos + 3 * n;
    HIPCHK3(g, hipMemcpyAsync(dpos, pos, bytes, hipMemcpyHostToDevice, g->stream));
    const int blocks = stream_blocks(n);
    if (mode == FS3D_SAMPLE_REFERENCE)
        hipLaunchKernelGGL((fs3d::vel_interp_kernel<double, true>), dim3(blocks), dim3(kBlock), 0, g->stream,
                           grid3(g), dpos, (long)n, dout, g->range_err);
    else
        hipLaunchKernelGGL((fs3d::vel_interp_kernel<double, false>), dim3(blocks), dim3(kBlock), 0, g->stream,
                           grid3(g), dpos, (long)n, dout, g->range_err);
    HIPCHK3(g, hipGetLastError());
    HIPCHK3(g, hipMemcpyAsync(out, dout, bytes, hipMemcpyDeviceToHost, g->stream));
    const long bad = fs3d_range_errors(g);  // waits for the stream
    if (bad < 0) return fail3(g, FS2D_ERR_DEVICE, "range-error counter unreadable");
    if (bad > 0)
        return fail3(g, FS2D_ERR_ARG,
                     "%ld of %zu samples read taps outside the array through the reference's (zp, yp, xp) "
                     "index; those samples are NaN",
                     bad, n);
    return FS2D_OK;
}

fs2d_status fs3d_vel_div_device(fs3d_grid* g, float* out) {
    if (!g) return FS2D_ERR_ARG;
    if (!out) return fail3(g, FS2D_ERR_ARG, "null device buffer");
    hipLaunchKernelGGL(fs3d::vel_div_kernel, dim3(stream_blocks(cells3(g))), dim3(kBlock), 0, g->stream, grid3(g),
                       out);
    HIPCHK3(g, hipGetLastError());
    return FS2D_OK;
}

fs2d_status fs3d_vel_cell_device(fs3d_grid* g, float* out) {
    if (!g) return FS2D_ERR_ARG;
    if (!out) return fail3(g, FS2D_ERR_ARG, "null device buffer");
    hipLaunchKernelGGL(fs3d::vel_cell_kernel, dim3(stream_blocks(cells3(g))), dim3(kBlock), 0, g->stream, grid3(g),
                       out);
    HIPCHK3(g, hipGetLastError());
    return FS2D_OK;
}

fs2d_status fs3d_vel_face_device(fs3d_grid* g, fs3d_field which, float* out) {
    if (!g) return FS2D_ERR_ARG;
    if ((int)which < 0 || (int)which > 2) return fail3(g, FS2D_ERR_ARG, "field %d", (int)which);
    if (!out) return fail3(g, FS2D_ERR_ARG, "null device buffer");
    const int blocks = stream_blocks(count3(g, (int)which));
    if (which == FS3D_U)
        hipLaunchKernelGGL(fs3d::vel_face_kernel<0>, dim3(blocks), dim3(kBlock), 0, g->stream, grid3(g), out);
    else if (which == FS3D_V)
        hipLaunchKernelGGL(fs3d::vel_face_kernel<1>, dim3(blocks), dim3(kBlock), 0, g->stream, grid3(g), out);
    else
        hipLaunchKernelGGL(fs3d::vel_face_kernel<2>, dim3(blocks), dim3(kBlock), 0, g->stream, grid3(g), out);
    HIPCHK3(g, hipGetLastError());
    return FS2D_OK;
}

fs2d_status fs3d_vel_div(fs3d_grid* g, fs2d_dtype t, void* host, size_t count) {
    fs2d_status st = check_out(g, t, host, count, g ? cells3(g) : 0);
    if (st != FS2D_OK) return st;
    if ((st = ensure_stage(g, count * sizeof(float))) != FS2D_OK) return st;
    if ((st = fs3d_vel_div_device(g, (float*)g->stage)) != FS2D_OK) return st;
    return download3(g, (const float*)g->stage, count, t, host);
}

fs2d_status fs3d_vel_cell(fs3d_grid* g, fs2d_dtype t, void* host, size_t count) {
    fs2d_status st = check_out(g, t, host, count, g ? 3 * cells3(g) : 0);
    if (st != FS2D_OK) return st;
    if ((st = ensure_stage(g, count * sizeof(float))) != FS2D_OK) return st;
    if ((st = fs3d_vel_cell_device(g, (float*)g->stage)) != FS2D_OK) return st;
    return download3(g, (const float*)g->stage, count, t, host);
}

fs2d_status fs3d_vel_face(fs3d_grid* g, fs3d_field which, fs2d_dtype t, void* host, size_t count) {
    if (!g) return FS2D_ERR_ARG;
    if ((int)which < 0 || (int)which > 2) return fail3(g, FS2D_ERR_ARG, "field %d", (int)which);
    fs2d_status st = check_out(g, t, host, count, 3 * count3(g, (int)which));
    if (st != FS2D_OK) return st;
    if ((st = ensure_stage(g, count * sizeof(float))) != FS2D_OK) return st;
    if ((st = fs3d_vel_face_device(g, which, (float*)g->stage)) != FS2D_OK) return st;
    return download3(g, (const float*)g->stage, count, t, host);
}

fs2d_status fs3d_array_inner_product(fs3d_grid* g, const float* a, const float* b, size_t n, double* out) {
    if (!g) return FS2D_ERR_ARG;
    if (!out || (n && (!a || !b))) return fail3(g, FS2D_ERR_ARG, "null buffer");
    hipLaunchKernelGGL(fs3d::array_reduce_kernel<false>, dim3(kRedBlocks), dim3(kBlock), 0, g->stream, a, b,
                       (long)n, g->partials);
    hipLaunchKernelGGL(fs3d::array_finalize_kernel<false>, dim3(1), dim3(64), 0, g->stream, g->partials,
                       kRedBlocks, g->red_out);
    HIPCHK3(g, hipGetLastError());
    HIPCHK3(g, hipMemcpyAsync(out, g->red_out, sizeof(double), hipMemcpyDeviceToHost, g->stream));
    HIPCHK3(g, hipStreamSynchronize(g->stream));
    return FS2D_OK;
}

fs2d_status fs3d_array_infinite_norm(fs3d_grid* g, const float* a, size_t n, double* out) {
    if (!g) return FS2D_ERR_ARG;
    if (!out || (n && !a)) return fail3(g, FS2D_ERR_ARG, "null buffer");
    hipLaunchKernelGGL(fs3d::array_reduce_kernel<true>, dim3(kRedBlocks), dim3(kBlock), 0, g->stream, a, a,
                       (long)n, g->partials);
    hipLaunchKernelGGL(fs3d::array_finalize_kernel<true>, dim3(1), dim3(64), 0, g->stream, g->partials,
                       kRedBlocks, g->red_out);
    HIPCHK3(g, hipGetLastError());
    HIPCHK3(g, hipMemcpyAsync(out, g->red_out, sizeof(double), hipMemcpyDeviceToHost, g->stream));
    HIPCHK3(g, hipStreamSynchronize(g->stream));
    return FS2D_OK;
}

fs2d_status fs3d_array_multiply_add(fs3d_grid* g, float* dst, const float* a, double b, const float* c, size_t n) {
    if (!g) return FS2D_ERR_ARG;
    if (n && (!dst || !a || !c)) return fail3(g, FS2D_ERR_ARG, "null buffer");
    if (n == 0) return FS2D_OK;
    hipLaunchKernelGGL(fs3d::array_multiply_add_kernel, dim3(stream_blocks(n)), dim3(kBlock), 0, g->stream, dst, a,
                       b, c, (long)n);
    HIPCHK3(g, hipGetLastError());
    return FS2D_OK;
}

}  // extern "C"
