// MACGrid3D.hpp -- the reference's MACGrid3D / Array3D API over the gfx950
// library (include/fs3d.h; link -lfs2d).
//
// Header-only.  Code written against include/MACGrid3D.h and include/Array3D.h
// compiles against this header instead:
//
//   MACGrid3D mac = MACGrid3D::create(nx, ny, nz);   // MACGrid3D.h:16-23
//   mac.u(i, j, k) = 1.0;                            // Array3D.h:32-37
//   vec3d v = mac.velInterp(vec3d{x, y, z});         // MACGrid3D.h:161-163
//   double d = mac.velDiv(i, j, k);                  // MACGrid3D.h:165-167
//   mac.free();                                      // MACGrid3D.h:25-29
//
// u, v and w are host mirrors of the device arrays: element access through
// operator() downloads the array the first time after it changed on the
// device; a non-const access marks it dirty, and dirty arrays are uploaded
// before the next device call.  Raw `.data` pointers bypass both: call
// pull() before reading through them and markDirty() after writing.
// vel / velU / velV / velW / velDiv at one index read a whole-grid result that
// the GPU computes once per state of u, v, w.  velInterp of one point is one
// GPU call; the batched overload samples many points in one launch.
// sampleMode selects the reference's tap order (default) or the velocity one
// (fs3d_sample in include/fs3d.h).  Errors throw std::runtime_error.
//
// Define FS3D_HAVE_VEC3D before including this header when another header
// (the reference's altmath) already defines `vec3d` with members x, y, z.
#ifndef FS3D_MACGRID3D_HPP
#define FS3D_MACGRID3D_HPP

#include <cstddef>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/fs3d.h"

#ifndef FS3D_HAVE_VEC3D
struct vec3d {
    double x, y, z;
};
#endif

namespace fs3d_detail {

inline void check(fs3d_grid* g, fs2d_status st, const char* what) {
    if (st == FS2D_OK) return;
    std::string msg = std::string(what) + ": " + fs2d_status_string(st);
    const char* e = fs3d_last_error(g);
    if (e && *e) msg += std::string(" (") + e + ")";
    throw std::runtime_error(msg);
}

// Shared state of one grid: the handle, the host mirrors and the cached
// whole-grid results.  Copies of a MACGrid3D alias it, as the reference's
// value type aliases its arrays.
struct State {
    fs3d_grid* g = nullptr;
    std::vector<double> host[3];
    bool valid[3] = {false, false, false}, dirty[3] = {false, false, false};
    unsigned long version = 0;  // bumped whenever u, v or w may change
    std::vector<double> cell, face[3], div;
    unsigned long cell_v = ~0ul, face_v[3] = {~0ul, ~0ul, ~0ul}, div_v = ~0ul;

    void read(int f) {
        if (!valid[f]) {
            check(g, fs3d_get_field(g, (fs3d_field)f, FS2D_F64, host[f].data(), host[f].size()), "fs3d_get_field");
            valid[f] = true;
        }
    }
    void write(int f) {
        read(f);
        dirty[f] = true;
        version++;
    }
    void push() {
        for (int f = 0; f < 3; f++)
            if (dirty[f]) {
                check(g, fs3d_set_field(g, (fs3d_field)f, FS2D_F64, host[f].data(), host[f].size()),
                      "fs3d_set_field");
                dirty[f] = false;
                valid[f] = false;  // the device holds FP32: re-read what it kept
            }
    }
    void changed_on_device(int f) {
        valid[f] = false;
        version++;
    }
};

}  // namespace fs3d_detail

// Array3D<double> (Array3D.h:15-127) as a mirror of u, v or w: element
// (i, j, k) at data[k*NY*NX + j*NX + i].
struct Array3DMirror {
    double* data = nullptr;
    int NX = 0, NY = 0, NZ = 0;
    int field = 0;
    std::shared_ptr<fs3d_detail::State> st;

    double& operator()(int i, int j, int k) {
        st->write(field);
        return data[((size_t)k * NY + j) * NX + i];
    }
    const double& operator()(int i, int j, int k) const {
        st->read(field);
        return data[((size_t)k * NY + j) * NX + i];
    }
    void pull() { st->read(field); }
    void markDirty() {
        st->valid[field] = true;
        st->dirty[field] = true;
        st->version++;
    }
    size_t size() const { return (size_t)NX * NY * NZ; }
    float* device() const { return fs3d_device_field(st->g, (fs3d_field)field); }

    // Array3D::innerProduct / infiniteNorm (Array3D.h:55-72) on the GPU
    double innerProduct(const Array3DMirror& rhs) const {
        if (rhs.size() != size()) throw std::runtime_error("innerProduct: arrays differ in size");
        st->push();
        rhs.st->push();
        double r = 0;
        fs3d_detail::check(st->g, fs3d_array_inner_product(st->g, device(), rhs.device(), size(), &r),
                           "fs3d_array_inner_product");
        return r;
    }
    double infiniteNorm() const {
        st->push();
        double r = 0;
        fs3d_detail::check(st->g, fs3d_array_infinite_norm(st->g, device(), size(), &r), "fs3d_array_infinite_norm");
        return r;
    }
    // Array3D::setMultiplyAdd (Array3D.h:48-53): this = a + b c, on the GPU
    void setMultiplyAdd(const Array3DMirror& a, double b, const Array3DMirror& c) {
        if (a.size() != size() || c.size() != size()) throw std::runtime_error("setMultiplyAdd: arrays differ in size");
        st->push();
        a.st->push();
        c.st->push();
        fs3d_detail::check(st->g, fs3d_array_multiply_add(st->g, device(), a.device(), b, c.device(), size()),
                           "fs3d_array_multiply_add");
        // the other grids' streams must see the inputs: they were uploaded synchronously
        st->changed_on_device(field);
    }
    // Array3D::operator+= (Array3D.h:39-46)
    Array3DMirror& operator+=(const Array3DMirror& rhs) {
        setMultiplyAdd(*this, 1.0, rhs);
        return *this;
    }
};

struct MACGrid3D {
    Array3DMirror u, v, w;
    int sizeX = 0, sizeY = 0, sizeZ = 0;
    fs3d_sample sampleMode = FS3D_SAMPLE_REFERENCE;

    // MACGrid3D::create (MACGrid3D.h:16-23); the arrays start at zero
    static MACGrid3D create(int nx, int ny, int nz) {
        auto st = std::make_shared<fs3d_detail::State>();
        fs3d_detail::check(nullptr, fs3d_create(nx, ny, nz, &st->g), "fs3d_create");
        MACGrid3D m;
        m.sizeX = nx;
        m.sizeY = ny;
        m.sizeZ = nz;
        Array3DMirror* arr[3] = {&m.u, &m.v, &m.w};
        for (int f = 0; f < 3; f++) {
            Array3DMirror& a = *arr[f];
            a.NX = nx + (f == 0);
            a.NY = ny + (f == 1);
            a.NZ = nz + (f == 2);
            a.field = f;
            a.st = st;
            st->host[f].assign(a.size(), 0.0);
            st->valid[f] = true;  // zero on both sides
            a.data = st->host[f].data();
        }
        return m;
    }
    // MACGrid3D::free (MACGrid3D.h:25-29)
    void free() {
        if (u.st && u.st->g) {
            fs3d_destroy(u.st->g);
            u.st->g = nullptr;
        }
    }

    // MACGrid3D::iterate* (MACGrid3D.h:31-71): host loops, as in the reference
    template <typename Fun>
    void iterateU(Fun f) {
        for (int k = 0; k < sizeZ; k++)
            for (int j = 0; j < sizeY; j++)
                for (int i = 0; i < sizeX + 1; i++) f(i, j, k);
    }
    template <typename Fun>
    void iterateV(Fun f) {
        for (int k = 0; k < sizeZ; k++)
            for (int j = 0; j < sizeY + 1; j++)
                for (int i = 0; i < sizeX; i++) f(i, j, k);
    }
    template <typename Fun>
    void iterateW(Fun f) {
        for (int k = 0; k < sizeZ + 1; k++)
            for (int j = 0; j < sizeY; j++)
                for (int i = 0; i < sizeX; i++) f(i, j, k);
    }
    template <typename Fun>
    void iterate(Fun f) {
        for (int k = 0; k < sizeZ; k++)
            for (int j = 0; j < sizeY; j++)
                for (int i = 0; i < sizeX; i++) f(i, j, k);
    }
    template <typename Fun>
    void iterateBackwards(Fun f) {
        for (int k = sizeZ; k-- > 0;)
            for (int j = sizeY; j-- > 0;)
                for (int i = sizeX; i-- > 0;) f(i, j, k);
    }

    // MACGrid3D::vel (:73-78), velU / velV / velW (:80-141), velDiv (:165-167)
    vec3d vel(int i, int j, int k) {
        fs3d_detail::State& s = *u.st;
        if (s.cell_v != s.version || s.dirty[0] || s.dirty[1] || s.dirty[2]) {
            s.push();
            s.cell.resize(3 * (size_t)sizeX * sizeY * sizeZ);
            fs3d_detail::check(s.g, fs3d_vel_cell(s.g, FS2D_F64, s.cell.data(), s.cell.size()), "fs3d_vel_cell");
            s.cell_v = s.version;
        }
        const double* o = &s.cell[3 * (((size_t)k * sizeY + j) * sizeX + i)];
        return vec3d{o[0], o[1], o[2]};
    }
    vec3d velU(int i, int j, int k) { return face(0, i, j, k); }
    vec3d velV(int i, int j, int k) { return face(1, i, j, k); }
    vec3d velW(int i, int j, int k) { return face(2, i, j, k); }
    double velDiv(int i, int j, int k) {
        fs3d_detail::State& s = *u.st;
        if (s.div_v != s.version || s.dirty[0] || s.dirty[1] || s.dirty[2]) {
            s.push();
            s.div.resize((size_t)sizeX * sizeY * sizeZ);
            fs3d_detail::check(s.g, fs3d_vel_div(s.g, FS2D_F64, s.div.data(), s.div.size()), "fs3d_vel_div");
            s.div_v = s.version;
        }
        return s.div[((size_t)k * sizeY + j) * sizeX + i];
    }

    // MACGrid3D::velInterpU/V/W and velInterp (MACGrid3D.h:143-163), grid units
    vec3d velInterp(vec3d pos) {
        vec3d out;
        velInterp(&pos, 1, &out);
        return out;
    }
    double velInterpU(vec3d pos) { return velInterp(pos).x; }
    double velInterpV(vec3d pos) { return velInterp(pos).y; }
    double velInterpW(vec3d pos) { return velInterp(pos).z; }
    // Many points in one launch.
    void velInterp(const vec3d* pos, size_t n, vec3d* out) {
        static_assert(sizeof(vec3d) == 3 * sizeof(double), "vec3d must be three packed doubles");
        u.st->push();
        fs3d_detail::check(u.st->g, fs3d_vel_interp(u.st->g, &pos->x, n, &out->x, sampleMode), "fs3d_vel_interp");
    }

   private:
    vec3d face(int f, int i, int j, int k) {
        fs3d_detail::State& s = *u.st;
        const int NX = sizeX + (f == 0), NY = sizeY + (f == 1), NZ = sizeZ + (f == 2);
        if (s.face_v[f] != s.version || s.dirty[0] || s.dirty[1] || s.dirty[2]) {
            s.push();
            s.face[f].resize(3 * (size_t)NX * NY * NZ);
            fs3d_detail::check(s.g, fs3d_vel_face(s.g, (fs3d_field)f, FS2D_F64, s.face[f].data(), s.face[f].size()),
                               "fs3d_vel_face");
            s.face_v[f] = s.version;
        }
        const double* o = &s.face[f][3 * (((size_t)k * NY + j) * NX + i)];
        return vec3d{o[0], o[1], o[2]};
    }
};

#endif  // FS3D_MACGRID3D_HPP
