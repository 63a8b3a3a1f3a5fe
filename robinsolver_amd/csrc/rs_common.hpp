// rs_common.hpp -- complex helpers and shared device-side types for the
// MI355X (gfx950) HPS factor/solve library.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace rs {

// Complex scalar stored as an interleaved (re, im) pair, matching Fortran
// complex(dp) / C99 double complex memory layout.
template <typename R> struct cpx;
template <> struct cpx<double> { using T = double2; };
template <> struct cpx<float> { using T = float2; };
template <typename R> using c_t = typename cpx<R>::T;

template <typename R> __host__ __device__ inline c_t<R> mk(R re, R im) { c_t<R> z; z.x = re; z.y = im; return z; }
template <typename R> __host__ __device__ inline c_t<R> cz() { return mk<R>(R(0), R(0)); }
template <typename V> __host__ __device__ inline V cadd(V a, V b) { V z; z.x = a.x + b.x; z.y = a.y + b.y; return z; }
template <typename V> __host__ __device__ inline V csub(V a, V b) { V z; z.x = a.x - b.x; z.y = a.y - b.y; return z; }
template <typename V> __host__ __device__ inline V cmul(V a, V b) {
    V z; z.x = a.x * b.x - a.y * b.y; z.y = a.x * b.y + a.y * b.x; return z;
}
// a + b*c
template <typename V> __host__ __device__ inline V cfma(V b, V c, V a) {
    V z; z.x = a.x + b.x * c.x - b.y * c.y; z.y = a.y + b.x * c.y + b.y * c.x; return z;
}
// a - b*c as four fused multiply-adds
template <typename V> __host__ __device__ inline V cfms4(V a, V b, V c) {
    V z; z.x = fma(b.y, c.y, fma(-b.x, c.x, a.x)); z.y = fma(-b.y, c.x, fma(-b.x, c.y, a.y)); return z;
}
template <typename V, typename S> __host__ __device__ inline V cscale(V a, S s) { V z; z.x = a.x * s; z.y = a.y * s; return z; }
template <typename V> __host__ __device__ inline V cinv(V a) {
    // 1/a with scaling to avoid overflow of |a|^2
    auto s = fabs(a.x) > fabs(a.y) ? fabs(a.x) : fabs(a.y);
    V b; b.x = a.x / s; b.y = a.y / s;
    auto den = (b.x * b.x + b.y * b.y) * s;
    V z; z.x = b.x / den; z.y = -b.y / den; return z;
}
template <typename V> __host__ __device__ inline auto cabs1(V a) { return fabs(a.x) + fabs(a.y); }

// Device pointers that the compiler cannot prove global (read from descriptors
// or column tables) would be accessed with flat instructions, which count
// against the LDS wait counter as well: every wait for an LDS result would also
// wait for outstanding prefetches.  Such pointers are cast to the global address
// space before use.
template <typename T> using gptr = T __attribute__((address_space(1)))*;
template <typename V> struct vec_of;
template <> struct vec_of<float2> { typedef float T __attribute__((ext_vector_type(2))); };
template <> struct vec_of<double2> { typedef double T __attribute__((ext_vector_type(2))); };
// complex load / store through the global address space
template <typename V> __device__ inline V gld(const V* p) {
    using T = typename vec_of<V>::T;
    const T t = *(gptr<const T>)p;
    V v; v.x = t.x; v.y = t.y; return v;
}
template <typename V> __device__ inline void gst(V* p, V v) {
    using T = typename vec_of<V>::T;
    T t; t.x = v.x; t.y = v.y;
    *(gptr<T>)p = t;
}
__device__ inline int gldi(const int* p) { return *(gptr<const int>)p; }

// Wave-wide reductions through DPP lane permutations (quad swaps, then the row
// half-mirror and mirror: every lane of a 16-lane row holds the row's result),
// the four rows combined through readlane.  Unlike a shuffle (ds_bpermute) no
// step goes through the LDS unit.
template <int CTRL> __device__ inline int dpp_i(int x) { return __builtin_amdgcn_update_dpp(x, x, CTRL, 0xf, 0xf, false); }
template <int CTRL> __device__ inline double dpp_d(double x) {
    const long long b = __builtin_bit_cast(long long, x);
    const int lo = dpp_i<CTRL>((int)(b & 0xffffffffLL)), hi = dpp_i<CTRL>((int)(b >> 32));
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}
__device__ inline double readlane_d(double x, int l) {
    const long long b = __builtin_bit_cast(long long, x);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}
// maximum over the wave (NaN values are passed over by fmax)
__device__ inline double wave_max(double v) {
    v = fmax(v, dpp_d<0xB1>(v));    // quad_perm [1,0,3,2]
    v = fmax(v, dpp_d<0x4E>(v));    // quad_perm [2,3,0,1]
    v = fmax(v, dpp_d<0x141>(v));   // row_half_mirror
    v = fmax(v, dpp_d<0x140>(v));   // row_mirror
    return fmax(fmax(readlane_d(v, 0), readlane_d(v, 16)), fmax(readlane_d(v, 32), readlane_d(v, 48)));
}
__device__ inline int wave_min(int v) {
    v = min(v, dpp_i<0xB1>(v));
    v = min(v, dpp_i<0x4E>(v));
    v = min(v, dpp_i<0x141>(v));
    v = min(v, dpp_i<0x140>(v));
    return min(min(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16)),
               min(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48)));
}
// argmax over the wave: the largest v, and the smallest index i among the lanes
// holding it; every lane gets both
__device__ inline void wave_argmax(double& v, int& i) {
    const double m = wave_max(v);
    i = wave_min(v == m ? i : 0x7fffffff);
    v = m;
}

template <typename Dst, typename Src> __host__ __device__ inline Dst ccast(Src a) {
    Dst z; z.x = (decltype(z.x))a.x; z.y = (decltype(z.y))a.y; return z;
}

// Geometry shared by every leaf of one tree (all leaves of a tree have the
// same grid; the host refuses trees where they do not).
struct LeafGeom {
    int d;            // 2 or 3
    int n[3];         // interior points per axis (ptbox)
    double h[3];      // grid spacing per axis
    int n_int;        // product of n
    int nb;           // boundary points
    int m;            // nb + n_int
    int s;            // block size: interior points per slow-axis layer
    int K;            // number of blocks = n[d-1]
    int pf[7];        // points per face, 1..6
    int foff[7];      // start of face f in the leaf's local boundary order (fgrid order)
    int noff[2][7];   // start of face f in the node order of an even / odd leaf
    double kb;        // opts%kb: alpha = +i kb (incoming), sigma = -i kb (outgoing)
};

}  // namespace rs
