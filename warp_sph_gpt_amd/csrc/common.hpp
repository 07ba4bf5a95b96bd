// common.hpp -- device types and helpers shared by the libsphdem kernels.
//
// All kernels are templates over the real type T (float: the reference's
// FP32 path; double: the FP64 shadow).  The library is compiled with
// -ffp-contract=off so every a*b+c below rounds twice, as the reference's
// expressions are written; divisions and square roots are the correctly
// rounded IEEE operations.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "../../include/sphdem.h"

namespace sphdem {

// ---------------------------------------------------------------- errors
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);
int check_launch(const char *what);

#define SPHDEM_REQUIRE(cond, msg)                                   \
    do {                                                            \
        if (!(cond)) return ::sphdem::fail(SPHDEM_E_ARG, (msg));    \
    } while (0)

// Device-wide exclusive scan (grid.hip); `scratch` holds scan_scratch(n) ints.
int scan_scratch(int n);
int exclusive_scan(const int *in, int *out, int n, int *scratch, hipStream_t st);
// Devices the library keeps per-device state for.
constexpr int kMaxDevices = 64;
// Device address of sphdem_exclusive_scan_noalloc's tile totals on the
// current device, resolved on the first call per device (grid.hip).
int noalloc_scan_scratch(int **out);

// Per-pass event timing (grid.hip); a no-op unless sphdem_profile_enable(1).
int prof_begin(int pass, hipStream_t st);
void prof_end(int slot, hipStream_t st);
struct ProfScope {
    int slot;
    hipStream_t st;
    ProfScope(int pass, hipStream_t s) : slot(prof_begin(pass, s)), st(s) {}
    ~ProfScope() { prof_end(slot, st); }
};

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }
// A stream for work that may run beside the caller's, with the events that
// fork it from and join it back to the caller's stream (all null: none).
struct Beside {
    hipStream_t s = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
};
inline int blocks_for(long n, int threads) { return (int)((n + threads - 1) / threads); }

// Buckets of a grid over all its instances.
inline int grid_cells(const sphdem_grid_t *g) {
    return g->dim_x * g->dim_y * g->dim_z * (g->n_inst > 1 ? g->n_inst : 1);
}

// Launch shapes set by sphdem_tuning_set (grid.hip).
const sphdem_tuning_t &tuning();
inline int tuned_block(int pass) { return tuning().block[pass]; }
inline int tuned_remap(int pass) { return tuning().xcd_remap[pass]; }

// Workgroup -> subject tile.  The dispatcher hands workgroup b to XCD b % 8;
// with `remap` set, the workgroups of one XCD take one contiguous run of
// tiles, so tiles of neighbouring (sorted, hence spatially close) subjects
// share an L2.  A bijection on [0, gridDim.x).
__device__ __forceinline__ int tile_id(int remap) {
    const int b = blockIdx.x;
    if (!remap) return b;
    const int G = gridDim.x, per = G >> 3, rem = G & 7, x = b & 7, i = b >> 3;
    return x < rem ? x * (per + 1) + i : rem * (per + 1) + (x - rem) * per + i;
}
__device__ __forceinline__ int subject_id(int remap) {
    return tile_id(remap) * blockDim.x + threadIdx.x;
}

// ---------------------------------------------------------------- vectors
// Overloads that keep float math in float on both compilation passes.
__device__ __forceinline__ float msqrt(float x) { return sqrtf(x); }
__device__ __forceinline__ double msqrt(double x) { return ::sqrt(x); }
// FP32 pow is evaluated in double and rounded once to float: the correctly
// rounded result (up to a double-rounding tie), the same value the CPU oracle
// forms, so the EOS and drag laws agree bit for bit across the two.
__device__ __forceinline__ float mpow(float a, float b) {
    return (float)::pow((double)a, (double)b);
}
__device__ __forceinline__ double mpow(double a, double b) { return ::pow(a, b); }
__device__ __forceinline__ float mabs(float x) { return fabsf(x); }
__device__ __forceinline__ double mabs(double x) { return ::fabs(x); }
// Fast-math forms: the hardware square root and reciprocal in FP32 (about
// one ulp), the IEEE operations in FP64.
template <typename T> __device__ __forceinline__ T fast_sqrt(T x);
template <> __device__ __forceinline__ float fast_sqrt<float>(float x) {
    return __builtin_amdgcn_sqrtf(x);
}
template <> __device__ __forceinline__ double fast_sqrt<double>(double x) { return ::sqrt(x); }
template <typename T> __device__ __forceinline__ T fast_rcp(T x);
template <> __device__ __forceinline__ float fast_rcp<float>(float x) {
    return __builtin_amdgcn_rcpf(x);
}
template <> __device__ __forceinline__ double fast_rcp<double>(double x) { return 1.0 / x; }

template <typename T> struct alignas(sizeof(T) * 4) V4 { T x, y, z, w; };
template <typename T> struct V3 { T x, y, z; };

template <typename T> __device__ __forceinline__ V3<T> mk3(T x, T y, T z) { return {x, y, z}; }
template <typename T> __device__ __forceinline__ V3<T> ld3(const T *a, int i) {
    return {a[3 * i], a[3 * i + 1], a[3 * i + 2]};
}
template <typename T> __device__ __forceinline__ void st3(T *a, int i, V3<T> v) {
    a[3 * i] = v.x; a[3 * i + 1] = v.y; a[3 * i + 2] = v.z;
}
template <typename T> __device__ __forceinline__ V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename T> __device__ __forceinline__ V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename T> __device__ __forceinline__ V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <typename T> __device__ __forceinline__ V3<T> operator*(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <typename T> __device__ __forceinline__ V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <typename T> __device__ __forceinline__ V3<T> operator/(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
// Warp's dot: (x*x' + y*y') + z*z'.
template <typename T> __device__ __forceinline__ T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename T> __device__ __forceinline__ T length(V3<T> a) { return msqrt(dot(a, a)); }
template <typename T> __device__ __forceinline__ V3<T> cross(V3<T> a, V3<T> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <typename T> __device__ __forceinline__ V3<T> xyz(V4<T> a) { return {a.x, a.y, a.z}; }
template <typename T> __device__ __forceinline__ T sgn(T x) { return x < (T)0 ? (T)-1 : (T)1; }
template <typename T> __device__ __forceinline__ T clampr(T x, T a, T b) {
    T r = x < a ? a : x;
    return r > b ? b : r;
}

// ---------------------------------------------------------------- physics
#define SPHDEM_PI 3.14159265358979323846

// kernel/Kernel_KNL.py:7-23, evaluated left to right as written there.
template <typename T> __device__ __forceinline__ T wendland_w(T r, T h) {
    T q = r / h;
    T w = (T)0;
    T u = (T)1.0 - (T)0.5 * q;
    if (q < (T)2.0)
        w = (T)21.0 * u * u * u * u * ((T)1.0 + (T)2.0 * q) / ((T)16.0 * (T)SPHDEM_PI * h * h * h);
    return w;
}
// kernel/Kernel_KNL.py:26-41
template <typename T> __device__ __forceinline__ T wendland_dw(T r, T h) {
    T q = r / h;
    T dw = (T)0;
    if (q < (T)2.0) {
        T u = (T)1.0 - (T)0.5 * q;
        dw = (T)21.0 * ((T)-5.0 * q * u * u * u) / ((T)16.0 * (T)SPHDEM_PI * h * h * h * h);
    }
    return dw;
}
// The same two functions with their divisions by the pass-wide constants h,
// 16 pi h^3 and 16 pi h^4 done through reciprocals held in double.  For FP32
// a and b, (float)((double)a * RN64(1/b)) is the correctly rounded a/b: an
// exact float quotient is never a midpoint between floats and lies at least
// 2^-50 (relative) from one, while the double product is within 2^-52 of a/b
// (normal quotients; these passes make no subnormal ones).  Three full-rate
// instructions replace a division sequence; the bits are the IEEE
// division's.  FP64 keeps the IEEE division.
template <typename T> struct WendlandK {
    T h, c3, c4;
    __device__ explicit WendlandK(T hh)
        : h(hh), c3((T)16.0 * (T)SPHDEM_PI * hh * hh * hh),
          c4((T)16.0 * (T)SPHDEM_PI * hh * hh * hh * hh) {}
    __device__ T div_h(T a) const { return a / h; }
    __device__ T div_c3(T a) const { return a / c3; }
    __device__ T div_c4(T a) const { return a / c4; }
};
template <> struct WendlandK<float> {
    double ih, ic3, ic4;
    __device__ explicit WendlandK(float h)
        : ih(1.0 / (double)h), ic3(1.0 / (double)(16.0f * (float)SPHDEM_PI * h * h * h)),
          ic4(1.0 / (double)(16.0f * (float)SPHDEM_PI * h * h * h * h)) {}
    __device__ float div_h(float a) const { return (float)((double)a * ih); }
    __device__ float div_c3(float a) const { return (float)((double)a * ic3); }
    __device__ float div_c4(float a) const { return (float)((double)a * ic4); }
};
template <typename T> __device__ __forceinline__ T wendland_w(T r, const WendlandK<T> &k) {
    T q = k.div_h(r);
    T w = (T)0;
    T u = (T)1.0 - (T)0.5 * q;
    if (q < (T)2.0) w = k.div_c3((T)21.0 * u * u * u * u * ((T)1.0 + (T)2.0 * q));
    return w;
}
template <typename T> __device__ __forceinline__ T wendland_dw(T r, const WendlandK<T> &k) {
    T q = k.div_h(r);
    T dw = (T)0;
    if (q < (T)2.0) {
        T u = (T)1.0 - (T)0.5 * q;
        dw = k.div_c4((T)21.0 * ((T)-5.0 * q * u * u * u));
    }
    return dw;
}
// kernel/KERNEL_pres.py:21-23
template <typename T> __device__ __forceinline__ T tait(T rho, T rho0, T c0, T gamma) {
    const T B = rho0 * c0 * c0 / gamma;
    T q = mpow(rho / rho0, gamma) - (T)1.0;
    return B * (q > (T)0 ? q : (T)0);
}

// ---------------------------------------------------------------- views
template <typename T> struct Prm {
    T h, support, rho0, c0, gamma, mu, g, dt, eps_min, eps_max;
    T radius_max, boundary_radius_max, h_max, h_uniform;
};
template <typename T> inline Prm<T> prm(const sphdem_params_t *p) {
    return {(T)p->h, (T)p->support, (T)p->rho0, (T)p->c0, (T)p->gamma, (T)p->mu, (T)p->g,
            (T)p->dt, (T)p->eps_min, (T)p->eps_max, (T)p->radius_max,
            (T)p->boundary_radius_max, (T)p->h_max, (T)p->h_uniform};
}

template <typename T> struct SphV {
    T *pos, *vel, *rho_raw, *rho, *pres, *acc, *m, *flt, *porosity, *pgf, *acc_dem;
    int n;
};
template <typename T> struct BndV { T *pos, *vel, *rho_raw, *rho, *pres, *acc, *m; int n; };
template <typename T> struct DemV {
    T *pos, *vel, *acc, *force, *omega, *torque, *radius, *rho, *m, *volume, *inertia, *K,
      *eta, *mu, *h, *flt_s, *porosity, *drag, *pressure_force;
    int *dem_off_old, *dem_id_old; T *dem_tang_old;
    int *dem_off_new, *dem_id_new; T *dem_tang_new;
    int *bnd_off_old, *bnd_id_old; T *bnd_tang_old;
    int *bnd_off_new, *bnd_id_new; T *bnd_tang_new;
    int cap_dem_new, cap_bnd_new, n;
};
template <typename T> struct WallV {
    T *pos, *vel, *omega, *radius, *rho, *m, *volume, *inertia, *K, *eta, *mu;
    int n;
};
// A grid as the kernels see it.  `off` is the first bucket of the instance
// being queried (instance * cells); set it with at_inst() before a query.
template <typename T> struct GridV {
    int dx, dy, dz, n;
    T inv;
    const int *ids, *lo, *hi;
    const V4<T> *s;
    const int *inst;  // instance of each point (input = sorted layout), or null
    int cells;        // buckets per instance
    int off;
};
template <typename T> __device__ __forceinline__ int inst_of(const GridV<T> &g, int i) {
    return g.inst ? g.inst[i] : 0;
}
template <typename T> __device__ __forceinline__ GridV<T> at_inst(GridV<T> g, int k) {
    g.off = k * g.cells;
    return g;
}
template <typename T> struct WorkV {
    T *fl_rho, *fl_pres, *fl_V;
    V4<T> *fl_vel, *fl_pgf;
    T *wl_rho, *wl_pres;
    V4<T> *wl_vel;
    int *scan_tmp, *flags;
};

template <typename T> inline SphV<T> view(const sphdem_sph_t *a) {
    return {(T *)a->pos, (T *)a->vel, (T *)a->rho_raw, (T *)a->rho, (T *)a->pres, (T *)a->acc,
            (T *)a->m, (T *)a->flt, (T *)a->porosity, (T *)a->pgf, (T *)a->acc_dem, a->n};
}
template <typename T> inline BndV<T> view(const sphdem_bnd_t *a) {
    return {(T *)a->pos, (T *)a->vel, (T *)a->rho_raw, (T *)a->rho, (T *)a->pres, (T *)a->acc,
            (T *)a->m, a->n};
}
template <typename T> inline DemV<T> view(const sphdem_dem_t *a) {
    return {(T *)a->pos, (T *)a->vel, (T *)a->acc, (T *)a->force, (T *)a->omega, (T *)a->torque,
            (T *)a->radius, (T *)a->rho, (T *)a->m, (T *)a->volume, (T *)a->inertia, (T *)a->K,
            (T *)a->eta, (T *)a->mu, (T *)a->h, (T *)a->flt_s, (T *)a->porosity, (T *)a->drag,
            (T *)a->pressure_force,
            a->contact_dem_offset_old, a->contact_dem_id_old, (T *)a->tang_dem_old,
            a->contact_dem_offset_new, a->contact_dem_id_new, (T *)a->tang_dem_new,
            a->contact_bnd_offset_old, a->contact_bnd_id_old, (T *)a->tang_bnd_old,
            a->contact_bnd_offset_new, a->contact_bnd_id_new, (T *)a->tang_bnd_new,
            a->cap_dem_new, a->cap_bnd_new, a->n};
}
template <typename T> inline WallV<T> view(const sphdem_dembnd_t *a) {
    return {(T *)a->pos, (T *)a->vel, (T *)a->omega, (T *)a->radius, (T *)a->rho, (T *)a->m,
            (T *)a->volume, (T *)a->inertia, (T *)a->K, (T *)a->eta, (T *)a->mu, a->n};
}
template <typename T> inline GridV<T> view(const sphdem_grid_t *g) {
    // Warp stores 1/cell_width rounded to the grid's float type.
    T inv = (T)1.0 / (T)g->cell_width;
    return {g->dim_x, g->dim_y, g->dim_z, g->n, inv, g->point_ids, g->cell_lo, g->cell_hi,
            (const V4<T> *)g->sorted, g->n_inst > 1 ? g->inst : nullptr,
            g->dim_x * g->dim_y * g->dim_z, 0};
}
template <typename T> inline WorkV<T> view(const sphdem_work_t *w) {
    return {(T *)w->fl_rho, (T *)w->fl_pres, (T *)w->fl_V, (V4<T> *)w->fl_vel,
            (V4<T> *)w->fl_pgf, (T *)w->wl_rho, (T *)w->wl_pres, (V4<T> *)w->wl_vel,
            w->scan_tmp, w->flags};
}

// ---------------------------------------------------------------- queries
constexpr int kOrigin = 1 << 20;

__device__ __forceinline__ int fold(int c, int dim) {
    c += kOrigin;
    return (c < 0 ? 0 : c) % dim;
}

// Visit the buckets of a HashGrid query in Warp's order -- z outer, y, x
// inner, cells [int((p-r)/w), int((p+r)/w)] capped at start+dim-1 -- as
// contiguous ranges [lo, hi) of sorted slots.  Consecutive x buckets of one
// (y, z) row are adjacent in sorted order, so a row is one range, or two
// when it wraps around the folded grid.  fn(lo, hi) sees the points in the
// same order as Warp's per-bucket iteration.
// The bucket box of a query: virtual cells [x0, x1] x [y0, y1] x [z0, z1],
// at most the grid's dims along each axis.
struct QueryBox {
    int x0, x1, y0, y1, z0, z1;
};
template <typename T>
__device__ __forceinline__ QueryBox query_box(const GridV<T> &g, T px, T py, T pz, T r) {
    const T inv = g.inv;
    QueryBox b;
    b.x0 = (int)((px - r) * inv);
    b.y0 = (int)((py - r) * inv);
    b.z0 = (int)((pz - r) * inv);
    b.x1 = min((int)((px + r) * inv), b.x0 + g.dx - 1);
    b.y1 = min((int)((py + r) * inv), b.y0 + g.dy - 1);
    b.z1 = min((int)((pz + r) * inv), b.z0 + g.dz - 1);
    return b;
}
// The slot ranges of one (y, z) row of a query box, in Warp's order.
template <typename T, typename Fn>
__device__ __forceinline__ void row_runs(const GridV<T> &g, const QueryBox &b, int y, int z,
                                         Fn &&fn) {
    const int *lo = g.lo + g.off, *hi = g.hi + g.off;
    const int len = b.x1 - b.x0 + 1;
    const int base = fold(z, g.dz) * (g.dx * g.dy) + fold(y, g.dy) * g.dx;
    if (b.x0 + kOrigin >= 0) {
        const int cx = (b.x0 + kOrigin) % g.dx;
        const int first = min(len, g.dx - cx);
        if (first > 0) fn(lo[base + cx], hi[base + cx + first - 1]);
        if (len > first) fn(lo[base], hi[base + len - first - 1]);
    } else {
        for (int x = b.x0; x <= b.x1; ++x) {
            const int c = base + fold(x, g.dx);
            fn(lo[c], hi[c]);
        }
    }
}
template <typename T, typename Fn>
__device__ __forceinline__ void query_runs(const GridV<T> &g, T px, T py, T pz, T r, Fn &&fn) {
    const QueryBox b = query_box(g, px, py, pz, r);
    if (g.n == 0) return;
    for (int z = b.z0; z <= b.z1; ++z)
        for (int y = b.y0; y <= b.y1; ++y) row_runs(g, b, y, z, fn);
}

// A lower bound of |x - p| over the points x of virtual cell c along one
// axis and the points p of [lo_p, hi_p].  Cell c holds int(x * inv) == c
// (truncation toward zero, so cell 0 spans (-w, w)), w = 1 / inv.  The
// bound is lowered by 1e-5 of the magnitudes involved, far more than the
// rounding of x * inv, of w and of this arithmetic.
template <typename T>
__device__ __forceinline__ T axis_gap(T lo_p, T hi_p, int c, T w) {
    const T lo = (T)(c <= 0 ? c - 1 : c) * w;
    const T hi = (T)(c >= 0 ? c + 1 : c) * w;
    T d = lo - hi_p;
    const T e = lo_p - hi;
    d = d > e ? d : e;
    d = d - (T)1.0e-5 * (mabs(lo) + mabs(hi) + mabs(lo_p) + mabs(hi_p));
    return d > (T)0 ? d : (T)0;
}

// Visit slots a, a+step, ... < b in order, loading the records of U slots
// before evaluating any of them, so that their loads are in flight together
// (the neighbour passes are bound by load latency); the evaluation order,
// hence every sum, is the plain loop's.
template <int U, typename Load, typename Eval>
__device__ __forceinline__ void each_slot(int a, int b, int step, Load &&load, Eval &&eval) {
    int s = a;
    for (; s + (U - 1) * step < b; s += U * step) {
        decltype(load(0)) c[U];
#pragma unroll
        for (int u = 0; u < U; ++u) c[u] = load(s + u * step);
#pragma unroll
        for (int u = 0; u < U; ++u) eval(s + u * step, c[u]);
    }
    for (; s < b; s += step) eval(s, load(s));
}

// ---------------------------------------------------------------- wave64 sums
template <typename T> __device__ __forceinline__ T wave_sum(T v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

}  // namespace sphdem
