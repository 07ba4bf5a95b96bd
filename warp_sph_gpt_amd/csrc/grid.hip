// grid.hip -- device-wide exclusive scan and the counting-sort hash grid.
//
// Replaces wp.HashGrid (build: main.py:66,73,85,87; queries in every
// neighbour kernel) and wp.utils.array_scan (source/Simulation.py:148,181).
//
// Build of n points into dim_x*dim_y*dim_z buckets:
//   1. key:     point_cells[i] = bucket(int(p*inv)); rank[i] = atomicAdd(count[key])
//   2. scan:    cell_lo = exclusive_scan(count), cell_hi = cell_lo + count
//               (the scan's last pass)
//   3. scatter: point_ids[cell_lo[key] + rank[i]] = i
//   4. order:   each bucket's ids sorted ascending, which makes the order
//               equal to a stable sort by key and independent of the order
//               the atomics were served in; rank[i] becomes i's sorted slot
//               (the inverse of point_ids)
//   5. sorted:  sorted[k] = (p[id], w[id]) for coalesced neighbour reads
#include "common.hpp"

#include <cstdio>
#include <mutex>
#include <vector>

namespace sphdem {

static thread_local std::string g_err;

void set_error(const std::string &msg) { g_err = msg; }

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

int check_launch(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(SPHDEM_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
    return SPHDEM_OK;
}

// ---------------------------------------------------------------- launch shapes
namespace {
sphdem_tuning_t default_tuning() {
    sphdem_tuning_t t;
    for (int i = 0; i < SPHDEM_NUM_PASSES; ++i) {
        t.block[i] = 256;
        t.xcd_remap[i] = 0;
    }
    t.por_x_tile = 4;
    t.dem_beside = 1;
    t.por_parts = 4;
    t.por_queue = 1;
    t.wall_beside = 1;
    t.force_beside = 1;
    return t;
}
sphdem_tuning_t g_tuning = default_tuning();
}  // namespace

const sphdem_tuning_t &tuning() { return g_tuning; }

// ---------------------------------------------------------------- timing
namespace {
struct Prof {
    bool on = false;
    std::vector<hipEvent_t> ev;  // start/stop pairs
    std::vector<int> pass;
    size_t used = 0;
};
Prof g_prof;
}  // namespace

int prof_begin(int pass, hipStream_t st) {
    if (!g_prof.on) return -1;
    const size_t k = g_prof.used;
    if (g_prof.pass.size() <= k) {
        hipEvent_t a, b;
        if (hipEventCreate(&a) != hipSuccess) return -1;
        if (hipEventCreate(&b) != hipSuccess) {
            (void)hipEventDestroy(a);
            return -1;
        }
        g_prof.ev.push_back(a);
        g_prof.ev.push_back(b);
        g_prof.pass.push_back(pass);
    } else {
        g_prof.pass[k] = pass;
    }
    if (hipEventRecord(g_prof.ev[2 * k], st) != hipSuccess) return -1;
    g_prof.used = k + 1;
    return (int)k;
}

void prof_end(int slot, hipStream_t st) {
    if (slot >= 0) (void)hipEventRecord(g_prof.ev[2 * slot + 1], st);
}

// ---------------------------------------------------------------- scan
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

// Exclusive scan of one tile (THREADS x kScanItems values) of `in` into
// `out`, offset by `carry`; returns the tile total to every thread.  With
// `vec` (both arrays 16-byte aligned) a thread's 8 values move as two int4
// when all of them are in range.
template <int THREADS>
__device__ int scan_tile(const int *in, int *out, int n, int tile_base, int carry, bool vec) {
    __shared__ int wsum[THREADS / 64];
    __shared__ int total;
    const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
    const int base = tile_base + t * kScanItems;
    const bool whole = vec && base + kScanItems <= n;
    int x[kScanItems];
    if (whole) {
        const int4 a = *reinterpret_cast<const int4 *>(in + base);
        const int4 b = *reinterpret_cast<const int4 *>(in + base + 4);
        x[0] = a.x; x[1] = a.y; x[2] = a.z; x[3] = a.w;
        x[4] = b.x; x[5] = b.y; x[6] = b.z; x[7] = b.w;
    } else {
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) x[k] = base + k < n ? in[base + k] : 0;
    }
    int v[kScanItems];
    int tsum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = tsum;
        tsum += x[k];
    }
    int incl = tsum;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(incl, off, 64);
        if (lane >= off) incl += y;
    }
    if (lane == 63) wsum[wid] = incl;
    __syncthreads();
    int wofs = 0;
    for (int w = 0; w < wid; ++w) wofs += wsum[w];
    const int excl = carry + wofs + incl - tsum;
    if (whole) {
        *reinterpret_cast<int4 *>(out + base) =
            int4{v[0] + excl, v[1] + excl, v[2] + excl, v[3] + excl};
        *reinterpret_cast<int4 *>(out + base + 4) =
            int4{v[4] + excl, v[5] + excl, v[6] + excl, v[7] + excl};
    } else {
#pragma unroll
        for (int k = 0; k < kScanItems; ++k)
            if (base + k < n) out[base + k] = v[k] + excl;
    }
    if (t == THREADS - 1) total = wofs + incl;
    __syncthreads();
    const int r = total;
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_tiles(const int *in, int *out, int n,
                                                             int *partial, bool vec) {
    const int tot = scan_tile<kScanThreads>(in, out, n, blockIdx.x * kScanTile, 0, vec);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// One block scans the tile totals in place (8,192 per pass, any count).
constexpr int kPartialThreads = 1024;
__global__ __launch_bounds__(kPartialThreads) void k_scan_partials(int *partial, int tiles,
                                                                   bool vec) {
    int carry = 0;
    for (int base = 0; base < tiles; base += kPartialThreads * kScanItems)
        carry += scan_tile<kPartialThreads>(partial, partial, tiles, base, carry, vec);
}

// out += the tile offsets, four values per thread; with `hi` (the counts on
// entry) also hi = out + count, the buckets' ends.
__global__ void k_scan_add(int *out, int n, const int *partial, bool add, int *hi, bool vec) {
    const int i0 = 4 * (blockIdx.x * blockDim.x + threadIdx.x);
    if (i0 >= n) return;
    const int p = add ? partial[i0 / kScanTile] : 0;
    if (vec && i0 + 4 <= n) {
        int4 o = *reinterpret_cast<const int4 *>(out + i0);
        o.x += p; o.y += p; o.z += p; o.w += p;
        *reinterpret_cast<int4 *>(out + i0) = o;
        if (hi) {
            int4 c = *reinterpret_cast<const int4 *>(hi + i0);
            c.x += o.x; c.y += o.y; c.z += o.z; c.w += o.w;
            *reinterpret_cast<int4 *>(hi + i0) = c;
        }
    } else {
        for (int i = i0; i < n && i < i0 + 4; ++i) {
            const int o = out[i] + p;
            out[i] = o;
            if (hi) hi[i] += o;
        }
    }
}

int scan_scratch(int n) { return blocks_for(n, kScanTile) + 1; }

// ---------------------------------------------------------------- scan without allocation
// The tile totals of sphdem_exclusive_scan_noalloc: one array per GPU, which
// the GPU holds until the process ends and every call on it shares.  In host
// code `noalloc_scratch` only names the symbol; its device address on each
// GPU comes from hipGetSymbolAddress, resolved once per device.
constexpr int kNoallocScanInts = (SPHDEM_SCAN_NOALLOC_MAX + kScanTile - 1) / kScanTile + 1;
static __device__ __align__(256) int noalloc_scratch[kNoallocScanInts];

namespace {
std::mutex g_noalloc_mu;
int *g_noalloc_ptr[kMaxDevices] = {};
}  // namespace

int noalloc_scan_scratch(int **out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
        return fail(SPHDEM_E_HIP, "no current HIP device for the scan scratch");
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(noalloc_scratch));
        if (e != hipSuccess || !p)
            return fail(SPHDEM_E_HIP, std::string("hipGetSymbolAddress(noalloc_scratch): ") +
                                          hipGetErrorString(e));
        g_noalloc_ptr[dev] = static_cast<int *>(p);
    }
    if (out) *out = g_noalloc_ptr[dev];
    return SPHDEM_OK;
}

static bool aligned16(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

// The scan; with `hi` (holding the counts that `in` also points to) the
// last pass writes the bucket ends hi = out + count as well.
static int scan_impl(const int *in, int *out, int n, int *scratch, hipStream_t st, int *hi) {
    if (n <= 0) return SPHDEM_OK;
    const int tiles = blocks_for(n, kScanTile);
    const bool vec = aligned16(in) && aligned16(out) && (!hi || aligned16(hi));
    hipLaunchKernelGGL(k_scan_tiles, dim3(tiles), dim3(kScanThreads), 0, st, in, out, n, scratch,
                       vec);
    if (tiles > 1)
        hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(kPartialThreads), 0, st, scratch, tiles,
                           aligned16(scratch));
    if (tiles > 1 || hi)
        hipLaunchKernelGGL(k_scan_add, dim3(blocks_for(n, 4 * 256)), dim3(256), 0, st, out, n,
                           (const int *)scratch, tiles > 1, hi, vec);
    return check_launch("exclusive_scan");
}

int exclusive_scan(const int *in, int *out, int n, int *scratch, hipStream_t st) {
    return scan_impl(in, out, n, scratch, st, nullptr);
}

// ---------------------------------------------------------------- grid build
__global__ void k_fill_int(int *a, int n, int v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

template <typename T>
__global__ void k_grid_key(const T *pos, int n, T inv, int dx, int dy, int dz, const int *inst,
                           int *cells, int *rank, int *count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = i < n;
    int key = -1;
    if (live) {
        const int x = (int)(pos[3 * i] * inv), y = (int)(pos[3 * i + 1] * inv),
                  z = (int)(pos[3 * i + 2] * inv);
        key = fold(z, dz) * (dx * dy) + fold(y, dy) * dx + fold(x, dx);
        if (inst) key += inst[i] * (dx * dy * dz);
        cells[i] = key;
    }
    // consecutive points of one bucket (runs of lanes with the same key)
    // take their ranks with one atomic per run: the first lane of the run
    // reserves the run's length and each lane takes base + its place in the
    // run.  Ranks only need to be distinct within a bucket (k_grid_order
    // puts every bucket in ascending point order afterwards).
    const int lane = threadIdx.x & 63;
    const int prev = __shfl_up(key, 1, 64);
    const unsigned long long starts = __ballot(live && (lane == 0 || prev != key));
    const unsigned long long upto = starts & (~0ull >> (63 - lane));
    const int head = 63 - __builtin_clzll(upto | 1ull);
    const unsigned long long later = starts & ~(~0ull >> (63 - lane));
    const unsigned long long lives = __ballot(live);
    const int end = later ? __builtin_ctzll(later) : 64 - __builtin_clzll(lives);
    int base = 0;
    if (live && lane == head) base = atomicAdd(&count[key], end - head);
    base = __shfl(base, head, 64);
    if (live) rank[i] = base + (lane - head);
}

__global__ void k_grid_scatter(const int *cells, const int *rank, const int *lo, int n, int *ids) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) ids[lo[cells[i]] + rank[i]] = i;
}

// Warp's order inside a bucket is ascending point index.  The scatter left
// each bucket's ids in atomic-arrival order in `tmp`; point i takes the slot
// after the ids of its bucket smaller than i, and records it in slot[i] (the
// arrival ranks there are no longer needed).
__global__ void k_grid_order(const int *cells, const int *lo, const int *hi, const int *tmp,
                             int n, int *ids, int *slot) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = cells[i], a = lo[c], e = hi[c];
    int pos = a, k = a;
    // eight independent loads in flight, not one load and wait per slot
    for (; k + 8 <= e; k += 8) {
        int t[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) t[u] = tmp[k + u];
#pragma unroll
        for (int u = 0; u < 8; ++u) pos += t[u] < i;
    }
    for (; k < e; ++k) pos += tmp[k] < i;
    ids[pos] = i;
    slot[i] = pos;
}

template <typename T>
__global__ void k_grid_sorted(const T *pos, const T *w, const int *ids, int n, V4<T> *s) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int i = ids[k];
    s[k] = V4<T>{pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], w ? w[i] : (T)0};
}

template <typename T>
static int grid_build(sphdem_grid_t *g, const void *points, int n, double radius, const void *w,
                      hipStream_t st) {
    ProfScope prof(SPHDEM_PASS_GRID, st);
    const int ncell = grid_cells(g);
    const int bt = tuned_block(SPHDEM_PASS_GRID);
    const int *inst = g->n_inst > 1 ? g->inst : nullptr;
    const T inv = (T)1.0 / (T)radius;
    g->cell_width = radius;
    g->n = n;
    hipLaunchKernelGGL(k_fill_int, dim3(blocks_for(ncell, bt)), dim3(bt), 0, st, g->cell_hi, ncell, 0);
    if (n > 0)
        hipLaunchKernelGGL(k_grid_key<T>, dim3(blocks_for(n, bt)), dim3(bt), 0, st,
                           (const T *)points, n, inv, g->dim_x, g->dim_y, g->dim_z, inst,
                           g->point_cells, g->rank, g->cell_hi);
    // cell_lo = exclusive scan of the counts in cell_hi, which become the
    // bucket ends cell_lo + count in the same pass
    int rc = scan_impl(g->cell_hi, g->cell_lo, ncell, g->scan_tmp, st, g->cell_hi);
    if (rc) return rc;
    if (n > 0) {
        // the sorted copy's storage (4n values, written last) holds the
        // buckets in arrival order meanwhile
        int *tmp = reinterpret_cast<int *>(g->sorted);
        hipLaunchKernelGGL(k_grid_scatter, dim3(blocks_for(n, bt)), dim3(bt), 0, st,
                           g->point_cells, g->rank, g->cell_lo, n, tmp);
        hipLaunchKernelGGL(k_grid_order, dim3(blocks_for(n, bt)), dim3(bt), 0, st,
                           g->point_cells, g->cell_lo, g->cell_hi, tmp, n, g->point_ids,
                           g->rank);
    }
    if (n > 0)
        hipLaunchKernelGGL(k_grid_sorted<T>, dim3(blocks_for(n, bt)), dim3(bt), 0, st,
                           (const T *)points, (const T *)w, g->point_ids, n, (V4<T> *)g->sorted);
    return check_launch("grid_build");
}

template <typename T>
__global__ void k_grid_set_w(const T *w, const int *ids, int n, V4<T> *s) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) s[k].w = w[ids[k]];
}

__global__ void k_warp_bounds(const int *lo, const int *hi, int ncell, int *starts, int *ends) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const bool any = hi[c] > lo[c];
    starts[c] = any ? lo[c] : 0;
    ends[c] = any ? hi[c] : 0;
}

template <typename T>
__global__ void k_neighbour_counts(GridV<T> gs, GridV<T> gn, T r, int *counts) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= gs.n) return;
    const V4<T> me = gs.s[k];
    const T rr = r * r;
    int c = 0;
    query_runs(at_inst(gn, inst_of(gs, k)), me.x, me.y, me.z, r, [&](int a, int b) {
        for (int s = a; s < b; ++s) {
            const V4<T> q = gn.s[s];
            const T dx = me.x - q.x, dy = me.y - q.y, dz = me.z - q.z;
            if (dx * dx + dy * dy + dz * dz < rr) ++c;
        }
    });
    counts[gs.ids[k]] = c;
}

static int check_grid(const sphdem_grid_t *g) {
    SPHDEM_REQUIRE(g != nullptr, "grid is NULL");
    SPHDEM_REQUIRE(g->precision == 4 || g->precision == 8, "grid precision must be 4 or 8");
    SPHDEM_REQUIRE(g->dim_x > 0 && g->dim_y > 0 && g->dim_z > 0, "grid dimensions must be positive");
    SPHDEM_REQUIRE(g->n_inst >= 0, "n_inst must be nonnegative");
    SPHDEM_REQUIRE((long)g->dim_x * g->dim_y * g->dim_z * (g->n_inst > 1 ? g->n_inst : 1) <
                       (1L << 30),
                   "grid has too many buckets");
    SPHDEM_REQUIRE(g->n_inst <= 1 || g->n == 0 || g->inst, "batched grid needs its inst array");
    SPHDEM_REQUIRE(g->cell_lo && g->cell_hi && g->scan_tmp, "grid bucket buffers are NULL");
    SPHDEM_REQUIRE(g->n == 0 || (g->point_ids && g->sorted), "grid point buffers are NULL");
    return SPHDEM_OK;
}

}  // namespace sphdem

using namespace sphdem;

extern "C" {

int sphdem_version(void) { return 1; }

int sphdem_profile_enable(int on) {
    g_prof.on = on != 0;
    if (g_prof.on) g_prof.used = 0;
    return SPHDEM_OK;
}

int sphdem_profile_read(double *ms, int *counts, int n) {
    SPHDEM_REQUIRE(n >= 0 && (n == 0 || (ms && counts)), "invalid profile buffers");
    for (int i = 0; i < n; ++i) {
        ms[i] = 0.0;
        counts[i] = 0;
    }
    for (size_t k = 0; k < g_prof.used; ++k) {
        if (hipEventSynchronize(g_prof.ev[2 * k + 1]) != hipSuccess)
            return fail(SPHDEM_E_HIP, "profile event wait failed");
        float t = 0.f;
        if (hipEventElapsedTime(&t, g_prof.ev[2 * k], g_prof.ev[2 * k + 1]) != hipSuccess)
            return fail(SPHDEM_E_HIP, "profile elapsed time failed");
        const int id = g_prof.pass[k];
        if (id >= 0 && id < n) {
            ms[id] += t;
            counts[id] += 1;
        }
    }
    return SPHDEM_OK;
}

const char *sphdem_last_error(void) { return g_err.c_str(); }

int sphdem_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int sphdem_scan_scratch(int n) { return scan_scratch(n); }

int sphdem_tuning_get(sphdem_tuning_t *t) {
    SPHDEM_REQUIRE(t, "tuning is NULL");
    *t = g_tuning;
    return SPHDEM_OK;
}

int sphdem_tuning_set(const sphdem_tuning_t *t) {
    SPHDEM_REQUIRE(t, "tuning is NULL");
    for (int i = 0; i < SPHDEM_NUM_PASSES; ++i) {
        SPHDEM_REQUIRE(t->block[i] == 64 || t->block[i] == 128 || t->block[i] == 256,
                       "block sizes must be 64, 128 or 256");
        SPHDEM_REQUIRE(t->xcd_remap[i] == 0 || t->xcd_remap[i] == 1, "xcd_remap must be 0 or 1");
    }
    SPHDEM_REQUIRE(t->por_x_tile >= 2 && t->por_x_tile <= 6, "por_x_tile must be in 2..6");
    SPHDEM_REQUIRE(t->dem_beside == 0 || t->dem_beside == 1, "dem_beside must be 0 or 1");
    SPHDEM_REQUIRE(t->por_parts >= 1 && t->por_parts <= 16, "por_parts must be in 1..16");
    SPHDEM_REQUIRE(t->por_queue == 0 || t->por_queue == 1, "por_queue must be 0 or 1");
    SPHDEM_REQUIRE(t->wall_beside == 0 || t->wall_beside == 1, "wall_beside must be 0 or 1");
    SPHDEM_REQUIRE(t->force_beside == 0 || t->force_beside == 1, "force_beside must be 0 or 1");
    g_tuning = *t;
    return SPHDEM_OK;
}

int sphdem_exclusive_scan(int *a, int n, int *scratch, void *stream) {
    SPHDEM_REQUIRE(n >= 0, "scan length must be nonnegative");
    SPHDEM_REQUIRE(n == 0 || (a && scratch), "scan buffers are NULL");
    return exclusive_scan(a, a, n, scratch, as_stream(stream));
}

int sphdem_exclusive_scan_noalloc(int *a, int n, void *stream) {
    SPHDEM_REQUIRE(n >= 0, "scan length must be nonnegative");
    SPHDEM_REQUIRE(n <= SPHDEM_SCAN_NOALLOC_MAX,
                   "scan length exceeds SPHDEM_SCAN_NOALLOC_MAX (" +
                       std::to_string(SPHDEM_SCAN_NOALLOC_MAX) + "): use sphdem_exclusive_scan");
    SPHDEM_REQUIRE(n == 0 || a, "scan buffer is NULL");
    if (n == 0) return SPHDEM_OK;
    int *tiles = nullptr;
    const int rc = noalloc_scan_scratch(&tiles);
    if (rc) return rc;
    return exclusive_scan(a, a, n, tiles, as_stream(stream));
}

int sphdem_grid_build(sphdem_grid_t *g, const void *points, int n, double radius, const void *w,
                      void *stream) {
    SPHDEM_REQUIRE(g != nullptr, "grid is NULL");
    SPHDEM_REQUIRE(n >= 0, "point count must be nonnegative");
    SPHDEM_REQUIRE(radius > 0.0, "grid radius must be positive");
    SPHDEM_REQUIRE(n == 0 || (points && g->point_cells && g->point_ids && g->rank && g->sorted),
                   "grid build buffers are NULL");
    g->n = n;
    int rc = check_grid(g);
    if (rc) return rc;
    if (g->precision == 4) return grid_build<float>(g, points, n, radius, w, as_stream(stream));
    return grid_build<double>(g, points, n, radius, w, as_stream(stream));
}

int sphdem_grid_set_w(sphdem_grid_t *g, const void *w, void *stream) {
    int rc = check_grid(g);
    if (rc) return rc;
    SPHDEM_REQUIRE(g->n == 0 || w, "w is NULL");
    if (g->n == 0) return SPHDEM_OK;
    if (g->precision == 4)
        hipLaunchKernelGGL(k_grid_set_w<float>, dim3(blocks_for(g->n, 256)), dim3(256), 0,
                           as_stream(stream), (const float *)w, g->point_ids, g->n,
                           (V4<float> *)g->sorted);
    else
        hipLaunchKernelGGL(k_grid_set_w<double>, dim3(blocks_for(g->n, 256)), dim3(256), 0,
                           as_stream(stream), (const double *)w, g->point_ids, g->n,
                           (V4<double> *)g->sorted);
    return check_launch("grid_set_w");
}

int sphdem_grid_warp_bounds(const sphdem_grid_t *g, int *cell_starts, int *cell_ends, void *stream) {
    int rc = check_grid(g);
    if (rc) return rc;
    SPHDEM_REQUIRE(cell_starts && cell_ends, "bounds buffers are NULL");
    const int ncell = grid_cells(g);
    hipLaunchKernelGGL(k_warp_bounds, dim3(blocks_for(ncell, 256)), dim3(256), 0, as_stream(stream),
                       g->cell_lo, g->cell_hi, ncell, cell_starts, cell_ends);
    return check_launch("grid_warp_bounds");
}

int sphdem_neighbour_counts(const sphdem_grid_t *gs, const sphdem_grid_t *gn, double radius,
                            int *counts, void *stream) {
    int rc = check_grid(gs);
    if (rc) return rc;
    rc = check_grid(gn);
    if (rc) return rc;
    SPHDEM_REQUIRE(gs->precision == gn->precision, "grids differ in precision");
    SPHDEM_REQUIRE(gs->n == 0 || counts, "counts buffer is NULL");
    if (gs->n == 0) return SPHDEM_OK;
    if (gs->precision == 4)
        hipLaunchKernelGGL(k_neighbour_counts<float>, dim3(blocks_for(gs->n, 256)), dim3(256), 0,
                           as_stream(stream), view<float>(gs), view<float>(gn), (float)radius, counts);
    else
        hipLaunchKernelGGL(k_neighbour_counts<double>, dim3(blocks_for(gs->n, 256)), dim3(256), 0,
                           as_stream(stream), view<double>(gs), view<double>(gn), radius, counts);
    return check_launch("neighbour_counts");
}

}  // extern "C"
