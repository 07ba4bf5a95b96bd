// eisph.hip -- the Eulerian ISPH lid-driven cavity (source/EISPH.py,
// kernel/KERNEL_EISPH_*.py) for gfx950.
//
// The points never move, so both HashGrids are walked once: the KGC pass
// (Kernel_prepare_kgc) and a two-pass build of a fixed neighbour list whose
// entries carry the neighbour id and the per-pair geometry every later pass
// needs (corrected gradient, Laplacian weight, neighbour volume), in the
// grid's visiting order.  A step then streams the list: each pass adds the
// same terms in the same order as the reference's grid walk, so the sums
// are the reference's sums.  The list is sliced by wavefront (64 rows, entry
// k of the 64 rows in consecutive slots), so every list load of a
// wavefront is one contiguous 256-byte (id) or 1-KiB (geo) access.
//
// Step (sphdem_eisph_step), five launches, one thread per point, ghosts
// after the fluid points:
//   advec    u* = u + dt (-(u.grad)u + nu lap u)
//   ppe      div u*, sum A_ij, sum A_ij p_j^t, b = rho0 div/dt, per-workgroup
//            FP64 sums of b; ghosts publish vel_star
//   solve    the workgroup sums added in a fixed shape (inline for small
//            clouds, else by a one-workgroup launch before it: the same
//            bits), b -= mean(b); p = (b + Aijpj)/Aij - gauge
//   correct  u = u* - dt grad p / rho; ghosts publish pres
//   finish   ghosts publish vel; fluid diagnostics on request
// A ghost's velocity 2 vel_bc - u[mirror] and pressure p[mirror] are formed
// where they are read: the same values the reference's Dirichlet and
// Neumann launches store before each read.
#include "common.hpp"

namespace sphdem {
namespace {

constexpr int kEisThreads = 256;

template <typename T> struct EisV {
    T *pos, *vel, *vel_star, *vel_bc, *acc, *kgc, *rho, *m, *pres, *Aij, *bi, *Aijpj, *div, *res;
    const int *mirror;
    int n;
};
template <typename T> EisV<T> eview(const sphdem_eisph_t *a) {
    return {(T *)a->pos, (T *)a->vel, (T *)a->vel_star, (T *)a->vel_bc, (T *)a->acc, (T *)a->kgc,
            (T *)a->rho, (T *)a->m, (T *)a->pres, (T *)a->Aij, (T *)a->bi, (T *)a->Aijpj,
            (T *)a->divergence, (T *)a->residual, a->mirror, a->n};
}
template <typename T> struct NbrV {
    const int *len, *off, *id;
    const V4<T> *geo;
};
template <typename T> NbrV<T> nview(const sphdem_eisph_nbr_t *L) {
    return {L->len, L->offset, L->id, (const V4<T> *)L->geo};
}
constexpr int kSlice = 64;
// Slot of entry k of row i.
__device__ __forceinline__ int slot0(const int *off, int i) {
    return off[i / kSlice] + (i % kSlice);
}

// The entries of a row from its first slot and length, the list loads of
// kU entries issued before their terms are added in order.
template <typename T, typename Fn>
__device__ __forceinline__ void row_batched(const NbrV<T> &L, int e, int len, Fn &&fn) {
    constexpr int kU = 4;
    for (int k = 0; k < len; k += kU, e += kU * kSlice) {
        int id[kU];
        V4<T> g[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            id[u] = 0;
            g[u] = V4<T>{(T)0, (T)0, (T)0, (T)0};
            if (k + u < len) {
                id[u] = L.id[e + u * kSlice];
                g[u] = L.geo[e + u * kSlice];
            }
        }
#pragma unroll
        for (int u = 0; u < kU; ++u)
            if (k + u < len) fn(id[u], g[u]);
    }
}

// Visit row i's entries in order: fn(id, geo).
template <typename T, typename Fn>
__device__ __forceinline__ void each_entry(const NbrV<T> &L, int i, Fn &&fn) {
    row_batched(L, slot0(L.off, i), L.len[i], fn);
}

// ---------------------------------------------------------------- kernels of the pair
// KERNEL_EISPH_KNL.py:34-49 -- 2-D Wendland C2 dW/dr, left to right.
template <typename T> __device__ __forceinline__ T wendland2_dw(T r, T h) {
    T q = r / h;
    T v = (T)0;
    if (q < (T)2.0) {
        T u = (T)1.0 - (T)0.5 * q;
        v = (T)-35.0 * q * u * u * u / ((T)4.0 * (T)SPHDEM_PI * h * h * h);
    }
    return v;
}
// KERNEL_EISPH_KNL.py:52-56
template <typename T> __device__ __forceinline__ V3<T> grad_w(V3<T> rij, T r2, T h) {
    T r = msqrt(r2);
    return (wendland2_dw(r, h) * rij) / r;
}
// KERNEL_EISPH_KNL.py:59-64
template <typename T> __device__ __forceinline__ V3<T> corrected(const T c[4], V3<T> g) {
    return mk3(c[0] * g.x + c[1] * g.z, (T)0.0, c[2] * g.x + c[3] * g.z);
}
template <typename T> __device__ __forceinline__ bool in_support(T r2, T support) {
    return r2 > (T)1.0e-12 && r2 < support * support;  // KERNEL_EISPH_KNL.py:10 R2_EPS
}

// Walk the fluid then the ghost grid around point i in Warp's order; fn(side,
// j, rij, r2) for every pair inside the support.
template <typename T, typename Fn>
__device__ __forceinline__ void walk_pairs(const GridV<T> &gs, const GridV<T> &gb, V3<T> ri,
                                           T support, Fn &&fn) {
    for (int side = 0; side < 2; ++side) {
        const GridV<T> &g = side == 0 ? gs : gb;
        query_runs(g, ri.x, ri.y, ri.z, support, [&](int a, int b) {
            for (int s = a; s < b; ++s) {
                const V4<T> q = g.s[s];
                const V3<T> rij = ri - xyz(q);
                const T r2 = dot(rij, rij);
                if (in_support(r2, support)) fn(side, g.ids[s], rij, r2);
            }
        });
    }
}

// KERNEL_EISPH_KNL.py:74-132 Kernel_prepare_kgc
template <typename T>
__global__ void k_eis_kgc(EisV<T> P, EisV<T> B, GridV<T> gs, GridV<T> gb, T support, T h) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= P.n) return;
    const int i = gs.ids[tid];
    const V3<T> ri = ld3(P.pos, i);
    T m00 = 0, m01 = 0, m10 = 0, m11 = 0;
    walk_pairs(gs, gb, ri, support, [&](int side, int j, V3<T> rij, T r2) {
        const EisV<T> &S = side == 0 ? P : B;
        const V3<T> g = grad_w(rij, r2, h);
        const T vol = S.m[j] / S.rho[j];
        m00 = m00 - vol * rij.x * g.x;
        m01 = m01 - vol * rij.x * g.z;
        m10 = m10 - vol * rij.z * g.x;
        m11 = m11 - vol * rij.z * g.z;
    });
    const T det = m00 * m11 - m01 * m10;
    T c[4] = {(T)1, (T)0, (T)0, (T)1};
    if (mabs(det) > (T)1.0e-8) {
        const T inv = (T)1.0 / det;
        c[0] = m11 * inv; c[1] = -m01 * inv; c[2] = -m10 * inv; c[3] = m00 * inv;
    }
    for (int k = 0; k < 4; ++k) P.kgc[4 * i + k] = c[k];
}

// Entries per fluid point (len[n] = 0).
template <typename T>
__global__ void k_eis_count(EisV<T> P, GridV<T> gs, GridV<T> gb, T support, int *len) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid == 0) len[P.n] = 0;
    if (tid >= P.n) return;
    const int i = gs.ids[tid];
    int c = 0;
    walk_pairs(gs, gb, ld3(P.pos, i), support, [&](int, int, V3<T>, T) { ++c; });
    len[i] = c;
}

// Slots of each slice: 64 x its longest row; width[slices] = 0 so the
// exclusive scan of slices+1 gives offset[slices] = total slots.
__global__ void k_eis_width(const int *len, int n, int slices, int *width) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s > slices) return;
    int w = 0;
    if (s < slices)
        for (int i = s * kSlice, e = min(n, i + kSlice); i < e; ++i) w = max(w, len[i]);
    width[s] = w * kSlice;
}

template <typename T>
__global__ void k_eis_fill(EisV<T> P, EisV<T> B, GridV<T> gs, GridV<T> gb, T support, T h,
                           const int *off, int *id, V4<T> *geo, int cap) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= P.n) return;
    const int i = gs.ids[tid];
    T c[4];
    for (int k = 0; k < 4; ++k) c[k] = P.kgc[4 * i + k];
    int e = slot0(off, i);
    walk_pairs(gs, gb, ld3(P.pos, i), support, [&](int side, int j, V3<T> rij, T r2) {
        const EisV<T> &S = side == 0 ? P : B;
        const V3<T> g = corrected(c, grad_w(rij, r2, h));
        const T vol = S.m[j] / S.rho[j];
        // KERNEL_EISPH_KNL.py:67-71 _laplace_weight
        const T lap = (T)2.0 * vol * dot(rij, g) / r2;
        if (e < cap) {
            id[e] = side == 0 ? j : -1 - j;
            geo[e] = V4<T>{g.x, g.z, lap, vol};
        }
        e += kSlice;
    });
}

// ---------------------------------------------------------------- the step's passes
// KERNEL_EISPH_BC.py:6-31: a ghost's velocity from its wall velocity and mirror.
template <typename T> __device__ __forceinline__ V3<T> ghost_of(const V3<T> &vbc, const V3<T> &vm) {
    return (T)2.0 * vbc - vm;
}
// A ghost's field as the pass reads it: stored in B (kFly false, the
// reference kernels' reads) or formed from its mirror (kFly true).
template <typename T, bool kFly>
__device__ __forceinline__ V3<T> ghost_vec(const EisV<T> &B, const T *fluid, const T *stored,
                                           int b) {
    if (kFly) return ghost_of(ld3(B.vel_bc, b), ld3(fluid, B.mirror[b]));
    return ld3(stored, b);
}
template <typename T> __device__ __forceinline__ V3<T> gvec(const V4<T> &G) {
    return mk3(G.x, (T)0.0, G.y);
}

// The per-entry updates and per-point results of the passes.  The pass
// kernels and the one-workgroup multi-step kernel both use these, so both
// evaluate the same operations in the same order (-ffp-contract=off).
// KERNEL_EISPH_force.py:12-72: convection and Laplacian terms.
template <typename T>
__device__ __forceinline__ void advec_term(V3<T> &conv, V3<T> &lap, const V3<T> &vi,
                                           const V3<T> &vj, const V4<T> &G) {
    const V3<T> dv = vj - vi;
    conv = conv + (G.w * dot(vi, gvec(G))) * dv;
    lap = lap - G.z * dv;
}
template <typename T>
__device__ __forceinline__ V3<T> advec_acc(const V3<T> &conv, const V3<T> &lap, T nu) {
    return -conv + nu * lap;
}
// KERNEL_EISPH_PPE.py:12-79 (and the divergence of KERNEL_EISPH_step.py:64-121).
template <typename T>
__device__ __forceinline__ void div_term(T &div, const V3<T> &vi, const V3<T> &vj,
                                         const V4<T> &G) {
    div = div + G.w * dot(vj - vi, gvec(G));
}
template <typename T>
__device__ __forceinline__ void ppe_term(T &div, T &A, T &Ap, const V3<T> &vi, const V3<T> &vj,
                                         T pj, const V4<T> &G) {
    div_term(div, vi, vj, G);
    A = A + G.z;
    Ap = Ap + G.z * pj;
}
// KERNEL_EISPH_PPE.py:100-125: the gauge from row 0 (its raw b, before the
// mean removal), then p = (b + sum A_ij p_j) / A_ii - gauge.
template <typename T> __device__ __forceinline__ T gauge_of(T b0_raw, T mean, T A0, T Ap0) {
    T gauge = 0;
    if (mabs(A0) > (T)1.0e-12) gauge = (b0_raw - mean + Ap0) / A0;
    return gauge;
}
template <typename T> __device__ __forceinline__ T solved_p(T b, T p, T A, T Ap, T gauge) {
    if (mabs(A) > (T)1.0e-12) p = (b + Ap) / A;
    return p - gauge;
}
// KERNEL_EISPH_step.py:12-61: the pressure gradient.
template <typename T>
__device__ __forceinline__ void correct_term(V3<T> &gp, T pi, T pj, const V4<T> &G) {
    gp = gp + (G.w * (pj - pi)) * gvec(G);
}
// KERNEL_EISPH_step.py:64-121: the Laplacian of p.
template <typename T> __device__ __forceinline__ void op_term(T &op, T pi, T pj, const V4<T> &G) {
    op = op + G.z * (pi - pj);
}

// KERNEL_EISPH_force.py:12-72 Kernel_advec_vis_eisph
template <typename T, bool kFly>
__global__ void __launch_bounds__(kEisThreads)
k_eis_advec(EisV<T> P, EisV<T> B, NbrV<T> L, T nu, T dt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    const V3<T> vi = ld3(P.vel, i);
    V3<T> conv = mk3((T)0, (T)0, (T)0), lap = conv;
    each_entry(L, i, [&](int j, const V4<T> &G) {
        const V3<T> vj = j >= 0 ? ld3(P.vel, j) : ghost_vec<T, kFly>(B, P.vel, B.vel, -1 - j);
        advec_term(conv, lap, vi, vj, G);
    });
    const V3<T> acc = advec_acc(conv, lap, nu);
    st3(P.acc, i, acc);
    st3(P.vel_star, i, vi + dt * acc);
}

// Sum of v over the workgroup (fixed order: wave trees, then waves in order).
__device__ __forceinline__ double block_sum(double v, double *lds) {
    v = wave_sum(v);
    const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if ((threadIdx.x & 63) == 0) lds[w] = v;
    __syncthreads();
    double s = 0.0;
    for (int k = 0; k < nw; ++k) s += lds[k];
    __syncthreads();
    return s;
}

// Total of the nb workgroup sums part[2..2+nb): each thread adds one run of
// consecutive sums, then block_sum.  The shape depends on nb and blockDim
// only, so every caller gets the same bits.
__device__ __forceinline__ double sum_partials(const double *part, int nb, double *lds) {
    const int per = (nb + blockDim.x - 1) / blockDim.x;
    const int a = threadIdx.x * per, z = min(nb, a + per);
    double acc = 0.0;
    for (int k = a; k < z; ++k) acc += part[2 + k];
    return block_sum(acc, lds);
}

// KERNEL_EISPH_PPE.py:12-79 Kernel_build_ppe_eisph, with the workgroup sums
// of the raw bi (for EISPH.py:72) in part[2 + workgroup] and the raw bi[0]
// in part[1]; with kFly the ghost threads publish vel_star
// (KERNEL_EISPH_BC.py:6-31, predicted).
template <typename T, bool kFly>
__global__ void __launch_bounds__(kEisThreads)
k_eis_ppe(EisV<T> P, EisV<T> B, NbrV<T> L, T rho0, T dt, double *part) {
    __shared__ double lds[kEisThreads / 64];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    double mine = 0.0;
    if (t < P.n) {
        const int i = t;
        const V3<T> vi = ld3(P.vel_star, i);
        T div = 0, A = 0, Ap = 0;
        each_entry(L, i, [&](int j, const V4<T> &G) {
            V3<T> vj;
            T pj;
            if (j >= 0) {
                vj = ld3(P.vel_star, j);
                pj = P.pres[j];
            } else {
                const int b = -1 - j;
                vj = ghost_vec<T, kFly>(B, P.vel_star, B.vel_star, b);
                pj = P.pres[B.mirror[b]];
            }
            ppe_term(div, A, Ap, vi, vj, pj, G);
        });
        const T bi = rho0 * div / dt;
        P.div[i] = div;
        P.Aij[i] = A;
        P.bi[i] = bi;
        P.Aijpj[i] = Ap;
        mine = (double)bi;
        if (i == 0) part[1] = (double)bi;
    } else if (kFly && t - P.n < B.n) {
        const int b = t - P.n;
        st3(B.vel_star, b, ghost_of(ld3(B.vel_bc, b), ld3(P.vel_star, B.mirror[b])));
    }
    const double s = block_sum(mine, lds);
    if (threadIdx.x == 0) part[2 + blockIdx.x] = s;
}

// EISPH.py:72 the right-hand-side sum: part[0] in FP64, residual[0] rounded
// to T, as the reference's array_sum leaves it.
template <typename T>
__global__ void __launch_bounds__(kEisThreads) k_eis_rhs_sum(EisV<T> P, double *part, int nb) {
    __shared__ double lds[kEisThreads / 64];
    const double total = sum_partials(part, nb, lds);
    if (threadIdx.x == 0) {
        part[0] = total;
        P.res[0] = (T)total;
    }
}

// KERNEL_EISPH_PPE.py:82-125 Kernel_remove_bi_mean and Kernel_solve_ppe_eisph.
// The mean is residual[0] / n with residual[0] the sum rounded to T, as in
// the reference; the gauge row 0 uses its own bi after the mean removal.
// kInline: every workgroup adds the nb workgroup sums itself (workgroup 0
// publishes them); otherwise k_eis_rhs_sum has left the total in part[0].
template <typename T, bool kInline>
__global__ void __launch_bounds__(kEisThreads)
k_eis_solve(EisV<T> P, double *part, int nb) {
    __shared__ double lds[kEisThreads / 64];
    double total;
    if (kInline) {
        total = sum_partials(part, nb, lds);
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            part[0] = total;
            P.res[0] = (T)total;
        }
    } else {
        total = part[0];
    }
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    const T mean = (T)total / (T)P.n;
    const T bi = P.bi[i] - mean;
    P.bi[i] = bi;
    const T gauge = gauge_of((T)part[1], mean, P.Aij[0], P.Aijpj[0]);
    P.pres[i] = solved_p(bi, P.pres[i], P.Aij[i], P.Aijpj[i], gauge);
}

// KERNEL_EISPH_step.py:12-61 Kernel_correct_eisph; with kFly the ghost
// threads publish pres (KERNEL_EISPH_BC.py:34-47).
template <typename T, bool kFly>
__global__ void __launch_bounds__(kEisThreads)
k_eis_correct(EisV<T> P, EisV<T> B, NbrV<T> L, T dt) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= P.n) {
        const int b = t - P.n;
        if (kFly && b < B.n) B.pres[b] = P.pres[B.mirror[b]];
        return;
    }
    const int i = t;
    const T pi = P.pres[i];
    V3<T> gp = mk3((T)0, (T)0, (T)0);
    each_entry(L, i, [&](int j, const V4<T> &G) {
        T pj;
        if (j >= 0) pj = P.pres[j];
        else pj = kFly ? P.pres[B.mirror[-1 - j]] : B.pres[-1 - j];
        correct_term(gp, pi, pj, G);
    });
    const V3<T> old = ld3(P.vel, i);
    const V3<T> v = ld3(P.vel_star, i) - (dt * gp) / P.rho[i];
    st3(P.vel, i, v);
    st3(P.acc, i, (v - old) / dt);
}

// KERNEL_EISPH_step.py:64-121 Kernel_diagnostics_eisph (kDiag); with kFly
// the ghost threads publish vel (KERNEL_EISPH_BC.py:6-31).
template <typename T, bool kFly, bool kDiag>
__global__ void __launch_bounds__(kEisThreads)
k_eis_finish(EisV<T> P, EisV<T> B, NbrV<T> L) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= P.n) {
        const int b = t - P.n;
        if (kFly && b < B.n)
            st3(B.vel, b, ghost_of(ld3(B.vel_bc, b), ld3(P.vel, B.mirror[b])));
        return;
    }
    if (!kDiag) return;
    const int i = t;
    const V3<T> vi = ld3(P.vel, i);
    const T pi = P.pres[i];
    T div = 0, op = 0;
    each_entry(L, i, [&](int j, const V4<T> &G) {
        V3<T> vj;
        T pj;
        if (j >= 0) {
            vj = ld3(P.vel, j);
            pj = P.pres[j];
        } else {
            const int b = -1 - j;
            vj = ghost_vec<T, kFly>(B, P.vel, B.vel, b);
            pj = kFly ? P.pres[B.mirror[b]] : B.pres[b];
        }
        div_term(div, vi, vj, G);
        op_term(op, pi, pj, G);
    });
    P.div[i] = div;
    P.res[i] = op - P.bi[i];
}

template <typename T>
__global__ void k_eis_dirichlet(EisV<T> P, EisV<T> B, int predicted) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B.n) return;
    const T *src = predicted ? P.vel_star : P.vel;
    T *dst = predicted ? B.vel_star : B.vel;
    st3(dst, b, ghost_of(ld3(B.vel_bc, b), ld3(src, B.mirror[b])));
}

template <typename T>
__global__ void k_eis_pressure_boundary(EisV<T> P, EisV<T> B) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < B.n) B.pres[b] = P.pres[B.mirror[b]];
}

// ---------------------------------------------------------------- many steps, one workgroup
// For a small cloud each launch of the five-launch step is a few
// microseconds of almost no work.  k_eis_steps runs K steps in one
// workgroup of 1024 threads: the fields a pass reads at other points (vel,
// vel_star, pres, the ghosts' vel_bc and mirror) and the own-point values
// carried between passes (rho, bi, Aij, Aijpj, first list slot, row length)
// live in LDS, only the neighbour list is read from memory, and workgroup
// barriers separate the passes.  The FP64 sum of b keeps the five-launch
// shape: the same 64-point wave trees, 256-point group sums, and the
// group sums added as sum_partials adds them with 256 threads; so every
// field ends as K calls of the five-launch step leave it.  Ghost fields and
// the write-only outputs (acc, div, bi, Aij, Aijpj, residual) are stored
// once, from the last step.
constexpr int kOneThreads = 1024;
constexpr int kOneMaxGroups = 64;
constexpr long kOneMaxLds = 60 * 1024;  // dynamic LDS; the static part is below 1 KiB

template <typename T> long one_group_lds(int nf, int ng) {
    return (long)nf * (11 * (long)sizeof(T) + 2 * (long)sizeof(int)) +
           (long)ng * (3 * (long)sizeof(T) + (long)sizeof(int));
}
// ... and with the volume table of the register-row kernel
template <typename T> long reg_group_lds(int nf, int ng) {
    return one_group_lds<T>(nf, ng) + ((long)nf + ng) * (long)sizeof(T);
}


// The body of both one-workgroup kernels.  kReg > 0: one point per thread
// (blockDim.x >= n), whose first kReg row entries stay in registers for the
// whole launch, so the passes read only LDS (and the list for longer rows).
template <typename T, int kReg>
__device__ __forceinline__ void steps_body(EisV<T> P, EisV<T> B, NbrV<T> L, T nu, T rho0, T dt,
                                           double *part, int nb, int steps, int diag,
                                           double *one_lds, double *wsum, double *gsum,
                                           T &b0_s) {
    const int n = P.n, ng = B.n, tid = threadIdx.x, nt = blockDim.x;
    constexpr int kWaves = kEisThreads / 64;  // waves of a five-launch workgroup
    T *vel = reinterpret_cast<T *>(one_lds);
    T *vst = vel + 3 * n;
    T *pres = vst + 3 * n;
    T *rho = pres + n;
    T *bi = rho + n;
    T *A = bi + n;
    T *Ap = A + n;
    T *vbc = Ap + n;
    // kReg > 0: V_j = m_j / rho_j of the fluid then the ghosts, the value
    // the list's geometry holds (the same IEEE division as k_eis_fill)
    T *vtab = vbc + 3 * ng;
    int *mir = reinterpret_cast<int *>(vtab + (kReg > 0 ? n + ng : 0));
    int *ent = mir + ng;
    int *cnt = ent + n;
    for (int i = tid; i < n; i += nt) {
        st3(vel, i, ld3(P.vel, i));
        st3(vst, i, ld3(P.vel_star, i));
        pres[i] = P.pres[i];
        rho[i] = P.rho[i];
        ent[i] = slot0(L.off, i);
        cnt[i] = L.len[i];
    }
    for (int b = tid; b < ng; b += nt) {
        st3(vbc, b, ld3(B.vel_bc, b));
        mir[b] = B.mirror[b];
    }
    if (kReg > 0) {
        for (int i = tid; i < n; i += nt) vtab[i] = P.m[i] / P.rho[i];
        for (int b = tid; b < ng; b += nt) vtab[n + b] = B.m[b] / B.rho[b];
    }
    __syncthreads();
    auto ghost = [&](const T *f, int b) { return ghost_of(ld3(vbc, b), ld3(f, mir[b])); };
    // the thread's own row in registers (kReg > 0: point tid): id, gx, gz,
    // A_ij of the first kReg entries; V_j comes from vtab
    constexpr int kC = kReg > 0 ? kReg : 1;
    int cid[kC];
    T cgx[kC], cgz[kC], cA[kC];
    int own_e = 0, own_len = 0;
    if (kReg > 0 && tid < n) {
        own_e = slot0(L.off, tid);
        own_len = L.len[tid];
#pragma unroll
        for (int k = 0; k < kC; ++k) {
            cid[k] = 0;
            cgx[k] = cgz[k] = cA[k] = (T)0;
            if (k < own_len) {
                const V4<T> G = L.geo[own_e + k * kSlice];
                cid[k] = L.id[own_e + k * kSlice];
                cgx[k] = G.x;
                cgz[k] = G.y;
                cA[k] = G.z;
            }
        }
    }
    // row i's entries in order (with kReg > 0, i is always tid)
    auto row = [&](int i, auto &&fn) {
        if constexpr (kReg > 0) {
#pragma unroll
            for (int k = 0; k < kC; ++k) {
                if (k < own_len) {
                    const int j = cid[k];
                    fn(j, V4<T>{cgx[k], cgz[k], cA[k], vtab[j >= 0 ? j : n - 1 - j]});
                }
                // keep the scheduler from hoisting every entry's LDS reads
                // at once (their live ranges would spill the cached row)
                if ((k & 3) == 3) __builtin_amdgcn_sched_barrier(0);
            }
            for (int k = kC, e = own_e + kC * kSlice; k < own_len; ++k, e += kSlice)
                fn(L.id[e], L.geo[e]);
        } else {
            row_batched(L, ent[i], cnt[i], fn);
        }
    };

    double total = 0.0;
    for (int s = 0; s < steps; ++s) {
        const bool last = s == steps - 1;
        // advec (KERNEL_EISPH_force.py:12-72)
        for (int i = tid; i < n; i += nt) {
            const V3<T> vi = ld3(vel, i);
            V3<T> conv = mk3((T)0, (T)0, (T)0), lap = conv;
            row(i, [&](int j, const V4<T> &G) {
                const V3<T> vj = j >= 0 ? ld3(vel, j) : ghost(vel, -1 - j);
                advec_term(conv, lap, vi, vj, G);
            });
            st3(vst, i, vi + dt * advec_acc(conv, lap, nu));
        }
        __syncthreads();
        // ppe (KERNEL_EISPH_PPE.py:12-79) over the five-launch thread numbering
        for (int t0 = 0; t0 < nb * kEisThreads; t0 += nt) {
            const int i = t0 + tid;
            double mine = 0.0;
            if (i < n) {
                const V3<T> vi = ld3(vst, i);
                T div = 0, a = 0, ap = 0;
                row(i, [&](int j, const V4<T> &G) {
                    V3<T> vj;
                    T pj;
                    if (j >= 0) {
                        vj = ld3(vst, j);
                        pj = pres[j];
                    } else {
                        const int b = -1 - j;
                        vj = ghost(vst, b);
                        pj = pres[mir[b]];
                    }
                    ppe_term(div, a, ap, vi, vj, pj, G);
                });
                const T b = rho0 * div / dt;
                bi[i] = b;
                A[i] = a;
                Ap[i] = ap;
                if (last && !diag) P.div[i] = div;
                mine = (double)b;
                if (i == 0) b0_s = b;
            }
            const double w = wave_sum(mine);
            if ((tid & 63) == 0) wsum[tid >> 6] = w;
            __syncthreads();
            if (tid < nt / kEisThreads) {
                const int g = t0 / kEisThreads + tid;
                if (g < nb) {
                    double acc = 0.0;
                    for (int k = 0; k < kWaves; ++k) acc += wsum[tid * kWaves + k];
                    gsum[g] = acc;
                }
            }
            __syncthreads();
        }
        // EISPH.py:72, in the shape of sum_partials over 256 threads
        if (tid < kEisThreads) {
            const int per = (nb + kEisThreads - 1) / kEisThreads;
            const int a = tid * per, z = min(nb, a + per);
            double acc = 0.0;
            for (int k = a; k < z; ++k) acc += gsum[k];
            const double w = wave_sum(acc);
            if ((tid & 63) == 0) wsum[tid >> 6] = w;
        }
        __syncthreads();
        total = 0.0;
        for (int k = 0; k < kWaves; ++k) total += wsum[k];
        // solve (KERNEL_EISPH_PPE.py:82-125)
        const T mean = (T)total / (T)n;
        const T gauge = gauge_of(b0_s, mean, A[0], Ap[0]);
        for (int i = tid; i < n; i += nt) {
            const T b = bi[i] - mean;
            bi[i] = b;
            pres[i] = solved_p(b, pres[i], A[i], Ap[i], gauge);
            if (last) {
                P.bi[i] = b;
                P.Aij[i] = A[i];
                P.Aijpj[i] = Ap[i];
            }
        }
        if (last && tid == 0) {
            P.res[0] = (T)total;
            part[0] = total;
            part[1] = (double)b0_s;
        }
        __syncthreads();
        // correct (KERNEL_EISPH_step.py:12-61)
        for (int i = tid; i < n; i += nt) {
            const T pi = pres[i];
            V3<T> gp = mk3((T)0, (T)0, (T)0);
            row(i, [&](int j, const V4<T> &G) {
                correct_term(gp, pi, j >= 0 ? pres[j] : pres[mir[-1 - j]], G);
            });
            const V3<T> old = ld3(vel, i);
            const V3<T> v = ld3(vst, i) - (dt * gp) / rho[i];
            st3(vel, i, v);
            if (last) st3(P.acc, i, (v - old) / dt);
        }
        __syncthreads();
    }
    // diagnostics of the last step (KERNEL_EISPH_step.py:64-121)
    if (diag) {
        for (int i = tid; i < n; i += nt) {
            const V3<T> vi = ld3(vel, i);
            const T pi = pres[i];
            T div = 0, op = 0;
            row(i, [&](int j, const V4<T> &G) {
                V3<T> vj;
                T pj;
                if (j >= 0) {
                    vj = ld3(vel, j);
                    pj = pres[j];
                } else {
                    const int b = -1 - j;
                    vj = ghost(vel, b);
                    pj = pres[mir[b]];
                }
                div_term(div, vi, vj, G);
                op_term(op, pi, pj, G);
            });
            P.div[i] = div;
            P.res[i] = op - bi[i];
        }
    }
    for (int i = tid; i < n; i += nt) {
        st3(P.vel, i, ld3(vel, i));
        st3(P.vel_star, i, ld3(vst, i));
        P.pres[i] = pres[i];
    }
    // KERNEL_EISPH_BC.py:6-47 from the final fields
    for (int b = tid; b < ng; b += nt) {
        st3(B.vel_star, b, ghost(vst, b));
        B.pres[b] = pres[mir[b]];
        st3(B.vel, b, ghost(vel, b));
    }
    if (tid < nb) part[2 + tid] = gsum[tid];
}

template <typename T>
__global__ void __launch_bounds__(kOneThreads)
k_eis_steps(EisV<T> P, EisV<T> B, NbrV<T> L, T nu, T rho0, T dt, double *part, int nb,
            int steps, int diag) {
    extern __shared__ double one_lds[];
    __shared__ double wsum[kOneThreads / 64];
    __shared__ double gsum[kOneMaxGroups];
    __shared__ T b0_s;
    steps_body<T, 0>(P, B, L, nu, rho0, dt, part, nb, steps, diag, one_lds, wsum, gsum, b0_s);
}

// FP32 clouds of at most kRegThreads points: one point per thread, rows of up
// to kRegRow entries held in registers.
constexpr int kRegThreads = 768;
constexpr int kRegRow = 20;
template <typename T>
__global__ void __launch_bounds__(kRegThreads)
k_eis_steps_reg(EisV<T> P, EisV<T> B, NbrV<T> L, T nu, T rho0, T dt, double *part, int nb,
                int steps, int diag) {
    extern __shared__ double one_lds[];
    __shared__ double wsum[kRegThreads / 64];
    __shared__ double gsum[kOneMaxGroups];
    __shared__ T b0_s;
    steps_body<T, kRegRow>(P, B, L, nu, rho0, dt, part, nb, steps, diag, one_lds, wsum, gsum,
                           b0_s);
}

// ---------------------------------------------------------------- checks
int check_points(const sphdem_eisph_t *S, const char *what, bool fluid) {
    SPHDEM_REQUIRE(S, std::string(what) + " is NULL");
    SPHDEM_REQUIRE(S->n >= 0, std::string(what) + ": negative count");
    if (S->n == 0) return SPHDEM_OK;
    SPHDEM_REQUIRE(S->pos && S->vel && S->vel_star && S->vel_bc && S->rho && S->m && S->pres &&
                       S->mirror,
                   std::string(what) + ": field pointer is NULL");
    if (fluid)
        SPHDEM_REQUIRE(S->acc && S->kgc && S->Aij && S->bi && S->Aijpj && S->divergence &&
                           S->residual,
                       std::string(what) + ": field pointer is NULL");
    return SPHDEM_OK;
}

int check_pair(const sphdem_eisph_params_t *p, const sphdem_eisph_t *P, const sphdem_eisph_t *B) {
    SPHDEM_REQUIRE(p, "params is NULL");
    SPHDEM_REQUIRE(p->precision == 4 || p->precision == 8, "precision must be 4 or 8");
    int rc = check_points(P, "fluid points", true);
    if (rc) return rc;
    return check_points(B, "ghost points", false);
}

int check_grids(const sphdem_eisph_params_t *p, const sphdem_eisph_t *P, const sphdem_eisph_t *B,
                const sphdem_grid_t *gs, const sphdem_grid_t *gb) {
    SPHDEM_REQUIRE(gs && gb, "grid is NULL");
    SPHDEM_REQUIRE(gs->precision == p->precision && gb->precision == p->precision,
                   "grid precision differs from the point precision");
    SPHDEM_REQUIRE(gs->n == P->n, "fluid grid was built from a different number of points");
    SPHDEM_REQUIRE(gb->n == B->n, "ghost grid was built from a different number of points");
    SPHDEM_REQUIRE(gs->n_inst <= 1 && gb->n_inst <= 1, "batched grids are not supported here");
    SPHDEM_REQUIRE(gs->n == 0 || (gs->point_ids && gs->sorted), "fluid grid is not built");
    SPHDEM_REQUIRE(gb->n == 0 || (gb->point_ids && gb->sorted), "ghost grid is not built");
    return SPHDEM_OK;
}

int check_list(const sphdem_eisph_t *P, const sphdem_eisph_nbr_t *L) {
    SPHDEM_REQUIRE(L, "neighbour list is NULL");
    SPHDEM_REQUIRE(L->n == P->n, "neighbour list was built for a different number of points");
    SPHDEM_REQUIRE(L->len && L->offset && (L->cap == 0 || (L->id && L->geo)),
                   "neighbour list is not built");
    return SPHDEM_OK;
}

int check_sums(const sphdem_eisph_nbr_t *L) {
    SPHDEM_REQUIRE(L->part, "neighbour list has no sum buffer");
    return SPHDEM_OK;
}

#define EIS_DISPATCH(prec, CALL)       \
    do {                               \
        if ((prec) == 4) {             \
            using T = float;           \
            CALL;                      \
        } else {                       \
            using T = double;          \
            CALL;                      \
        }                              \
    } while (0)

inline dim3 fluid_grid(int n) { return dim3(blocks_for(n, kEisThreads)); }
inline dim3 all_grid(int nf, int nb) { return dim3(blocks_for((long)nf + nb, kEisThreads)); }

// Workgroups of a ppe launch over nf fluid (+ nb ghost) threads.
inline int ppe_groups(int nf, int nb) { return blocks_for((long)nf + nb, kEisThreads); }
// Below this many workgroup sums the solve kernel adds them itself.
constexpr int kInlineSums = 64;

template <typename T, bool kFly>
void launch_ppe(const sphdem_eisph_params_t *p, sphdem_eisph_t *P, const sphdem_eisph_t *B,
                const sphdem_eisph_nbr_t *L, hipStream_t st) {
    const int nb_ghost = kFly ? B->n : 0;
    hipLaunchKernelGGL((k_eis_ppe<T, kFly>), dim3(ppe_groups(P->n, nb_ghost)),
                       dim3(kEisThreads), 0, st, eview<T>(P), eview<T>(B), nview<T>(L),
                       (T)p->rho0, (T)p->dt, L->part);
}

}  // namespace
}  // namespace sphdem

using namespace sphdem;

extern "C" {

int sphdem_eisph_slices(int n) { return n < 0 ? -1 : blocks_for(n, kSlice); }

long sphdem_eisph_part_size(int n_fluid, int n_ghost) {
    if (n_fluid < 0 || n_ghost < 0) return -1;
    return 2 + (long)blocks_for((long)n_fluid + n_ghost, kEisThreads);
}

int sphdem_eisph_kgc(const sphdem_eisph_params_t *p, sphdem_eisph_t *P, const sphdem_eisph_t *B,
                     const sphdem_grid_t *gs, const sphdem_grid_t *gb, void *stream) {
    int rc = check_pair(p, P, B);
    if (!rc) rc = check_grids(p, P, B, gs, gb);
    if (rc) return rc;
    if (P->n == 0) return SPHDEM_OK;
    EIS_DISPATCH(p->precision,
                 hipLaunchKernelGGL(k_eis_kgc<T>, fluid_grid(P->n), dim3(kEisThreads), 0,
                                    as_stream(stream), eview<T>(P), eview<T>(B), view<T>(gs),
                                    view<T>(gb), (T)p->support, (T)p->h));
    return check_launch("eisph_kgc");
}

int sphdem_eisph_nbr_count(const sphdem_eisph_params_t *p, const sphdem_eisph_t *P,
                           const sphdem_eisph_t *B, const sphdem_grid_t *gs,
                           const sphdem_grid_t *gb, sphdem_eisph_nbr_t *L, void *stream) {
    int rc = check_pair(p, P, B);
    if (!rc) rc = check_grids(p, P, B, gs, gb);
    if (rc) return rc;
    SPHDEM_REQUIRE(L && L->len && L->offset && L->width && L->scan_tmp,
                   "neighbour list buffers are NULL");
    SPHDEM_REQUIRE(L->n == P->n, "neighbour list was sized for a different number of points");
    hipStream_t st = as_stream(stream);
    const int slices = blocks_for(P->n, kSlice);
    EIS_DISPATCH(p->precision,
                 hipLaunchKernelGGL(k_eis_count<T>, dim3(blocks_for(P->n + 1, kEisThreads)),
                                    dim3(kEisThreads), 0, st, eview<T>(P), view<T>(gs),
                                    view<T>(gb), (T)p->support, L->len));
    hipLaunchKernelGGL(k_eis_width, dim3(blocks_for(slices + 1, kEisThreads)), dim3(kEisThreads),
                       0, st, L->len, P->n, slices, L->width);
    rc = check_launch("eisph_nbr_count");
    if (rc) return rc;
    return exclusive_scan(L->width, L->offset, slices + 1, L->scan_tmp, st);
}

int sphdem_eisph_nbr_fill(const sphdem_eisph_params_t *p, const sphdem_eisph_t *P,
                          const sphdem_eisph_t *B, const sphdem_grid_t *gs,
                          const sphdem_grid_t *gb, sphdem_eisph_nbr_t *L, void *stream) {
    int rc = check_pair(p, P, B);
    if (!rc) rc = check_grids(p, P, B, gs, gb);
    if (!rc) rc = check_list(P, L);
    if (rc) return rc;
    SPHDEM_REQUIRE(L->cap >= 0, "negative neighbour list capacity");
    if (P->n == 0) return SPHDEM_OK;
    EIS_DISPATCH(p->precision,
                 hipLaunchKernelGGL(k_eis_fill<T>, fluid_grid(P->n), dim3(kEisThreads), 0,
                                    as_stream(stream), eview<T>(P), eview<T>(B), view<T>(gs),
                                    view<T>(gb), (T)p->support, (T)p->h, L->offset, L->id,
                                    (V4<T> *)L->geo, L->cap));
    return check_launch("eisph_nbr_fill");
}

int sphdem_eisph_dirichlet(int precision, const sphdem_eisph_t *P, sphdem_eisph_t *B,
                           int predicted, void *stream) {
    SPHDEM_REQUIRE(precision == 4 || precision == 8, "precision must be 4 or 8");
    int rc = check_points(P, "fluid points", false);
    if (!rc) rc = check_points(B, "ghost points", false);
    if (rc) return rc;
    if (B->n == 0) return SPHDEM_OK;
    EIS_DISPATCH(precision,
                 hipLaunchKernelGGL(k_eis_dirichlet<T>, dim3(blocks_for(B->n, kEisThreads)),
                                    dim3(kEisThreads), 0, as_stream(stream), eview<T>(P),
                                    eview<T>(B), predicted));
    return check_launch("eisph_dirichlet");
}

int sphdem_eisph_pressure_boundary(int precision, const sphdem_eisph_t *P, sphdem_eisph_t *B,
                                   void *stream) {
    SPHDEM_REQUIRE(precision == 4 || precision == 8, "precision must be 4 or 8");
    int rc = check_points(P, "fluid points", false);
    if (!rc) rc = check_points(B, "ghost points", false);
    if (rc) return rc;
    if (B->n == 0) return SPHDEM_OK;
    EIS_DISPATCH(precision,
                 hipLaunchKernelGGL(k_eis_pressure_boundary<T>, dim3(blocks_for(B->n, kEisThreads)),
                                    dim3(kEisThreads), 0, as_stream(stream), eview<T>(P),
                                    eview<T>(B)));
    return check_launch("eisph_pressure_boundary");
}

int sphdem_eisph_advec_vis(const sphdem_eisph_params_t *p, sphdem_eisph_t *P,
                           const sphdem_eisph_t *B, const sphdem_eisph_nbr_t *L, void *stream) {
    int rc = check_pair(p, P, B);
    if (!rc) rc = check_list(P, L);
    if (rc) return rc;
    if (P->n == 0) return SPHDEM_OK;
    EIS_DISPATCH(p->precision,
                 hipLaunchKernelGGL((k_eis_advec<T, false>), fluid_grid(P->n), dim3(kEisThreads),
                                    0, as_stream(stream), eview<T>(P), eview<T>(B), nview<T>(L),
                                    (T)p->nu, (T)p->dt));
    return check_launch("eisph_advec_vis");
}

int sphdem_eisph_build_ppe(const sphdem_eisph_params_t *p, sphdem_eisph_t *P,
                           const sphdem_eisph_t *B, const sphdem_eisph_nbr_t *L, void *stream) {
    int rc = check_pair(p, P, B);
    if (!rc) rc = check_list(P, L);
    if (!rc) rc = check_sums(L);
    if (rc) return rc;
    SPHDEM_REQUIRE(P->n > 0, "the pressure equation needs fluid points");
    hipStream_t st = as_stream(stream);
    EIS_DISPATCH(p->precision, ({
        launch_ppe<T, false>(p, P, B, L, st);
        hipLaunchKernelGGL(k_eis_rhs_sum<T>, dim3(1), dim3(kEisThreads), 0, st, eview<T>(P),
                           L->part, ppe_groups(P->n, 0));
    }));
    return check_launch("eisph_build_ppe");
}

int sphdem_eisph_solve_ppe(const sphdem_eisph_params_t *p, sphdem_eisph_t *P,
                           const sphdem_eisph_nbr_t *L, void *stream) {
    SPHDEM_REQUIRE(p && (p->precision == 4 || p->precision == 8), "invalid params");
    int rc = check_points(P, "fluid points", true);
    if (!rc) rc = check_list(P, L);
    if (!rc) rc = check_sums(L);
    if (rc) return rc;
    SPHDEM_REQUIRE(P->n > 0, "the pressure equation needs fluid points");
    EIS_DISPATCH(p->precision,
                 hipLaunchKernelGGL((k_eis_solve<T, false>), fluid_grid(P->n), dim3(kEisThreads),
                                    0, as_stream(stream), eview<T>(P), L->part, 0));
    return check_launch("eisph_solve_ppe");
}

int sphdem_eisph_correct(const sphdem_eisph_params_t *p, sphdem_eisph_t *P,
                         const sphdem_eisph_t *B, const sphdem_eisph_nbr_t *L, void *stream) {
    int rc = check_pair(p, P, B);
    if (!rc) rc = check_list(P, L);
    if (rc) return rc;
    if (P->n == 0) return SPHDEM_OK;
    EIS_DISPATCH(p->precision,
                 hipLaunchKernelGGL((k_eis_correct<T, false>), fluid_grid(P->n),
                                    dim3(kEisThreads), 0, as_stream(stream), eview<T>(P),
                                    eview<T>(B), nview<T>(L), (T)p->dt));
    return check_launch("eisph_correct");
}

int sphdem_eisph_diagnostics(const sphdem_eisph_params_t *p, sphdem_eisph_t *P,
                             const sphdem_eisph_t *B, const sphdem_eisph_nbr_t *L, void *stream) {
    int rc = check_pair(p, P, B);
    if (!rc) rc = check_list(P, L);
    if (rc) return rc;
    if (P->n == 0) return SPHDEM_OK;
    EIS_DISPATCH(p->precision,
                 hipLaunchKernelGGL((k_eis_finish<T, false, true>), fluid_grid(P->n),
                                    dim3(kEisThreads), 0, as_stream(stream), eview<T>(P),
                                    eview<T>(B), nview<T>(L)));
    return check_launch("eisph_diagnostics");
}

int sphdem_eisph_step(const sphdem_eisph_params_t *p, sphdem_eisph_t *P, sphdem_eisph_t *B,
                      const sphdem_eisph_nbr_t *L, int diagnostics, void *stream) {
    int rc = check_pair(p, P, B);
    if (!rc) rc = check_list(P, L);
    if (!rc) rc = check_sums(L);
    if (rc) return rc;
    SPHDEM_REQUIRE(P->n > 0, "the pressure equation needs fluid points");
    hipStream_t st = as_stream(stream);
    const dim3 blk(kEisThreads), gf = fluid_grid(P->n), ga = all_grid(P->n, B->n);
    EIS_DISPATCH(p->precision, ({
        const EisV<T> Pv = eview<T>(P), Bv = eview<T>(B);
        const NbrV<T> Lv = nview<T>(L);
        {
            ProfScope ps(SPHDEM_PASS_EISPH_ADVEC, st);
            hipLaunchKernelGGL((k_eis_advec<T, true>), gf, blk, 0, st, Pv, Bv, Lv, (T)p->nu,
                               (T)p->dt);
        }
        {
            ProfScope ps(SPHDEM_PASS_EISPH_PPE, st);
            launch_ppe<T, true>(p, P, B, L, st);
        }
        {
            ProfScope ps(SPHDEM_PASS_EISPH_SOLVE, st);
            const int nb = ppe_groups(P->n, B->n);
            if (nb <= kInlineSums) {
                hipLaunchKernelGGL((k_eis_solve<T, true>), gf, blk, 0, st, Pv, L->part, nb);
            } else {
                hipLaunchKernelGGL(k_eis_rhs_sum<T>, dim3(1), blk, 0, st, Pv, L->part, nb);
                hipLaunchKernelGGL((k_eis_solve<T, false>), gf, blk, 0, st, Pv, L->part, nb);
            }
        }
        {
            ProfScope ps(SPHDEM_PASS_EISPH_CORRECT, st);
            hipLaunchKernelGGL((k_eis_correct<T, true>), ga, blk, 0, st, Pv, Bv, Lv, (T)p->dt);
        }
        ProfScope ps(SPHDEM_PASS_EISPH_FINISH, st);
        if (diagnostics)
            hipLaunchKernelGGL((k_eis_finish<T, true, true>), ga, blk, 0, st, Pv, Bv, Lv);
        else
            hipLaunchKernelGGL((k_eis_finish<T, true, false>), ga, blk, 0, st, Pv, Bv, Lv);
    }));
    return check_launch("eisph_step");
}

int sphdem_eisph_one_group(int precision, int n_fluid, int n_ghost) {
    if (n_fluid <= 0 || n_ghost < 0) return 0;
    const long bytes = precision == 8 ? one_group_lds<double>(n_fluid, n_ghost)
                                      : one_group_lds<float>(n_fluid, n_ghost);
    return bytes <= kOneMaxLds && ppe_groups(n_fluid, n_ghost) <= kOneMaxGroups;
}

int sphdem_eisph_steps(const sphdem_eisph_params_t *p, sphdem_eisph_t *P, sphdem_eisph_t *B,
                       const sphdem_eisph_nbr_t *L, int steps, int diagnostics, void *stream) {
    int rc = check_pair(p, P, B);
    if (!rc) rc = check_list(P, L);
    if (!rc) rc = check_sums(L);
    if (rc) return rc;
    SPHDEM_REQUIRE(steps >= 0, "negative step count");
    SPHDEM_REQUIRE(P->n > 0, "the pressure equation needs fluid points");
    if (steps == 0) return SPHDEM_OK;
    if (!sphdem_eisph_one_group(p->precision, P->n, B->n)) {
        for (int s = 0; s < steps; ++s) {
            rc = sphdem_eisph_step(p, P, B, L, diagnostics && s == steps - 1, stream);
            if (rc) return rc;
        }
        return SPHDEM_OK;
    }
    hipStream_t st = as_stream(stream);
    const int nb = ppe_groups(P->n, B->n);
    ProfScope ps(SPHDEM_PASS_EISPH_STEPS, st);
    if (p->precision == 4 && P->n <= kRegThreads &&
        reg_group_lds<float>(P->n, B->n) <= kOneMaxLds) {
        // one point per thread, rows in registers
        const int nt = kEisThreads * blocks_for(P->n, kEisThreads);
        hipLaunchKernelGGL(k_eis_steps_reg<float>, dim3(1), dim3(nt),
                           (size_t)reg_group_lds<float>(P->n, B->n), st, eview<float>(P),
                           eview<float>(B), nview<float>(L), (float)p->nu, (float)p->rho0,
                           (float)p->dt, L->part, nb, steps, diagnostics ? 1 : 0);
        return check_launch("eisph_steps");
    }
    EIS_DISPATCH(p->precision,
                 hipLaunchKernelGGL(k_eis_steps<T>, dim3(1), dim3(kOneThreads),
                                    (size_t)one_group_lds<T>(P->n, B->n), st, eview<T>(P),
                                    eview<T>(B), nview<T>(L), (T)p->nu, (T)p->rho0, (T)p->dt,
                                    L->part, nb, steps, diagnostics ? 1 : 0));
    return check_launch("eisph_steps");
}

}  // extern "C"
