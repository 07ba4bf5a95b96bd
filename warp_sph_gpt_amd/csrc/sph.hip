// sph.hip -- WCSPH neighbour passes on the hash grid.
//
// Reference kernels: kernel/KERNEL_rho.py (Shepard, fluid and wall density),
// kernel/KERNEL_pres.py (Tait EOS), kernel/KERNEL_force.py (pressure +
// Morris viscosity + gravity), kernel/KERNEL_step.py (semi-implicit Euler),
// and the SPH-scale pressure-gradient half of Kernel_prep_sphdem
// (kernel/KERNEL_SPHDEM_interaction.py:98-113).
//
// One thread per subject, subjects in grid-sorted order (the reference's
// hash_grid_point_id(grid, tid)), neighbours read from the grid's sorted
// copy (x, y, z, m) in Warp's bucket order, so every per-subject sum adds
// the same terms in the same order as the reference's loop.  Per-step
// working arrays (work.fl_*, work.wl_*) hold densities, pressures and
// velocities in sorted order, so the force pass reads neighbours
// contiguously instead of gathering by particle index.
#include "common.hpp"

#include <limits>

namespace sphdem {

constexpr int kThreads = 256;

// The fast form of the Wendland C2 gradient (kernel/Kernel_KNL.py:26-41)
// for r^2 < (2h)^2: (dW/dr) / r = c' u^3 with u = 1 - r/2h and
// c' = -105 / (16 pi h^5), so no division by r is needed.
template <typename T> struct FastDW {
    T k1, c1;
    __device__ explicit FastDW(T h) {
        const T inv_h = (T)1.0 / h;
        k1 = (T)-0.5 * inv_h;
        c1 = (T)-105.0 / ((T)16.0 * (T)SPHDEM_PI * h * h * h) * inv_h * inv_h;
    }
    __device__ T over_r(T td) const {
        const T u = fma(k1, fast_sqrt(td), (T)1.0);
        return c1 * (u * u * u);
    }
};

// Fluid subjects: Shepard filter and/or density (KERNEL_rho.py:6-82).
//  SHEP: flt_i = sum (m_j/rho_j) W over fluid and wall, stored in P.flt.
//  DENS: rho_i = sum m_j W / flt_i.
//  WORK: the fused step -- new density, Tait pressure, velocity and volume
//        m/rho go to the sorted working arrays (P.rho is still read as the old density by
//        the Shepard sum of other threads, so it is published by the force
//        pass); P.pres is written here.  Without WORK density goes to P.rho.
//  The density keeps the reference's arithmetic in both math modes: the
//  Tait pressure multiplies its relative error by gamma B / p.
template <typename T, bool SHEP, bool DENS, bool WORK>
__global__ __launch_bounds__(kThreads) void k_fluid_density(Prm<T> p, SphV<T> P, BndV<T> B,
                                                            GridV<T> gs, GridV<T> gb,
                                                            WorkV<T> w, int remap) {
    const int k = subject_id(remap);
    if (k >= gs.n) return;
    const V4<T> me = gs.s[k];
    const int i = gs.ids[k];
    const int ik = inst_of(gs, k);
    gs = at_inst(gs, ik);
    gb = at_inst(gb, ik);
    const T sup = p.support, ss = sup * sup, h = p.h;
    const WendlandK<T> wk(h);
    T flt = (T)0, rho = (T)0;
    query_runs(gs, me.x, me.y, me.z, sup, [&](int a, int b) {
        each_slot<4>(a, b, 1, [&](int s) { return gs.s[s]; }, [&](int s, const V4<T> &q) {
            const T dx = me.x - q.x, dy = me.y - q.y, dz = me.z - q.z;
            const T td = dx * dx + dy * dy + dz * dz;
            if (td < ss) {
                const T wv = wendland_w(msqrt(td), wk);
                if (SHEP) flt = flt + (q.w / P.rho[gs.ids[s]]) * wv;
                if (DENS) rho = rho + q.w * wv;
            }
        });
    });
    query_runs(gb, me.x, me.y, me.z, sup, [&](int a, int b) {
        each_slot<4>(a, b, 1, [&](int s) { return gb.s[s]; }, [&](int s, const V4<T> &q) {
            const T dx = me.x - q.x, dy = me.y - q.y, dz = me.z - q.z;
            const T td = dx * dx + dy * dy + dz * dz;
            if (td < ss) {
                const T wv = wendland_w(msqrt(td), wk);
                if (SHEP) flt = flt + (q.w / B.rho[gb.ids[s]]) * wv;
                if (DENS) rho = rho + q.w * wv;
            }
        });
    });
    if (SHEP) P.flt[i] = flt;
    if (DENS) {
        const T r = rho / (SHEP ? flt : P.flt[i]);
        if (WORK) {
            const T pr = tait(r, p.rho0, p.c0, p.gamma);
            w.fl_rho[k] = r;
            w.fl_pres[k] = pr;
            P.pres[i] = pr;
            w.fl_vel[k] = V4<T>{P.vel[3 * i], P.vel[3 * i + 1], P.vel[3 * i + 2], r};
            w.fl_V[k] = me.w / r;
        } else {
            P.rho[i] = r;
        }
    }
}

// Wall subjects: raw density (KERNEL_rho.py:87-126); with WORK also the EOS
// and the sorted copies the force pass reads.  The reference's arithmetic in
// both math modes, as for the fluid density.
//  wall_only (or null): per wall slot, the sum over wall neighbours alone,
//  NaN until first needed.  A subject whose fluid sum is exactly +0 then has
//  rho = 0 + (the same wall terms in the same order), a function of the
//  static wall positions and masses, so it is taken from there (and stored
//  on first use).  Valid while the wall grid, B.pos and B.m are unchanged.
template <typename T, bool WORK, int U>
__global__ __launch_bounds__(kThreads) void k_wall_density(Prm<T> p, SphV<T> P, BndV<T> B,
                                                           GridV<T> gs, GridV<T> gb, WorkV<T> w,
                                                           T *wall_only, const T *fbox,
                                                           int remap) {
    const int k = subject_id(remap);
    if (k >= gb.n) return;
    const V4<T> me = gb.s[k];
    const int bi = gb.ids[k];
    const int ik = inst_of(gb, k);
    gs = at_inst(gs, ik);
    gb = at_inst(gb, ik);
    const T sup = p.support, ss = sup * sup, h = p.h;
    const WendlandK<T> wk(h);
    T rho = (T)0;
    auto acc = [&](const GridV<T> &g) {
        query_runs(g, me.x, me.y, me.z, sup, [&](int a, int b) {
            each_slot<U>(a, b, 1, [&](int s) { return g.s[s]; }, [&](int, const V4<T> &q) {
                const T dx = me.x - q.x, dy = me.y - q.y, dz = me.z - q.z;
                const T td = dx * dx + dy * dy + dz * dz;
                if (td < ss) rho = rho + q.w * wendland_w(msqrt(td), wk);
            });
        });
    };
    // beyond the fluid box by more than the support along some axis, every
    // fluid candidate fails td < ss (|d| >= 1.001 support on that axis, the
    // box bound rounded by ~1e-7 m): the fluid sum is +0 without the query
    bool far = false;
    if (fbox) {
        const T m = sup * (T)1.001;
        far = me.x < fbox[0] - m || me.y < fbox[1] - m || me.z < fbox[2] - m ||
              me.x > fbox[3] + m || me.y > fbox[4] + m || me.z > fbox[5] + m;
    }
    if (!far) acc(gs);
    if (wall_only && rho == (T)0) {
        const T c = wall_only[k];
        if (c == c) {
            rho = c;
        } else {
            acc(gb);
            wall_only[k] = rho;
        }
    } else {
        acc(gb);
    }
    B.rho[bi] = rho;
    if (WORK) {
        const T pr = tait(rho, p.rho0, p.c0, p.gamma);
        B.pres[bi] = pr;
        w.wl_rho[k] = rho;
        w.wl_pres[k] = pr;
        w.wl_vel[k] = V4<T>{B.vel[3 * bi], B.vel[3 * bi + 1], B.vel[3 * bi + 2], rho};
    }
}

template <typename T>
__global__ void k_pres(Prm<T> p, const T *rho, T *pres, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) pres[i] = tait(rho[i], p.rho0, p.c0, p.gamma);
}

// Fluid subjects: acceleration (KERNEL_force.py:29-71) and/or the SPH-scale
// -grad p of Kernel_prep_sphdem (KERNEL_SPHDEM_interaction.py:98-113).  Both
// loops keep 1e-12 < r^2 < support^2 with the same support.
//  PUBLISH: write the sorted density back to P.rho (fused step).
//  PGF also leaves (pgf, rho) and m/rho in sorted order for the coupling.
//  FAST: dW/dr * rij / r is c' u^3 rij with u = 1 - r/2h (FastDW: one
//  hardware square root, no division by r), the divisions by rho_j and
//  by rho_i rho_j (r^2 + 0.01 h^2) are hardware reciprocals, and r^2 and
//  the sums are formed with fused multiply-adds.
template <typename T, bool ACC, bool PGF, bool PUBLISH, bool FAST>
__global__ __launch_bounds__(kThreads) void k_fluid_force(Prm<T> p, SphV<T> P, BndV<T> B,
                                                          GridV<T> gs, GridV<T> gb, WorkV<T> w,
                                                          int remap) {
    const int k = subject_id(remap);
    if (k >= gs.n) return;
    const V4<T> me = gs.s[k];
    const int i = gs.ids[k];
    const int ik = inst_of(gs, k);
    gs = at_inst(gs, ik);
    gb = at_inst(gb, ik);
    const T rhoi = w.fl_rho[k], presi = w.fl_pres[k];
    const V3<T> vi = xyz(w.fl_vel[k]);
    const T sup = p.support, ss = sup * sup, h = p.h, r2min = (T)1.0e-12;
    const T ai = presi / (rhoi * rhoi);
    const T two_mu = (T)2.0 * p.mu, reg = (T)0.01 * h * h;
    const FastDW<T> fdw(h);
    const WendlandK<T> wk(h);
    V3<T> acc = mk3<T>(0, 0, 0), pgf = mk3<T>(0, 0, 0);
    // neighbours' (velocity, density) come in one load from the sorted
    // (vel, rho) records the density passes write
    auto body = [&](const GridV<T> &g, const T *npres, const V4<T> *nvel) {
        query_runs(g, me.x, me.y, me.z, sup, [&](int a, int b) {
            each_slot<4>(a, b, 1, [&](int s) { return g.s[s]; }, [&](int s, const V4<T> &q) {
                const V3<T> rij = mk3(me.x - q.x, me.y - q.y, me.z - q.z);
                if constexpr (FAST) {
                    // fused multiply-adds throughout; grad W = gr rij enters
                    // each sum as one scalar times rij
                    const T td = fma(rij.z, rij.z, fma(rij.y, rij.y, rij.x * rij.x));
                    if (td > r2min && td < ss) {
                        const T gr = fdw.over_r(td);  // dW/dr / r
                        const V4<T> vr = nvel[s];   // (velocity, density)
                        const T rhoj = vr.w, presj = npres[s];
                        const T irho = fast_rcp(rhoj);
                        if (ACC) {
                            const T a = -(q.w * fma(presj * irho, irho, ai)) * gr;
                            acc.x = fma(a, rij.x, acc.x);
                            acc.y = fma(a, rij.y, acc.y);
                            acc.z = fma(a, rij.z, acc.z);
                            const T b = q.w * (two_mu * gr * td) *
                                        fast_rcp(rhoi * rhoj * (td + reg));
                            acc.x = fma(b, vi.x - vr.x, acc.x);
                            acc.y = fma(b, vi.y - vr.y, acc.y);
                            acc.z = fma(b, vi.z - vr.z, acc.z);
                        }
                        if (PGF) {
                            const T c = -((q.w * irho) * (presj - presi)) * gr;
                            pgf.x = fma(c, rij.x, pgf.x);
                            pgf.y = fma(c, rij.y, pgf.y);
                            pgf.z = fma(c, rij.z, pgf.z);
                        }
                    }
                    return;
                }
                const T td = dot(rij, rij);
                if (td > r2min && td < ss) {
                    const T r = msqrt(td);
                    const T dwv = wendland_dw(r, wk);
                    const V4<T> vr = nvel[s];   // (velocity, density)
                    const T rhoj = vr.w, presj = npres[s];
                    if (ACC) {
                        const V3<T> dwij = dwv * (rij / r);
                        acc = acc - (q.w * (ai + presj / (rhoj * rhoj))) * dwij;
                        const T visc = two_mu * dot(rij, dwij) / (rhoi * rhoj * (td + reg));
                        acc = acc + (q.w * visc) * (vi - xyz(vr));
                    }
                    if (PGF) {
                        const V3<T> dwp = (dwv * rij) / r;
                        pgf = pgf - ((q.w / rhoj) * (presj - presi)) * dwp;
                    }
                }
            });
        });
    };
    body(gs, w.fl_pres, w.fl_vel);
    body(gb, w.wl_pres, w.wl_vel);
    if (ACC) st3(P.acc, i, acc + mk3<T>(0, 0, -p.g));
    if (PUBLISH) P.rho[i] = rhoi;
    if (PGF) {
        st3(P.pgf, i, pgf);
        w.fl_pgf[k] = V4<T>{pgf.x, pgf.y, pgf.z, rhoi};
    }
}

// Sorted working copies of the current fluid / wall fields, for the
// single-kernel entry points that start from the particle arrays.
template <typename T>
__global__ void k_gather_fluid(SphV<T> P, GridV<T> gs, WorkV<T> w) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= gs.n) return;
    const int i = gs.ids[k];
    const T r = P.rho[i];
    w.fl_rho[k] = r;
    w.fl_pres[k] = P.pres[i];
    w.fl_vel[k] = V4<T>{P.vel[3 * i], P.vel[3 * i + 1], P.vel[3 * i + 2], r};
    w.fl_V[k] = P.m[i] / r;
    w.fl_pgf[k] = V4<T>{P.pgf[3 * i], P.pgf[3 * i + 1], P.pgf[3 * i + 2], r};
}

template <typename T>
__global__ void k_gather_wall(BndV<T> B, GridV<T> gb, WorkV<T> w) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= gb.n) return;
    const int bi = gb.ids[k];
    w.wl_rho[k] = B.rho[bi];
    w.wl_pres[k] = B.pres[bi];
    w.wl_vel[k] = V4<T>{B.vel[3 * bi], B.vel[3 * bi + 1], B.vel[3 * bi + 2], B.rho[bi]};
}

// KERNEL_step.py:6-22 (tid order; elementwise).
template <typename T>
__global__ void k_step_sph(SphV<T> P, T dt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    const V3<T> v = ld3(P.vel, i) + ld3(P.acc, i) * dt;
    st3(P.vel, i, v);
    st3(P.pos, i, ld3(P.pos, i) + v * dt);
}

template <typename T>
__global__ void k_wendland(const T *r, int n, T h, T *w, T *dw) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const WendlandK<T> wk(h);
    w[i] = wendland_w(r[i], wk);
    dw[i] = wendland_dw(r[i], wk);
}

// ---------------------------------------------------------------- host side
int check_sph_pair(const sphdem_params_t *p, const sphdem_sph_t *P, const sphdem_bnd_t *B,
                   const sphdem_grid_t *gs, const sphdem_grid_t *gb) {
    SPHDEM_REQUIRE(p && P && B && gs && gb, "NULL argument");
    SPHDEM_REQUIRE(p->precision == 4 || p->precision == 8, "precision must be 4 or 8");
    SPHDEM_REQUIRE(gs->precision == p->precision && gb->precision == p->precision,
                   "grid precision differs from the particle precision");
    SPHDEM_REQUIRE(gs->n == P->n, "fluid grid was built from a different number of points");
    SPHDEM_REQUIRE(gb->n == B->n, "wall grid was built from a different number of points");
    SPHDEM_REQUIRE(P->n == 0 || (P->pos && P->vel && P->rho && P->pres && P->acc && P->m && P->flt),
                   "fluid field pointer is NULL");
    SPHDEM_REQUIRE(B->n == 0 || (B->pos && B->vel && B->rho && B->pres && B->m),
                   "wall field pointer is NULL");
    SPHDEM_REQUIRE(gs->n == 0 || (gs->point_ids && gs->sorted), "fluid grid is not built");
    SPHDEM_REQUIRE(gb->n == 0 || (gb->point_ids && gb->sorted), "wall grid is not built");
    return SPHDEM_OK;
}

int check_work(const sphdem_work_t *w, int nf, int nb) {
    SPHDEM_REQUIRE(w != nullptr, "work is NULL");
    SPHDEM_REQUIRE(w->cap_f >= nf && w->cap_b >= nb, "work arrays are smaller than the particle sets");
    SPHDEM_REQUIRE(nf == 0 || (w->fl_rho && w->fl_pres && w->fl_vel && w->fl_V && w->fl_pgf),
                   "fluid work pointer is NULL");
    SPHDEM_REQUIRE(nb == 0 || (w->wl_rho && w->wl_pres && w->wl_vel), "wall work pointer is NULL");
    return SPHDEM_OK;
}

// The bounding box of points s[0, n) (x, y, z of each V4): 256 partial
// boxes from a grid-stride pass, then one workgroup folds them into
// box[1536 .. 1541] = (min x, y, z, max x, y, z).  min/max are exact, so
// the box does not depend on the launch shape.
constexpr int kBoxParts = 256;
constexpr int kBoxMinWalls = 1 << 22;
template <typename T>
__global__ __launch_bounds__(256) void k_box_parts(const V4<T> *s, int n, T *box) {
    __shared__ T red[6][4];
    const T big = std::numeric_limits<T>::max();
    T v[6] = {big, big, big, -big, -big, -big};
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += kBoxParts * 256) {
        const V4<T> q = s[i];
        v[0] = q.x < v[0] ? q.x : v[0];
        v[1] = q.y < v[1] ? q.y : v[1];
        v[2] = q.z < v[2] ? q.z : v[2];
        v[3] = q.x > v[3] ? q.x : v[3];
        v[4] = q.y > v[4] ? q.y : v[4];
        v[5] = q.z > v[5] ? q.z : v[5];
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        for (int off = 32; off > 0; off >>= 1) {
            const T o = __shfl_xor(v[c], off, 64);
            v[c] = c < 3 ? (o < v[c] ? o : v[c]) : (o > v[c] ? o : v[c]);
        }
        if ((threadIdx.x & 63) == 0) red[c][threadIdx.x >> 6] = v[c];
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int c = threadIdx.x;
        T r = red[c][0];
        for (int k = 1; k < 4; ++k) r = c < 3 ? (red[c][k] < r ? red[c][k] : r)
                                              : (red[c][k] > r ? red[c][k] : r);
        box[6 * blockIdx.x + c] = r;
    }
}
template <typename T> __global__ __launch_bounds__(64) void k_box_fold(T *box) {
    const int c = threadIdx.x;
    if (c >= 6) return;
    T r = box[c];
    for (int k = 1; k < kBoxParts; ++k) {
        const T o = box[6 * k + c];
        r = c < 3 ? (o < r ? o : r) : (o > r ? o : r);
    }
    box[6 * kBoxParts + c] = r;
}

template <typename T, bool SHEP, bool DENS, bool WORK>
static void launch_fluid_density(const sphdem_params_t *p, const sphdem_sph_t *P,
                                 const sphdem_bnd_t *B, const sphdem_grid_t *gs,
                                 const sphdem_grid_t *gb, const sphdem_work_t *w, hipStream_t st) {
    if (gs->n == 0) return;
    ProfScope prof(SPHDEM_PASS_FLUID_DENSITY, st);
    WorkV<T> wv{};
    if (w) wv = view<T>(w);
    const int bt = tuned_block(SPHDEM_PASS_FLUID_DENSITY);
    hipLaunchKernelGGL((k_fluid_density<T, SHEP, DENS, WORK>), dim3(blocks_for(gs->n, bt)),
                       dim3(bt), 0, st, prm<T>(p), view<T>(P), view<T>(B), view<T>(gs),
                       view<T>(gb), wv, tuned_remap(SPHDEM_PASS_FLUID_DENSITY));
}

template <typename T, bool WORK>
static void launch_wall_density(const sphdem_params_t *p, const sphdem_sph_t *P,
                                const sphdem_bnd_t *B, const sphdem_grid_t *gs,
                                const sphdem_grid_t *gb, const sphdem_work_t *w, hipStream_t st) {
    if (gb->n == 0) return;
    ProfScope prof(SPHDEM_PASS_WALL_DENSITY, st);
    WorkV<T> wv{};
    if (w) wv = view<T>(w);
    const int bt = tuned_block(SPHDEM_PASS_WALL_DENSITY);
    T *wall_only = (w && w->wl_static && w->cap_b >= gb->n) ? (T *)w->wl_static : nullptr;
    // the two box launches cost ~15 us; they pay for themselves on large
    // wall sets (11.5M walls: wall pass 1.33 -> 1.09 ms) but not on 1M
    // (no change) or 50k (+0.015 ms)
    const T *fbox = nullptr;
    if (w && w->fl_box && gs->n > 0 && gs->sorted && gb->n >= kBoxMinWalls) {
        T *box = (T *)w->fl_box;
        hipLaunchKernelGGL(k_box_parts<T>, dim3(kBoxParts), dim3(256), 0, st,
                           (const V4<T> *)gs->sorted, gs->n, box);
        hipLaunchKernelGGL(k_box_fold<T>, dim3(1), dim3(64), 0, st, box);
        fbox = box + 6 * kBoxParts;
    }
    // below that size the pass is set by the chains of the walls near the
    // fluid, which are few waves: eight candidate loads in flight instead of
    // four (1.08M walls: 0.615 -> 0.540 ms; 11.5M walls: 1.094 -> 1.113 ms)
    if (gb->n >= kBoxMinWalls)
        hipLaunchKernelGGL((k_wall_density<T, WORK, 4>), dim3(blocks_for(gb->n, bt)), dim3(bt), 0,
                           st, prm<T>(p), view<T>(P), view<T>(B), view<T>(gs), view<T>(gb), wv,
                           wall_only, fbox, tuned_remap(SPHDEM_PASS_WALL_DENSITY));
    else
        hipLaunchKernelGGL((k_wall_density<T, WORK, 8>), dim3(blocks_for(gb->n, bt)), dim3(bt), 0,
                           st, prm<T>(p), view<T>(P), view<T>(B), view<T>(gs), view<T>(gb), wv,
                           wall_only, fbox, tuned_remap(SPHDEM_PASS_WALL_DENSITY));
}

template <typename T, bool ACC, bool PGF, bool PUBLISH>
static void launch_fluid_force(const sphdem_params_t *p, const sphdem_sph_t *P,
                               const sphdem_bnd_t *B, const sphdem_grid_t *gs,
                               const sphdem_grid_t *gb, const sphdem_work_t *w, hipStream_t st) {
    if (gs->n == 0) return;
    ProfScope prof(SPHDEM_PASS_FLUID_FORCE, st);
    const int bt = tuned_block(SPHDEM_PASS_FLUID_FORCE);
    const dim3 grid(blocks_for(gs->n, bt));
    if (p->fast_math)
        hipLaunchKernelGGL((k_fluid_force<T, ACC, PGF, PUBLISH, true>), grid, dim3(bt), 0, st,
                           prm<T>(p), view<T>(P), view<T>(B), view<T>(gs), view<T>(gb),
                           view<T>(w), tuned_remap(SPHDEM_PASS_FLUID_FORCE));
    else
        hipLaunchKernelGGL((k_fluid_force<T, ACC, PGF, PUBLISH, false>), grid, dim3(bt), 0, st,
                           prm<T>(p), view<T>(P), view<T>(B), view<T>(gs), view<T>(gb),
                           view<T>(w), tuned_remap(SPHDEM_PASS_FLUID_FORCE));
}

template <typename T>
void launch_gather(const sphdem_sph_t *P, const sphdem_bnd_t *B, const sphdem_grid_t *gs,
                   const sphdem_grid_t *gb, const sphdem_work_t *w, hipStream_t st) {
    ProfScope prof(SPHDEM_PASS_OTHER, st);
    if (P && gs && gs->n > 0)
        hipLaunchKernelGGL(k_gather_fluid<T>, dim3(blocks_for(gs->n, kThreads)), dim3(kThreads), 0,
                           st, view<T>(P), view<T>(gs), view<T>(w));
    if (B && gb && gb->n > 0)
        hipLaunchKernelGGL(k_gather_wall<T>, dim3(blocks_for(gb->n, kThreads)), dim3(kThreads), 0,
                           st, view<T>(B), view<T>(gb), view<T>(w));
}
template void launch_gather<float>(const sphdem_sph_t *, const sphdem_bnd_t *,
                                   const sphdem_grid_t *, const sphdem_grid_t *,
                                   const sphdem_work_t *, hipStream_t);
template void launch_gather<double>(const sphdem_sph_t *, const sphdem_bnd_t *,
                                    const sphdem_grid_t *, const sphdem_grid_t *,
                                    const sphdem_work_t *, hipStream_t);

template <typename T>
void launch_step_sph(const sphdem_params_t *p, const sphdem_sph_t *P, hipStream_t st) {
    if (P->n == 0) return;
    ProfScope prof(SPHDEM_PASS_FLUID_FINAL, st);
    hipLaunchKernelGGL(k_step_sph<T>, dim3(blocks_for(P->n, kThreads)), dim3(kThreads), 0, st,
                       view<T>(P), (T)p->dt);
}

// SPH part of one step (Simulation.py:17-91), fused: P1 fluid Shepard +
// density + EOS, P1b wall density + EOS, P2 force (+ -grad p when coupled).
// `beside` (stream and two events, or nulls): off a Shepard step the wall
// pass runs on that stream beside the fluid pass -- the fluid density then
// reads positions and masses only, the wall pass writes only wall fields --
// and the force pass waits for both.  On a Shepard step the fluid pass reads
// the old wall densities, so the wall pass follows it on `st`.  `force`
// (the same, or nulls): the force pass runs on its stream after both density
// passes, and the caller makes every reader of the force's outputs wait for
// `force.join` (recorded whenever `force.s` is set, also on an error).
template <typename T>
int sph_stage(const sphdem_params_t *p, const sphdem_sph_t *P, const sphdem_bnd_t *B,
              const sphdem_grid_t *gs, const sphdem_grid_t *gb, int step, bool coupled,
              const sphdem_work_t *w, hipStream_t st, const Beside &beside,
              const Beside &force) {
    const bool shep = step == 0 || (p->shepard_step > 0 && step % p->shepard_step == 0);
    const bool apart = !shep && beside.s && beside.fork && beside.join;
    if (apart) {
        (void)hipEventRecord(beside.fork, st);
        (void)hipStreamWaitEvent(beside.s, beside.fork, 0);
        launch_wall_density<T, true>(p, P, B, gs, gb, w, beside.s);
        (void)hipEventRecord(beside.join, beside.s);
    }
    if (shep)
        launch_fluid_density<T, true, true, true>(p, P, B, gs, gb, w, st);
    else
        launch_fluid_density<T, false, true, true>(p, P, B, gs, gb, w, st);
    if (apart)
        (void)hipStreamWaitEvent(st, beside.join, 0);
    else
        launch_wall_density<T, true>(p, P, B, gs, gb, w, st);
    hipStream_t fs = st;
    if (force.s) {
        (void)hipEventRecord(force.fork, st);
        (void)hipStreamWaitEvent(force.s, force.fork, 0);
        fs = force.s;
    }
    if (coupled)
        launch_fluid_force<T, true, true, true>(p, P, B, gs, gb, w, fs);
    else
        launch_fluid_force<T, true, false, true>(p, P, B, gs, gb, w, fs);
    if (force.s) (void)hipEventRecord(force.join, force.s);
    return check_launch("sph_stage");
}
template int sph_stage<float>(const sphdem_params_t *, const sphdem_sph_t *, const sphdem_bnd_t *,
                              const sphdem_grid_t *, const sphdem_grid_t *, int, bool,
                              const sphdem_work_t *, hipStream_t, const Beside &,
                              const Beside &);
template int sph_stage<double>(const sphdem_params_t *, const sphdem_sph_t *, const sphdem_bnd_t *,
                               const sphdem_grid_t *, const sphdem_grid_t *, int, bool,
                               const sphdem_work_t *, hipStream_t, const Beside &,
                              const Beside &);

// SPH-scale -grad p alone, from the particle arrays (prep_sphdem entry).
template <typename T>
void launch_pgf_only(const sphdem_params_t *p, const sphdem_sph_t *P, const sphdem_bnd_t *B,
                     const sphdem_grid_t *gs, const sphdem_grid_t *gb, const sphdem_work_t *w,
                     hipStream_t st) {
    launch_gather<T>(P, B, gs, gb, w, st);
    launch_fluid_force<T, false, true, false>(p, P, B, gs, gb, w, st);
}
template void launch_pgf_only<float>(const sphdem_params_t *, const sphdem_sph_t *,
                                     const sphdem_bnd_t *, const sphdem_grid_t *,
                                     const sphdem_grid_t *, const sphdem_work_t *, hipStream_t);
template void launch_pgf_only<double>(const sphdem_params_t *, const sphdem_sph_t *,
                                      const sphdem_bnd_t *, const sphdem_grid_t *,
                                      const sphdem_grid_t *, const sphdem_work_t *, hipStream_t);

}  // namespace sphdem

using namespace sphdem;

#define DISPATCH(prec, CALL)                 \
    do {                                     \
        if ((prec) == 4) {                   \
            using T = float;                 \
            CALL;                            \
        } else {                             \
            using T = double;                \
            CALL;                            \
        }                                    \
    } while (0)

extern "C" {

int sphdem_wendland(int precision, const void *r, int n, double h, void *w, void *dw,
                    void *stream) {
    SPHDEM_REQUIRE(precision == 4 || precision == 8, "precision must be 4 or 8");
    SPHDEM_REQUIRE(n >= 0 && (n == 0 || (r && w && dw)), "invalid Wendland buffers");
    if (n == 0) return SPHDEM_OK;
    DISPATCH(precision, hipLaunchKernelGGL(k_wendland<T>, dim3(blocks_for(n, kThreads)),
                                           dim3(kThreads), 0, as_stream(stream), (const T *)r, n,
                                           (T)h, (T *)w, (T *)dw));
    return check_launch("wendland");
}

int sphdem_shepard_sph(const sphdem_params_t *p, sphdem_sph_t *P, sphdem_bnd_t *B,
                       const sphdem_grid_t *gs, const sphdem_grid_t *gb, void *stream) {
    int rc = check_sph_pair(p, P, B, gs, gb);
    if (rc) return rc;
    DISPATCH(p->precision, (launch_fluid_density<T, true, false, false>(p, P, B, gs, gb, nullptr,
                                                                         as_stream(stream))));
    return check_launch("shepard_sph");
}

int sphdem_density_sph(const sphdem_params_t *p, sphdem_sph_t *P, sphdem_bnd_t *B,
                       const sphdem_grid_t *gs, const sphdem_grid_t *gb, void *stream) {
    int rc = check_sph_pair(p, P, B, gs, gb);
    if (rc) return rc;
    DISPATCH(p->precision, (launch_fluid_density<T, false, true, false>(p, P, B, gs, gb, nullptr,
                                                                         as_stream(stream))));
    return check_launch("density_sph");
}

int sphdem_density_bnd(const sphdem_params_t *p, sphdem_sph_t *P, sphdem_bnd_t *B,
                       const sphdem_grid_t *gs, const sphdem_grid_t *gb, void *stream) {
    int rc = check_sph_pair(p, P, B, gs, gb);
    if (rc) return rc;
    DISPATCH(p->precision,
             (launch_wall_density<T, false>(p, P, B, gs, gb, nullptr, as_stream(stream))));
    return check_launch("density_bnd");
}

int sphdem_pres(const sphdem_params_t *p, const void *rho, void *pres, int n, void *stream) {
    SPHDEM_REQUIRE(p && (p->precision == 4 || p->precision == 8), "invalid params");
    SPHDEM_REQUIRE(n >= 0 && (n == 0 || (rho && pres)), "invalid pressure buffers");
    if (n == 0) return SPHDEM_OK;
    DISPATCH(p->precision, hipLaunchKernelGGL(k_pres<T>, dim3(blocks_for(n, kThreads)),
                                              dim3(kThreads), 0, as_stream(stream), prm<T>(p),
                                              (const T *)rho, (T *)pres, n));
    return check_launch("pres");
}

int sphdem_force_sph(const sphdem_params_t *p, sphdem_sph_t *P, sphdem_bnd_t *B,
                     const sphdem_grid_t *gs, const sphdem_grid_t *gb, const sphdem_work_t *w,
                     void *stream) {
    int rc = check_sph_pair(p, P, B, gs, gb);
    if (rc) return rc;
    rc = check_work(w, P->n, B->n);
    if (rc) return rc;
    hipStream_t st = as_stream(stream);
    DISPATCH(p->precision, (launch_gather<T>(P, B, gs, gb, w, st),
                            launch_fluid_force<T, true, false, false>(p, P, B, gs, gb, w, st)));
    return check_launch("force_sph");
}

int sphdem_step_sph(const sphdem_params_t *p, sphdem_sph_t *P, void *stream) {
    SPHDEM_REQUIRE(p && P && (p->precision == 4 || p->precision == 8), "invalid arguments");
    SPHDEM_REQUIRE(P->n == 0 || (P->pos && P->vel && P->acc), "fluid field pointer is NULL");
    DISPATCH(p->precision, launch_step_sph<T>(p, P, as_stream(stream)));
    return check_launch("step_sph");
}

int sphdem_sph_step(const sphdem_params_t *p, sphdem_sph_t *P, sphdem_bnd_t *B,
                    const sphdem_grid_t *gs, const sphdem_grid_t *gb, int step, int integrate,
                    const sphdem_work_t *w, void *stream) {
    int rc = check_sph_pair(p, P, B, gs, gb);
    if (rc) return rc;
    rc = check_work(w, P->n, B->n);
    if (rc) return rc;
    SPHDEM_REQUIRE(step >= 0, "step must be nonnegative");
    hipStream_t st = as_stream(stream);
    DISPATCH(p->precision, rc = sph_stage<T>(p, P, B, gs, gb, step, false, w, st, Beside{}, Beside{}));
    if (rc) return rc;
    if (integrate) DISPATCH(p->precision, launch_step_sph<T>(p, P, st));
    return check_launch("sph_step");
}

}  // extern "C"
