// dem.hip -- DEM contacts with CSR tangential history, SPH<->DEM coupling,
// the fused coupled step, and FP64 diagnostics.
//
// Reference kernels: kernel/KERNEL_DEM_force.py, kernel/KERNEL_DEM_BC.py,
// kernel/KERNEL_DEM_step.py, kernel/KERNEL_SPHDEM_interaction.py;
// orchestration source/Simulation.py:102-221.
#include "common.hpp"

namespace sphdem {

constexpr int kThreads = 256;
// Threads of the one-workgroup-per-DEM fluid search: few DEMs have fluid
// near, and each of those walks ~2e5 candidates, so the search of one DEM
// wants many waves in flight.
constexpr int kDemThreads = 1024;

// Declared in sph.hip.
template <typename T>
void launch_gather(const sphdem_sph_t *P, const sphdem_bnd_t *B, const sphdem_grid_t *gs,
                   const sphdem_grid_t *gb, const sphdem_work_t *w, hipStream_t st);
template <typename T>
void launch_step_sph(const sphdem_params_t *p, const sphdem_sph_t *P, hipStream_t st);
template <typename T>
int sph_stage(const sphdem_params_t *p, const sphdem_sph_t *P, const sphdem_bnd_t *B,
              const sphdem_grid_t *gs, const sphdem_grid_t *gb, int step, bool coupled,
              const sphdem_work_t *w, hipStream_t st, const Beside &beside,
              const Beside &force);
template <typename T>
void launch_pgf_only(const sphdem_params_t *p, const sphdem_sph_t *P, const sphdem_bnd_t *B,
                     const sphdem_grid_t *gs, const sphdem_grid_t *gb, const sphdem_work_t *w,
                     hipStream_t st);
int check_sph_pair(const sphdem_params_t *p, const sphdem_sph_t *P, const sphdem_bnd_t *B,
                   const sphdem_grid_t *gs, const sphdem_grid_t *gb);
int check_work(const sphdem_work_t *w, int nf, int nb);
// porosity.hip: mark / compact / LDS-tiled coupling-scale sums (uniform h).
template <typename T>
int launch_porosity_tiled(const sphdem_params_t *p, const sphdem_sph_t *P, const sphdem_dem_t *D,
                          const sphdem_grid_t *gs, const sphdem_grid_t *gd,
                          const sphdem_work_t *w, double csup, double hu, int *dem_near,
                          hipStream_t st, hipEvent_t marked);

// ---------------------------------------------------------------- contact law
// kernel/KERNEL_DEM_force.py:25-74 DEM_contact: linear spring-dashpot with
// unprojected tangential history and a componentwise Coulomb limit.
template <typename T>
__device__ __forceinline__ void dem_contact(V3<T> normal, T overlap, T ra, T rb, V3<T> va,
                                            V3<T> vb, V3<T> wa, V3<T> wb, V3<T> tangent, T K,
                                            T eta, T mu, T dt, V3<T> &f, V3<T> &disp,
                                            V3<T> &tq) {
    const V3<T> rc_a = (ra - (T)0.5 * overlap) * normal;
    const V3<T> rc_b = (-(rb - (T)0.5 * overlap)) * normal;
    const V3<T> urel = ((vb + cross(wb, rc_b)) - va) - cross(wa, rc_a);
    const V3<T> un = dot(urel, normal) * normal;
    const V3<T> us = urel - un;
    disp = tangent + us * dt;
    const V3<T> dir = us / (length(us) + (T)1.0e-20);
    const V3<T> fn = (-K * overlap) * normal + eta * un;
    const V3<T> fss = K * disp + eta * us;
    const V3<T> fsf = (mu * K * overlap) * dir;
    V3<T> fs = fss;
    if (length(fss) >= length(fsf))
        fs = mk3(sgn(fss.x) * mabs(fsf.x), sgn(fss.y) * mabs(fsf.y), sgn(fss.z) * mabs(fsf.z));
    f = fn + fs;
    tq = cross(rc_a, fs);
}

// KERNEL_DEM_force.py:9-22 and KERNEL_DEM_BC.py:10-24: overlap counts per row.
// WALL selects the fixed-sphere set (no self exclusion).
template <typename T, bool WALL>
__global__ void k_count_contacts(DemV<T> D, GridV<T> g, const int *dinst, const T *nb_radius,
                                 T rmax, int *off) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a == 0) off[D.n] = 0;
    if (a >= D.n) return;
    if (dinst) g = at_inst(g, dinst[a]);
    const V3<T> ra = ld3(D.pos, a);
    const T Ra = D.radius[a];
    int count = 0;
    query_runs(g, ra.x, ra.y, ra.z, Ra + rmax, [&](int lo, int hi) {
        for (int s = lo; s < hi; ++s) {
            const int b = g.ids[s];
            if (!WALL && b == a) continue;
            const V4<T> q = g.s[s];
            const T dist = length(mk3(q.x - ra.x, q.y - ra.y, q.z - ra.z));
            if (Ra + nb_radius[b] - dist > (T)0) ++count;
        }
    });
    off[a] = count;
}

// KERNEL_DEM_force.py:77-135 (WALL = false) and KERNEL_DEM_BC.py:27-84
// (WALL = true): loads plus the new CSR row; history looked up by stable id
// in the old row (last match wins).  Rows past `cap` set flags bit 0.
template <typename T, bool WALL>
__global__ void k_contact_force(DemV<T> D, GridV<T> g, const int *dinst, const T *nb_radius,
                                const T *nb_vel, const T *nb_omega, T rmax, T grav, T dt,
                                int *flags) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= D.n) return;
    if (dinst) g = at_inst(g, dinst[a]);
    const int *off_old = WALL ? D.bnd_off_old : D.dem_off_old;
    const int *id_old = WALL ? D.bnd_id_old : D.dem_id_old;
    const T *tang_old = WALL ? D.bnd_tang_old : D.dem_tang_old;
    const int *off_new = WALL ? D.bnd_off_new : D.dem_off_new;
    int *id_new = WALL ? D.bnd_id_new : D.dem_id_new;
    T *tang_new = WALL ? D.bnd_tang_new : D.dem_tang_new;
    const int cap = WALL ? D.cap_bnd_new : D.cap_dem_new;
    const V3<T> ra = ld3(D.pos, a);
    const T Ra = D.radius[a];
    const V3<T> va = ld3(D.vel, a), wa = ld3(D.omega, a);
    const T K = D.K[a], eta = D.eta[a], mu = D.mu[a];
    V3<T> force = WALL ? mk3<T>(0, 0, 0) : mk3<T>(0, 0, -D.m[a] * grav);
    V3<T> torque = mk3<T>(0, 0, 0);
    int out = off_new[a];
    const int o0 = off_old[a], o1 = off_old[a + 1];
    query_runs(g, ra.x, ra.y, ra.z, Ra + rmax, [&](int lo, int hi) {
        for (int s = lo; s < hi; ++s) {
            const int b = g.ids[s];
            if (!WALL && b == a) continue;
            const V4<T> q = g.s[s];
            const V3<T> rab = mk3(q.x - ra.x, q.y - ra.y, q.z - ra.z);
            const T dist = length(rab);
            const T overlap = Ra + nb_radius[b] - dist;
            if (overlap > (T)0) {
                V3<T> tangent = mk3<T>(0, 0, 0);
                for (int k = o0; k < o1; ++k)
                    if (id_old[k] == b) tangent = ld3(tang_old, k);
                V3<T> normal = WALL ? mk3<T>(0, 0, -1) : mk3<T>(1, 0, 0);
                if (!WALL && a > b) normal = -normal;
                if (dist > (T)1.0e-12) normal = rab / dist;
                V3<T> f, disp, tq;
                dem_contact(normal, overlap, Ra, nb_radius[b], va, ld3(nb_vel, b), wa,
                            ld3(nb_omega, b), tangent, K, eta, mu, dt, f, disp, tq);
                force = force + f;
                torque = torque + tq;
                if (out < cap) {
                    id_new[out] = b;
                    st3(tang_new, out, disp);
                } else {
                    atomicOr(flags, 1);
                }
                ++out;
            }
        }
    });
    if (WALL) {
        st3(D.force, a, ld3(D.force, a) + force);
        st3(D.torque, a, ld3(D.torque, a) + torque);
    } else {
        st3(D.force, a, force);
        st3(D.torque, a, torque);
    }
}

// KERNEL_DEM_step.py:6-30
template <typename T>
__global__ void k_step_dem(DemV<T> D, T dt) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= D.n) return;
    const V3<T> acc = ld3(D.force, a) / D.m[a];
    const V3<T> vel = ld3(D.vel, a) + acc * dt;
    st3(D.acc, a, acc);
    st3(D.vel, a, vel);
    st3(D.pos, a, ld3(D.pos, a) + vel * dt);
    st3(D.omega, a, ld3(D.omega, a) + (ld3(D.torque, a) / D.inertia[a]) * dt);
}

// Copy a finished CSR (offsets n+1, E = off[n] edges) into the read side:
// the no-round-trip step keeps buffer addresses fixed instead of swapping.
// Offsets are clamped to the capacity, so after an overflow (flags bit 0,
// which the caller must act on) the next step's history lookups stay inside
// the cap entries the read side holds.
template <typename T>
__global__ void k_copy_csr(const int *off_s, const int *id_s, const T *tang_s, int *off_d,
                           int *id_d, T *tang_d, int n, int cap) {
    const int E = off_s[n];
    const int lim = E < cap ? E : cap;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < (n + 1 > cap ? n + 1 : cap);
         k += gridDim.x * blockDim.x) {
        if (k <= n) off_d[k] = min(off_s[k], cap);
        if (k < lim) {
            id_d[k] = id_s[k];
            tang_d[3 * k] = tang_s[3 * k];
            tang_d[3 * k + 1] = tang_s[3 * k + 1];
            tang_d[3 * k + 2] = tang_s[3 * k + 2];
        }
    }
}

// ---------------------------------------------------------------- coupling
// kernel/KERNEL_SPHDEM_interaction.py:11-40
template <typename T>
__device__ __forceinline__ T drag_beta(T eps, T rho, T mu, T d, T speed) {
    T beta = (T)0;
    const T re = rho * d * speed / (mu + (T)1.0e-20);
    if (eps <= (T)0.8) {
        beta = (T)150.0 * ((T)1.0 - eps) / eps * mu / (d * d) + (T)1.75 * rho * speed / d;
    } else {
        T cd_speed = (T)0.44 * speed;
        if (re <= (T)1000.0)
            cd_speed = (T)24.0 * mu / (rho * d) * ((T)1.0 + (T)0.15 * mpow(re, (T)0.687));
        beta = (T)0.75 * cd_speed * rho / d * mpow(eps, (T)-1.65);
    }
    return beta;
}

template <typename T>
__global__ void k_drag_beta(const T *eps, const T *rho, const T *mu, const T *d, const T *speed,
                            int n, T *beta) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) beta[i] = drag_beta(eps[i], rho[i], mu[i], d[i], speed[i]);
}

// Porosity half of Kernel_prep_sphdem (KERNEL_SPHDEM_interaction.py:74-96,
// :111).  A subject with no DEM within 2*h_b has solid = 0, so its porosity
// is clamp(1) whatever its coupling-scale fluid sum is; that sum is a local
// the reference never stores, and it is only evaluated where a DEM is near.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_porosity(Prm<T> p, SphV<T> P, DemV<T> D,
                                                       GridV<T> gs, GridV<T> gd, WorkV<T> w,
                                                       T csup, T hu) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= gs.n) return;
    const V4<T> me = gs.s[k];
    const int i = gs.ids[k];
    const int ik = inst_of(gs, k);
    gs = at_inst(gs, ik);
    gd = at_inst(gd, ik);
    bool near = false;
    query_runs(gd, me.x, me.y, me.z, csup, [&](int lo, int hi) {
        for (int s = lo; s < hi; ++s) {
            const V4<T> q = gd.s[s];
            const T dist = length(mk3(me.x - q.x, me.y - q.y, me.z - q.z));
            if (dist < (T)2.0 * D.h[gd.ids[s]]) near = true;
        }
    });
    T solid = (T)0;
    if (near) {
        T flt = (T)0;
        if (hu > (T)0) {
            query_runs(gs, me.x, me.y, me.z, csup, [&](int lo, int hi) {
                for (int s = lo; s < hi; ++s) {
                    const V4<T> q = gs.s[s];
                    const T dist = length(mk3(me.x - q.x, me.y - q.y, me.z - q.z));
                    if (dist < csup) flt = flt + w.fl_V[s] * wendland_w(dist, hu);
                }
            });
        }
        query_runs(gd, me.x, me.y, me.z, csup, [&](int lo, int hi) {
            for (int s = lo; s < hi; ++s) {
                const int b = gd.ids[s];
                const V4<T> q = gd.s[s];
                const T hb = D.h[b];
                const T dist = length(mk3(me.x - q.x, me.y - q.y, me.z - q.z));
                if (dist < (T)2.0 * hb) {
                    T lflt = flt;
                    if (hu <= (T)0) {
                        lflt = (T)0;
                        const T r2 = (T)2.0 * hb;
                        query_runs(gs, me.x, me.y, me.z, r2, [&](int lo2, int hi2) {
                            for (int s2 = lo2; s2 < hi2; ++s2) {
                                const V4<T> q2 = gs.s[s2];
                                const T fd = length(mk3(me.x - q2.x, me.y - q2.y, me.z - q2.z));
                                if (fd < r2) lflt = lflt + w.fl_V[s2] * wendland_w(fd, hb);
                            }
                        });
                    }
                    solid = solid + D.volume[b] * wendland_w(dist, hb) / (lflt + (T)1.0e-20);
                }
            }
        });
    }
    P.porosity[i] = clampr((T)1.0 - solid, p.eps_min, p.eps_max);
}

// kernel/KERNEL_SPHDEM_interaction.py:118-187, one workgroup per DEM
// subject: the fluid sums are split over the workgroup's threads and added
// in a fixed order (wave shuffles, then the waves in turn).
// With `near` (the porosity mark's per-DEM flags of this step): a DEM whose
// flag is 0 has no fluid point within 2 h, since the mark tested exactly
// this distance between the same stored positions (the DEM's grid copy is
// checked to equal D.pos), so its search would add no term; it gets the
// no-fluid outputs at once.  The flags read are cleared for the next step.
template <typename T>
__global__ __launch_bounds__(kDemThreads) void k_interaction_dem(Prm<T> p, DemV<T> D, GridV<T> gs,
                                                              GridV<T> gd, WorkV<T> w,
                                                              int *near) {
    const int t = blockIdx.x;
    if (t >= gd.n) return;
    const int a = gd.ids[t];
    const int ia = inst_of(gd, a);
    gs = at_inst(gs, ia);
    gd = at_inst(gd, ia);
    const V3<T> ra = ld3(D.pos, a);
    const T h = D.h[a], sup = (T)2.0 * h;
    if (near) {
        const V4<T> own = gd.s[t];
        if (near[a] == 0 && own.x == ra.x && own.y == ra.y && own.z == ra.z) {
            if (threadIdx.x == 0) {
                D.flt_s[a] = (T)0;
                st3(D.drag, a, mk3<T>(0, 0, 0));
                st3(D.pressure_force, a, mk3<T>(0, 0, 0));
                D.porosity[a] = p.eps_max;
            }
            return;
        }
    }
    T acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // flt, rho_f, vel_f xyz, pgf xyz
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    // Each wave takes whole rows of the query box (wave w: rows w, w + 16,
    // ...), its lanes striding each row's slots, so the rows' bucket-bound
    // loads overlap instead of forming one chain.
    const QueryBox qb = query_box(gs, ra.x, ra.y, ra.z, sup);
    const int ny = qb.y1 - qb.y0 + 1, rows = gs.n ? (qb.z1 - qb.z0 + 1) * ny : 0;
    for (int r = wid; r < rows; r += kDemThreads / 64) {
        row_runs(gs, qb, qb.y0 + r % ny, qb.z0 + r / ny, [&](int lo, int hi) {
            for (int s = lo + lane; s < hi; s += 64) {
                const V4<T> q = gs.s[s];
                const T dist = length(mk3(ra.x - q.x, ra.y - q.y, ra.z - q.z));
                if (dist < sup) {
                    const T vw = w.fl_V[s] * wendland_w(dist, h);
                    const V4<T> pr = w.fl_pgf[s];
                    const V4<T> v = w.fl_vel[s];
                    acc[0] = acc[0] + vw;
                    acc[1] = acc[1] + pr.w * vw;
                    acc[2] = acc[2] + v.x * vw;
                    acc[3] = acc[3] + v.y * vw;
                    acc[4] = acc[4] + v.z * vw;
                    acc[5] = acc[5] + pr.x * vw;
                    acc[6] = acc[6] + pr.y * vw;
                    acc[7] = acc[7] + pr.z * vw;
                }
            }
        });
    }
    __shared__ T red[kDemThreads / 64][8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const T v = wave_sum(acc[c]);
        if (lane == 0) red[wid][c] = v;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    T tot[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        T s = red[0][c];
        for (int q = 1; q < kDemThreads / 64; ++q) s = s + red[q][c];
        tot[c] = s;
    }
    const T flt = tot[0], rho_f = tot[1];
    const V3<T> vel_f = mk3(tot[2], tot[3], tot[4]), pgf = mk3(tot[5], tot[6], tot[7]);
    T solid = (T)0;
    query_runs(gd, ra.x, ra.y, ra.z, sup, [&](int lo, int hi) {
        for (int s = lo; s < hi; ++s) {
            const V4<T> q = gd.s[s];
            const T dist = length(mk3(ra.x - q.x, ra.y - q.y, ra.z - q.z));
            if (dist < sup) solid = solid + D.volume[gd.ids[s]] * wendland_w(dist, h);
        }
    });
    if (near) near[a] = 0;
    D.flt_s[a] = flt;
    st3(D.drag, a, mk3<T>(0, 0, 0));
    st3(D.pressure_force, a, mk3<T>(0, 0, 0));
    D.porosity[a] = p.eps_max;
    if (flt > (T)1.0e-12) {
        const T eps = clampr((T)1.0 - solid / flt, p.eps_min, p.eps_max);
        const V3<T> urel = vel_f / flt - ld3(D.vel, a);
        const T beta = drag_beta(eps, rho_f / flt, p.mu, (T)2.0 * D.radius[a], length(urel));
        const T coef = beta * D.volume[a] / D.m[a];
        const V3<T> drag = (D.m[a] * coef / ((T)1.0 + p.dt * coef)) * urel;
        const V3<T> pf = ((D.m[a] / D.rho[a]) * pgf) / flt;
        D.porosity[a] = eps;
        st3(D.drag, a, drag);
        st3(D.pressure_force, a, pf);
        st3(D.force, a, (ld3(D.force, a) + drag) + pf);
    }
}

// kernel/KERNEL_SPHDEM_interaction.py:190-226 (INTER) and
// kernel/KERNEL_step.py:6-22 (STEP): both only touch the subject's own
// fields.  Two launches: the drag reaction of the near-DEM subjects in
// sorted order, whose lanes are neighbours and walk the same DEM buckets,
// leaving acc_dem; then the update in original particle order, whose reads
// and writes of the particle arrays are contiguous.
// near (or null): the porosity mark's near-DEM flag of each sorted fluid
// slot, from this step's positions.  The mark ran this very DEM query and
// tested the same distance, so an unflagged subject has no term here: its
// fd is +0 and so is fd / (rho eps), which the update uses without a query.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_fluid_drag(Prm<T> p, SphV<T> P, DemV<T> D,
                                                         GridV<T> gs, GridV<T> gd, T csup,
                                                         const int *near, int remap) {
    const int k = subject_id(remap);
    if (k >= gs.n || (near && !near[k])) return;
    const int i = gs.ids[k];
    gd = at_inst(gd, inst_of(gs, k));
    const V4<T> me = gs.s[k];
    V3<T> fd = mk3<T>(0, 0, 0);
    query_runs(gd, me.x, me.y, me.z, csup, [&](int lo, int hi) {
        for (int s = lo; s < hi; ++s) {
            const int b = gd.ids[s];
            const V4<T> q = gd.s[s];
            const T dist = length(mk3(me.x - q.x, me.y - q.y, me.z - q.z));
            const T hb = D.h[b];
            if (dist < (T)2.0 * hb && D.flt_s[b] > (T)1.0e-12)
                fd = fd - (ld3(D.drag, b) * wendland_w(dist, hb)) / D.flt_s[b];
        }
    });
    st3(P.acc_dem, i, fd / (P.rho[i] * P.porosity[i]));
}
// slot[i]: original point i's sorted slot (the grid build's inverse order).
template <typename T, bool INTER, bool STEP>
__global__ __launch_bounds__(kThreads) void k_fluid_final(Prm<T> p, SphV<T> P, const int *slot,
                                                          int n, const int *near, int remap) {
    const int i = subject_id(remap);
    if (i >= n) return;
    V3<T> acc = ld3(P.acc, i);
    if (INTER) {
        V3<T> ad = mk3<T>(0, 0, 0);
        if (!near || near[slot[i]])
            ad = ld3(P.acc_dem, i);
        else
            st3(P.acc_dem, i, ad);
        acc = acc + ad;
        st3(P.acc, i, acc);
    }
    if (STEP) {
        const V3<T> v = ld3(P.vel, i) + acc * p.dt;
        st3(P.vel, i, v);
        st3(P.pos, i, ld3(P.pos, i) + v * p.dt);
    }
}

// ---------------------------------------------------------------- diagnostics
template <typename T>
__global__ __launch_bounds__(1024) void k_reduce(const T *pos, const T *vel, const T *m, int nf,
                                                 const T *dvel, const T *dw, const T *dm,
                                                 const T *dI, int nd, double *out) {
    double s_m = 0, s_kf = 0, s_kd = 0, s_md = 0, front = -1.0e300;
    for (int i = threadIdx.x; i < nf; i += blockDim.x) {
        const double mi = (double)m[i];
        const double vx = vel[3 * i], vy = vel[3 * i + 1], vz = vel[3 * i + 2];
        s_m += mi;
        s_kf += 0.5 * mi * (vx * vx + vy * vy + vz * vz);
        const double x = (double)pos[3 * i];
        front = x > front ? x : front;
    }
    for (int a = threadIdx.x; a < nd; a += blockDim.x) {
        const double ma = (double)dm[a], Ia = (double)dI[a];
        const double vx = dvel[3 * a], vy = dvel[3 * a + 1], vz = dvel[3 * a + 2];
        const double wx = dw[3 * a], wy = dw[3 * a + 1], wz = dw[3 * a + 2];
        s_md += ma;
        s_kd += 0.5 * ma * (vx * vx + vy * vy + vz * vz) + 0.5 * Ia * (wx * wx + wy * wy + wz * wz);
    }
    __shared__ double red[16][5];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    double v[4] = {s_m, s_kf, s_kd, s_md};
    for (int c = 0; c < 4; ++c) {
        const double r = wave_sum(v[c]);
        if (lane == 0) red[wid][c] = r;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(front, off, 64);
        front = o > front ? o : front;
    }
    if (lane == 0) red[wid][4] = front;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t[5] = {0, 0, 0, 0, -1.0e300};
        for (int q = 0; q < (int)(blockDim.x / 64); ++q) {
            for (int c = 0; c < 4; ++c) t[c] += red[q][c];
            t[4] = red[q][4] > t[4] ? red[q][4] : t[4];
        }
        out[0] = t[0];
        out[1] = t[1];
        out[2] = t[2];
        out[3] = t[3];
        out[4] = t[4];
    }
}

// ---------------------------------------------------------------- host side
static int check_dem(const sphdem_params_t *p, const sphdem_dem_t *D, const sphdem_grid_t *gd) {
    SPHDEM_REQUIRE(p && D && gd, "NULL argument");
    SPHDEM_REQUIRE(p->precision == 4 || p->precision == 8, "precision must be 4 or 8");
    SPHDEM_REQUIRE(gd->precision == p->precision, "grid precision differs from the DEM precision");
    SPHDEM_REQUIRE(D->n >= 0, "DEM count must be nonnegative");
    SPHDEM_REQUIRE(D->n == 0 || (D->pos && D->vel && D->force && D->omega && D->torque &&
                                 D->radius && D->m && D->inertia && D->K && D->eta && D->mu &&
                                 D->h && D->volume && D->rho),
                   "DEM field pointer is NULL");
    SPHDEM_REQUIRE(gd->n == 0 || (gd->point_ids && gd->sorted), "DEM grid is not built");
    return SPHDEM_OK;
}

static int check_csr(const sphdem_dem_t *D, bool wall) {
    const int *oo = wall ? D->contact_bnd_offset_old : D->contact_dem_offset_old;
    const int *on = wall ? D->contact_bnd_offset_new : D->contact_dem_offset_new;
    SPHDEM_REQUIRE(oo && on, "CSR offset buffer is NULL");
    const int cap = wall ? D->cap_bnd_new : D->cap_dem_new;
    SPHDEM_REQUIRE(cap >= 0, "CSR capacity must be nonnegative");
    SPHDEM_REQUIRE(cap == 0 || (wall ? (D->contact_bnd_id_new && D->tang_bnd_new)
                                     : (D->contact_dem_id_new && D->tang_dem_new)),
                   "CSR edge buffer is NULL");
    return SPHDEM_OK;
}

// The instance of each DEM, when the DEM grid batches several scenes.
static const int *dem_inst(const sphdem_grid_t *gd) {
    return gd && gd->n_inst > 1 ? gd->inst : nullptr;
}

template <typename T, bool WALL>
static void launch_count(const sphdem_dem_t *D, const sphdem_grid_t *g, const int *dinst,
                         const void *nb_radius, double rmax, hipStream_t st) {
    int *off = WALL ? D->contact_bnd_offset_new : D->contact_dem_offset_new;
    ProfScope prof(SPHDEM_PASS_DEM_CSR, st);
    hipLaunchKernelGGL((k_count_contacts<T, WALL>), dim3(blocks_for(D->n + 1, 64)), dim3(64), 0, st,
                       view<T>(D), view<T>(g), dinst, (const T *)nb_radius, (T)rmax, off);
}

template <typename T, bool WALL>
static void launch_contact(const sphdem_dem_t *D, const sphdem_grid_t *g, const int *dinst,
                           const void *nb_radius, const void *nb_vel, const void *nb_omega,
                           double rmax, double grav, double dt, int *flags, hipStream_t st) {
    if (D->n == 0) return;
    ProfScope prof(SPHDEM_PASS_DEM_CONTACT, st);
    hipLaunchKernelGGL((k_contact_force<T, WALL>), dim3(blocks_for(D->n, 64)), dim3(64), 0, st,
                       view<T>(D), view<T>(g), dinst, (const T *)nb_radius, (const T *)nb_vel,
                       (const T *)nb_omega, (T)rmax, (T)grav, (T)dt, flags);
}

template <typename T>
static void launch_copy_csr(const sphdem_dem_t *D, bool wall, hipStream_t st) {
    const int cap = wall ? D->cap_bnd_new : D->cap_dem_new;
    const int span = (D->n + 1 > cap ? D->n + 1 : cap);
    ProfScope prof(SPHDEM_PASS_DEM_CONTACT, st);
    if (wall)
        hipLaunchKernelGGL(k_copy_csr<T>, dim3(blocks_for(span, 256)), dim3(256), 0, st,
                           D->contact_bnd_offset_new, D->contact_bnd_id_new,
                           (const T *)D->tang_bnd_new, D->contact_bnd_offset_old,
                           D->contact_bnd_id_old, (T *)D->tang_bnd_old, D->n, cap);
    else
        hipLaunchKernelGGL(k_copy_csr<T>, dim3(blocks_for(span, 256)), dim3(256), 0, st,
                           D->contact_dem_offset_new, D->contact_dem_id_new,
                           (const T *)D->tang_dem_new, D->contact_dem_offset_old,
                           D->contact_dem_id_old, (T *)D->tang_dem_old, D->n, cap);
}

template <typename T>
static void launch_porosity(const sphdem_params_t *p, const sphdem_sph_t *P,
                            const sphdem_dem_t *D, const sphdem_grid_t *gs,
                            const sphdem_grid_t *gd, const sphdem_work_t *w, double csup,
                            double hu, hipStream_t st) {
    if (gs->n == 0) return;
    ProfScope prof(SPHDEM_PASS_POROSITY, st);
    hipLaunchKernelGGL(k_porosity<T>, dim3(blocks_for(gs->n, kThreads)), dim3(kThreads), 0, st,
                       prm<T>(p), view<T>(P), view<T>(D), view<T>(gs), view<T>(gd), view<T>(w),
                       (T)csup, (T)hu);
}

// Uniform DEM h: the tiled pipeline of porosity.hip.  Heterogeneous h needs
// a fluid sum per DEM neighbour and h_b (KERNEL_SPHDEM_interaction.py:90-95):
// one thread per subject as above.
template <typename T>
static int porosity(const sphdem_params_t *p, const sphdem_sph_t *P, const sphdem_dem_t *D,
                    const sphdem_grid_t *gs, const sphdem_grid_t *gd, const sphdem_work_t *w,
                    double csup, double hu, hipStream_t st, int *dem_near = nullptr,
                    hipEvent_t marked = nullptr) {
    if (hu > 0.0) {
        SPHDEM_REQUIRE(w->fl_flag && w->fl_list && w->fl_count && w->fl_scan,
                       "porosity work buffers are NULL");
        return launch_porosity_tiled<T>(p, P, D, gs, gd, w, csup, hu, dem_near, st, marked);
    }
    launch_porosity<T>(p, P, D, gs, gd, w, csup, hu, st);
    return check_launch("porosity");
}

// The per-DEM flags of the tiled porosity mark for this step, or null: they
// need the work array and a mark query (radius csup) that spans fewer DEM
// grid buckets than the grid holds along every axis, so that it visits every
// DEM within csup.
static int *dem_near_flags(const sphdem_params_t *p, const sphdem_dem_t *D,
                           const sphdem_grid_t *gd, const sphdem_work_t *w, double csup) {
    if (!(p->h_uniform > 0.0) || !w->dem_near || w->cap_d < D->n || D->n == 0) return nullptr;
    const double span = 2.0 * csup / gd->cell_width + 3.0;
    if (span > gd->dim_x || span > gd->dim_y || span > gd->dim_z) return nullptr;
    return w->dem_near;
}

template <typename T>
static void launch_interaction_dem(const sphdem_params_t *p, const sphdem_dem_t *D,
                                   const sphdem_grid_t *gs, const sphdem_grid_t *gd,
                                   const sphdem_work_t *w, hipStream_t st,
                                   int *near = nullptr) {
    if (gd->n == 0) return;
    ProfScope prof(SPHDEM_PASS_INTERACTION_DEM, st);
    hipLaunchKernelGGL(k_interaction_dem<T>, dim3(gd->n), dim3(kDemThreads), 0, st, prm<T>(p),
                       view<T>(D), view<T>(gs), view<T>(gd), view<T>(w), near);
}

template <typename T, bool INTER, bool STEP>
static void launch_fluid_final(const sphdem_params_t *p, const sphdem_sph_t *P,
                               const sphdem_dem_t *D, const sphdem_grid_t *gs,
                               const sphdem_grid_t *gd, double csup, hipStream_t st,
                               const int *near = nullptr) {
    if (gs->n == 0) return;
    ProfScope prof(SPHDEM_PASS_FLUID_FINAL, st);
    const int bt = tuned_block(SPHDEM_PASS_FLUID_FINAL);
    const int rm = tuned_remap(SPHDEM_PASS_FLUID_FINAL);
    if (INTER)
        hipLaunchKernelGGL(k_fluid_drag<T>, dim3(blocks_for(gs->n, bt)), dim3(bt), 0, st,
                           prm<T>(p), view<T>(P), view<T>(D), view<T>(gs), view<T>(gd), (T)csup,
                           near, rm);
    hipLaunchKernelGGL((k_fluid_final<T, INTER, STEP>), dim3(blocks_for(gs->n, bt)), dim3(bt), 0,
                       st, prm<T>(p), view<T>(P), (const int *)gs->rank, gs->n, near, rm);
}

template <typename T>
static void launch_step_dem(const sphdem_params_t *p, const sphdem_dem_t *D, hipStream_t st) {
    if (D->n == 0) return;
    ProfScope prof(SPHDEM_PASS_STEP_DEM, st);
    hipLaunchKernelGGL(k_step_dem<T>, dim3(blocks_for(D->n, 64)), dim3(64), 0, st, view<T>(D),
                       (T)p->dt);
}

// A second stream per device for the DEM contact work of a fused step, with
// the events that fork it from and join it to the caller's stream.  Created
// on the first fused step that is not under graph capture (capture then
// records the fork and join); until they exist the step runs in one stream.
// Default priority, non-blocking, events without timing.
// Per device: the DEM side stream (s, fork, join, marked) and the wall
// pass's stream (wall.s, wall.fork, wall.join), all at the default priority.
struct SideStream {
    hipStream_t s = nullptr;
    hipEvent_t fork = nullptr, join = nullptr, marked = nullptr;
    Beside wall, force;   // the force pass runs on the wall pass's stream
};
static SideStream *side_stream(hipStream_t caller) {
    static SideStream per_device[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
    SideStream &r = per_device[dev];
    auto ready = [&]() {
        return r.s && r.fork && r.join && r.marked && r.wall.s && r.wall.fork && r.wall.join &&
               r.force.s && r.force.fork && r.force.join;
    };
    if (ready()) return &r;
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(caller, &cap) != hipSuccess || cap != hipStreamCaptureStatusNone)
        return nullptr;
    auto stream = [](hipStream_t &x) {
        if (!x && hipStreamCreateWithFlags(&x, hipStreamNonBlocking) != hipSuccess) x = nullptr;
    };
    auto event = [](hipEvent_t &e) {
        if (!e && hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) e = nullptr;
    };
    stream(r.s);
    stream(r.wall.s);
    r.force.s = r.wall.s;
    for (hipEvent_t *e : {&r.fork, &r.join, &r.marked, &r.wall.fork, &r.wall.join, &r.force.fork,
                          &r.force.join})
        event(*e);
    return ready() ? &r : nullptr;
}
// The wall pass's stream when the setting asks for it and the streams exist.
static Beside wall_beside(SideStream *side) {
    return (side && tuning().wall_beside) ? side->wall : Beside{};
}
// The force pass's stream, likewise.
static Beside force_beside(SideStream *side) {
    return (side && tuning().force_beside) ? side->force : Beside{};
}

// The DEM-DEM and DEM-wall contact work of a fused step (counts, scans,
// contact laws, CSR copies): it reads only DEM state, the DEM grids and the
// old CSR, and writes the new CSR, D.force/torque and the flags.
template <typename T>
static int dem_contacts(const sphdem_params_t *p, const sphdem_dem_t *D, const sphdem_dembnd_t *W,
                        const sphdem_grid_t *gd, const sphdem_grid_t *gw,
                        const sphdem_work_t *w, hipStream_t st) {
    launch_count<T, false>(D, gd, dem_inst(gd), D->radius, p->radius_max, st);
    int rc;
    {
        ProfScope prof(SPHDEM_PASS_DEM_CSR, st);
        rc = exclusive_scan(D->contact_dem_offset_new, D->contact_dem_offset_new, D->n + 1,
                            w->scan_tmp, st);
    }
    if (rc) return rc;
    launch_contact<T, false>(D, gd, dem_inst(gd), D->radius, D->vel, D->omega, p->radius_max,
                             p->g, p->dt, w->flags, st);
    launch_copy_csr<T>(D, false, st);
    launch_count<T, true>(D, gw, dem_inst(gd), W->radius, p->boundary_radius_max, st);
    {
        ProfScope prof(SPHDEM_PASS_DEM_CSR, st);
        rc = exclusive_scan(D->contact_bnd_offset_new, D->contact_bnd_offset_new, D->n + 1,
                            w->scan_tmp, st);
    }
    if (rc) return rc;
    launch_contact<T, true>(D, gw, dem_inst(gd), W->radius, W->vel, W->omega,
                            p->boundary_radius_max, 0.0, p->dt, w->flags, st);
    launch_copy_csr<T>(D, true, st);
    return SPHDEM_OK;
}

// Simulation.py:102-221; see sphdem_coupled_step in sphdem.h for the stages.
// The fused step (stage -1) runs the DEM contact work on the side stream,
// beside the SPH passes and the porosity (which touch none of its arrays),
// and joins it before the DEM interpolation adds the coupling forces.
template <typename T>
static int coupled_step(const sphdem_params_t *p, const sphdem_sph_t *P, const sphdem_bnd_t *B,
                        const sphdem_dem_t *D, const sphdem_dembnd_t *W, const sphdem_grid_t *gs,
                        const sphdem_grid_t *gb, const sphdem_grid_t *gd,
                        const sphdem_grid_t *gw, int step, int stage, const sphdem_work_t *w,
                        hipStream_t st) {
    const bool all = stage < 0;
    int rc;
    if (all) {
        SideStream *side = side_stream(st);
        hipStream_t ds = side ? side->s : st;
        auto join = [&]() {
            if (side) {
                (void)hipEventRecord(side->join, ds);
                (void)hipStreamWaitEvent(st, side->join, 0);
            }
        };
        if (side) {
            (void)hipEventRecord(side->fork, st);
            (void)hipStreamWaitEvent(ds, side->fork, 0);
        }
        // beside the porosity (which reads positions, m/rho from the
        // density pass and DEM state) the force pass writes accelerations,
        // densities and pressure gradients; the DEM interpolation and the
        // final update wait for it
        const Beside fb = force_beside(side);
        auto force_joined = [&](hipStream_t s) {
            if (fb.s) (void)hipStreamWaitEvent(s, fb.join, 0);
        };
        rc = dem_contacts<T>(p, D, W, gd, gw, w, ds);
        if (!rc) rc = sph_stage<T>(p, P, B, gs, gb, step, true, w, st, wall_beside(side), fb);
        if (rc) {
            force_joined(st);
            join();
            return rc;
        }
        const double csup = 2.0 * p->h_max;
        int *near = dem_near_flags(p, D, gd, w, csup);
        // with the tiled porosity the DEM interpolation needs only the mark's
        // flags: it runs on the side stream beside the porosity stream kernel
        const bool beside = side && p->h_uniform > 0.0 && gs->n > 0 && tuning().dem_beside;
        rc = porosity<T>(p, P, D, gs, gd, w, csup, p->h_uniform, st, near,
                         beside ? side->marked : nullptr);
        if (rc) {
            force_joined(st);
            join();
            return rc;
        }
        if (beside) {
            (void)hipStreamWaitEvent(ds, side->marked, 0);
            force_joined(ds);
            launch_interaction_dem<T>(p, D, gs, gd, w, ds, near);
            join();
        } else {
            join();
            force_joined(st);
            launch_interaction_dem<T>(p, D, gs, gd, w, st, near);
        }
        launch_fluid_final<T, true, true>(p, P, D, gs, gd, csup, st, near ? w->fl_flag : nullptr);
        launch_step_dem<T>(p, D, st);
        return check_launch("coupled_step");
    }
    if (stage == 0) {
        rc = sph_stage<T>(p, P, B, gs, gb, step, true, w, st, Beside{}, Beside{});
        if (rc) return rc;
        launch_count<T, false>(D, gd, dem_inst(gd), D->radius, p->radius_max, st);
        {
            ProfScope prof(SPHDEM_PASS_DEM_CSR, st);
            rc = exclusive_scan(D->contact_dem_offset_new, D->contact_dem_offset_new, D->n + 1,
                                w->scan_tmp, st);
        }
        if (rc) return rc;
    }
    if (stage == 1) {
        launch_contact<T, false>(D, gd, dem_inst(gd), D->radius, D->vel, D->omega,
                                 p->radius_max, p->g, p->dt, w->flags, st);
        launch_count<T, true>(D, gw, dem_inst(gd), W->radius, p->boundary_radius_max, st);
        {
            ProfScope prof(SPHDEM_PASS_DEM_CSR, st);
            rc = exclusive_scan(D->contact_bnd_offset_new, D->contact_bnd_offset_new, D->n + 1,
                                w->scan_tmp, st);
        }
        if (rc) return rc;
    }
    if (stage == 2) {
        launch_contact<T, true>(D, gw, dem_inst(gd), W->radius, W->vel, W->omega,
                                p->boundary_radius_max, 0.0, p->dt, w->flags, st);
        const double csup = 2.0 * p->h_max;
        int *near = dem_near_flags(p, D, gd, w, csup);
        rc = porosity<T>(p, P, D, gs, gd, w, csup, p->h_uniform, st, near);
        if (rc) return rc;
        launch_interaction_dem<T>(p, D, gs, gd, w, st, near);
        // the mark's per-fluid flags of this step go with its per-DEM flags
        launch_fluid_final<T, true, true>(p, P, D, gs, gd, csup, st, near ? w->fl_flag : nullptr);
        launch_step_dem<T>(p, D, st);
    }
    return check_launch("coupled_step");
}

// ---------------------------------------------------------------- surrogate support
// DEM loads of a substep plus the coupling force frozen at a surrogate
// call's start (drag and pressure force are zero for a dry sphere).
template <typename T>
__global__ void k_add_coupling(DemV<T> D) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= D.n) return;
    st3(D.force, a, (ld3(D.force, a) + ld3(D.drag, a)) + ld3(D.pressure_force, a));
}

// Start of a surrogate call: the SPH passes (density, EOS, acceleration with
// -grad p), the DEM coupling sums (drag, pressure force) and the drag
// reaction added to the fluid acceleration (with the porosity last computed
// by a plain step).  Leaves P.acc = the acceleration the network reads.
template <typename T>
static int surrogate_prep(const sphdem_params_t *p, const sphdem_sph_t *P, const sphdem_bnd_t *B,
                          const sphdem_dem_t *D, const sphdem_grid_t *gs, const sphdem_grid_t *gb,
                          const sphdem_grid_t *gd, int step, const sphdem_work_t *w,
                          hipStream_t st) {
    const bool coupled = D && D->n > 0;
    int rc = sph_stage<T>(p, P, B, gs, gb, step, coupled, w, st, Beside{}, Beside{});
    if (rc || !coupled) return rc;
    launch_interaction_dem<T>(p, D, gs, gd, w, st);
    launch_fluid_final<T, true, false>(p, P, D, gs, gd, 2.0 * p->h_max, st);
    return check_launch("surrogate_prep");
}

// n_sub DEM substeps of a surrogate call: per substep the DEM-DEM contacts
// (candidates from the DEM grid built at the call's start, queried with the
// radius widened by `skin` and read at their current positions), the
// DEM-wall contacts (static walls), the coupling force held from the call's
// start, then the integration of Kernel_step_dem.  Contact histories live in
// fixed-width rows (kEll per DEM and kind) in `scratch` and go back to the
// CSR (old and new, as stage -1 leaves them) at the end.  Loads are the same
// as the plain DEM stage's (dem_contact, gravity, last-match history) and are
// summed in the same candidate order; flags bit 0: CSR capacity exceeded,
// bit 1: a DEM moved farther than `skin` from where the grid was built, bit
// 2: a history row exceeded kEll contacts.
// Launches: init; per substep loads (one wave per DEM: the lanes test 64
// candidates at a time, the contacts are compacted in candidate order and
// lane 0 adds them in that order) and integration; the CSR write-back.
constexpr int kEll = 64;
constexpr int kSubThreads = 1024;

template <typename T>
__device__ __forceinline__ void block_scan_rows(const int *cnt, int n, int *off, int *total) {
    __shared__ int part[kSubThreads];
    const int t = threadIdx.x;
    const int chunk = (n + kSubThreads - 1) / kSubThreads;
    const int a0 = min(n, t * chunk), a1 = min(n, a0 + chunk);
    int sum = 0;
    for (int a = a0; a < a1; ++a) sum += cnt[a];
    part[t] = sum;
    __syncthreads();
    for (int d = 1; d < kSubThreads; d <<= 1) {
        const int v = t >= d ? part[t - d] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int run = part[t] - sum;
    for (int a = a0; a < a1; ++a) {
        off[a] = run;
        run += cnt[a];
    }
    if (t == kSubThreads - 1) *total = part[t];
    __syncthreads();
}

// Substep scratch: cnt/ids/tang rows [side 0/1][kind 0 = dem, 1 = wall][n][kEll].
template <typename T> struct SubRows {
    T *pos0;
    int *cnt, *ids;
    T *tang;
    int n;
    __device__ int row(int side, int kind, int a) const { return (side * 2 + kind) * n + a; }
};

// CSR (old) -> rows of side 0; the positions the grid was built from.
template <typename T>
__global__ void k_sub_init(DemV<T> D, SubRows<T> R, int *flags) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= R.n) return;
    st3(R.pos0, a, ld3(D.pos, a));
    for (int kind = 0; kind < 2; ++kind) {
        const int *off = kind ? D.bnd_off_old : D.dem_off_old;
        const int *id = kind ? D.bnd_id_old : D.dem_id_old;
        const T *tg = kind ? D.bnd_tang_old : D.dem_tang_old;
        const int o0 = off[a], o1 = off[a + 1];
        int c = 0;
        for (int k = o0; k < o1; ++k) {
            if (c == kEll) {
                atomicOr(flags, 4);
                break;
            }
            const int r = R.row(0, kind, a) * kEll + c;
            R.ids[r] = id[k];
            st3(R.tang, r, ld3(tg, k));
            ++c;
        }
        R.cnt[R.row(0, kind, a)] = c;
    }
}

// Loads of one substep: reads the neighbours' current state and the rows of
// side `cur`, writes D.force/torque and the rows of the other side.
constexpr int kSubWaves = 4;
template <typename T>
__global__ __launch_bounds__(64 * kSubWaves) void k_sub_loads(Prm<T> p, DemV<T> D, WallV<T> W,
                                                              GridV<T> gd, GridV<T> gw,
                                                              const int *dinst, T skin, int cur,
                                                              SubRows<T> R, int *flags) {
    __shared__ V3<T> fbuf[kSubWaves][64], tbuf[kSubWaves][64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int a = blockIdx.x * kSubWaves + wv;
    if (a >= R.n) return;
    const int nxt = cur ^ 1;
    const int ia = dinst ? dinst[a] : 0;
    const V3<T> ra = ld3(D.pos, a);
    const T Ra = D.radius[a];
    const V3<T> va = ld3(D.vel, a), wa = ld3(D.omega, a);
    const T K = D.K[a], eta = D.eta[a], mu = D.mu[a];
    // DEM-DEM loads start from gravity; the wall loads are summed apart and
    // added, as the two plain kernels do
    V3<T> fk[2] = {mk3<T>(0, 0, -D.m[a] * p.g), mk3<T>(0, 0, 0)};
    V3<T> tk[2] = {mk3<T>(0, 0, 0), mk3<T>(0, 0, 0)};
    for (int kind = 0; kind < 2; ++kind) {
        const GridV<T> g = at_inst(kind ? gw : gd, ia);
        const T *nr = kind ? W.radius : D.radius;
        const T *nv = kind ? W.vel : D.vel;
        const T *nw = kind ? W.omega : D.omega;
        const int rc = R.row(cur, kind, a), rn = R.row(nxt, kind, a);
        const int c0 = R.cnt[rc];
        int out = 0;
        V3<T> force = fk[kind], torque = tk[kind];
        const T reach = Ra + (kind ? p.boundary_radius_max : p.radius_max + skin);
        query_runs(g, ra.x, ra.y, ra.z, reach, [&](int lo, int hi) {
            for (int base = lo; base < hi; base += 64) {
                const int s = base + lane;
                bool hit = false;
                int b = 0;
                V3<T> f = mk3<T>(0, 0, 0), disp = f, tq = f;
                if (s < hi) {
                    b = g.ids[s];
                    if (kind || b != a) {
                        V3<T> rb;
                        if (kind) {
                            const V4<T> q = g.s[s];
                            rb = mk3(q.x, q.y, q.z);
                        } else {
                            rb = ld3(D.pos, b);
                        }
                        const V3<T> rab = rb - ra;
                        const T dist = length(rab);
                        const T overlap = Ra + nr[b] - dist;
                        if (overlap > (T)0) {
                            hit = true;
                            V3<T> tangent = mk3<T>(0, 0, 0);
                            for (int k = 0; k < c0; ++k)
                                if (R.ids[rc * kEll + k] == b)
                                    tangent = ld3(R.tang, rc * kEll + k);
                            V3<T> normal = kind ? mk3<T>(0, 0, -1) : mk3<T>(1, 0, 0);
                            if (!kind && a > b) normal = -normal;
                            if (dist > (T)1.0e-12) normal = rab / dist;
                            dem_contact(normal, overlap, Ra, nr[b], va, ld3(nv, b), wa,
                                        ld3(nw, b), tangent, K, eta, mu, p.dt, f, disp, tq);
                        }
                    }
                }
                const unsigned long long m = __ballot(hit);
                const int before = (int)__builtin_amdgcn_mbcnt_hi(
                    (unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
                if (hit) {
                    const int o = out + before;
                    if (o < kEll) {
                        R.ids[rn * kEll + o] = b;
                        st3(R.tang, rn * kEll + o, disp);
                    } else {
                        atomicOr(flags, 4);
                    }
                    fbuf[wv][before] = f;
                    tbuf[wv][before] = tq;
                }
                const int nh = __popcll(m);
                __builtin_amdgcn_wave_barrier();
                if (lane == 0)
                    for (int j = 0; j < nh; ++j) {
                        force = force + fbuf[wv][j];
                        torque = torque + tbuf[wv][j];
                    }
                __builtin_amdgcn_wave_barrier();
                out += nh;
            }
        });
        fk[kind] = force;
        tk[kind] = torque;
        if (lane == 0) R.cnt[rn] = out < kEll ? out : kEll;
    }
    if (lane == 0) {
        const V3<T> fc = ((fk[0] + fk[1]) + ld3(D.drag, a)) + ld3(D.pressure_force, a);
        st3(D.force, a, fc);
        st3(D.torque, a, tk[0] + tk[1]);
    }
}

// Integration of one substep (Kernel_step_dem) and the skin check.
template <typename T>
__global__ void k_sub_step(DemV<T> D, SubRows<T> R, T dt, T skin, int *flags) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= R.n) return;
    const V3<T> acc = ld3(D.force, a) / D.m[a];
    const V3<T> vel = ld3(D.vel, a) + acc * dt;
    const V3<T> pos = ld3(D.pos, a) + vel * dt;
    st3(D.acc, a, acc);
    st3(D.vel, a, vel);
    st3(D.pos, a, pos);
    st3(D.omega, a, ld3(D.omega, a) + (ld3(D.torque, a) / D.inertia[a]) * dt);
    if (length(pos - ld3(R.pos0, a)) > skin) atomicOr(flags, 2);
}

// Rows of the last substep (side fin) -> CSR, written to both old and new.
template <typename T>
__global__ __launch_bounds__(kSubThreads) void k_sub_final(DemV<T> D, SubRows<T> R, int fin,
                                                           int *flags) {
    const int n = R.n;
    for (int kind = 0; kind < 2; ++kind) {
        int *off_o = kind ? D.bnd_off_old : D.dem_off_old;
        int *off_n = kind ? D.bnd_off_new : D.dem_off_new;
        int *id_o = kind ? D.bnd_id_old : D.dem_id_old;
        int *id_n = kind ? D.bnd_id_new : D.dem_id_new;
        T *tg_o = kind ? D.bnd_tang_old : D.dem_tang_old;
        T *tg_n = kind ? D.bnd_tang_new : D.dem_tang_new;
        const int cap = kind ? D.cap_bnd_new : D.cap_dem_new;
        __shared__ int total;
        block_scan_rows<T>(R.cnt + R.row(fin, kind, 0), n, off_n, &total);
        // the read side's offsets are clamped to the capacity (as k_copy_csr
        // clamps them), so an overflow leaves the next lookups in bounds
        for (int a = threadIdx.x; a < n; a += kSubThreads) {
            const int o = off_n[a];
            off_o[a] = min(o, cap);
            const int r = R.row(fin, kind, a);
            for (int k = 0; k < R.cnt[r]; ++k) {
                if (o + k >= cap) {
                    atomicOr(flags, 1);
                    break;
                }
                const int id = R.ids[r * kEll + k];
                const V3<T> tv = ld3(R.tang, r * kEll + k);
                id_o[o + k] = id;
                id_n[o + k] = id;
                st3(tg_o, o + k, tv);
                st3(tg_n, o + k, tv);
            }
        }
        if (threadIdx.x == 0) {
            off_o[n] = min(total, cap);
            off_n[n] = total;
        }
        __syncthreads();
    }
}

template <typename T>
static int dem_substeps(const sphdem_params_t *p, const sphdem_dem_t *D, const sphdem_dembnd_t *W,
                        const sphdem_grid_t *gd, const sphdem_grid_t *gw, int n_sub, double skin,
                        void *scratch, const sphdem_work_t *w, hipStream_t st) {
    const int n = D->n;
    char *sp = (char *)scratch;
    SubRows<T> R;
    R.n = n;
    R.pos0 = (T *)sp;
    sp += sizeof(T) * 3 * (size_t)n;
    R.cnt = (int *)sp;
    sp += sizeof(int) * 4 * (size_t)n;
    R.ids = (int *)sp;
    sp += sizeof(int) * 4 * (size_t)n * kEll;
    R.tang = (T *)sp;
    ProfScope prof(SPHDEM_PASS_DEM_CONTACT, st);
    const DemV<T> dv = view<T>(D);
    const dim3 g1(blocks_for(n, 256)), b1(256);
    hipLaunchKernelGGL(k_sub_init<T>, g1, b1, 0, st, dv, R, w->flags);
    for (int sub = 0; sub < n_sub; ++sub) {
        hipLaunchKernelGGL(k_sub_loads<T>, dim3(blocks_for(n, kSubWaves)), dim3(64 * kSubWaves),
                           0, st, prm<T>(p), dv, view<T>(W), view<T>(gd), view<T>(gw),
                           dem_inst(gd), (T)skin, sub & 1, R, w->flags);
        hipLaunchKernelGGL(k_sub_step<T>, g1, b1, 0, st, dv, R, (T)p->dt, (T)skin, w->flags);
    }
    hipLaunchKernelGGL(k_sub_final<T>, dim3(1), dim3(kSubThreads), 0, st, dv, R, n_sub & 1,
                       w->flags);
    return check_launch("dem_substeps");
}

size_t dem_substeps_scratch(int n, int precision) {
    const size_t rs = precision == 8 ? 8 : 4;
    return rs * 3 * (size_t)n + 4 * 4 * (size_t)n + 4 * 4 * (size_t)n * kEll +
           rs * 3 * 4 * (size_t)n * kEll + 256;
}

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

int sphdem_drag_beta(int precision, const void *eps, const void *rho, const void *mu,
                     const void *diameter, const void *speed, int n, void *beta, void *stream) {
    SPHDEM_REQUIRE(precision == 4 || precision == 8, "precision must be 4 or 8");
    SPHDEM_REQUIRE(n >= 0 && (n == 0 || (eps && rho && mu && diameter && speed && beta)),
                   "invalid drag buffers");
    if (n == 0) return SPHDEM_OK;
    DISPATCH(precision, hipLaunchKernelGGL(k_drag_beta<T>, dim3(blocks_for(n, 256)), dim3(256), 0,
                                           as_stream(stream), (const T *)eps, (const T *)rho,
                                           (const T *)mu, (const T *)diameter, (const T *)speed, n,
                                           (T *)beta));
    return check_launch("drag_beta");
}

int sphdem_count_dem_contacts(const sphdem_params_t *p, sphdem_dem_t *D, const sphdem_grid_t *gd,
                              double radius_max, void *stream) {
    int rc = check_dem(p, D, gd);
    if (rc) return rc;
    SPHDEM_REQUIRE(D->contact_dem_offset_new, "contact_dem_offset_new is NULL");
    SPHDEM_REQUIRE(gd->n == D->n, "DEM grid was built from a different number of points");
    DISPATCH(p->precision, (launch_count<T, false>(D, gd, dem_inst(gd), D->radius, radius_max,
                                                   as_stream(stream))));
    return check_launch("count_dem_contacts");
}

int sphdem_force_dem(const sphdem_params_t *p, sphdem_dem_t *D, const sphdem_grid_t *gd,
                     double radius_max, double g, int *flags, void *stream) {
    int rc = check_dem(p, D, gd);
    if (rc) return rc;
    rc = check_csr(D, false);
    if (rc) return rc;
    SPHDEM_REQUIRE(flags, "flags buffer is NULL");
    SPHDEM_REQUIRE(gd->n == D->n, "DEM grid was built from a different number of points");
    DISPATCH(p->precision, (launch_contact<T, false>(D, gd, dem_inst(gd), D->radius, D->vel,
                                                     D->omega, radius_max, g, p->dt, flags,
                                                     as_stream(stream))));
    return check_launch("force_dem");
}

int sphdem_count_bnd_contacts(const sphdem_params_t *p, sphdem_dem_t *D, const sphdem_dembnd_t *W,
                              const sphdem_grid_t *gw, double radius_max, void *stream) {
    int rc = check_dem(p, D, gw);
    if (rc) return rc;
    SPHDEM_REQUIRE(W && D->contact_bnd_offset_new, "NULL argument");
    SPHDEM_REQUIRE(gw->n == W->n, "DEM-wall grid was built from a different number of points");
    SPHDEM_REQUIRE(gw->n_inst <= 1, "a batched DEM-wall grid needs sphdem_coupled_step");
    DISPATCH(p->precision, (launch_count<T, true>(D, gw, nullptr, W->radius, radius_max,
                                                  as_stream(stream))));
    return check_launch("count_bnd_contacts");
}

int sphdem_bc_dem(const sphdem_params_t *p, sphdem_dem_t *D, const sphdem_dembnd_t *W,
                  const sphdem_grid_t *gw, double radius_max, int *flags, void *stream) {
    int rc = check_dem(p, D, gw);
    if (rc) return rc;
    rc = check_csr(D, true);
    if (rc) return rc;
    SPHDEM_REQUIRE(W && flags, "NULL argument");
    SPHDEM_REQUIRE(gw->n == W->n, "DEM-wall grid was built from a different number of points");
    SPHDEM_REQUIRE(gw->n_inst <= 1, "a batched DEM-wall grid needs sphdem_coupled_step");
    DISPATCH(p->precision, (launch_contact<T, true>(D, gw, nullptr, W->radius, W->vel, W->omega,
                                                    radius_max, 0.0, p->dt, flags,
                                                    as_stream(stream))));
    return check_launch("bc_dem");
}

int sphdem_step_dem(const sphdem_params_t *p, sphdem_dem_t *D, void *stream) {
    SPHDEM_REQUIRE(p && D && (p->precision == 4 || p->precision == 8), "invalid arguments");
    SPHDEM_REQUIRE(D->n == 0 || (D->pos && D->vel && D->acc && D->force && D->omega && D->torque &&
                                 D->m && D->inertia),
                   "DEM field pointer is NULL");
    DISPATCH(p->precision, launch_step_dem<T>(p, D, as_stream(stream)));
    return check_launch("step_dem");
}

int sphdem_prep_sphdem(const sphdem_params_t *p, sphdem_sph_t *P, sphdem_bnd_t *B,
                       sphdem_dem_t *D, const sphdem_grid_t *gs, const sphdem_grid_t *gb,
                       const sphdem_grid_t *gd, double support, double h,
                       const sphdem_work_t *w, void *stream) {
    int rc = check_sph_pair(p, P, B, gs, gb);
    if (rc) return rc;
    rc = check_dem(p, D, gd);
    if (rc) return rc;
    rc = check_work(w, P->n, B->n);
    if (rc) return rc;
    SPHDEM_REQUIRE(P->n == 0 || (P->porosity && P->pgf), "fluid coupling field is NULL");
    SPHDEM_REQUIRE(gd->n == D->n, "DEM grid was built from a different number of points");
    hipStream_t st = as_stream(stream);
    DISPATCH(p->precision, (launch_pgf_only<T>(p, P, B, gs, gb, w, st),
                            rc = porosity<T>(p, P, D, gs, gd, w, support, h, st)));
    if (rc) return rc;
    return check_launch("prep_sphdem");
}

int sphdem_interaction_dem(const sphdem_params_t *p, sphdem_sph_t *P, sphdem_dem_t *D,
                           const sphdem_grid_t *gs, const sphdem_grid_t *gd,
                           const sphdem_work_t *w, void *stream) {
    int rc = check_dem(p, D, gd);
    if (rc) return rc;
    SPHDEM_REQUIRE(P && gs && gs->n == P->n, "fluid grid was built from a different number of points");
    SPHDEM_REQUIRE(gs->precision == p->precision, "grid precision differs");
    rc = check_work(w, P->n, 0);
    if (rc) return rc;
    SPHDEM_REQUIRE(gd->n == D->n, "DEM grid was built from a different number of points");
    SPHDEM_REQUIRE(D->n == 0 || (D->flt_s && D->porosity && D->drag && D->pressure_force),
                   "DEM coupling field is NULL");
    hipStream_t st = as_stream(stream);
    DISPATCH(p->precision, (launch_gather<T>(P, nullptr, gs, nullptr, w, st),
                            launch_interaction_dem<T>(p, D, gs, gd, w, st)));
    return check_launch("interaction_dem");
}

int sphdem_interaction_sph(const sphdem_params_t *p, sphdem_sph_t *P, sphdem_dem_t *D,
                           const sphdem_grid_t *gs, const sphdem_grid_t *gd, double support,
                           void *stream) {
    int rc = check_dem(p, D, gd);
    if (rc) return rc;
    SPHDEM_REQUIRE(P && gs && gs->n == P->n, "fluid grid was built from a different number of points");
    SPHDEM_REQUIRE(gs->precision == p->precision, "grid precision differs");
    SPHDEM_REQUIRE(P->n == 0 || (P->acc && P->acc_dem && P->rho && P->porosity),
                   "fluid field pointer is NULL");
    SPHDEM_REQUIRE(gd->n == D->n, "DEM grid was built from a different number of points");
    DISPATCH(p->precision,
             (launch_fluid_final<T, true, false>(p, P, D, gs, gd, support, as_stream(stream))));
    return check_launch("interaction_sph");
}

int sphdem_device_init(void *stream) {
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(as_stream(stream), &cap) != hipSuccess)
        return fail(SPHDEM_E_HIP, "stream capture status query failed");
    SPHDEM_REQUIRE(cap == hipStreamCaptureStatusNone,
                   "sphdem_device_init must not run under graph capture");
    if (!side_stream(as_stream(stream)))
        return fail(SPHDEM_E_HIP, "could not create the side stream and its events");
    return noalloc_scan_scratch(nullptr);
}

int sphdem_coupled_step(const sphdem_params_t *p, sphdem_sph_t *P, sphdem_bnd_t *B,
                        sphdem_dem_t *D, sphdem_dembnd_t *W, const sphdem_grid_t *gs,
                        const sphdem_grid_t *gb, const sphdem_grid_t *gd,
                        const sphdem_grid_t *gw, int step, int stage, const sphdem_work_t *w,
                        void *stream) {
    int rc = check_sph_pair(p, P, B, gs, gb);
    if (rc) return rc;
    rc = check_dem(p, D, gd);
    if (rc) return rc;
    rc = check_dem(p, D, gw);
    if (rc) return rc;
    rc = check_work(w, P->n, B->n);
    if (rc) return rc;
    rc = check_csr(D, false);
    if (rc) return rc;
    rc = check_csr(D, true);
    if (rc) return rc;
    SPHDEM_REQUIRE(W != nullptr, "DEM wall is NULL");
    SPHDEM_REQUIRE(gd->n == D->n, "DEM grid was built from a different number of points");
    SPHDEM_REQUIRE(gw->n == W->n, "DEM-wall grid was built from a different number of points");
    SPHDEM_REQUIRE(D->n > 0, "coupled step needs DEM particles");
    SPHDEM_REQUIRE(stage >= -1 && stage <= 2, "stage must be -1, 0, 1 or 2");
    SPHDEM_REQUIRE(step >= 0, "step must be nonnegative");
    SPHDEM_REQUIRE(w->scan_tmp && w->flags, "work scan/flags buffer is NULL");
    SPHDEM_REQUIRE(P->n == 0 || (P->porosity && P->pgf && P->acc_dem), "fluid coupling field is NULL");
    if (stage < 0) {
        SPHDEM_REQUIRE(D->contact_dem_id_old && D->tang_dem_old && D->contact_bnd_id_old &&
                           D->tang_bnd_old,
                       "old CSR edge buffer is NULL");
    }
    DISPATCH(p->precision,
             rc = coupled_step<T>(p, P, B, D, W, gs, gb, gd, gw, step, stage, w, as_stream(stream)));
    return rc;
}

int sphdem_surrogate_prep(const sphdem_params_t *p, sphdem_sph_t *P, sphdem_bnd_t *B,
                          sphdem_dem_t *D, const sphdem_grid_t *gs, const sphdem_grid_t *gb,
                          const sphdem_grid_t *gd, int step, const sphdem_work_t *w,
                          void *stream) {
    int rc = check_sph_pair(p, P, B, gs, gb);
    if (rc) return rc;
    rc = check_work(w, P->n, B->n);
    if (rc) return rc;
    SPHDEM_REQUIRE(step >= 0, "step must be nonnegative");
    if (D && D->n > 0) {
        rc = check_dem(p, D, gd);
        if (rc) return rc;
        SPHDEM_REQUIRE(gd->n == D->n, "DEM grid was built from a different number of points");
        SPHDEM_REQUIRE(P->n == 0 || (P->porosity && P->pgf && P->acc_dem),
                       "fluid coupling field is NULL");
        SPHDEM_REQUIRE(D->flt_s && D->porosity && D->drag && D->pressure_force,
                       "DEM coupling field is NULL");
    }
    DISPATCH(p->precision,
             rc = surrogate_prep<T>(p, P, B, D, gs, gb, gd, step, w, as_stream(stream)));
    return rc;
}

long sphdem_dem_substeps_scratch(int n, int precision) {
    if (n < 0 || (precision != 4 && precision != 8)) return -1;
    return (long)dem_substeps_scratch(n, precision);
}

int sphdem_dem_substeps(const sphdem_params_t *p, sphdem_dem_t *D, sphdem_dembnd_t *W,
                        const sphdem_grid_t *gd, const sphdem_grid_t *gw, int n_sub, double skin,
                        void *scratch, const sphdem_work_t *w, void *stream) {
    int rc = check_dem(p, D, gd);
    if (rc) return rc;
    rc = check_csr(D, false);
    if (rc) return rc;
    rc = check_csr(D, true);
    if (rc) return rc;
    SPHDEM_REQUIRE(W && gw && gw->n == W->n, "DEM-wall grid was built from a different number of points");
    SPHDEM_REQUIRE(gd->n == D->n, "DEM grid was built from a different number of points");
    SPHDEM_REQUIRE(n_sub >= 0 && skin >= 0.0, "invalid substep count or skin");
    SPHDEM_REQUIRE(w && w->flags && scratch, "flags or scratch buffer is NULL");
    SPHDEM_REQUIRE(D->contact_dem_id_old && D->tang_dem_old && D->contact_bnd_id_old &&
                       D->tang_bnd_old,
                   "old CSR edge buffer is NULL");
    if (D->n == 0) return SPHDEM_OK;
    DISPATCH(p->precision,
             rc = dem_substeps<T>(p, D, W, gd, gw, n_sub, skin, scratch, w, as_stream(stream)));
    return rc;
}

int sphdem_reduce_f64(int precision, const void *sph_pos, const void *sph_vel, const void *sph_m,
                      int nf, const void *dem_vel, const void *dem_omega, const void *dem_m,
                      const void *dem_inertia, int nd, double *out, void *stream) {
    SPHDEM_REQUIRE(precision == 4 || precision == 8, "precision must be 4 or 8");
    SPHDEM_REQUIRE(out && nf >= 0 && nd >= 0, "invalid reduction arguments");
    SPHDEM_REQUIRE(nf == 0 || (sph_pos && sph_vel && sph_m), "fluid buffer is NULL");
    SPHDEM_REQUIRE(nd == 0 || (dem_vel && dem_omega && dem_m && dem_inertia), "DEM buffer is NULL");
    DISPATCH(precision, hipLaunchKernelGGL(k_reduce<T>, dim3(1), dim3(1024), 0, as_stream(stream),
                                           (const T *)sph_pos, (const T *)sph_vel,
                                           (const T *)sph_m, nf, (const T *)dem_vel,
                                           (const T *)dem_omega, (const T *)dem_m,
                                           (const T *)dem_inertia, nd, out));
    return check_launch("reduce_f64");
}

}  // extern "C"
