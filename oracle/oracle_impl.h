/*
 * oracle_impl.h -- body of the CPU oracle, included once per precision by
 * oracle.c with REAL (float|double) and SFX (_f32|_f64) defined.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Subjects are processed in the order the reference launches them (tid, or
 * the grid's sorted order where the kernel calls hash_grid_point_id), and
 * neighbours are visited in the HashGrid query order: cells z-outer, y, x,
 * then bucket order.  Sums therefore accumulate in the same order as the
 * reference's per-thread loops.  An OpenMP loop over subjects changes no
 * result: each subject only writes its own slots.
 *
 * How each part is pinned (tests/ hold the checks):
 *   wendland_w / wendland_dw  closed form; unit integral over the 2h ball;
 *                             finite-difference slope; end values.
 *   grid_build / query        a hand-derived table of bucket keys, stable
 *                             order and bounds (truncation toward zero, the
 *                             2^20 origin on dims that do not divide it, the
 *                             fold and the clamp; tests/pins.py); brute-force
 *                             O(N^2) neighbour sets on small clouds;
 *                             stable-permutation and partition invariants.
 *                             The convention restates Warp's hashgrid
 *                             (third-party, absent): that Warp itself keys
 *                             and orders this way stays unverified.
 *   pres (Tait EOS)           the closed form evaluated exactly with
 *                             rationals over [0.98, 1.05] rho0 (gamma 7 and
 *                             1), rho = rho0 and the clamped side, and hand
 *                             values (tests/test_oracle_pins.py).
 *   density / shepard / force brute-force FP64 sums on small scenes; lattice
 *                             and hydrostatic invariants; gravity-at-rest.
 *   DEM contact / CSR         closed-form contact tests and CSR sequences of
 *                             the reference's tests/test_dem.py, test_3d.py.
 *   coupling                  drag branches, buoyancy, momentum balance and
 *                             variable-h brute force of the reference tests.
 */

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(name, SFX)

typedef struct { REAL x, y, z; } FN(v3);
#define V3 FN(v3)

static inline V3 FN(ld3)(const REAL *a, int i) {
    V3 r = {a[3 * i], a[3 * i + 1], a[3 * i + 2]};
    return r;
}
static inline void FN(st3)(REAL *a, int i, V3 v) {
    a[3 * i] = v.x; a[3 * i + 1] = v.y; a[3 * i + 2] = v.z;
}
static inline V3 FN(mk3)(REAL x, REAL y, REAL z) { V3 r = {x, y, z}; return r; }
static inline V3 FN(add3)(V3 a, V3 b) { return FN(mk3)(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V3 FN(sub3)(V3 a, V3 b) { return FN(mk3)(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V3 FN(mul3)(REAL s, V3 a) { return FN(mk3)(s * a.x, s * a.y, s * a.z); }
static inline V3 FN(mul3r)(V3 a, REAL s) { return FN(mk3)(a.x * s, a.y * s, a.z * s); }
static inline V3 FN(div3)(V3 a, REAL s) { return FN(mk3)(a.x / s, a.y / s, a.z / s); }
static inline V3 FN(neg3)(V3 a) { return FN(mk3)(-a.x, -a.y, -a.z); }
/* Warp dot: (x*x' + y*y') + z*z', no fused multiply-add. */
static inline REAL FN(dot3)(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline REAL FN(len3)(V3 a) { return SQRT(FN(dot3)(a, a)); }
static inline V3 FN(cross3)(V3 a, V3 b) {
    return FN(mk3)(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline REAL FN(sign)(REAL x) { return x < (REAL)0 ? (REAL)-1 : (REAL)1; }
static inline REAL FN(clampr)(REAL x, REAL a, REAL b) {
    REAL r = x < a ? a : x;
    return r > b ? b : r;
}

/* kernel/Kernel_KNL.py:7-23 -- Wendland C2, support 2h, evaluated left to right. */
REAL FN(wendland_w)(REAL r, REAL h) {
    REAL q = r / h;
    REAL w = (REAL)0;
    REAL u = (REAL)1.0 - (REAL)0.5 * q;
    if (q < (REAL)2.0)
        w = (REAL)21.0 * u * u * u * u * ((REAL)1.0 + (REAL)2.0 * q) /
            ((REAL)16.0 * (REAL)PI_D * h * h * h);
    return w;
}

/* kernel/Kernel_KNL.py:26-41 -- dW/dr. */
REAL FN(wendland_dw)(REAL r, REAL h) {
    REAL q = r / h;
    REAL dw = (REAL)0;
    if (q < (REAL)2.0) {
        REAL u = (REAL)1.0 - (REAL)0.5 * q;
        dw = (REAL)21.0 * ((REAL)-5.0 * q * u * u * u) /
             ((REAL)16.0 * (REAL)PI_D * h * h * h * h);
    }
    return dw;
}
#define WW FN(wendland_w)
#define DW FN(wendland_dw)

/* ------------------------------------------------------------------------ */
/* Hash grid.  Restates Warp's HashGrid (third-party warp-lang >= 1.17,
 * native/hashgrid.h; call sites main.py:66,73,85,87): bucket coordinates are
 * int(p * (1/w)) truncated toward zero with 1/w rounded to REAL, offset by
 * 2^20, clamped at 0, folded modulo the grid dimensions;
 * key = cz*dx*dy + cy*dx + cx.  Points are sorted stably by key. */
static inline int FN(bucket)(const int dims[3], int x, int y, int z) {
    const int origin = 1 << 20;
    x += origin; y += origin; z += origin;
    if (x < 0) x = 0;
    if (y < 0) y = 0;
    if (z < 0) z = 0;
    int cx = x % dims[0], cy = y % dims[1], cz = z % dims[2];
    return cz * (dims[0] * dims[1]) + cy * dims[0] + cx;
}

void FN(ora_grid_build)(int dim_x, int dim_y, int dim_z, REAL radius,
                        const REAL *pos, int n, REAL *cell_width_inv_out,
                        int *point_cells, int *point_ids,
                        int *cell_starts, int *cell_ends) {
    const int dims[3] = {dim_x, dim_y, dim_z};
    const REAL inv = (REAL)1.0 / radius;
    const int ncell = dim_x * dim_y * dim_z;
    int *key = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
    int *cnt = (int *)calloc((size_t)ncell + 1, sizeof(int));
    for (int i = 0; i < n; ++i) {
        key[i] = FN(bucket)(dims, (int)(pos[3 * i] * inv), (int)(pos[3 * i + 1] * inv),
                            (int)(pos[3 * i + 2] * inv));
        cnt[key[i] + 1]++;
    }
    for (int c = 0; c < ncell; ++c) cnt[c + 1] += cnt[c];
    /* stable counting sort: ascending original index inside each key */
    for (int i = 0; i < n; ++i) {
        int slot = cnt[key[i]]++;
        point_ids[slot] = i;
        point_cells[slot] = key[i];
    }
    for (int c = 0; c < ncell; ++c) { cell_starts[c] = 0; cell_ends[c] = 0; }
    for (int k = 0; k < n; ++k) {
        int c = point_cells[k];
        if (k == 0 || c != point_cells[k - 1]) cell_starts[c] = k;
        if (k == n - 1 || c != point_cells[k + 1]) cell_ends[c] = k + 1;
    }
    *cell_width_inv_out = inv;
    free(key);
    free(cnt);
}

/* Query iterator over the cell range [int((p-r)/w), int((p+r)/w)] per axis,
 * each end capped at start+dim-1; cells visited z-outer, y, x. */
typedef struct {
    const FN(ora_grid) *g;
    int x0, y0, z0, x1, y1, z1, x, y, z, cur, end;
} FN(query);

static inline void FN(q_cell)(FN(query) *q) {
    const int dims[3] = {q->g->dim_x, q->g->dim_y, q->g->dim_z};
    int c = FN(bucket)(dims, q->x, q->y, q->z);
    q->cur = q->g->cell_starts[c];
    q->end = q->g->cell_ends[c];
}

static inline void FN(q_init)(FN(query) *q, const FN(ora_grid) *g, V3 p, REAL r) {
    const REAL inv = g->cell_width_inv;
    q->g = g;
    q->x0 = (int)((p.x - r) * inv);
    q->y0 = (int)((p.y - r) * inv);
    q->z0 = (int)((p.z - r) * inv);
    int xe = (int)((p.x + r) * inv), ye = (int)((p.y + r) * inv), ze = (int)((p.z + r) * inv);
    q->x1 = xe < q->x0 + g->dim_x - 1 ? xe : q->x0 + g->dim_x - 1;
    q->y1 = ye < q->y0 + g->dim_y - 1 ? ye : q->y0 + g->dim_y - 1;
    q->z1 = ze < q->z0 + g->dim_z - 1 ? ze : q->z0 + g->dim_z - 1;
    q->x = q->x0; q->y = q->y0; q->z = q->z0;
    FN(q_cell)(q);
}

static inline int FN(q_next)(FN(query) *q, int *idx) {
    if (q->g->n == 0) return 0;
    for (;;) {
        if (q->cur < q->end) { *idx = q->g->point_ids[q->cur++]; return 1; }
        q->x++;
        if (q->x > q->x1) { q->x = q->x0; q->y++; }
        if (q->y > q->y1) { q->y = q->y0; q->z++; }
        if (q->z > q->z1) return 0;
        FN(q_cell)(q);
    }
}

/* Count true neighbours (r^2 < radius^2, self included) of every subject of
 * grid gs against grid gn: the neighbour-count parity view.  Output is at the
 * subject's original index. */
void FN(ora_neighbour_counts)(const FN(ora_grid) *gs, const REAL *spos,
                              const FN(ora_grid) *gn, const REAL *npos, REAL radius,
                              int *counts) {
    #pragma omp parallel for schedule(dynamic, 64)
    for (int tid = 0; tid < gs->n; ++tid) {
        int i = gs->point_ids[tid];
        V3 ri = FN(ld3)(spos, i);
        FN(query) q; FN(q_init)(&q, gn, ri, radius);
        int j, c = 0;
        while (FN(q_next)(&q, &j)) {
            V3 d = FN(sub3)(ri, FN(ld3)(npos, j));
            if (FN(dot3)(d, d) < radius * radius) c++;
        }
        counts[i] = c;
    }
}

/* ------------------------------------------------------------------------ */
/* kernel/KERNEL_rho.py:6-40 Kernel_shepard_sph */
void FN(ora_shepard_sph)(FN(ora_sph) *P, const FN(ora_bnd) *B, const FN(ora_grid) *gs,
                         const FN(ora_grid) *gb, REAL support, REAL h) {
    #pragma omp parallel for schedule(dynamic, 64)
    for (int tid = 0; tid < P->n; tid += ora_stride) {
        int i = gs->point_ids[tid];
        V3 ri = FN(ld3)(P->pos, i);
        REAL flt = 0;
        FN(query) q; int j;
        FN(q_init)(&q, gs, ri, support);
        while (FN(q_next)(&q, &j)) {
            V3 rij = FN(sub3)(ri, FN(ld3)(P->pos, j));
            REAL td = FN(dot3)(rij, rij);
            if (td < support * support) flt = flt + (P->m[j] / P->rho[j]) * WW(SQRT(td), h);
        }
        FN(q_init)(&q, gb, ri, support);
        while (FN(q_next)(&q, &j)) {
            V3 rij = FN(sub3)(ri, FN(ld3)(B->pos, j));
            REAL td = FN(dot3)(rij, rij);
            if (td < support * support) flt = flt + (B->m[j] / B->rho[j]) * WW(SQRT(td), h);
        }
        P->flt[i] = flt;
    }
}

/* kernel/KERNEL_rho.py:45-82 Kernel_density_sph */
void FN(ora_density_sph)(FN(ora_sph) *P, const FN(ora_bnd) *B, const FN(ora_grid) *gs,
                         const FN(ora_grid) *gb, REAL support, REAL h) {
    #pragma omp parallel for schedule(dynamic, 64)
    for (int tid = 0; tid < P->n; tid += ora_stride) {
        int i = gs->point_ids[tid];
        V3 ri = FN(ld3)(P->pos, i);
        REAL rho = 0;
        FN(query) q; int j;
        FN(q_init)(&q, gs, ri, support);
        while (FN(q_next)(&q, &j)) {
            V3 rij = FN(sub3)(ri, FN(ld3)(P->pos, j));
            REAL td = FN(dot3)(rij, rij);
            if (td < support * support) rho = rho + P->m[j] * WW(SQRT(td), h);
        }
        FN(q_init)(&q, gb, ri, support);
        while (FN(q_next)(&q, &j)) {
            V3 rij = FN(sub3)(ri, FN(ld3)(B->pos, j));
            REAL td = FN(dot3)(rij, rij);
            if (td < support * support) rho = rho + B->m[j] * WW(SQRT(td), h);
        }
        P->rho[i] = rho / P->flt[i];
    }
}

/* kernel/KERNEL_rho.py:87-126 Kernel_density_bnd (raw sum, no Shepard). */
void FN(ora_density_bnd)(const FN(ora_sph) *P, FN(ora_bnd) *B, const FN(ora_grid) *gs,
                         const FN(ora_grid) *gb, REAL support, REAL h) {
    #pragma omp parallel for schedule(dynamic, 64)
    for (int tid = 0; tid < B->n; tid += ora_stride) {
        int bi = gb->point_ids[tid];
        V3 rb = FN(ld3)(B->pos, bi);
        REAL rho = 0;
        FN(query) q; int j;
        FN(q_init)(&q, gs, rb, support);
        while (FN(q_next)(&q, &j)) {
            V3 d = FN(sub3)(rb, FN(ld3)(P->pos, j));
            REAL td = FN(dot3)(d, d);
            if (td < support * support) rho = rho + P->m[j] * WW(SQRT(td), h);
        }
        FN(q_init)(&q, gb, rb, support);
        while (FN(q_next)(&q, &j)) {
            V3 d = FN(sub3)(rb, FN(ld3)(B->pos, j));
            REAL td = FN(dot3)(d, d);
            if (td < support * support) rho = rho + B->m[j] * WW(SQRT(td), h);
        }
        B->rho[bi] = rho;
    }
}

/* kernel/KERNEL_pres.py:6-23 / 26-45 -- Tait EOS clamped at zero. */
void FN(ora_pres)(const REAL *rho, REAL *pres, int n, REAL rho0, REAL c0, REAL gamma) {
    const REAL B = rho0 * c0 * c0 / gamma;
    for (int i = 0; i < n; ++i) {
        REAL q = POW(rho[i] / rho0, gamma) - (REAL)1.0;
        pres[i] = B * (q > (REAL)0 ? q : (REAL)0);
    }
}

/* kernel/KERNEL_force.py:7-71 Kernel_force_sph */
void FN(ora_force_sph)(FN(ora_sph) *P, const FN(ora_bnd) *B, const FN(ora_grid) *gs,
                       const FN(ora_grid) *gb, REAL support, REAL h, REAL mu, REAL g) {
    const REAL r2min = (REAL)1.0e-12;
    #pragma omp parallel for schedule(dynamic, 64)
    for (int tid = 0; tid < P->n; tid += ora_stride) {
        int i = gs->point_ids[tid];
        V3 ri = FN(ld3)(P->pos, i), vi = FN(ld3)(P->vel, i);
        REAL rhoi = P->rho[i], presi = P->pres[i];
        V3 acc = FN(mk3)(0, 0, 0);
        FN(query) q; int j;
        FN(q_init)(&q, gs, ri, support);
        while (FN(q_next)(&q, &j)) {
            V3 rij = FN(sub3)(ri, FN(ld3)(P->pos, j));
            REAL td = FN(dot3)(rij, rij);
            if (td > r2min && td < support * support) {
                REAL r = SQRT(td);
                V3 dwij = FN(mul3)(DW(r, h), FN(div3)(rij, r));
                REAL rhoj = P->rho[j], presj = P->pres[j];
                acc = FN(sub3)(acc, FN(mul3)(P->m[j] * (presi / (rhoi * rhoi) + presj / (rhoj * rhoj)), dwij));
                REAL visc = (REAL)2.0 * mu * FN(dot3)(rij, dwij) /
                            (rhoi * rhoj * (td + (REAL)0.01 * h * h));
                acc = FN(add3)(acc, FN(mul3)(P->m[j] * visc, FN(sub3)(vi, FN(ld3)(P->vel, j))));
            }
        }
        FN(q_init)(&q, gb, ri, support);
        while (FN(q_next)(&q, &j)) {
            V3 rij = FN(sub3)(ri, FN(ld3)(B->pos, j));
            REAL td = FN(dot3)(rij, rij);
            if (td > r2min && td < support * support) {
                REAL r = SQRT(td);
                V3 dwij = FN(mul3)(DW(r, h), FN(div3)(rij, r));
                REAL rhoj = B->rho[j], presj = B->pres[j];
                acc = FN(sub3)(acc, FN(mul3)(B->m[j] * (presi / (rhoi * rhoi) + presj / (rhoj * rhoj)), dwij));
                REAL visc = (REAL)2.0 * mu * FN(dot3)(rij, dwij) /
                            (rhoi * rhoj * (td + (REAL)0.01 * h * h));
                acc = FN(add3)(acc, FN(mul3)(B->m[j] * visc, FN(sub3)(vi, FN(ld3)(B->vel, j))));
            }
        }
        FN(st3)(P->acc, i, FN(add3)(acc, FN(mk3)(0, 0, -g)));
    }
}

/* kernel/KERNEL_step.py:6-22 Kernel_step_sph */
void FN(ora_step_sph)(FN(ora_sph) *P, REAL dt) {
    for (int i = 0; i < P->n; ++i) {
        V3 v = FN(add3)(FN(ld3)(P->vel, i), FN(mul3r)(FN(ld3)(P->acc, i), dt));
        FN(st3)(P->vel, i, v);
        FN(st3)(P->pos, i, FN(add3)(FN(ld3)(P->pos, i), FN(mul3r)(v, dt)));
    }
}

/* ------------------------------------------------------------------------ */
/* kernel/KERNEL_DEM_force.py:25-74 DEM_contact */
static void FN(dem_contact)(V3 normal, REAL overlap, REAL ra, REAL rb, V3 va, V3 vb,
                            V3 wa, V3 wb, V3 tangent, REAL K, REAL eta, REAL mu, REAL dt,
                            V3 *f_out, V3 *disp_out, V3 *tq_out) {
    V3 rc_a = FN(mul3)(ra - (REAL)0.5 * overlap, normal);
    V3 rc_b = FN(mul3)(-(rb - (REAL)0.5 * overlap), normal);
    V3 urel = FN(sub3)(FN(sub3)(FN(add3)(vb, FN(cross3)(wb, rc_b)), va), FN(cross3)(wa, rc_a));
    V3 un = FN(mul3)(FN(dot3)(urel, normal), normal);
    V3 us = FN(sub3)(urel, un);
    V3 disp = FN(add3)(tangent, FN(mul3r)(us, dt));
    V3 dir = FN(div3)(us, FN(len3)(us) + (REAL)1.0e-20);
    V3 fn = FN(add3)(FN(mul3)(-K * overlap, normal), FN(mul3)(eta, un));
    V3 fss = FN(add3)(FN(mul3)(K, disp), FN(mul3)(eta, us));
    V3 fsf = FN(mul3)(mu * K * overlap, dir);
    V3 fs = fss;
    if (FN(len3)(fss) >= FN(len3)(fsf))
        fs = FN(mk3)(FN(sign)(fss.x) * FABS(fsf.x), FN(sign)(fss.y) * FABS(fsf.y),
                     FN(sign)(fss.z) * FABS(fsf.z));
    *f_out = FN(add3)(fn, fs);
    *disp_out = disp;
    *tq_out = FN(cross3)(rc_a, fs);
}

/* kernel/KERNEL_DEM_force.py:9-22 Kernel_count_dem_contacts */
void FN(ora_count_dem_contacts)(FN(ora_dem) *D, const FN(ora_grid) *gd, REAL radius_max) {
    #pragma omp parallel for schedule(dynamic, 16)
    for (int a = 0; a < D->n; ++a) {
        V3 ra = FN(ld3)(D->pos, a);
        int count = 0, b;
        FN(query) q; FN(q_init)(&q, gd, ra, D->radius[a] + radius_max);
        while (FN(q_next)(&q, &b)) {
            if (b != a) {
                REAL dist = FN(len3)(FN(sub3)(FN(ld3)(D->pos, b), ra));
                if (D->radius[a] + D->radius[b] - dist > (REAL)0) count++;
            }
        }
        D->contact_dem_offset_new[a] = count;
    }
}

/* kernel/KERNEL_DEM_force.py:77-135 Kernel_force_dem */
void FN(ora_force_dem)(FN(ora_dem) *D, const FN(ora_grid) *gd, REAL radius_max, REAL g, REAL dt) {
    #pragma omp parallel for schedule(dynamic, 16)
    for (int a = 0; a < D->n; ++a) {
        V3 ra = FN(ld3)(D->pos, a);
        V3 force = FN(mk3)(0, 0, -D->m[a] * g), torque = FN(mk3)(0, 0, 0);
        int out = D->contact_dem_offset_new[a], b;
        FN(query) q; FN(q_init)(&q, gd, ra, D->radius[a] + radius_max);
        while (FN(q_next)(&q, &b)) {
            if (b == a) continue;
            V3 rab = FN(sub3)(FN(ld3)(D->pos, b), ra);
            REAL dist = FN(len3)(rab);
            REAL overlap = D->radius[a] + D->radius[b] - dist;
            if (overlap > (REAL)0) {
                V3 tangent = FN(mk3)(0, 0, 0);
                for (int k = D->contact_dem_offset_old[a]; k < D->contact_dem_offset_old[a + 1]; ++k)
                    if (D->contact_dem_id_old[k] == b) tangent = FN(ld3)(D->tang_dem_old, k);
                V3 normal = FN(mk3)(1, 0, 0);
                if (a > b) normal = FN(neg3)(normal);
                if (dist > (REAL)1.0e-12) normal = FN(div3)(rab, dist);
                V3 f, disp, tq;
                FN(dem_contact)(normal, overlap, D->radius[a], D->radius[b], FN(ld3)(D->vel, a),
                                FN(ld3)(D->vel, b), FN(ld3)(D->omega, a), FN(ld3)(D->omega, b),
                                tangent, D->K[a], D->eta[a], D->mu[a], dt, &f, &disp, &tq);
                force = FN(add3)(force, f);
                torque = FN(add3)(torque, tq);
                D->contact_dem_id_new[out] = b;
                FN(st3)(D->tang_dem_new, out, disp);
                out++;
            }
        }
        FN(st3)(D->force, a, force);
        FN(st3)(D->torque, a, torque);
    }
}

/* kernel/KERNEL_DEM_BC.py:10-24 Kernel_count_bnd_contacts */
void FN(ora_count_bnd_contacts)(FN(ora_dem) *D, const FN(ora_dembnd) *W, const FN(ora_grid) *gw,
                                REAL radius_max) {
    #pragma omp parallel for schedule(dynamic, 16)
    for (int a = 0; a < D->n; ++a) {
        V3 ra = FN(ld3)(D->pos, a);
        int count = 0, b;
        FN(query) q; FN(q_init)(&q, gw, ra, D->radius[a] + radius_max);
        while (FN(q_next)(&q, &b)) {
            REAL dist = FN(len3)(FN(sub3)(FN(ld3)(W->pos, b), ra));
            if (D->radius[a] + W->radius[b] - dist > (REAL)0) count++;
        }
        D->contact_bnd_offset_new[a] = count;
    }
}

/* kernel/KERNEL_DEM_BC.py:27-84 Kernel_bc_dem */
void FN(ora_bc_dem)(FN(ora_dem) *D, const FN(ora_dembnd) *W, const FN(ora_grid) *gw,
                    REAL radius_max, REAL dt) {
    #pragma omp parallel for schedule(dynamic, 16)
    for (int a = 0; a < D->n; ++a) {
        V3 ra = FN(ld3)(D->pos, a);
        V3 force = FN(mk3)(0, 0, 0), torque = FN(mk3)(0, 0, 0);
        int out = D->contact_bnd_offset_new[a], b;
        FN(query) q; FN(q_init)(&q, gw, ra, D->radius[a] + radius_max);
        while (FN(q_next)(&q, &b)) {
            V3 rab = FN(sub3)(FN(ld3)(W->pos, b), ra);
            REAL dist = FN(len3)(rab);
            REAL overlap = D->radius[a] + W->radius[b] - dist;
            if (overlap > (REAL)0) {
                V3 tangent = FN(mk3)(0, 0, 0);
                for (int k = D->contact_bnd_offset_old[a]; k < D->contact_bnd_offset_old[a + 1]; ++k)
                    if (D->contact_bnd_id_old[k] == b) tangent = FN(ld3)(D->tang_bnd_old, k);
                V3 normal = FN(mk3)(0, 0, -1);
                if (dist > (REAL)1.0e-12) normal = FN(div3)(rab, dist);
                V3 f, disp, tq;
                FN(dem_contact)(normal, overlap, D->radius[a], W->radius[b], FN(ld3)(D->vel, a),
                                FN(ld3)(W->vel, b), FN(ld3)(D->omega, a), FN(ld3)(W->omega, b),
                                tangent, D->K[a], D->eta[a], D->mu[a], dt, &f, &disp, &tq);
                force = FN(add3)(force, f);
                torque = FN(add3)(torque, tq);
                D->contact_bnd_id_new[out] = b;
                FN(st3)(D->tang_bnd_new, out, disp);
                out++;
            }
        }
        FN(st3)(D->force, a, FN(add3)(FN(ld3)(D->force, a), force));
        FN(st3)(D->torque, a, FN(add3)(FN(ld3)(D->torque, a), torque));
    }
}

/* kernel/KERNEL_DEM_step.py:6-30 Kernel_step_dem */
void FN(ora_step_dem)(FN(ora_dem) *D, REAL dt) {
    for (int a = 0; a < D->n; ++a) {
        V3 acc = FN(div3)(FN(ld3)(D->force, a), D->m[a]);
        V3 vel = FN(add3)(FN(ld3)(D->vel, a), FN(mul3r)(acc, dt));
        FN(st3)(D->acc, a, acc);
        FN(st3)(D->vel, a, vel);
        FN(st3)(D->pos, a, FN(add3)(FN(ld3)(D->pos, a), FN(mul3r)(vel, dt)));
        FN(st3)(D->omega, a, FN(add3)(FN(ld3)(D->omega, a),
                                      FN(mul3r)(FN(div3)(FN(ld3)(D->torque, a), D->inertia[a]), dt)));
    }
}

/* Exclusive scan in place over n+1 entries (wp.utils.array_scan,
 * Simulation.py:148-149): out[k] = sum(in[0..k-1]). */
void FN(ora_exclusive_scan)(int *a, int n) {
    int run = 0;
    for (int k = 0; k < n; ++k) { int v = a[k]; a[k] = run; run += v; }
}

/* ------------------------------------------------------------------------ */
/* kernel/KERNEL_SPHDEM_interaction.py:11-40 DEM_drag_beta */
REAL FN(ora_drag_beta)(REAL eps, REAL rho, REAL mu, REAL d, REAL speed) {
    REAL beta = 0;
    REAL re = rho * d * speed / (mu + (REAL)1.0e-20);
    if (eps <= (REAL)0.8) {
        beta = (REAL)150.0 * ((REAL)1.0 - eps) / eps * mu / (d * d) + (REAL)1.75 * rho * speed / d;
    } else {
        REAL cd_speed = (REAL)0.44 * speed;
        if (re <= (REAL)1000.0)
            cd_speed = (REAL)24.0 * mu / (rho * d) * ((REAL)1.0 + (REAL)0.15 * POW(re, (REAL)0.687));
        beta = (REAL)0.75 * cd_speed * rho / d * POW(eps, (REAL)-1.65);
    }
    return beta;
}

/* kernel/KERNEL_SPHDEM_interaction.py:43-115 Kernel_prep_sphdem.
 * Restatement note: when no DEM lies within 2*h_b of the subject, solid = 0
 * and the porosity is clamp(1) whatever the fluid sum is; that sum is a local
 * never stored, so it is skipped then.  Outputs are unchanged (checked by a
 * test that runs both ways, skip_fluid_sum = 0/1). */
void FN(ora_prep_sphdem)(FN(ora_sph) *P, const FN(ora_bnd) *B, const FN(ora_dem) *D,
                         const FN(ora_grid) *gs, const FN(ora_grid) *gb, const FN(ora_grid) *gd,
                         REAL support, REAL h, REAL sph_support, REAL sph_h,
                         REAL eps_min, REAL eps_max, int skip_fluid_sum) {
    const REAL r2min = (REAL)1.0e-12;
    #pragma omp parallel for schedule(dynamic, 32)
    for (int tid = 0; tid < P->n; tid += ora_stride) {
        int i = gs->point_ids[tid];
        V3 ri = FN(ld3)(P->pos, i);
        REAL flt = 0, solid = 0;
        V3 pgf = FN(mk3)(0, 0, 0);
        FN(query) q; int j, b;
        int near = 1;
        if (skip_fluid_sum) {
            near = 0;
            FN(q_init)(&q, gd, ri, support);
            while (FN(q_next)(&q, &b)) {
                REAL dist = FN(len3)(FN(sub3)(ri, FN(ld3)(D->pos, b)));
                if (dist < (REAL)2.0 * D->h[b]) near = 1;
            }
        }
        if (near) {
            if (h > (REAL)0) {
                FN(q_init)(&q, gs, ri, support);
                while (FN(q_next)(&q, &j)) {
                    REAL dist = FN(len3)(FN(sub3)(ri, FN(ld3)(P->pos, j)));
                    if (dist < support) flt = flt + P->m[j] / P->rho[j] * WW(dist, h);
                }
            }
            FN(q_init)(&q, gd, ri, support);
            while (FN(q_next)(&q, &b)) {
                REAL hb = D->h[b];
                REAL dist = FN(len3)(FN(sub3)(ri, FN(ld3)(D->pos, b)));
                if (dist < (REAL)2.0 * hb) {
                    REAL lflt = flt;
                    if (h <= (REAL)0) {
                        lflt = 0;
                        FN(query) q2; FN(q_init)(&q2, gs, ri, (REAL)2.0 * hb);
                        while (FN(q_next)(&q2, &j)) {
                            REAL fd = FN(len3)(FN(sub3)(ri, FN(ld3)(P->pos, j)));
                            if (fd < (REAL)2.0 * hb) lflt = lflt + P->m[j] / P->rho[j] * WW(fd, hb);
                        }
                    }
                    solid = solid + D->volume[b] * WW(dist, hb) / (lflt + (REAL)1.0e-20);
                }
            }
        }
        FN(q_init)(&q, gs, ri, sph_support);
        while (FN(q_next)(&q, &j)) {
            V3 rij = FN(sub3)(ri, FN(ld3)(P->pos, j));
            REAL r2 = FN(dot3)(rij, rij);
            if (r2 > r2min && r2 < sph_support * sph_support) {
                REAL dist = SQRT(r2);
                V3 dwij = FN(div3)(FN(mul3)(DW(dist, sph_h), rij), dist);
                pgf = FN(sub3)(pgf, FN(mul3)(P->m[j] / P->rho[j] * (P->pres[j] - P->pres[i]), dwij));
            }
        }
        FN(q_init)(&q, gb, ri, sph_support);
        while (FN(q_next)(&q, &j)) {
            V3 rij = FN(sub3)(ri, FN(ld3)(B->pos, j));
            REAL r2 = FN(dot3)(rij, rij);
            if (r2 > r2min && r2 < sph_support * sph_support) {
                REAL dist = SQRT(r2);
                V3 dwij = FN(div3)(FN(mul3)(DW(dist, sph_h), rij), dist);
                pgf = FN(sub3)(pgf, FN(mul3)(B->m[j] / B->rho[j] * (B->pres[j] - P->pres[i]), dwij));
            }
        }
        P->porosity[i] = FN(clampr)((REAL)1.0 - solid, eps_min, eps_max);
        FN(st3)(P->pgf, i, pgf);
    }
}

/* kernel/KERNEL_SPHDEM_interaction.py:118-187 Kernel_interaction_dem */
void FN(ora_interaction_dem)(const FN(ora_sph) *P, FN(ora_dem) *D, const FN(ora_grid) *gs,
                             const FN(ora_grid) *gd, REAL mu, REAL eps_min, REAL eps_max,
                             REAL dt) {
    #pragma omp parallel for schedule(dynamic, 1)
    for (int tid = 0; tid < D->n; ++tid) {
        int a = gd->point_ids[tid];
        V3 ra = FN(ld3)(D->pos, a);
        REAL h = D->h[a], support = (REAL)2.0 * h;
        REAL flt = 0, solid = 0, rho_f = 0;
        V3 vel_f = FN(mk3)(0, 0, 0), pgf = FN(mk3)(0, 0, 0);
        FN(query) q; int j, b;
        FN(q_init)(&q, gs, ra, support);
        while (FN(q_next)(&q, &j)) {
            REAL dist = FN(len3)(FN(sub3)(ra, FN(ld3)(P->pos, j)));
            if (dist < support) {
                REAL vw = P->m[j] / P->rho[j] * WW(dist, h);
                flt = flt + vw;
                rho_f = rho_f + P->rho[j] * vw;
                vel_f = FN(add3)(vel_f, FN(mul3r)(FN(ld3)(P->vel, j), vw));
                pgf = FN(add3)(pgf, FN(mul3r)(FN(ld3)(P->pgf, j), vw));
            }
        }
        FN(q_init)(&q, gd, ra, support);
        while (FN(q_next)(&q, &b)) {
            REAL dist = FN(len3)(FN(sub3)(ra, FN(ld3)(D->pos, b)));
            if (dist < support) solid = solid + D->volume[b] * WW(dist, h);
        }
        D->flt_s[a] = flt;
        FN(st3)(D->drag, a, FN(mk3)(0, 0, 0));
        FN(st3)(D->pressure_force, a, FN(mk3)(0, 0, 0));
        D->porosity[a] = eps_max;
        if (flt > (REAL)1.0e-12) {
            REAL eps = FN(clampr)((REAL)1.0 - solid / flt, eps_min, eps_max);
            V3 urel = FN(sub3)(FN(div3)(vel_f, flt), FN(ld3)(D->vel, a));
            REAL beta = FN(ora_drag_beta)(eps, rho_f / flt, mu, (REAL)2.0 * D->radius[a], FN(len3)(urel));
            REAL coef = beta * D->volume[a] / D->m[a];
            V3 drag = FN(mul3)(D->m[a] * coef / ((REAL)1.0 + dt * coef), urel);
            V3 pf = FN(div3)(FN(mul3)(D->m[a] / D->rho[a], pgf), flt);
            D->porosity[a] = eps;
            FN(st3)(D->drag, a, drag);
            FN(st3)(D->pressure_force, a, pf);
            FN(st3)(D->force, a, FN(add3)(FN(add3)(FN(ld3)(D->force, a), drag), pf));
        }
    }
}

/* kernel/KERNEL_SPHDEM_interaction.py:190-226 Kernel_interaction_sph */
void FN(ora_interaction_sph)(FN(ora_sph) *P, const FN(ora_dem) *D, const FN(ora_grid) *gs,
                             const FN(ora_grid) *gd, REAL support) {
    #pragma omp parallel for schedule(dynamic, 64)
    for (int tid = 0; tid < P->n; tid += ora_stride) {
        int i = gs->point_ids[tid];
        V3 ri = FN(ld3)(P->pos, i);
        V3 fd = FN(mk3)(0, 0, 0);
        FN(query) q; int b;
        FN(q_init)(&q, gd, ri, support);
        while (FN(q_next)(&q, &b)) {
            REAL dist = FN(len3)(FN(sub3)(ri, FN(ld3)(D->pos, b)));
            if (dist < (REAL)2.0 * D->h[b] && D->flt_s[b] > (REAL)1.0e-12)
                fd = FN(sub3)(fd, FN(div3)(FN(mul3r)(FN(ld3)(D->drag, b), WW(dist, D->h[b])), D->flt_s[b]));
        }
        V3 acc = FN(div3)(fd, P->rho[i] * P->porosity[i]);
        FN(st3)(P->acc_dem, i, acc);
        FN(st3)(P->acc, i, FN(add3)(FN(ld3)(P->acc, i), acc));
    }
}

/* The Eulerian ISPH cavity (same helpers, same precision). */
#include "eisph_impl.h"

#undef V3
#undef WW
#undef DW
#undef FN
#undef CAT
#undef CAT_
