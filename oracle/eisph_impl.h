/*
 * eisph_impl.h -- CPU oracle of the Eulerian ISPH cavity step, included at
 * the end of oracle_impl.h, once per precision (REAL, SFX, V3 and the grid
 * query are defined there).  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates kernel/KERNEL_EISPH_{KNL,BC,force,PPE,step}.py: fixed points, so
 * every neighbour sum walks the same HashGrid query as the reference (fluid
 * grid first, then the ghost grid), subjects in the grid's sorted order,
 * expressions evaluated left to right without fused multiply-add.
 *
 * How each part is pinned (tests/test_eisph_oracle.py):
 *   wendland2_w / _dw   unit integral over the 2h disk, finite-difference
 *                       slope, closed-form values.
 *   eisph_kgc           first-order consistency: with the correction, the
 *                       gradient of any linear pressure field is exact at
 *                       every fluid point (ghosts complete the support).
 *   laplace weight      the reference's sign case (A_ij < 0 for an inward
 *                       gradient) and the -4 nu diffusion of u = -r^2 at the
 *                       cavity centre (test_eisph.py:144-181).
 *   build/solve PPE     b = rho0 div/dt, constant-pressure Aijpj = p Aij,
 *                       the one-pass diagonal update and its shift
 *                       invariance (test_eisph.py:112-142, 266-320).
 *   whole step          Ghia et al. (1982) Re=100 centre lines after 5000
 *                       steps, and the figures the reference publishes for
 *                       that run (README.md:79-83).
 */

/* kernel/KERNEL_EISPH_KNL.py:14-31 -- 2-D Wendland C2, support 2h. */
REAL FN(wendland2_w)(REAL r, REAL h) {
    REAL q = r / h;
    REAL v = (REAL)0;
    if (q < (REAL)2.0) {
        REAL u = (REAL)1.0 - (REAL)0.5 * q;
        v = (REAL)7.0 * u * u * u * u * ((REAL)1.0 + (REAL)2.0 * q) /
            ((REAL)4.0 * (REAL)PI_D * h * h);
    }
    return v;
}

/* kernel/KERNEL_EISPH_KNL.py:34-49 -- dW/dr. */
REAL FN(wendland2_dw)(REAL r, REAL h) {
    REAL q = r / h;
    REAL v = (REAL)0;
    if (q < (REAL)2.0) {
        REAL u = (REAL)1.0 - (REAL)0.5 * q;
        v = (REAL)-35.0 * q * u * u * u / ((REAL)4.0 * (REAL)PI_D * h * h * h);
    }
    return v;
}

/* KERNEL_EISPH_KNL.py:10 R2_EPS: the subject itself is excluded. */
#define E_R2_EPS ((REAL)1.0e-12)

/* KERNEL_EISPH_KNL.py:52-56 -- uncorrected grad_i W_ij. */
static inline V3 FN(e_grad)(V3 rij, REAL r2, REAL h) {
    REAL r = SQRT(r2);
    return FN(div3)(FN(mul3)(FN(wendland2_dw)(r, h), rij), r);
}

/* KERNEL_EISPH_KNL.py:59-64 -- the subject's 2x2 correction, x-z plane. */
static inline V3 FN(e_corr)(const REAL *c, V3 g) {
    return FN(mk3)(c[0] * g.x + c[1] * g.z, (REAL)0.0, c[2] * g.x + c[3] * g.z);
}

/* KERNEL_EISPH_KNL.py:67-71 -- signed PDF Laplacian weight A_ij. */
static inline REAL FN(e_lap)(REAL vol, V3 rij, V3 g, REAL r2) {
    return (REAL)2.0 * vol * FN(dot3)(rij, g) / r2;
}

typedef struct {
    V3 rij;
    REAL r2;
    int ok;
} FN(e_pair);

static inline FN(e_pair) FN(e_pair_of)(V3 ri, const REAL *pos, int j, REAL support) {
    FN(e_pair) e;
    e.rij = FN(sub3)(ri, FN(ld3)(pos, j));
    e.r2 = FN(dot3)(e.rij, e.rij);
    e.ok = e.r2 > E_R2_EPS && e.r2 < support * support;
    return e;
}

/* KERNEL_EISPH_KNL.py:74-132 Kernel_prepare_kgc */
void FN(ora_eisph_kgc)(FN(ora_eisph) *P, const FN(ora_eisph) *B, const FN(ora_grid) *gs,
                       const FN(ora_grid) *gb, REAL support, REAL h) {
    #pragma omp parallel for schedule(dynamic, 64)
    for (int tid = 0; tid < P->n; ++tid) {
        int i = gs->point_ids[tid];
        V3 ri = FN(ld3)(P->pos, i);
        REAL m00 = 0, m01 = 0, m10 = 0, m11 = 0;
        FN(query) q; int j;
        for (int side = 0; side < 2; ++side) {
            const FN(ora_eisph) *S = side == 0 ? (const FN(ora_eisph) *)P : B;
            FN(q_init)(&q, side == 0 ? gs : gb, ri, support);
            while (FN(q_next)(&q, &j)) {
                FN(e_pair) e = FN(e_pair_of)(ri, S->pos, j, support);
                if (!e.ok) continue;
                V3 g = FN(e_grad)(e.rij, e.r2, h);
                REAL vol = S->m[j] / S->rho[j];
                m00 = m00 - vol * e.rij.x * g.x;
                m01 = m01 - vol * e.rij.x * g.z;
                m10 = m10 - vol * e.rij.z * g.x;
                m11 = m11 - vol * e.rij.z * g.z;
            }
        }
        REAL det = m00 * m11 - m01 * m10;
        REAL c[4] = {1, 0, 0, 1};
        if (FABS(det) > (REAL)1.0e-8) {
            REAL inv = (REAL)1.0 / det;
            c[0] = m11 * inv; c[1] = -m01 * inv; c[2] = -m10 * inv; c[3] = m00 * inv;
        }
        for (int k = 0; k < 4; ++k) P->kgc[4 * i + k] = c[k];
    }
}

/* KERNEL_EISPH_BC.py:6-31 Kernel_Dirichlet_BC (predicted selects vel_star). */
void FN(ora_eisph_dirichlet)(const FN(ora_eisph) *P, FN(ora_eisph) *B, int predicted) {
    const REAL *src = predicted ? P->vel_star : P->vel;
    REAL *dst = predicted ? B->vel_star : B->vel;
    for (int b = 0; b < B->n; ++b) {
        V3 v = FN(sub3)(FN(mul3)((REAL)2.0, FN(ld3)(B->vel_bc, b)), FN(ld3)(src, B->mirror[b]));
        FN(st3)(dst, b, v);
    }
}

/* KERNEL_EISPH_BC.py:34-47 Kernel_pressure_boundary_eisph */
void FN(ora_eisph_pressure_boundary)(const FN(ora_eisph) *P, FN(ora_eisph) *B) {
    for (int b = 0; b < B->n; ++b) B->pres[b] = P->pres[B->mirror[b]];
}

/* KERNEL_EISPH_force.py:12-72 Kernel_advec_vis_eisph */
void FN(ora_eisph_advec_vis)(FN(ora_eisph) *P, const FN(ora_eisph) *B, const FN(ora_grid) *gs,
                             const FN(ora_grid) *gb, REAL support, REAL h, REAL nu, REAL dt) {
    #pragma omp parallel for schedule(dynamic, 64)
    for (int tid = 0; tid < P->n; ++tid) {
        int i = gs->point_ids[tid];
        V3 ri = FN(ld3)(P->pos, i), vi = FN(ld3)(P->vel, i);
        const REAL *C = P->kgc + 4 * i;
        V3 conv = FN(mk3)(0, 0, 0), lap = FN(mk3)(0, 0, 0);
        FN(query) q; int j;
        for (int side = 0; side < 2; ++side) {
            const FN(ora_eisph) *S = side == 0 ? (const FN(ora_eisph) *)P : B;
            FN(q_init)(&q, side == 0 ? gs : gb, ri, support);
            while (FN(q_next)(&q, &j)) {
                FN(e_pair) e = FN(e_pair_of)(ri, S->pos, j, support);
                if (!e.ok) continue;
                V3 g = FN(e_corr)(C, FN(e_grad)(e.rij, e.r2, h));
                REAL vol = S->m[j] / S->rho[j];
                V3 dv = FN(sub3)(FN(ld3)(S->vel, j), vi);
                conv = FN(add3)(conv, FN(mul3)(vol * FN(dot3)(vi, g), dv));
                REAL coef = FN(e_lap)(vol, e.rij, g, e.r2);
                lap = FN(sub3)(lap, FN(mul3)(coef, dv));
            }
        }
        V3 acc = FN(add3)(FN(neg3)(conv), FN(mul3)(nu, lap));
        FN(st3)(P->acc, i, acc);
        FN(st3)(P->vel_star, i, FN(add3)(vi, FN(mul3)(dt, acc)));
    }
}

/* KERNEL_EISPH_PPE.py:12-79 Kernel_build_ppe_eisph */
void FN(ora_eisph_build_ppe)(FN(ora_eisph) *P, const FN(ora_eisph) *B, const FN(ora_grid) *gs,
                             const FN(ora_grid) *gb, REAL support, REAL h, REAL rho0, REAL dt) {
    #pragma omp parallel for schedule(dynamic, 64)
    for (int tid = 0; tid < P->n; ++tid) {
        int i = gs->point_ids[tid];
        V3 ri = FN(ld3)(P->pos, i), vi = FN(ld3)(P->vel_star, i);
        const REAL *C = P->kgc + 4 * i;
        REAL div = 0, A = 0, Ap = 0;
        FN(query) q; int j;
        for (int side = 0; side < 2; ++side) {
            const FN(ora_eisph) *S = side == 0 ? (const FN(ora_eisph) *)P : B;
            FN(q_init)(&q, side == 0 ? gs : gb, ri, support);
            while (FN(q_next)(&q, &j)) {
                FN(e_pair) e = FN(e_pair_of)(ri, S->pos, j, support);
                if (!e.ok) continue;
                V3 g = FN(e_corr)(C, FN(e_grad)(e.rij, e.r2, h));
                REAL vol = S->m[j] / S->rho[j];
                div = div + vol * FN(dot3)(FN(sub3)(FN(ld3)(S->vel_star, j), vi), g);
                REAL coef = FN(e_lap)(vol, e.rij, g, e.r2);
                A = A + coef;
                /* a ghost contributes its mirrored fluid point's p^t */
                REAL pj = side == 0 ? P->pres[j] : P->pres[B->mirror[j]];
                Ap = Ap + coef * pj;
            }
        }
        P->divergence[i] = div;
        P->Aij[i] = A;
        P->bi[i] = rho0 * div / dt;
        P->Aijpj[i] = Ap;
    }
}

/* EISPH.py:70-73 the RHS sum (in FP64, index order) into residual[0], then
 * KERNEL_EISPH_PPE.py:82-96 Kernel_remove_bi_mean. */
void FN(ora_eisph_remove_bi_mean)(FN(ora_eisph) *P) {
    double s = 0.0;
    for (int i = 0; i < P->n; ++i) s += (double)P->bi[i];
    P->residual[0] = (REAL)s;
    REAL mean = P->residual[0] / (REAL)P->n;
    for (int i = 0; i < P->n; ++i) P->bi[i] = P->bi[i] - mean;
}

/* KERNEL_EISPH_PPE.py:99-125 Kernel_solve_ppe_eisph */
void FN(ora_eisph_solve_ppe)(FN(ora_eisph) *P) {
    REAL gauge = 0;
    REAL gA = P->Aij[0];
    if (FABS(gA) > (REAL)1.0e-12) gauge = (P->bi[0] + P->Aijpj[0]) / gA;
    for (int i = 0; i < P->n; ++i) {
        REAL p = P->pres[i];
        REAL A = P->Aij[i];
        if (FABS(A) > (REAL)1.0e-12) p = (P->bi[i] + P->Aijpj[i]) / A;
        P->pres[i] = p - gauge;
    }
}

/* KERNEL_EISPH_step.py:12-61 Kernel_correct_eisph */
void FN(ora_eisph_correct)(FN(ora_eisph) *P, const FN(ora_eisph) *B, const FN(ora_grid) *gs,
                           const FN(ora_grid) *gb, REAL support, REAL h, REAL dt) {
    #pragma omp parallel for schedule(dynamic, 64)
    for (int tid = 0; tid < P->n; ++tid) {
        int i = gs->point_ids[tid];
        V3 ri = FN(ld3)(P->pos, i);
        REAL pi = P->pres[i];
        const REAL *C = P->kgc + 4 * i;
        V3 gp = FN(mk3)(0, 0, 0);
        FN(query) q; int j;
        for (int side = 0; side < 2; ++side) {
            const FN(ora_eisph) *S = side == 0 ? (const FN(ora_eisph) *)P : B;
            FN(q_init)(&q, side == 0 ? gs : gb, ri, support);
            while (FN(q_next)(&q, &j)) {
                FN(e_pair) e = FN(e_pair_of)(ri, S->pos, j, support);
                if (!e.ok) continue;
                V3 g = FN(e_corr)(C, FN(e_grad)(e.rij, e.r2, h));
                REAL vol = S->m[j] / S->rho[j];
                gp = FN(add3)(gp, FN(mul3)(vol * (S->pres[j] - pi), g));
            }
        }
        V3 old = FN(ld3)(P->vel, i);
        V3 v = FN(sub3)(FN(ld3)(P->vel_star, i), FN(div3)(FN(mul3)(dt, gp), P->rho[i]));
        FN(st3)(P->vel, i, v);
        FN(st3)(P->acc, i, FN(div3)(FN(sub3)(v, old), dt));
    }
}

/* KERNEL_EISPH_step.py:64-121 Kernel_diagnostics_eisph */
void FN(ora_eisph_diagnostics)(FN(ora_eisph) *P, const FN(ora_eisph) *B, const FN(ora_grid) *gs,
                               const FN(ora_grid) *gb, REAL support, REAL h) {
    #pragma omp parallel for schedule(dynamic, 64)
    for (int tid = 0; tid < P->n; ++tid) {
        int i = gs->point_ids[tid];
        V3 ri = FN(ld3)(P->pos, i), vi = FN(ld3)(P->vel, i);
        REAL pi = P->pres[i];
        const REAL *C = P->kgc + 4 * i;
        REAL div = 0, op = 0;
        FN(query) q; int j;
        for (int side = 0; side < 2; ++side) {
            const FN(ora_eisph) *S = side == 0 ? (const FN(ora_eisph) *)P : B;
            FN(q_init)(&q, side == 0 ? gs : gb, ri, support);
            while (FN(q_next)(&q, &j)) {
                FN(e_pair) e = FN(e_pair_of)(ri, S->pos, j, support);
                if (!e.ok) continue;
                V3 g = FN(e_corr)(C, FN(e_grad)(e.rij, e.r2, h));
                REAL vol = S->m[j] / S->rho[j];
                div = div + vol * FN(dot3)(FN(sub3)(FN(ld3)(S->vel, j), vi), g);
                REAL coef = FN(e_lap)(vol, e.rij, g, e.r2);
                op = op + coef * (pi - S->pres[j]);
            }
        }
        P->divergence[i] = div;
        P->residual[i] = op - P->bi[i];
    }
}

/* source/EISPH.py:26-90 EISPH_OneStep */
void FN(ora_eisph_step)(FN(ora_eisph) *P, FN(ora_eisph) *B, const FN(ora_grid) *gs,
                        const FN(ora_grid) *gb, REAL support, REAL h, REAL rho0, REAL nu,
                        REAL dt, int diagnostics) {
    FN(ora_eisph_dirichlet)(P, B, 0);
    FN(ora_eisph_advec_vis)(P, B, gs, gb, support, h, nu, dt);
    FN(ora_eisph_dirichlet)(P, B, 1);
    FN(ora_eisph_build_ppe)(P, B, gs, gb, support, h, rho0, dt);
    FN(ora_eisph_remove_bi_mean)(P);
    FN(ora_eisph_solve_ppe)(P);
    FN(ora_eisph_pressure_boundary)(P, B);
    FN(ora_eisph_correct)(P, B, gs, gb, support, h, dt);
    FN(ora_eisph_dirichlet)(P, B, 0);
    if (diagnostics) FN(ora_eisph_diagnostics)(P, B, gs, gb, support, h);
}

#undef E_R2_EPS
