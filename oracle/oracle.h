/*
 * CPU oracle for the coupled WCSPH-DEM step.  TEST INFRASTRUCTURE ONLY.
 *
 * This library restates, in plain C, the arithmetic of the reference's Warp
 * kernels (the kernel/ directory) and of the Warp HashGrid they query.  It exists to
 * check the HIP implementation: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product path never links or
 * calls it.
 *
 * Every function is compiled twice from oracle_impl.h: with REAL=float
 * (suffix _f32, the reference's precision) and REAL=double (suffix _f64, the
 * FP64 shadow).  FP32 arithmetic follows the reference expression order
 * operand by operand; the build uses -ffp-contract=off so no a*b+c is fused.
 *
 * Parity status of each part (how it is pinned) is listed in oracle_impl.h
 * beside each function.
 */
#ifndef SPHDEM_ORACLE_H
#define SPHDEM_ORACLE_H

#define ORACLE_DECLARE(REAL, SFX)                                              \
typedef struct {                                                               \
    REAL *pos, *vel, *rho_raw, *rho, *pres, *acc, *m, *flt, *porosity, *pgf,   \
         *acc_dem;                                                             \
    int n;                                                                     \
} ora_sph##SFX;                                                                \
typedef struct {                                                               \
    REAL *pos, *vel, *rho_raw, *rho, *pres, *acc, *m;                          \
    int n;                                                                     \
} ora_bnd##SFX;                                                                \
typedef struct {                                                               \
    REAL *pos, *vel, *acc, *force, *omega, *torque, *radius, *rho, *m,         \
         *volume, *inertia, *K, *eta, *mu, *h, *flt_s, *porosity, *drag,       \
         *pressure_force;                                                      \
    int *contact_dem_offset_old, *contact_dem_id_old;                          \
    REAL *tang_dem_old;                                                        \
    int *contact_dem_offset_new, *contact_dem_id_new;                          \
    REAL *tang_dem_new;                                                        \
    int *contact_bnd_offset_old, *contact_bnd_id_old;                          \
    REAL *tang_bnd_old;                                                        \
    int *contact_bnd_offset_new, *contact_bnd_id_new;                          \
    REAL *tang_bnd_new;                                                        \
    int n;                                                                     \
} ora_dem##SFX;                                                                \
typedef struct {                                                               \
    REAL *pos, *vel, *omega, *radius, *rho, *m, *volume, *inertia, *K, *eta,   \
         *mu;                                                                  \
    int n;                                                                     \
} ora_dembnd##SFX;                                                             \
typedef struct {                                                               \
    REAL *pos, *vel, *vel_star, *vel_bc, *acc, *kgc, *rho, *m, *pres, *Aij,    \
         *bi, *Aijpj, *divergence, *residual;                                  \
    int *mirror;                                                               \
    int n;                                                                     \
} ora_eisph##SFX;                                                              \
typedef struct {                                                               \
    int dim_x, dim_y, dim_z;                                                   \
    REAL cell_width_inv;                                                       \
    int n;                                                                     \
    const int *point_ids;                                                      \
    const int *cell_starts;                                                    \
    const int *cell_ends;                                                      \
} ora_grid##SFX;

ORACLE_DECLARE(float, _f32)
ORACLE_DECLARE(double, _f64)

#endif
